"""Device drivers (BASS, FixedBasisSimulator) end to end vs the CPU oracle.

Workloads here have generic amplitudes (random fields / Haar gates), so no
|a|^2 ties sit at a truncation threshold and the device (ties -> smaller key)
and the oracle with the same rule follow the same trajectory.
"""

import numpy as np
import pytest

from _parity import assert_same_entries
from conftest import pkg_module
from oracle import circuits as OC
from oracle import drivers as OD
from oracle import exact as OE

pytestmark = pytest.mark.gpu


def _circuits():
    rc = pkg_module("utils.random_circuits")
    out = []
    rng = np.random.default_rng(20260520 * 10 ** 6 + 1)
    out.append(("tfim-disordered", 12, rc.generate_tfim_circuit(12, 8, J=1.0, h=1.0, dt=0.1, rng=rng,
                                                                  disorder_strength_J=0.3,
                                                                  disorder_strength_h=0.5), None))
    circ, ends = rc.brickwork_haar_circuit(12, 6, np.random.default_rng(7))
    out.append(("brickwork-haar", 12, circ, ends))
    out.append(("rfim-random-h", 10, rc.generate_tfim_circuit(10, 6, J=1.0, h=1.2, dt=0.2,
                                                               rng=np.random.default_rng(3),
                                                               disorder_strength_h=0.7), None))
    return out


@pytest.mark.parametrize("idx", [0, 1, 2])
@pytest.mark.parametrize("K", [64, 256])
def test_bass_end_to_end(gpu_ctx, idx, K):
    name, N, circ, ends = _circuits()[idx]
    print(f"[case] bass {name} N={N} K={K} gates={len(circ)}", flush=True)
    B = pkg_module("simulation.bass_simulator")
    kw = dict(optimize_every=5, buffer_factor=1)
    if ends is not None:
        kw = dict(optimize_at=ends)
    td, to = [], []
    sim = B.BASS(N, K, device="cuda:0", trace=td, **kw)
    st = sim.simulate(circ)
    ref = OD.BASS(N, K, selector=OD.key_tiebreak_selector, trace=to, accept_rtol=sim.accept_rtol, **kw)
    rs = ref.simulate(circ)
    # same decisions in the same order: truncation steps, rotations tried and accepted
    def decisions(tr):
        return [(e[0], e[1], e[2] if e[0] == "rot" else None) for e in tr]

    assert decisions(td) == decisions(to), "decision logs differ"
    sv = OE.exact_statevector(N, circ)
    f_dev = OE.fidelity_bass(N, st.x[: st.nnz], st.alpha[: st.nnz], sim.U, sv)
    f_ora = OE.fidelity_bass(N, rs.x[: rs.nnz], rs.alpha[: rs.nnz], ref.U, sv)
    print(f"  F dev={f_dev:.15f} oracle={f_ora:.15f} gamma2 {st.gamma ** 2:.15f} {rs.gamma ** 2:.15f} "
          f"counters={sim.counters}", flush=True)
    assert st.nnz == rs.nnz
    assert abs(f_dev - f_ora) <= 1e-10
    assert abs(st.gamma ** 2 - rs.gamma ** 2) <= 1e-10
    # The frames U_j are fixed up to a phase per eigenvector column when a
    # single-qubit RDM is degenerate to rounding (|rho01|^2 at the 1e-28 branch
    # of eigh_2x2), so compare frame-independent quantities: the kept keys,
    # |a|, and the reconstructed state vector.
    gx, ga = st.x[: st.nnz], st.alpha[: st.nnz]
    assert set(gx.tolist()) == set(rs.x[: rs.nnz].tolist())
    assert_same_entries(gx, np.abs(ga), rs.x[: rs.nnz], np.abs(rs.alpha[: rs.nnz]), rtol=1e-10)
    psi_dev = OE.to_statevector(N, gx, ga, sim.U)
    psi_ora = OE.to_statevector(N, rs.x[: rs.nnz], rs.alpha[: rs.nnz], ref.U)
    np.testing.assert_allclose(psi_dev, psi_ora, rtol=0, atol=1e-10)
    same_frame = all(np.allclose(sim.U[j], ref.U[j], rtol=0, atol=1e-10) for j in range(N))
    if same_frame:
        assert_same_entries(gx, ga, rs.x[: rs.nnz], rs.alpha[: rs.nnz], rtol=1e-10)


@pytest.mark.parametrize("dedup", ["table", "sort"])
def test_fixed_basis_end_to_end(gpu_ctx, dedup):
    N, K = 12, 200
    print(f"[case] fixed N={N} K={K} dedup={dedup}", flush=True)
    circ = OC.tfim(N, 6, J=1.0, h=1.0, dt=0.15, rng=np.random.default_rng(11), dJ=0.4, dh=0.6)
    S = pkg_module("simulation.simulator")
    sim = S.FixedBasisSimulator(N, K, verbose=False, device="cuda:0", dedup1q=dedup)
    st = sim.simulate(circ)
    ref = OD.FixedBasisSimulator(N, K, selector=OD.key_tiebreak_selector)
    rs = ref.simulate(circ, seed=1)
    assert st.nnz == rs.nnz
    assert abs(st.gamma ** 2 - rs.gamma ** 2) <= 1e-10
    assert_same_entries(st.x[: st.nnz], st.alpha[: st.nnz], rs.x[: rs.nnz], rs.alpha[: rs.nnz], rtol=1e-10)
    np.testing.assert_allclose(sim.get_metrics()["gamma_squared"], ref.gamma_history ** 2, rtol=0, atol=1e-10)
