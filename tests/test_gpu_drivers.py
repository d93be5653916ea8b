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
    # random pairing (runner.py:162-165): 2-qubit gates on arbitrary, mostly
    # distant qubit pairs
    out.append(("haar-random-pairing", 10, pkg_module("experimentation.runner").make_haar(
        10, 4, np.random.default_rng(11)), None))
    return out


@pytest.mark.parametrize("idx", [0, 1, 2, 3])
@pytest.mark.parametrize("K", [64, 256])
def test_bass_end_to_end(gpu_ctx, idx, K):
    name, N, circ, ends = _circuits()[idx]
    print(f"[case] bass {name} N={N} K={K} gates={len(circ)}", flush=True)
    B = pkg_module("simulation.bass_simulator")
    kw = dict(optimize_every=5, buffer_factor=1)
    if ends is not None:
        kw = dict(optimize_at=ends)
    td, to = [], []
    sim = B.BASS(N, K, device="cuda:0", trace=td, accept_rtol=1e-10, **kw)
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


@pytest.mark.parametrize("idx", [0, 1])
def test_bass_2qubit_pass(gpu_ctx, idx):
    """use_2qubit_rotations=True (the brick-wall pass after the 1-qubit
    passes, bass_simulator.py:679-728, :820-824): the same decisions as the
    oracle (1-qubit rotations and 2-qubit pairs, accepted or reverted), F and
    gamma^2 within 1e-10."""
    name, N, circ, ends = _circuits()[idx]
    K = 128
    print(f"[case] bass 2q pass {name} N={N} K={K} gates={len(circ)}", flush=True)
    B = pkg_module("simulation.bass_simulator")
    kw = dict(optimize_every=5, use_2qubit_rotations=True)
    if ends is not None:
        kw = dict(optimize_at=ends, use_2qubit_rotations=True)
    td, to = [], []
    sim = B.BASS(N, K, device="cuda:0", trace=td, accept_rtol=1e-10, **kw)
    st = sim.simulate(circ)
    ref = OD.BASS(N, K, selector=OD.key_tiebreak_selector, trace=to, accept_rtol=sim.accept_rtol, **kw)
    rs = ref.simulate(circ)

    def decisions(tr):
        return [(e[0], e[1], e[2] if e[0] in ("rot", "rot2") else None) for e in tr]

    n2 = sum(e[0] == "rot2" for e in to)
    acc2 = sum(e[0] == "rot2" and e[2] for e in to)
    print(f"  2-qubit pairs tried {n2}, accepted {acc2}, counters={sim.counters}", flush=True)
    assert n2 > 0
    assert decisions(td) == decisions(to), "decision logs differ"
    sv = OE.exact_statevector(N, circ)
    f_dev = OE.fidelity_bass(N, st.x[: st.nnz], st.alpha[: st.nnz], sim.U, sv)
    f_ora = OE.fidelity_bass(N, rs.x[: rs.nnz], rs.alpha[: rs.nnz], ref.U, sv)
    print(f"  F dev={f_dev:.15f} oracle={f_ora:.15f} gamma2 {st.gamma ** 2:.15f} {rs.gamma ** 2:.15f}", flush=True)
    assert st.nnz == rs.nnz
    assert abs(f_dev - f_ora) <= 1e-10
    assert abs(st.gamma ** 2 - rs.gamma ** 2) <= 1e-10
    psi_dev = OE.to_statevector(N, st.x[: st.nnz], st.alpha[: st.nnz], sim.U)
    psi_ora = OE.to_statevector(N, rs.x[: rs.nnz], rs.alpha[: rs.nnz], ref.U)
    np.testing.assert_allclose(psi_dev, psi_ora, rtol=0, atol=1e-10)


def test_tfim_config0_lockstep(gpu_ctx):
    """Uniform TFIM n=20, J=h=1, dt=0.05, 40 steps, K=2^14, optimize_every=39.
    From |0...0> with uniform couplings the |a|^2 values tie exactly across
    symmetry classes and nearly (to an ulp) across others, so a renormalisation
    divisor summed in a different order flips near-ties at the threshold and
    the trajectories of any two implementations drift apart.  The contract is
    therefore checked per gate, in lock step: before each gate the oracle takes
    the device's state, applies the same gate and truncation rule
    (oracle/drivers.py BASS.simulate in the identity frame: the optimiser
    never rotates here, PR stays below k/4, the reference's skip rule,
    bass_simulator.py:731), and the device's result must match: kept keys
    identical except within 1e-12 (relative) of the k-th |a|^2, amplitudes
    within 1e-12, kept/total within 1e-12."""
    print("[case] tfim config0 n=20 K=2^14 lock step", flush=True)
    from oracle import kernels as OK
    from _parity import assert_topk_sets

    rc = pkg_module("utils.random_circuits")
    B = pkg_module("simulation.bass_simulator")
    N, K = 20, 1 << 14
    circ = rc.generate_tfim_circuit(N, 40, J=1.0, h=1.0, dt=0.05, rng=np.random.default_rng(0))
    assert len(circ) == 1560
    td = []
    sim = B.BASS(N, K, device="cuda:0", trace=td, optimize_every=39, accept_rtol=1e-10)
    sim.begin(len(circ))
    ref = OD.BASS(N, K, selector=OD.key_tiebreak_selector, optimize_every=39)
    internal_k = min(max(2 * K, 8 * K), 2 ** N)
    flips = truncs = 0
    for i, g in enumerate(circ):
        st = sim.state
        pk, pa = st.keys().copy(), st.alpha[: st.nnz].copy()
        sim.step(i, g)
        assert all(np.array_equal(u, np.eye(2)) for u in sim.U), "config 0 never rotates"
        o = OD.SparseState(N, internal_k, selector=OD.key_tiebreak_selector)
        o.x[: len(pk)], o.alpha[: len(pk)], o.nnz = pk, pa, len(pk)
        if g.n_qubits == 1:
            o.apply_gate(OD._G(g.qubits, g.matrix), "top-k")
        else:
            assert OK.is_gate_diagonal(g.matrix)
            OK.apply_2qubit_diagonal_kernel(o.x, o.alpha, o.nnz, g.qubits[0], g.qubits[1],
                                            np.diag(g.matrix).astype(np.complex128))
        st = sim.state
        gk, ga = st.keys(), st.alpha[: st.nnz]
        if g.n_qubits == 1 and o.nnz > K:
            truncs += 1
            cx, ca = o.x[: o.nnz].copy(), o.alpha[: o.nnz].copy()
            p = np.abs(ca) ** 2
            top = ref._select(o, p)
            kept, total = np.sum(p[top]), np.sum(p)
            e = td[-1]
            assert e[0] == "trunc" and e[1] == i
            assert abs(e[2] / e[3] - kept / total) <= 1e-12
            flips += assert_topk_sets(gk, cx[top], cx, ca, K, band=1e-12) > 0
            o.x[:K], o.alpha[:K], o.nnz = cx[top], ca[top], K
            ref._renorm(o)
        common, gi, oi = np.intersect1d(gk, o.x[: o.nnz], return_indices=True)
        assert len(common) >= len(gk) - 64
        np.testing.assert_allclose(ga[gi], o.alpha[: o.nnz][oi], rtol=1e-12, atol=1e-15)
    print(f"  truncations {truncs}, steps with a near-tie swap at the threshold {flips}", flush=True)
    assert truncs == sum(e[0] == "trunc" for e in td) and truncs > 700


@pytest.mark.parametrize("idx,K", [(0, 64), (1, 256), (1, 64), (2, 32)])
def test_device_counts_match_host_counts(gpu_ctx, idx, K):
    """Counted 1-qubit gates (input count kept on the device, op log read at
    the driver's decision points) and the per-gate host round trip produce the
    same state bit for bit, the same frame, gamma and operation counters.  The
    byte counters of selects and diagonal gates follow the passes each mode
    makes (the counted mode fuses runs of diagonal gates and lets the lift run
    the select's digit-1 pass), so they may only be lower there."""
    name, N, circ, ends = _circuits()[idx]
    print(f"[case] device counts {name} N={N} K={K}", flush=True)
    B = pkg_module("simulation.bass_simulator")
    dev = pkg_module("_device")
    res = []
    for counted in (False, True):
        kw = {"optimize_at": ends} if ends is not None else {"optimize_every": 5}
        # start from an empty workspace: the first counted gate grows it
        gpu_ctx.ws, gpu_ctx.ws_bytes = None, 0
        sim = B.BASS(N, K, device_counts=counted, **kw)
        st = sim.simulate(circ)
        res.append((dev.to_host_keys(st._buf.x, st.nnz).copy(), dev.to_host_amps(st._buf.a, st.nnz).copy(),
                    st.gamma, [u.copy() for u in sim.U], st.pos.copy(), dict(sim.counters)))
    (x0, a0, g0, u0, p0, c0), (x1, a1, g1, u1, p1, c1) = res
    assert np.array_equal(x0, x1) and np.array_equal(a0.view(np.float64), a1.view(np.float64))
    assert g0 == g1 and np.array_equal(p0, p1) and all(np.array_equal(a, b) for a, b in zip(u0, u1))
    by_pass = {"bytes_select", "bytes_diag2q"}
    # speculation hits and pruned-lift relifts count how each mode predicted
    # thresholds, not what it computed
    same = lambda c: {k: v for k, v in c.items()
                      if k not in by_pass and not k.startswith("spec_hit") and k != "relift"}
    assert same(c0) == same(c1)
    for key in ("bytes_select", "bytes_diag2q"):
        assert c1.get(key, 0) <= c0.get(key, 0), key
    assert c0.get("spec_hit", 0) == 0
    assert c1.get("truncate", 0) > 0 and c1.get("gate1q", 0) > 0
