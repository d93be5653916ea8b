"""Dense state vectors on the device (exact simulation, reconstruction of a
BASS state, fidelities) against the oracle's numpy restatement
(oracle/exact.py, exact_simulator.py:66-144, bass_simulator.py:608-621,
runner.py:615-626)."""

import numpy as np
import pytest

from conftest import pkg_module, random_unitary
from oracle import exact as OE

pytestmark = pytest.mark.gpu


def _circuits():
    rc = pkg_module("utils.random_circuits")
    G = pkg_module("core.gates")
    rng = np.random.default_rng(17)
    out = [("tfim", 10, rc.generate_tfim_circuit(10, 4, J=1.0, h=0.8, dt=0.2, rng=np.random.default_rng(5),
                                                 disorder_strength_h=0.4))]
    circ, _ = rc.brickwork_haar_circuit(9, 3, np.random.default_rng(6))
    out.append(("brickwork", 9, circ))
    # generic 2-qubit gates in both qubit orders, next to 1-qubit Haar gates
    gen = []
    for t in range(12):
        q1, q2 = (int(v) for v in rng.choice(8, size=2, replace=False))
        gen.append(G.QuantumGate([q1, q2], random_unitary(4, rng)))
        gen.append(G.QuantumGate([int(rng.integers(8))], random_unitary(2, rng)))
    out.append(("generic", 8, gen))
    return out


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_exact_simulator_matches_oracle(gpu_ctx, idx):
    name, N, circ = _circuits()[idx]
    print(f"[case] exact {name} N={N} gates={len(circ)}", flush=True)
    E = pkg_module("simulation.exact_simulator")
    got = E.ExactSimulator(N, verbose=False, device="cuda:0").simulate(circ)
    ref = OE.exact_statevector(N, circ)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)
    assert abs(np.vdot(got, got).real - 1.0) < 1e-12


def test_vdot_known_answers(gpu_ctx):
    import torch

    ops = pkg_module("_ops")
    rng = np.random.default_rng(2)
    for n in (1, 5, 4096, 4097, 300000):
        print(f"[case] vdot n={n}", flush=True)
        u = rng.normal(size=n) + 1j * rng.normal(size=n)
        v = rng.normal(size=n) + 1j * rng.normal(size=n)
        tu, tv = torch.from_numpy(u).to("cuda:0"), torch.from_numpy(v).to("cuda:0")
        z = ops.dense_vdot(gpu_ctx, tu, tv)
        ref = np.vdot(u, v)
        assert abs(z - ref) <= 1e-12 * max(1.0, np.abs(u).sum() * np.abs(v).max())
        zz = ops.dense_vdot(gpu_ctx, tu, tu)
        assert abs(zz.imag) < 1e-9 and abs(zz.real - np.vdot(u, u).real) <= 1e-12 * zz.real


@pytest.mark.parametrize("two_q", [False, True])
def test_bass_statevector_and_fidelity(gpu_ctx, two_q):
    """Device reconstruction (scatter through the key map + frame rotations)
    and fidelity of a BASS run with accepted rotations, against the oracle's
    host reconstruction of the same stored state."""
    rc = pkg_module("utils.random_circuits")
    B = pkg_module("simulation.bass_simulator")
    Fm = pkg_module("benchmarking.fidelity")
    N, K = 12, 128
    circ, ends = rc.brickwork_haar_circuit(N, 6, np.random.default_rng(7))
    print(f"[case] statevector brickwork N={N} K={K} 2q={two_q}", flush=True)
    sim = B.BASS(N, K, device="cuda:0", optimize_at=ends, use_2qubit_rotations=two_q)
    st = sim.simulate(circ)
    assert sim.counters.get("accept", 0) > 0 and not np.array_equal(st.pos, np.arange(N))
    keys, amps = st.keys(), st.alpha[: st.nnz]
    psi = sim.to_statevector(st)
    ref = OE.to_statevector(N, keys, amps, sim.U)
    np.testing.assert_allclose(psi, ref, rtol=0, atol=1e-13)
    sv = OE.exact_statevector(N, circ)
    f_dev = Fm.fidelity_bass(sim, st, sv)
    f_ref = OE.fidelity_bass(N, keys, amps, sim.U, sv)
    print(f"  F device {f_dev:.15f} oracle {f_ref:.15f}", flush=True)
    assert abs(f_dev - f_ref) <= 1e-12
    # compute_fidelity reads keys through the stored-bit map
    E = pkg_module("simulation.exact_simulator")
    sv_dev = E.exact_statevector_device(N, circ, device="cuda:0")
    assert abs(Fm.fidelity_bass(sim, st, sv_dev) - f_dev) <= 1e-12
    assert abs(Fm.compute_fidelity(st, sv) - OE.compute_fidelity(keys, amps, sv)) <= 1e-12


def test_config0_exact_statevector(gpu_ctx):
    """Uniform TFIM n=20, J=h=1, dt=0.05 (5 of the 40 steps): the dense
    device simulation against the oracle's einsum restatement."""
    rc = pkg_module("utils.random_circuits")
    E = pkg_module("simulation.exact_simulator")
    circ = rc.generate_tfim_circuit(20, 5, J=1.0, h=1.0, dt=0.05, rng=np.random.default_rng(0))
    print(f"[case] exact tfim n=20 gates={len(circ)}", flush=True)
    got = E.exact_statevector(20, circ, device="cuda:0")
    ref = OE.exact_statevector(20, circ)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)
