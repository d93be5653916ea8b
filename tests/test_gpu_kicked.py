"""The periodic kicked-Ising chain (each step RZZ(pi/4) on every bond, then
RX(0.3 + 0.01 s) on every qubit) through the device BASS driver in its
default counted mode (the state count and the truncation log stay on the
device between decision points), against the oracle.

From |0...0> with uniform couplings the |a|^2 values tie across whole
symmetry classes (equal Hamming-weight patterns), exactly or to an ulp, so
the top-k threshold bucket is huge: these are the massive-tie selects that the
multi-block stages of the truncation resolve.  The contract is checked gate by
gate, in lock step: before each gate the oracle takes the device's state and
applies the same gate and truncation rule (largest |a|^2, ties to the smaller
key; src/simulation/bass_simulator.py:899-916 with the key tie-break); the
device's result must match -- kept keys identical except within 1e-12
(relative) of the k-th |a|^2, amplitudes within 1e-12, kept/total within
1e-12."""

import time

import numpy as np
import pytest

from _parity import assert_topk_sets
from conftest import pkg_module
from oracle import drivers as OD
from oracle import kernels as OK

pytestmark = pytest.mark.gpu


def _flush_state(sim):
    """Host copy (keys, amplitudes) of the driver's state after booking the
    pending counted gates."""
    sim._counted_flush()
    st = sim.state
    return st.keys().copy(), st.alpha[: st.nnz].copy()


@pytest.mark.parametrize("N,K,steps", [(14, 512, 4), (16, 4096, 3), (22, 1 << 19, 1)])
def test_kicked_ising_counted_lockstep(gpu_ctx, N, K, steps):
    """Kicked Ising, default BASS keywords (counted gates; the optimiser never
    rotates: PR stays below k/4, the reference's skip rule :731).  At N=22,
    K=2^19 the first truncations select 2^19 of 2^20 entries whose threshold
    class holds C(20,10) = 184756 entries: more than one block resolves."""
    print(f"[case] kicked ising counted lock step N={N} K={K} steps={steps}", flush=True)
    rc = pkg_module("utils.random_circuits")
    B = pkg_module("simulation.bass_simulator")
    circ, _ = rc.kicked_ising_circuit(N, steps)
    sim = B.BASS(N, K, device="cuda:0")
    assert sim.device_counts and sim.trace is None
    sim.begin(len(circ))
    ref = OD.BASS(N, K, selector=OD.key_tiebreak_selector)
    internal_k = min(max(2 * K, 8 * K), 2 ** N)
    truncs = flips = 0
    t0 = time.time()
    for i, g in enumerate(circ):
        pk, pa = _flush_state(sim)
        g_before = sim._gamma_running
        sim.step(i, g)
        gk, ga = _flush_state(sim)
        assert all(np.array_equal(u, np.eye(2)) for u in sim.U), "the optimiser never rotates here"
        o = OD.SparseState(N, internal_k, selector=OD.key_tiebreak_selector)
        o.x[: len(pk)], o.alpha[: len(pk)], o.nnz = pk, pa, len(pk)
        if g.n_qubits == 1:
            o.apply_gate(OD._G(g.qubits, g.matrix), "top-k")
        else:
            assert OK.is_gate_diagonal(g.matrix)
            OK.apply_2qubit_diagonal_kernel(o.x, o.alpha, o.nnz, g.qubits[0], g.qubits[1],
                                            np.diag(g.matrix).astype(np.complex128))
        if g.n_qubits == 1 and o.nnz > K:
            truncs += 1
            cx, ca = o.x[: o.nnz].copy(), o.alpha[: o.nnz].copy()
            p = np.abs(ca) ** 2
            top = ref._select(o, p)
            kept, total = np.sum(p[top]), np.sum(p)
            ratio = sim._gamma_running / g_before
            assert abs(ratio - kept / total) <= 1e-12 * (kept / total)
            flips += assert_topk_sets(gk, cx[top], cx, ca, K, band=1e-12) > 0
            o.x[:K], o.alpha[:K], o.nnz = cx[top], ca[top], K
            ref._renorm(o)
        else:
            assert sim._gamma_running == g_before
        assert len(gk) == o.nnz
        common, gi, oi = np.intersect1d(gk, o.x[: o.nnz], return_indices=True)
        assert len(common) >= len(gk) - 64
        np.testing.assert_allclose(ga[gi], o.alpha[: o.nnz][oi], rtol=1e-12, atol=1e-15)
    print(f"  gates {len(circ)}, truncations {truncs}, steps with a near-tie swap {flips}, "
          f"{time.time() - t0:.1f} s", flush=True)
    assert truncs > 0 and sim.counters.get("truncate", 0) == truncs


def _host_topk(x, a, K):
    """Indices of the K largest |a|^2 (re*re + im*im), ties to the smaller
    key, by selection (linear time, for 10^8 entries)."""
    p = a.real * a.real + a.imag * a.imag
    m = len(p)
    thr = np.partition(p, m - K)[m - K]
    above = np.flatnonzero(p > thr)
    eq = np.flatnonzero(p == thr)
    need = K - len(above)
    assert 0 < need <= len(eq)
    if need < len(eq):
        part = np.argpartition(x[eq], need - 1)[:need]
        eq = eq[part]
    return np.concatenate([above, eq]), p


@pytest.mark.timeout(900)
def test_kicked_ising_config2_window(gpu_ctx):
    """Kicked Ising at its configured size, n=56, K=2^26 (1.6 GB of state),
    default counted driver: the first two steps (the tie-heavy start, where
    every truncation's threshold class holds millions of entries) and the
    fourth.  After every step: exactly K entries, stored words strictly
    increasing, keys unique, sum |a|^2 = 1.  Three 1-qubit gates (the first
    truncation, one in step 2, one in step 4) are checked against the
    oracle's single-thread C++ restatement of the reference's gate (emit,
    16-bit radix sort, merge) on the device's input state: the kept set
    against a host top-K of that output (ties to the smaller key) except
    within 1e-12 of the threshold, kept/total within 1e-12, amplitudes
    within 1e-12."""
    from oracle import port

    print("[case] kicked ising n=56 K=2^26 window: steps 1, 2, 4", flush=True)
    rc = pkg_module("utils.random_circuits")
    B = pkg_module("simulation.bass_simulator")
    N, K, S = 56, 1 << 26, 4
    circ, ends = rc.kicked_ising_circuit(N, S)
    per = 2 * N
    check = {N + 26, per + N + 40, 3 * per + N + 11}
    sim = B.BASS(N, K, device="cuda:0", optimize_at=ends)
    sim.begin(len(circ))
    t0 = time.time()
    for i, g in enumerate(circ):
        if i in check:
            px, pa = _flush_state(sim)
            n = len(px)
            G = sim.U[g.qubits[0]].conj().T @ g.matrix @ sim.U[g.qubits[0]]
            g_before = sim._gamma_running
            sim.step(i, g)
            gk, ga = _flush_state(sim)
            ratio = sim._gamma_running / g_before
            xo = np.zeros(2 * n, np.uint64)
            ao = np.zeros(2 * n, np.complex128)
            m = port.gate_1q(px, pa, n, g.qubits[0], G, xo, ao)
            del px, pa
            xo, ao = xo[:m], ao[:m]
            assert m > K
            idx, p = _host_topk(xo, ao, K)
            kept, total = float(np.sum(p[idx])), float(np.sum(p))
            print(f"  gate {i}: {m} entries -> {K}, kept/total {kept / total:.15f}, device {ratio:.15f}",
                  flush=True)
            assert abs(ratio - kept / total) <= 1e-12 * (kept / total)
            hk = xo[idx]
            ha = ao[idx] / np.sqrt(kept)
            o = np.argsort(hk)
            hk, ha = hk[o], ha[o]
            og = np.argsort(gk)
            gks, gas = gk[og], ga[og]
            assert len(gks) == K and np.all(gks[1:] > gks[:-1])
            only_dev = np.setdiff1d(gks, hk, assume_unique=True)
            only_host = np.setdiff1d(hk, gks, assume_unique=True)
            assert len(only_dev) == len(only_host)
            if len(only_dev):
                thr = float(p[idx].min())
                pos = np.searchsorted(xo, np.concatenate([only_dev, only_host]))
                assert np.all(np.abs(p[pos] - thr) <= 1e-12 * thr), "kept sets differ outside the tie band"
            common, gi, hi = np.intersect1d(gks, hk, assume_unique=True, return_indices=True)
            err = np.abs(gas[gi] - ha[hi])
            assert np.all(err <= 1e-12 * np.abs(ha[hi]) + 1e-15 * np.abs(ha).max())
            del xo, ao, p, idx
        else:
            sim.step(i, g)
        if i in ends:
            s = ends.index(i)
            gk, ga = _flush_state(sim)
            w = sim.state._buf.x[: sim.state.nnz].cpu().numpy().view(np.uint64)
            print(f"  step {s + 1}: nnz {len(gk)}, {time.time() - t0:.1f} s", flush=True)
            assert len(gk) == K
            assert np.all(w[1:] > w[:-1]), "stored words strictly increasing"
            assert len(np.unique(gk)) == K
            s2 = float(np.sum(ga.real * ga.real + ga.imag * ga.imag))
            assert abs(s2 - 1.0) <= 1e-12
    assert sim.counters.get("truncate", 0) > 0
