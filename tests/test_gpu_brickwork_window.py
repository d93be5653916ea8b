"""The brickwork random circuit at its configured size -- n=40, 30 layers, a
Haar-random U(2) on every qubit then CZ on alternating bonds, K=2^22, the
adaptive basis optimisation after every layer -- through the device BASS
driver in its default counted mode, checked in lock step against the oracle
over one whole layer after the state has saturated (the window the benchmark
times):

* every 1-qubit gate against the oracle's single-thread C++ restatement of
  the reference's gate (emit, 16-bit radix sort, merge) on the device's input
  state, then a host top-K (ties to the smaller key): kept keys identical
  except within 1e-12 of the threshold, kept/total within 1e-12, amplitudes
  within 1e-12;
* every CZ in the current frame ((U1 (x) U2)^dagger CZ (U1 (x) U2), a general
  2-qubit gate once rotations were accepted) against the numpy oracle's
  apply_2qubit_gate_ht (or its diagonal kernel), then the same top-K rule;
* the optimiser pass at the end of the layer: the oracle's _optimize_basis
  from the device's state and frame (src/simulation/bass_simulator.py:730-827)
  must take the same decisions on the first trials, with PR before / after and
  the kept fraction within 1e-12.
Reference semantics: src/simulation/bass_simulator.py:847-929 (gates,
truncation), :730-827 (optimiser)."""

import time

import numpy as np
import pytest

from conftest import pkg_module
from oracle import drivers as OD
from oracle import kernels as OK

pytestmark = pytest.mark.gpu

SEED = 20260520 * 10 ** 6 + 1  # the benchmark's brickwork circuit
TRIALS = 8


class _Stop(Exception):
    pass


class _LimitedTrace(list):
    """A decision log that ends the oracle's optimiser after `limit` trials."""

    def __init__(self, limit):
        super().__init__()
        self.limit = limit

    def append(self, rec):
        super().append(rec)
        if len(self) >= self.limit:
            raise _Stop()


def _flush_state(sim):
    sim._counted_flush()
    st = sim.state
    return st.keys().copy(), st.alpha[: st.nnz].copy()


def _host_topk(x, a, K):
    """Indices of the K largest |a|^2 (re*re + im*im), ties to the smaller key."""
    p = a.real * a.real + a.imag * a.imag
    m = len(p)
    thr = np.partition(p, m - K)[m - K]
    above = np.flatnonzero(p > thr)
    eq = np.flatnonzero(p == thr)
    need = K - len(above)
    if need < len(eq):
        eq = eq[np.argpartition(x[eq], need - 1)[:need]]
    return np.concatenate([above, eq]), p


def _compare(gk, ga, hx, ha, K, ratio, where):
    """Device state (gk, ga) against the oracle's gate output (hx, ha):
    truncated to its top K (renormalised) when it holds more than K."""
    if len(hx) > K:
        idx, p = _host_topk(hx, ha, K)
        kept, total = float(np.sum(p[idx])), float(np.sum(p))
        assert abs(ratio - kept / total) <= 1e-12 * (kept / total), where
        thr = float(p[idx].min())
        hk, hv = hx[idx], ha[idx] / np.sqrt(kept)
    else:
        assert ratio == 1.0, where
        hk, hv, thr, p = hx, ha, None, None
    o, og = np.argsort(hk), np.argsort(gk)
    hk, hv, gks, gvs = hk[o], hv[o], gk[og], ga[og]
    assert len(gks) == len(hk) and np.all(gks[1:] > gks[:-1]), where
    only_dev = np.setdiff1d(gks, hk, assume_unique=True)
    only_host = np.setdiff1d(hk, gks, assume_unique=True)
    assert len(only_dev) == len(only_host), where
    if len(only_dev):
        ox = np.argsort(hx)
        pos = ox[np.searchsorted(hx, np.concatenate([only_dev, only_host]), sorter=ox)]
        assert np.all(np.abs(p[pos] - thr) <= 1e-12 * thr), f"{where}: kept sets differ outside the tie band"
    common, gi, hi = np.intersect1d(gks, hk, assume_unique=True, return_indices=True)
    err = np.abs(gvs[gi] - hv[hi])
    assert np.all(err <= 1e-12 * np.abs(hv[hi]) + 1e-15 * np.abs(hv).max()), where
    return len(only_dev)


@pytest.mark.timeout(1500)
def test_brickwork_config1_window(gpu_ctx):
    from oracle import port

    rc = pkg_module("utils.random_circuits")
    B = pkg_module("simulation.bass_simulator")
    N, K, W = 40, 1 << 22, 2
    circ, ends = rc.brickwork_haar_circuit(N, 30, np.random.default_rng(SEED))
    starts = [0] + [e + 1 for e in ends[:-1]]
    print(f"[case] brickwork n={N} K=2^22 lock-step window: layer {W}", flush=True)
    sim = B.BASS(N, K, device="cuda:0", optimize_at=ends, accept_rtol=1e-10)
    sim.begin(len(circ))
    for i in range(starts[W]):
        sim.step(i, circ[i])
    t0 = time.time()
    last = ends[W]
    sim.optimize_at = frozenset(e for e in ends if e != last)  # the window's pass runs below
    checked = swaps = 0
    for i in range(starts[W], last + 1):
        g = circ[i]
        px, pa = _flush_state(sim)
        assert len(px) == K
        U = [u.copy() for u in sim.U]
        g_before = sim._gamma_running
        sim.step(i, g)
        gk, ga = _flush_state(sim)
        ratio = sim._gamma_running / g_before
        if g.n_qubits == 1:
            q = g.qubits[0]
            G = U[q].conj().T @ g.matrix @ U[q]
            xo = np.zeros(2 * K, np.uint64)
            ao = np.zeros(2 * K, np.complex128)
            m = port.gate_1q(px, pa, K, q, G, xo, ao)
            hx, ha = xo[:m], ao[:m]
        else:
            q1, q2 = g.qubits
            Ut = OK.manual_kron_2x2(U[q1], U[q2])
            G = Ut.conj().T @ g.matrix @ Ut
            if OK.is_gate_diagonal(G):
                hx, ha = px.copy(), pa.copy()
                OK.apply_2qubit_diagonal_kernel(hx, ha, K, q1, q2, np.diag(G).astype(np.complex128))
            else:
                xo = np.zeros(4 * K, np.uint64)
                ao = np.zeros(4 * K, np.complex128)
                m = OK.apply_2qubit_gate_ht(px, pa, K, xo, ao, G, q1, q2)
                hx, ha = xo[:m], ao[:m]
        swaps += _compare(gk, ga, hx, ha, K, ratio, f"gate {i} ({g.n_qubits}q)")
        checked += 1
    print(f"  {checked} gates in lock step ({swaps} near-tie swaps at the threshold), {time.time() - t0:.1f} s",
          flush=True)
    # the optimiser pass at the layer end, from the device's state
    px, pa = _flush_state(sim)
    U = [u.copy() for u in sim.U]
    assert sim._trunc_active and sim.optimize_trigger(sim.state)
    sim.trace = []
    sim._optimize_basis(sim.state)
    td = sim.trace
    ref = OD.BASS(N, K, selector=OD.key_tiebreak_selector, accept_rtol=1e-10, trace=_LimitedTrace(TRIALS))
    ref.U = U
    ref._gamma_running = 1.0
    o = OD.SparseState(N, 2 * K, selector=OD.key_tiebreak_selector)
    o.x[:K], o.alpha[:K], o.nnz = px, pa, K
    with pytest.raises(_Stop):
        ref._optimize_basis(o)
    to = list(ref.trace)
    assert len(td) >= TRIALS and len(to) == TRIALS
    for d, r in zip(td[:TRIALS], to):
        assert d[0] == r[0] == "rot" and d[1] == r[1] and d[2] == r[2], (d, r)
        for a, b in zip(d[3:], r[3:]):
            assert abs(a - b) <= 1e-12 * abs(b), (d, r)
    acc = sum(1 for e in td if e[2])
    print(f"  optimiser: {len(td)} trials on the device ({acc} accepted), first {TRIALS} decided as the "
          f"oracle, {time.time() - t0:.1f} s", flush=True)
