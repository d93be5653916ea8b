"""Key-ordered gates (bass_lift) and the key-map tie-break of bass_truncate
against the oracle, through the C ABI.

States are given a random qubit -> bit placement, stored as permuted words
sorted ascending, and the device results are mapped back to keys."""

import numpy as np
import pytest
import torch

from _parity import assert_same_entries, numpy_uses_fma_cmul
from conftest import pkg_module, random_state, random_unitary
from oracle import drivers as OD
from oracle import kernels as OK

pytestmark = pytest.mark.gpu


def _mods():
    return pkg_module("_ops"), pkg_module("_device"), pkg_module("core.sparse_state"), pkg_module("_lib")


def _to_stored(x, pos):
    out = np.zeros_like(x)
    for q, p in enumerate(pos):
        out |= ((x >> np.uint64(q)) & np.uint64(1)) << np.uint64(p)
    return out


def _ordered(x, a, pos):
    w = _to_stored(x, pos)
    o = np.argsort(w)
    return w[o], a[o]


def _upload(dev, ctx, w, a):
    n = len(w)
    b = dev.Buffers(max(n, 1), ctx.device)
    if n:
        b.x[:n] = torch.from_numpy(w.view(np.int64)).to(ctx.device)
        b.a[:n] = torch.from_numpy(a).to(ctx.device)
    return b


def _gate_with_zero(rng, d):
    G = random_unitary(d, rng)
    G[0, d - 1] = 0.0  # exercise the exact-zero emission rule
    return G


def _oracle_1q(x, a, G, q):
    n = len(x)
    ox, oa = np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.complex128)
    m = OK.apply_1qubit_gate_kernel_opt(x, a, n, ox, oa, G, q)
    OK.radix_sort_uint64(ox, oa, m)
    m = OK.merge_sorted_kernel(ox, oa, m)
    return ox[:m], oa[:m]


def _run_lift(ctx, w, a, j, G, nq, mode):
    ops, dev, _, L = _mods()
    n = len(w)
    src = _upload(dev, ctx, w, a)
    seg = dev.Buffers(1, ctx.device)
    m, tiles = ops.lift(ctx, src, n, j, seg, G=G, nq=nq, mode=mode)
    out = dev.Buffers(max(m, 1), ctx.device)
    if m:
        ops.densify(ctx, seg, tiles, m, out)
    return dev.to_host_keys(out.x, m), dev.to_host_amps(out.a, m)


@pytest.mark.parametrize("n,nbits,mode", [(1, 5, 0), (2, 3, 0), (777, 12, 0), (30000, 18, 0), (70000, 24, 0),
                                          (300000, 22, 0), (300000, 40, 1),
                                          (5000, 14, 1), (40000, 20, 1)])
def test_lift_1q(gpu_ctx, n, nbits, mode):
    _, _, SS, L = _mods()
    rng = np.random.default_rng(n + nbits + 100 * mode)
    x, a = random_state(n, nbits, seed=n + 11)
    pos = rng.permutation(nbits)
    for q in sorted({0, nbits - 1, int(rng.integers(nbits))}):
        print(f"[case] lift1q n={n} nbits={nbits} mode={mode} q={q} bit={pos[q]}", flush=True)
        w, aw = _ordered(x, a, pos)
        G = _gate_with_zero(rng, 2) if mode == 0 else random_unitary(2, rng)
        gw, ga = _run_lift(gpu_ctx, w, aw, int(pos[q]), G, 1, mode)
        assert np.all(gw[1:] > gw[:-1]), "output not sorted by stored word"
        npos = SS.lifted_positions(pos, int(pos[q]), 0)
        assert npos[q] == 0
        gx = SS.stored_to_keys(gw, npos)
        if mode == 0:
            ox, oa = _oracle_1q(x, a, G, q)
        else:
            ox, oa = np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.complex128)
            ox[:n], oa[:n] = x, a
            m = OK.apply_1qubit_rotation_sparse(ox, oa, n, q, G)
            ox, oa = ox[:m], oa[:m]
        # one or two products per output: the sum does not depend on order
        assert_same_entries(gx, ga, ox, oa, exact=numpy_uses_fma_cmul())


@pytest.mark.parametrize("n,nbits,mode", [(3, 3, 0), (5000, 13, 0), (40000, 18, 0), (6000, 14, 1),
                                          (200000, 20, 0), (200000, 36, 1)])
def test_lift_2q(gpu_ctx, n, nbits, mode):
    ops, dev, SS, L = _mods()
    rng = np.random.default_rng(n + 7 * mode)
    x, a = random_state(n, nbits, seed=n + 7)
    pos = rng.permutation(nbits)
    q1, q2 = (int(v) for v in rng.choice(nbits, size=2, replace=False))
    print(f"[case] lift2q n={n} nbits={nbits} mode={mode} q=({q1},{q2}) bits=({pos[q1]},{pos[q2]})", flush=True)
    w, aw = _ordered(x, a, pos)
    # bring q2 to bit 0 with a move, then lift q1's bit to bit 1 with the gate
    src = _upload(dev, gpu_ctx, w, aw)
    mv = dev.Buffers(n, gpu_ctx.device)
    ops.lift(gpu_ctx, src, n, int(pos[q2]), mv, mode=L.MODE_MOVE)
    p1 = SS.lifted_positions(pos, int(pos[q2]), 0)
    mw = dev.to_host_keys(mv.x, n)
    assert np.all(mw[1:] > mw[:-1]), "move output not sorted"
    assert np.array_equal(np.sort(SS.stored_to_keys(mw, p1)), np.sort(x)), "move lost entries"
    G = _gate_with_zero(rng, 4) if mode == 0 else random_unitary(4, rng)
    gw, ga = _run_lift(gpu_ctx, mw, dev.to_host_amps(mv.a, n), int(p1[q1]), G, 2, mode)
    assert np.all(gw[1:] > gw[:-1]), "output not sorted by stored word"
    p2 = SS.lifted_positions(p1, int(p1[q1]), 1)
    assert p2[q1] == 1 and p2[q2] == 0
    gx = SS.stored_to_keys(gw, p2)
    ox, oa = np.zeros(4 * n, np.uint64), np.zeros(4 * n, np.complex128)
    if mode == 0:
        m = OK.apply_2qubit_gate_ht(x, a, n, ox, oa, G, q1, q2)
    else:
        ox[:n], oa[:n] = x, a
        m = OK.apply_2qubit_rotation_sparse(ox, oa, n, q1, q2, G)
    # rotations sum in matrix order (bitwise); gate products are summed in
    # stored order, which may differ from the oracle's input order
    assert_same_entries(gx, ga, ox[:m], oa[:m], rtol=1e-13, exact=(mode == 1 and numpy_uses_fma_cmul()))


@pytest.mark.parametrize("n,nbits", [(1, 4), (5, 3), (3000, 13), (50000, 30), (400000, 21)])
def test_lift_move(gpu_ctx, n, nbits):
    ops, dev, SS, L = _mods()
    rng = np.random.default_rng(n)
    x, a = random_state(n, nbits, seed=n + 5)
    pos = rng.permutation(nbits)
    w, aw = _ordered(x, a, pos)
    for j in sorted({0, nbits - 1, int(rng.integers(nbits))}):
        print(f"[case] move n={n} nbits={nbits} bit={j}", flush=True)
        src = _upload(dev, gpu_ctx, w, aw)
        out = dev.Buffers(n, gpu_ctx.device)
        ops.lift(gpu_ctx, src, n, j, out, mode=L.MODE_MOVE)
        npos = SS.lifted_positions(pos, j, 0)
        exp_w = _to_stored(x, npos)
        o = np.argsort(exp_w)
        assert np.array_equal(dev.to_host_keys(out.x, n), exp_w[o])
        assert np.array_equal(dev.to_host_amps(out.a, n), a[o])


@pytest.mark.parametrize("n,k", [(4096, 1000), (30000, 20000)])
def test_truncate_key_map_ties(gpu_ctx, n, k):
    """Top-k of an ordered state: ties broken by the key, not the stored word."""
    ops, dev, SS, L = _mods()
    print(f"[case] truncate key map n={n} k={k}", flush=True)
    nbits = 24
    rng = np.random.default_rng(k)
    x, a = random_state(n, nbits, seed=n + k, ties=True)
    a[: n // 2] = a[0]  # a large block of exactly equal |a|^2 around the threshold
    pos = rng.permutation(nbits)
    w, aw = _ordered(x, a, pos)
    src = _upload(dev, gpu_ctx, w, aw)
    dst = dev.Buffers(k, gpu_ctx.device)
    st = ops.truncate(gpu_ctx, src, n, k, dst, renorm=False, key_map=ops.key_map_arg(pos))
    gw = dev.to_host_keys(dst.x, k)
    assert np.all(gw[1:] > gw[:-1]), "truncation must preserve order"
    gx = SS.stored_to_keys(gw, pos)
    ox, oa = x.copy(), a.copy()
    OK.top_k_truncate_kernel_opt(ox, oa, n, k, selector=OD.key_tiebreak_selector)
    assert set(gx.tolist()) == set(ox[:k].tolist())
    assert_same_entries(gx, dev.to_host_amps(dst.a, k), ox[:k], oa[:k], exact=True)
    assert st.kept > 0


@pytest.mark.parametrize("n", [1, 3000])
def test_truncate_segmented_keeps_all_up_to_k(gpu_ctx, n):
    """A segmented gate output with <= k entries comes back whole, in order and
    not renormalised (stats.flags bit 0 clear); with > k it is truncated."""
    ops, dev, SS, L = _mods()
    print(f"[case] keep-all n={n}", flush=True)
    rng = np.random.default_rng(n + 1)
    x, a = random_state(n, 16, seed=n + 2)
    pos = rng.permutation(16)
    w, aw = _ordered(x, a, pos)
    G = random_unitary(2, rng)
    ew, ea = _run_lift(gpu_ctx, w, aw, int(pos[3]), G, 1, 0)
    m = len(ew)
    km = ops.key_map_arg(SS.lifted_positions(pos, int(pos[3]), 0))
    for k, trunc in [(m + 5, False), (m, False), (max(1, m - 7), m > 1)]:
        src = _upload(dev, gpu_ctx, w, aw)
        seg = dev.Buffers(1, gpu_ctx.device)
        none, tiles = ops.lift(gpu_ctx, src, n, int(pos[3]), seg, G=G, nq=1, count=False)
        assert none is None
        dst = dev.Buffers(k, gpu_ctx.device)
        st = ops.truncate(gpu_ctx, seg, None, k, dst, renorm=True, hist_tiles=tiles, segmented=True, key_map=km)
        assert st.n_in == m
        assert bool(st.flags & 1) == trunc
        if not trunc:
            assert st.nnz == m and st.norm == 1.0
            assert np.array_equal(dev.to_host_keys(dst.x, m), ew)
            assert np.array_equal(dev.to_host_amps(dst.a, m).view(np.float64), ea.view(np.float64))
        else:
            assert st.nnz == k
            assert abs(np.sum(np.abs(dev.to_host_amps(dst.a, k)) ** 2) - 1.0) < 1e-12


@pytest.mark.parametrize("n,nbits", [(1, 5), (7, 4), (5000, 14), (70000, 40), (300000, 28)])
def test_rdm_ordered(gpu_ctx, n, nbits):
    """RDMs of every stored bit of an ordered state vs the oracle's (hash) RDMs."""
    ops, dev, SS, L = _mods()
    print(f"[case] rdm ordered n={n} nbits={nbits}", flush=True)
    rng = np.random.default_rng(n + nbits)
    if nbits <= 16:
        x, a = random_state(n, nbits, seed=n + 3)
    else:
        # clustered keys so that many pairs exist on every bit
        base = rng.choice(2 ** 20, size=n // 8, replace=False).astype(np.uint64)
        hi = rng.choice(2 ** (nbits - 20), size=8, replace=False).astype(np.uint64) << np.uint64(20)
        x = np.unique((base[:, None] | hi[None, :]).ravel())[:n]
        n = len(x)
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        a /= np.linalg.norm(a)
    pos = rng.permutation(nbits)
    w, aw = _ordered(x, a, pos)
    src = _upload(dev, gpu_ctx, w, aw)
    rho_st = ops.rdm1q(gpu_ctx, src, n, nbits, ordered=True)
    rho = rho_st[pos]
    ref = OK.compute_all_1qubit_rdms(x, a, n, nbits)
    np.testing.assert_allclose(rho, ref, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("n,k,ordered", [(300000, 123457, False), (200000, 150001, True)])
def test_truncate_massive_ties(gpu_ctx, n, k, ordered):
    """|a|^2 with three distinct values: the threshold bucket holds far more
    than one block resolves, so the call repeats with the multi-block path."""
    ops, dev, SS, L = _mods()
    print(f"[case] massive ties n={n} k={k} ordered={ordered}", flush=True)
    nbits = 40
    rng = np.random.default_rng(n)
    x = np.unique(rng.integers(0, 2 ** nbits, size=n + 1000, dtype=np.uint64))[:n]
    rng.shuffle(x)
    vals = np.array([0.5, 0.25 + 0.25j, 0.125], dtype=np.complex128)
    a = vals[rng.integers(0, 3, size=n)]
    a[: n // 2] = vals[1]  # the threshold lands inside this tie class
    if ordered:
        pos = rng.permutation(nbits)
        w, aw = _ordered(x, a, pos)
        km = ops.key_map_arg(pos)
    else:
        w, aw, km = x, a, None
    src = _upload(dev, gpu_ctx, w, aw)
    dst = dev.Buffers(k, gpu_ctx.device)
    st = ops.truncate(gpu_ctx, src, n, k, dst, renorm=False, key_map=km)
    assert st.nnz == k and st.flags & 1
    gw = dev.to_host_keys(dst.x, k)
    gx = SS.stored_to_keys(gw, pos) if ordered else gw
    ox, oa = x.copy(), a.copy()
    OK.top_k_truncate_kernel_opt(ox, oa, n, k, selector=OD.key_tiebreak_selector)
    assert set(gx.tolist()) == set(ox[:k].tolist())
    assert_same_entries(gx, dev.to_host_amps(dst.a, k), ox[:k], oa[:k], exact=True)


@pytest.mark.parametrize("n,nbits", [(1, 3), (7, 3), (5000, 13), (60000, 20), (300000, 22)])
def test_rdm2q_ordered(gpu_ctx, n, nbits):
    """4x4 RDM of a qubit pair of a key-ordered state (second qubit moved to
    bit 0, first at bit j) against compute_2qubit_rdm, both index orders."""
    ops, dev, SS, L = _mods()
    rng = np.random.default_rng(n + 3 * nbits)
    x, a = random_state(n, nbits, seed=n + 31)
    pos = rng.permutation(nbits)
    q1, q2 = (int(v) for v in rng.choice(nbits, size=2, replace=False))
    for f, s in ((q1, q2), (q2, q1)):
        print(f"[case] rdm2q n={n} nbits={nbits} pair=({f},{s}) bits=({pos[f]},{pos[s]})", flush=True)
        w, aw = _ordered(x, a, pos)
        src = _upload(dev, gpu_ctx, w, aw)
        p1 = pos
        if pos[s] != 0:
            mv = dev.Buffers(n, gpu_ctx.device)
            ops.lift(gpu_ctx, src, n, int(pos[s]), mv, mode=L.MODE_MOVE)
            src, p1 = mv, SS.lifted_positions(pos, int(pos[s]), 0)
        assert p1[s] == 0 and p1[f] >= 1
        rho = ops.rdm2q(gpu_ctx, src, n, int(p1[f]))
        ref = OK.compute_2qubit_rdm(x, a, n, f, s)
        np.testing.assert_allclose(rho, ref, rtol=0, atol=1e-13 * max(1.0, np.abs(ref).max()))
        assert np.allclose(rho, rho.conj().T, rtol=0, atol=0)


@pytest.mark.parametrize("n,nbits", [(1, 4), (9, 4), (5000, 13), (200000, 30)])
def test_compute_2qubit_rdm_module(gpu_ctx, n, nbits):
    """The module-level compute_2qubit_rdm (an unordered state, as the
    reference's kernel takes it) against the oracle: pairs with qubit 0 in
    either place and both orders, device tensors and NumPy arrays."""
    B = pkg_module("simulation.bass_simulator")
    rng = np.random.default_rng(7 * n + nbits)
    x, a = random_state(n, nbits, seed=n + 5)
    perm = rng.permutation(n)
    x, a = x[perm], a[perm]  # input order is not key order
    pairs = [(0, nbits - 1), (nbits - 1, 0), tuple(int(v) for v in rng.choice(nbits, size=2, replace=False))]
    xt = torch.from_numpy(x.view(np.int64)).to(gpu_ctx.device)
    at = torch.from_numpy(a).to(gpu_ctx.device)
    for q1, q2 in pairs:
        print(f"[case] compute_2qubit_rdm n={n} nbits={nbits} pair=({q1},{q2})", flush=True)
        ref = OK.compute_2qubit_rdm(x, a, n, q1, q2)
        tol = 1e-13 * max(1.0, np.abs(ref).max())
        np.testing.assert_allclose(B.compute_2qubit_rdm(xt, at, n, q1, q2), ref, rtol=0, atol=tol)
        np.testing.assert_allclose(B.compute_2qubit_rdm(x, a, n, q1, q2), ref, rtol=0, atol=tol)
    # the inputs are left as they were
    assert np.array_equal(xt.cpu().numpy().view(np.uint64), x)


@pytest.mark.parametrize("n,k,ordered", [(300000, 123457, False), (200000, 150001, True), (2100000, 1048577, True)])
def test_truncate_counted_massive_ties(gpu_ctx, n, k, ordered):
    """Counted truncation (the driver's gate -> truncate pair with the count
    kept on the device) of a gate output whose threshold bucket holds far more
    than FIN_MAX = 65536 exactly tied entries: the threshold and the key
    tie-break are resolved inside the truncate's own launches (multi-block
    stages), with no hand-off to the host.  Against the key-tie-break oracle:
    identical kept keys (bit for bit: every tie is exact), amplitudes within
    1e-12 after the renormalisation; the op log records the truncation and the
    state-count slot holds k."""
    ops, dev, SS, L = _mods()
    print(f"[case] counted massive ties n={n} k={k} ordered={ordered}", flush=True)
    nbits = 40
    rng = np.random.default_rng(n + 17)
    x = np.unique(rng.integers(0, 2 ** nbits, size=n + 1000, dtype=np.uint64))[:n]
    vals = np.array([0.5, 0.25 + 0.25j, 0.125], dtype=np.complex128)
    a = vals[rng.integers(0, 3, size=n)]
    a[rng.permutation(n)[: n // 2]] = vals[1]  # the threshold lands inside this tie class
    pos = rng.permutation(nbits) if ordered else np.arange(nbits)
    w, aw = _ordered(x, a, pos)
    q = int(rng.integers(nbits))
    j = int(pos[q])
    # diag(1, i): every output is its input times 1 or i, exactly (ties stay exact)
    G = np.diag([1.0, 1j]).astype(np.complex128)
    src = _upload(dev, gpu_ctx, w, aw)
    seg = dev.Buffers(1, gpu_ctx.device)
    dst = dev.Buffers(k, gpu_ctx.device)
    ops.counted_begin(gpu_ctx, n, n)
    tiles = ops.lift_counted(gpu_ctx, src, n, j, seg, G=G, nq=1)
    npos = SS.lifted_positions(pos, j, 0)
    ops.truncate_counted(gpu_ctx, seg, k, dst, tiles, key_map=ops.key_map_arg(npos))
    log = ops.counted_read(gpu_ctx, 8)
    assert len(log) == 1
    st = log[0]
    assert st.err == 0 and st.nnz == k and st.n_in == n and (st.flags & 1) and not (st.flags & 2)
    gw = dev.to_host_keys(dst.x, k)
    assert np.all(gw[1:] > gw[:-1]), "truncation must preserve order"
    gx = SS.stored_to_keys(gw, npos)
    ga = dev.to_host_amps(dst.a, k)
    # oracle: the same diagonal gate (exact), key-tie-break top-k, renormalised
    ox = x.copy()
    bit = (ox >> np.uint64(q)) & np.uint64(1)
    oa = np.where(bit == 1, a * 1j, a)
    OK.top_k_truncate_kernel_opt(ox, oa, n, k, selector=OD.key_tiebreak_selector)
    kx, ka = ox[:k].copy(), oa[:k].copy()
    kept = np.sum(np.abs(ka) ** 2)
    assert abs(st.kept - kept) <= 1e-12 * kept
    assert set(gx.tolist()) == set(kx.tolist())
    assert_same_entries(gx, ga, kx, ka / np.sqrt(kept), rtol=1e-12)


def _counted_pair(ops, dev, ctx, w, aw, n, j, G, nq, k, km, spec):
    """One counted lift -> truncate of (w, aw); returns (words, amps, stats)."""
    src = _upload(dev, ctx, w, aw)
    seg = dev.Buffers(1, ctx.device)
    dst = dev.Buffers(k, ctx.device)
    ops.counted_begin(ctx, n, n)
    tiles = ops.lift_counted(ctx, src, n, j, seg, G=G, nq=nq, spec=spec)
    ops.truncate_counted(ctx, seg, k, dst, tiles, key_map=km, spec=spec)
    log = ops.counted_read(ctx, 4)
    assert len(log) == 1 and log[0].err == 0
    m = int(log[0].nnz)
    return dev.to_host_keys(dst.x, m).copy(), dev.to_host_amps(dst.a, m).copy(), log[0]


@pytest.mark.parametrize("n,nq,k", [(70000, 1, 50000), (300000, 1, 123457), (60000, 2, 90001), (5000, 2, 30000)])
def test_speculative_digit1_pass(gpu_ctx, n, nq, k):
    """A gate lift that runs the truncate's digit-1 pass against the last
    select's bucket (bass_lift mode | 32, bass_truncate in_flags bit 3) gives
    the plain pair's output and stats bit for bit: when the prediction holds
    (the same input again: stats flags bit 2 set), when it does not (a select
    on the state scaled by 2^-8 first: bit 2 clear), and when the truncate
    keeps everything (k >= output count)."""
    ops, dev, SS, L = _mods()
    print(f"[case] speculative digit-1 pass n={n} nq={nq} k={k}", flush=True)
    nbits = 36
    rng = np.random.default_rng(n + nq)
    x, a = random_state(n, nbits, seed=n + 3)
    pos = rng.permutation(nbits)
    w, aw = _ordered(x, a, pos)
    G = random_unitary(1 << nq, rng)
    if nq == 1:
        j = int(pos[int(rng.integers(nbits))])
        npos = SS.lifted_positions(pos, j, 0)
    else:
        # the 2-qubit lift needs its second qubit at bit 0
        q2 = int(np.flatnonzero(pos == 0)[0])
        q1 = int(rng.choice([q for q in range(nbits) if q != q2]))
        j = int(pos[q1])
        npos = SS.lifted_positions(pos, j, 1)
    km = ops.key_map_arg(npos)
    ref = _counted_pair(ops, dev, gpu_ctx, w, aw, n, j, G, nq, k, km, spec=False)
    # prediction = the plain pair's bucket
    hit = _counted_pair(ops, dev, gpu_ctx, w, aw, n, j, G, nq, k, km, spec=True)
    # prediction from a state whose |a|^2 sit 16 exponents lower
    _counted_pair(ops, dev, gpu_ctx, w, aw * 2.0 ** -8, n, j, G, nq, k, km, spec=False)
    miss = _counted_pair(ops, dev, gpu_ctx, w, aw, n, j, G, nq, k, km, spec=True)
    keepall = (ref[2].flags & 1) == 0
    assert (hit[2].flags & 4) == (0 if keepall else 4)
    assert (miss[2].flags & 4) == 0
    for got in (hit, miss):
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1].view(np.float64), ref[1].view(np.float64))
        for f in ("nnz", "n_in", "total", "kept", "norm", "sum2", "sum4"):
            assert getattr(got[2], f) == getattr(ref[2], f), f
        assert (got[2].flags & 3) == (ref[2].flags & 3)


@pytest.mark.parametrize("family", ["kicked", "brickwork"])
def test_driver_speculation_changes_nothing(gpu_ctx, family):
    """The counted driver with and without the speculative digit-1 pass:
    identical keys, amplitudes, gamma and frames, and the speculation is used
    on most truncations."""
    print(f"[case] driver speculation on/off: {family}", flush=True)
    rc = pkg_module("utils.random_circuits")
    B = pkg_module("simulation.bass_simulator")
    if family == "kicked":
        N, K = 18, 1 << 14
        circ, ends = rc.kicked_ising_circuit(N, 4)
    else:
        N, K = 16, 1 << 11
        circ, ends = rc.brickwork_haar_circuit(N, 8, np.random.default_rng(77))
    out = []
    for spec in (False, True):
        sim = B.BASS(N, K, device="cuda:0", optimize_at=ends, accept_rtol=1e-10)
        sim.speculate = spec
        st = sim.simulate(circ)
        out.append((st.keys().copy(), st.alpha[: st.nnz].copy(), st.gamma, [u.copy() for u in sim.U],
                    dict(sim.counters)))
    (k0, a0, g0, u0, c0), (k1, a1, g1, u1, c1) = out
    assert np.array_equal(k0, k1) and np.array_equal(a0.view(np.float64), a1.view(np.float64))
    assert g0 == g1 and all(np.array_equal(p, q) for p, q in zip(u0, u1))
    assert c0.get("spec_hit", 0) == 0
    print(f"  truncations {c1.get('truncate', 0)}, speculation used {c1.get('spec_hit', 0)}", flush=True)
    assert c1.get("spec_hit", 0) >= c1.get("truncate", 0) // 2
