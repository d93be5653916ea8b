"""Device kernels vs the CPU oracle on seeded inputs (one op at a time)."""

import numpy as np
import pytest
import torch

from _parity import assert_same_entries, assert_topk_sets, numpy_uses_fma_cmul
from conftest import pkg_module, random_state, random_unitary
from oracle import drivers as OD
from oracle import kernels as OK

pytestmark = pytest.mark.gpu

EXACT = None


def exact():
    global EXACT
    if EXACT is None:
        EXACT = numpy_uses_fma_cmul()
    return EXACT


def ss():
    return pkg_module("core.sparse_state")


def bs():
    return pkg_module("simulation.bass_simulator")


def tr():
    return pkg_module("core.truncation")


def _gate_with_zero(rng, d):
    G = random_unitary(d, rng)
    G[0, d - 1] = 0.0  # exercise the exact-zero emission rule
    return G


@pytest.mark.parametrize("n,nbits,q", [(1, 4, 0), (777, 12, 3), (20000, 20, 19), (5000, 16, 0)])
def test_emit_1q_reference_order(gpu_ctx, n, nbits, q):
    print(f"[case] emit1q n={n} nbits={nbits} q={q}", flush=True)
    rng = np.random.default_rng(n + q)
    x, a = random_state(n, nbits, seed=n)
    G = _gate_with_zero(rng, 2)
    ox, oa = np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.complex128)
    m = OK.apply_1qubit_gate_kernel_opt(x, a, n, ox, oa, G, q)
    gx, ga = np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.complex128)
    g = ss().apply_1qubit_gate_kernel_opt(x, a, n, gx, ga, G, q)
    assert g == m
    assert np.array_equal(gx[:m], ox[:m]), "emission order differs"
    np.testing.assert_allclose(ga[:m], oa[:m], rtol=1e-12, atol=1e-15)
    if exact():
        assert np.array_equal(ga[:m].view(np.float64), oa[:m].view(np.float64))


def test_emit_2q_reference_order_and_overflow(gpu_ctx):
    print("[case] emit2q", flush=True)
    rng = np.random.default_rng(5)
    n = 3000
    x, a = random_state(n, 14, seed=11)
    G = _gate_with_zero(rng, 4)
    ox, oa = np.zeros(4 * n, np.uint64), np.zeros(4 * n, np.complex128)
    m = OK.apply_2qubit_gate_kernel_opt(x, a, n, ox, oa, G, 5, 2)
    gx, ga = np.zeros(4 * n, np.uint64), np.zeros(4 * n, np.complex128)
    g = ss().apply_2qubit_gate_kernel_opt(x, a, n, gx, ga, G, 5, 2)
    assert g == m
    assert np.array_equal(gx[:m], ox[:m])
    np.testing.assert_allclose(ga[:m], oa[:m], rtol=1e-12, atol=1e-15)
    small_x, small_a = np.zeros(m - 1, np.uint64), np.zeros(m - 1, np.complex128)
    assert ss().apply_2qubit_gate_kernel_opt(x, a, n, small_x, small_a, G, 5, 2) == -1
    assert OK.apply_2qubit_gate_kernel_opt(x, a, n, small_x, small_a, G, 5, 2) == -1


def _oracle_1q_dedup(x, a, n, G, q):
    ox, oa = np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.complex128)
    m = OK.apply_1qubit_gate_kernel_opt(x, a, n, ox, oa, G, q)
    OK.radix_sort_uint64(ox, oa, m)
    m = OK.merge_sorted_kernel(ox, oa, m)
    return ox[:m], oa[:m]


@pytest.mark.parametrize("n,nbits,q", [(2, 3, 1), (4096, 12, 0), (30000, 18, 7), (30000, 15, 14)])
def test_gate_1q_dedup(gpu_ctx, n, nbits, q):
    print(f"[case] gate1q-dedup n={n} nbits={nbits} q={q}", flush=True)
    rng = np.random.default_rng(q)
    x, a = random_state(n, nbits, seed=3 * n + q)
    G = _gate_with_zero(rng, 2)
    okx, oka = _oracle_1q_dedup(x, a, n, G, q)
    ops = pkg_module("_ops")
    dev = pkg_module("_device")
    src = dev.Buffers(n, gpu_ctx.device)
    src.x[:n] = torch.from_numpy(x.view(np.int64)).to(gpu_ctx.device)
    src.a[:n] = torch.from_numpy(a).to(gpu_ctx.device)
    dst = dev.Buffers(2 * n, gpu_ctx.device)
    m, _ = ops.gate(gpu_ctx, src, n, G, (q,), dst, dedup=True)
    assert m == len(okx)
    gx = dev.to_host_keys(dst.x, m)
    ga = dev.to_host_amps(dst.a, m)
    assert_same_entries(gx, ga, okx, oka, exact=exact())


@pytest.mark.parametrize("nq,n,k", [(1, 1, 1), (1, 3000, 2000), (1, 70000, 50000), (2, 5000, 3000),
                                    (2, 40000, 60000)])
def test_segmented_gate_matches_dense(gpu_ctx, nq, n, k):
    print(f"[case] segmented nq={nq} n={n} k={k}", flush=True)
    ops = pkg_module("_ops")
    dev = pkg_module("_device")
    rng = np.random.default_rng(n + nq)
    x, a = random_state(n, 20, seed=n + 3)
    G = _gate_with_zero(rng, 2 ** nq)
    qs = (5,) if nq == 1 else (5, 12)
    src = dev.Buffers(n, gpu_ctx.device)
    src.x[:n] = torch.from_numpy(x.view(np.int64)).to(gpu_ctx.device)
    src.a[:n] = torch.from_numpy(a).to(gpu_ctx.device)
    dense = dev.Buffers(1, gpu_ctx.device)
    m, _ = ops.gate(gpu_ctx, src, n, G, qs, dense, dedup=True)
    segb = dev.Buffers(1, gpu_ctx.device)
    m2, tiles = ops.gate(gpu_ctx, src, n, G, qs, segb, dedup=True, segmented=True)
    assert m2 == m
    out = dev.Buffers(1, gpu_ctx.device)
    ops.densify(gpu_ctx, segb, tiles, m2, out)
    assert np.array_equal(dev.to_host_keys(out.x, m), dev.to_host_keys(dense.x, m)), "order must match"
    assert np.array_equal(dev.to_host_amps(out.a, m), dev.to_host_amps(dense.a, m))
    if m > k:
        t1 = dev.Buffers(k, gpu_ctx.device)
        s1 = ops.truncate(gpu_ctx, dense, m, k, t1, renorm=True)
        m3, tiles = ops.gate(gpu_ctx, src, n, G, qs, segb, dedup=True, segmented=True)
        t2 = dev.Buffers(k, gpu_ctx.device)
        s2 = ops.truncate(gpu_ctx, segb, m3, k, t2, renorm=True, hist_tiles=tiles, segmented=True)
        assert np.array_equal(dev.to_host_keys(t1.x, k), dev.to_host_keys(t2.x, k))
        # the two inputs are tiled differently, so the kept sum (and the
        # renormalisation) may differ in the last bit
        np.testing.assert_allclose(dev.to_host_amps(t1.a, k), dev.to_host_amps(t2.a, k), rtol=1e-15, atol=0)
        assert abs(s1.kept - s2.kept) <= 1e-15 * s1.kept and abs(s1.total - s2.total) <= 1e-15 * s1.total


@pytest.mark.parametrize("n,nbits,q1,q2", [(3, 3, 0, 2), (5000, 13, 4, 1), (40000, 18, 17, 0)])
def test_gate_2q_ht(gpu_ctx, n, nbits, q1, q2):
    print(f"[case] gate2q-ht n={n} q=({q1},{q2})", flush=True)
    rng = np.random.default_rng(n)
    x, a = random_state(n, nbits, seed=n + 7)
    G = _gate_with_zero(rng, 4)
    ox, oa = np.zeros(4 * n, np.uint64), np.zeros(4 * n, np.complex128)
    m = OK.apply_2qubit_gate_ht(x, a, n, ox, oa, G, q1, q2)
    gx, ga = np.zeros(4 * n, np.uint64), np.zeros(4 * n, np.complex128)
    g = bs().apply_2qubit_gate_ht(x, a, n, gx, ga, G, q1, q2, None, None)
    assert g == m
    assert_same_entries(gx[:g], ga[:g], ox[:m], oa[:m], exact=exact())


@pytest.mark.parametrize("nq", [1, 2])
def test_rotation(gpu_ctx, nq):
    print(f"[case] rotation nq={nq}", flush=True)
    rng = np.random.default_rng(40 + nq)
    n = 6000
    x, a = random_state(n, 14, seed=99)
    U = random_unitary(2 ** nq, rng)
    ox, oa = np.zeros(4 * n, np.uint64), np.zeros(4 * n, np.complex128)
    ox[:n], oa[:n] = x, a
    gx, ga = ox.copy(), oa.copy()
    if nq == 1:
        m = OK.apply_1qubit_rotation_sparse(ox, oa, n, 9, U)
        g = bs().apply_1qubit_rotation_sparse(gx, ga, n, 9, U)
    else:
        m = OK.apply_2qubit_rotation_sparse(ox, oa, n, 3, 11, U)
        g = bs().apply_2qubit_rotation_sparse(gx, ga, n, 3, 11, U)
    assert g == m
    assert_same_entries(gx[:g], ga[:g], ox[:m], oa[:m], exact=exact())


def test_diagonal(gpu_ctx):
    print("[case] diag", flush=True)
    n = 10000
    x, a = random_state(n, 16, seed=5)
    d = np.exp(1j * np.array([0.1, -0.4, 0.7, 2.0]))
    oa = a.copy()
    OK.apply_2qubit_diagonal_kernel(x, oa, n, 6, 2, d)
    ga = a.copy()
    ss().apply_2qubit_diagonal_kernel(x, ga, n, 6, 2, d)
    np.testing.assert_allclose(ga, oa, rtol=1e-12, atol=0)
    if exact():
        assert np.array_equal(ga.view(np.float64), oa.view(np.float64))


@pytest.mark.parametrize("n,k,ties", [(10, 3, False), (5000, 2500, False), (70000, 30000, False),
                                      (4096, 1000, True), (200000, 65536, False)])
def test_top_k(gpu_ctx, n, k, ties):
    print(f"[case] topk n={n} k={k} ties={ties}", flush=True)
    x, a = random_state(n, 24, seed=n + k, ties=ties)
    gx, ga = x.copy(), a.copy()
    m, gsq = tr().top_k_truncate_kernel_opt(gx, ga, n, k)
    ox, oa = x.copy(), a.copy()
    m2, osq = OK.top_k_truncate_kernel_opt(ox, oa, n, k, selector=OD.key_tiebreak_selector)
    assert m == m2 == k
    assert set(gx[:k].tolist()) == set(ox[:k].tolist()), "key tie-break selection differs"
    assert_same_entries(gx[:k], ga[:k], ox[:k], oa[:k], exact=True)
    assert abs(gsq - osq) <= 1e-12 * osq
    # against the reference's argpartition rule: differences only inside the band
    rx, ra = x.copy(), a.copy()
    OK.top_k_truncate_kernel_opt(rx, ra, n, k)
    assert_topk_sets(gx[:k], rx[:k], x, a, k)


def test_top_k_all_equal(gpu_ctx):
    print("[case] topk all-equal", flush=True)
    n, k = 4000, 1234
    rng = np.random.default_rng(1)
    x = rng.choice(2 ** 30, size=n, replace=False).astype(np.uint64)
    a = np.full(n, 1 / np.sqrt(n), dtype=np.complex128)
    gx, ga = x.copy(), a.copy()
    tr().top_k_truncate_kernel_opt(gx, ga, n, k)
    assert set(gx[:k].tolist()) == set(np.sort(x)[:k].tolist())


def test_truncate_renorm_stats(gpu_ctx):
    print("[case] truncate renorm", flush=True)
    ops = pkg_module("_ops")
    dev = pkg_module("_device")
    n, k = 50000, 12345
    x, a = random_state(n, 22, seed=8)
    src = dev.Buffers(n, gpu_ctx.device)
    src.x[:n] = torch.from_numpy(x.view(np.int64)).to(gpu_ctx.device)
    src.a[:n] = torch.from_numpy(a).to(gpu_ctx.device)
    dst = dev.Buffers(k, gpu_ctx.device)
    st = ops.truncate(gpu_ctx, src, n, k, dst, renorm=True)
    sel = OD.key_tiebreak_selector(x, a, None, k)
    kept = np.sum(np.abs(a[sel]) ** 2)
    assert abs(st.total - np.sum(np.abs(a) ** 2)) < 1e-13
    assert abs(st.kept - kept) < 1e-13
    ga = dev.to_host_amps(dst.a, k)
    gx = dev.to_host_keys(dst.x, k)
    assert np.array_equal(gx, x[sel]), "compaction must keep input order"
    np.testing.assert_allclose(ga, a[sel] / np.sqrt(st.kept), rtol=1e-15, atol=0)
    assert np.array_equal(ga, a[sel] / np.sqrt(st.kept)), "renormalisation must round like numpy"
    assert abs(st.sum4 - np.sum(np.abs(ga) ** 4)) < 1e-12 * st.sum4


@pytest.mark.parametrize("n,k", [(9_000_001, 4_000_000), (34_000_001, 15_000_000)])
def test_truncate_chunked(gpu_ctx, n, k):
    """More than 4096 tiles of 2048 entries: the per-tile bookkeeping (kept
    offsets, kept / input / output sums) reduces by chunk in two levels.  From
    16384 tiles on (the second case) the first candidate pass takes a tile
    count per block chosen to fill whole rounds of resident blocks."""
    print(f"[case] truncate chunked n={n} k={k}", flush=True)
    ops = pkg_module("_ops")
    dev = pkg_module("_device")
    x, a = random_state(n, 40, seed=9)
    src = dev.Buffers(n, gpu_ctx.device)
    src.x[:n] = torch.from_numpy(x.view(np.int64)).to(gpu_ctx.device)
    src.a[:n] = torch.from_numpy(a).to(gpu_ctx.device)
    dst = dev.Buffers(k, gpu_ctx.device)
    st = ops.truncate(gpu_ctx, src, n, k, dst, renorm=True)
    sel = OD.key_tiebreak_selector(x, a, None, k)
    p = a.real * a.real + a.imag * a.imag
    assert st.nnz == k
    assert abs(st.total - p.sum()) < 1e-12 and abs(st.kept - p[sel].sum()) < 1e-12
    gx = dev.to_host_keys(dst.x, k)
    ga = dev.to_host_amps(dst.a, k)
    assert np.array_equal(gx, x[sel]), "compaction must keep input order"
    assert np.array_equal(ga, a[sel] / np.sqrt(st.kept)), "renormalisation must round like numpy"
    q = ga.real * ga.real + ga.imag * ga.imag
    assert abs(st.sum2 - q.sum()) < 1e-12 and abs(st.sum4 - np.sum(q * q)) < 1e-12 * st.sum4


def test_segmented_gate_chunked(gpu_ctx):
    """A 1-qubit gate on 5M entries: more than 4096 output tiles, so the
    segment scan, densify and the truncation of the segmented output run their
    chunked forms; all must agree with the dense path."""
    n, k = 5_000_000, 4_500_000
    print(f"[case] segmented chunked n={n} k={k}", flush=True)
    ops = pkg_module("_ops")
    dev = pkg_module("_device")
    rng = np.random.default_rng(5)
    x, a = random_state(n, 36, seed=11)
    G = _gate_with_zero(rng, 2)
    src = dev.Buffers(n, gpu_ctx.device)
    src.x[:n] = torch.from_numpy(x.view(np.int64)).to(gpu_ctx.device)
    src.a[:n] = torch.from_numpy(a).to(gpu_ctx.device)
    dense = dev.Buffers(1, gpu_ctx.device)
    m, _ = ops.gate(gpu_ctx, src, n, G, (7,), dense, dedup=True)
    segb = dev.Buffers(1, gpu_ctx.device)
    m2, tiles = ops.gate(gpu_ctx, src, n, G, (7,), segb, dedup=True, segmented=True)
    assert tiles > 4096 and m2 == m
    out = dev.Buffers(1, gpu_ctx.device)
    ops.densify(gpu_ctx, segb, tiles, m2, out)
    assert torch.equal(out.x[:m], dense.x[:m]) and torch.equal(out.a[:m], dense.a[:m])
    t1 = dev.Buffers(k, gpu_ctx.device)
    s1 = ops.truncate(gpu_ctx, dense, m, k, t1, renorm=True)
    m3, tiles = ops.gate(gpu_ctx, src, n, G, (7,), segb, dedup=True, segmented=True)
    t2 = dev.Buffers(k, gpu_ctx.device)
    s2 = ops.truncate(gpu_ctx, segb, m3, k, t2, renorm=True, hist_tiles=tiles, segmented=True)
    assert torch.equal(t1.x[:k], t2.x[:k])
    np.testing.assert_allclose(dev.to_host_amps(t1.a, k), dev.to_host_amps(t2.a, k), rtol=1e-15, atol=0)
    assert abs(s1.kept - s2.kept) <= 1e-14 * s1.kept and abs(s1.total - s2.total) <= 1e-14 * s1.total
    # against the oracle's selection on the dense gate output
    hx, ha = dev.to_host_keys(dense.x, m), dev.to_host_amps(dense.a, m)
    sel = OD.key_tiebreak_selector(hx, ha, None, k)
    assert np.array_equal(dev.to_host_keys(t2.x, k), hx[sel])


def test_rdm(gpu_ctx):
    print("[case] rdm", flush=True)
    n, N = 20000, 18
    x, a = random_state(n, N, seed=21)
    ref = OK.compute_all_1qubit_rdms(x, a, n, N)
    got = bs().compute_all_1qubit_rdms(x, a, n, N)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("n,bits", [(2, 5), (3000, 12), (100000, 40), (70001, 64)])
def test_radix_sort_and_merge(gpu_ctx, n, bits):
    print(f"[case] radix n={n} bits={bits}", flush=True)
    rng = np.random.default_rng(n)
    hi = 2 ** bits if bits < 64 else 2 ** 64
    x = rng.integers(0, hi // 2, size=n, dtype=np.uint64) * np.uint64(2) if bits < 64 else \
        rng.integers(0, 2 ** 63, size=n, dtype=np.uint64) * np.uint64(2)
    x[: n // 3] = x[n // 3: 2 * (n // 3)]  # duplicates
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    gx, ga = x.copy(), a.copy()
    ss().radix_sort_uint64(gx, ga, n)
    order = np.argsort(x, kind="stable")
    assert np.array_equal(gx, x[order])
    assert np.array_equal(ga, a[order]), "payload not stable"
    ox, oa = gx.copy(), ga.copy()
    m = OK.merge_sorted_kernel(ox, oa, n)
    g = ss().merge_sorted_kernel(gx, ga, n)
    assert g == m
    assert np.array_equal(gx[:m], ox[:m])
    np.testing.assert_allclose(ga[:m], oa[:m], rtol=1e-15, atol=0)


@pytest.mark.parametrize("n,bits,digit,kind", [
    (4095, 24, 0, "random"), (4096, 20, 0, "random"), (4097, 40, 0, "random"),
    (300000, 40, 8, "random"), (300000, 40, 10, "random"), (300000, 56, 0, "random"),
    (250001, 64, 10, "random"), (250001, 64, 0, "random"), (123457, 33, 0, "random"),
    (200000, 40, 0, "few"), (200000, 40, 10, "equal"), (200000, 48, 0, "sorted"),
    (200000, 40, 0, "reversed"), (1 << 23, 40, 0, "pairs")])
def test_radix_sort_one_sweep(gpu_ctx, n, bits, digit, kind):
    """Stable sort against numpy's stable argsort, for both digit widths, tile
    edges (4096 entries per tile), long runs of one digit, presorted input and
    the emitted pair layout of a 1-qubit gate (k & ~m, k | m)."""
    print(f"[case] one-sweep sort n={n} bits={bits} digit={digit} kind={kind}", flush=True)
    rng = np.random.default_rng(n + bits)
    top = (1 << bits) if bits < 64 else (1 << 64)
    if kind == "few":
        x = rng.choice(rng.integers(0, top, size=7, dtype=np.uint64), size=n)
    elif kind == "equal":
        x = np.full(n, top - 1, np.uint64)
    elif kind == "pairs":
        k = rng.integers(0, top, size=n // 2, dtype=np.uint64)
        m = np.uint64(1 << 17)
        x = np.stack([k & ~m, k | m], axis=1).reshape(-1)
    else:
        x = rng.integers(0, top - 1, size=n, dtype=np.uint64, endpoint=True) if bits == 64 else \
            rng.integers(0, top, size=n, dtype=np.uint64)
        if kind == "sorted":
            x = np.sort(x)
        elif kind == "reversed":
            x = np.sort(x)[::-1].copy()
    a = (np.arange(n) + 1j * rng.standard_normal(n)).astype(np.complex128)  # real part = input index
    lib = gpu_ctx.lib
    old = lib.bass_sort_digit_bits(digit)
    try:
        gx, ga = x.copy(), a.copy()
        ss().radix_sort_uint64(gx, ga, n, bits)
    finally:
        lib.bass_sort_digit_bits(old)
    order = np.argsort(x, kind="stable")
    assert np.array_equal(gx, x[order])
    assert np.array_equal(ga, a[order]), "payload not stable"


def test_hash_table(gpu_ctx):
    print("[case] hash", flush=True)
    n = 50000
    x, _ = random_state(n, 40, seed=2)
    cap = 1 << 17
    hk = torch.empty(cap, dtype=torch.int64, device=gpu_ctx.device)
    hv = torch.empty(cap, dtype=torch.int32, device=gpu_ctx.device)
    B = bs()
    B.ht_clear(hk)
    xt = torch.from_numpy(x.view(np.int64)).to(gpu_ctx.device)
    B.ht_build(xt, n, hk, hv)
    q = np.concatenate([x[::7], x[::11] ^ np.uint64(1 << 41)])
    got = B.ht_lookup_batch(hk, hv, torch.from_numpy(q.view(np.int64)).to(gpu_ctx.device)).cpu().numpy()
    pos = {int(k): i for i, k in enumerate(x.tolist())}
    exp = np.array([pos.get(int(k), -1) for k in q.tolist()])
    assert np.array_equal(got, exp)
    assert B.ht_lookup(hk, hv, int(x[123])) == 123
    assert B.ht_lookup(hk, hv, int(x[123]) ^ (1 << 45)) == -1


def test_random_k(gpu_ctx):
    print("[case] random-k", flush=True)
    n, k = 3000, 1000
    x, a = random_state(n, 20, seed=4)
    gx, ga = x.copy(), a.copy()
    m, gsq = tr().random_k_truncate_kernel_opt(gx, ga, n, k, 12345)
    ox, oa = x.copy(), a.copy()
    m2, osq = OK.random_k_truncate_kernel_opt(ox, oa, n, k, 12345)
    assert m == m2 == k
    assert np.array_equal(gx[:k], ox[:k])
    assert np.array_equal(ga[:k], oa[:k])
    assert abs(gsq - osq) < 1e-13


@pytest.mark.parametrize("dedup", ["sort", "table"])
@pytest.mark.parametrize("n,q1,q2", [(40, 5, 1), (3000, 2, 9), (20000, 11, 0)])
def test_sparse_state_2q_paths(gpu_ctx, dedup, n, q1, q2):
    """SparseState.apply_gate for a 2-qubit gate (sparse_state.py:303-377):
    emit, stable radix sort, merge (dedup="sort"; the reference's path, key
    order and sums) or the group table (dedup="table", same entries and sums)
    against the oracle's SparseState, with and without truncation."""
    print(f"[case] sparse state 2q dedup={dedup} n={n} q=({q1},{q2})", flush=True)
    G = pkg_module("core.gates")
    rng = np.random.default_rng(n + q1)
    N = 16
    x, a = random_state(n, N, seed=n + 5)
    U = _gate_with_zero(rng, 4)
    for k in (4 * n, n // 2 + 1):
        st = ss().SparseState(N, k, device="cuda:0", dedup=dedup, capacity=4 * n)
        st._buf.ensure(n)
        st._buf.x[:n] = torch.from_numpy(x.view(np.int64)).to("cuda:0")
        st._buf.a[:n] = torch.from_numpy(a).to("cuda:0")
        st.nnz = n
        st.apply_gate(G.QuantumGate([q1, q2], U), "top-k")
        o = OD.SparseState(N, k, selector=OD.key_tiebreak_selector)
        o._ensure_temp(4 * n)
        o.x, o.alpha = np.zeros(4 * n, np.uint64), np.zeros(4 * n, np.complex128)
        o.x[:n], o.alpha[:n], o.nnz = x, a, n
        o.apply_gate(OD._G((q1, q2), U), "top-k")
        gx, ga = st.x[: st.nnz], st.alpha[: st.nnz]
        assert st.nnz == o.nnz
        if dedup == "sort" and k >= 4 * n:
            assert np.array_equal(gx, o.x[: o.nnz]), "sort path: reference key order"
        assert_same_entries(gx, ga, o.x[: o.nnz], o.alpha[: o.nnz],
                            exact=(k >= 4 * n and exact()), rtol=1e-13)
        assert abs(st.gamma - o.gamma) <= 1e-13


@pytest.mark.parametrize("dedup", ["sort", "table"])
def test_sparse_state_random_k_sequence(gpu_ctx, dedup):
    """SparseState.apply_gate(..., "random-k", seed) over a sequence of 1- and
    2-qubit gates against the oracle's SparseState: the reference draws the
    kept entries by array index from its sort-and-merge output (ascending
    keys, src/core/sparse_state.py:268-301, :303-377, :402-415), so the
    device's state must hold the same keys in the same order after every
    gate, whichever way it deduplicates; amplitudes and gamma within 1e-12."""
    print(f"[case] sparse state random-k sequence dedup={dedup}", flush=True)
    rng = np.random.default_rng(77)
    N, k = 12, 150
    st = ss().SparseState(N, k, device="cuda:0", dedup=dedup)
    o = OD.SparseState(N, k)
    draws = 0
    for i in range(24):
        if i % 4 == 3:
            q1, q2 = (int(v) for v in rng.choice(N, size=2, replace=False))
            g = OD._G((q1, q2), random_unitary(4, rng))
        else:
            g = OD._G((int(rng.integers(N)),), random_unitary(2, rng))
        seed = 1000 + 7 * i
        g_before = o.gamma
        st.apply_gate(g, "random-k", seed)
        o.apply_gate(g, "random-k", seed)
        assert st.nnz == o.nnz
        gk, ga = st.keys(), st.alpha[: st.nnz]
        if o.gamma != g_before:
            # a draw happened: the kept entries in draw order, as the reference's
            draws += 1
            assert np.array_equal(gk, o.x[: o.nnz]), f"gate {i}: kept keys or their order differ"
            np.testing.assert_allclose(ga, o.alpha[: o.nnz], rtol=1e-12, atol=1e-15)
        else:
            # no truncation: same entries (the order of a gate's output is free)
            og, oo = np.argsort(gk), np.argsort(o.x[: o.nnz])
            assert np.array_equal(gk[og], o.x[: o.nnz][oo])
            np.testing.assert_allclose(ga[og], o.alpha[: o.nnz][oo], rtol=1e-12, atol=1e-15)
        assert abs(st.gamma - o.gamma) <= 1e-12 * o.gamma
    assert o.nnz == k and draws >= 10


def test_tuning_knobs_keep_results(gpu_ctx):
    """Every value of every tuning knob (table capacity factor, emission tile
    size, radix digit width; bass_tune) leaves the results bit for bit
    unchanged: a sequence of 1- and 2-qubit gates with truncations through
    SparseState (both dedup modes) at the size class of its state."""
    print("[case] tuning knobs keep results", flush=True)
    L = pkg_module("_lib")
    T = pkg_module("tuning")
    lib = L.load()
    rng = np.random.default_rng(31)
    N, n0, k = 30, 30000, 20000
    x0, a0 = random_state(n0, N, seed=9)
    gates = []
    for i in range(6):
        if i % 3 == 2:
            q1, q2 = (int(v) for v in rng.choice(N, size=2, replace=False))
            gates.append(OD._G((q1, q2), random_unitary(4, rng)))
        else:
            gates.append(OD._G((int(rng.integers(N)),), random_unitary(2, rng)))

    def run(dedup):
        st = ss().SparseState(N, k, device="cuda:0", dedup=dedup, capacity=4 * n0)
        st._buf.ensure(n0)
        st._buf.x[:n0] = torch.from_numpy(x0.view(np.int64)).to("cuda:0")
        st._buf.a[:n0] = torch.from_numpy(a0).to("cuda:0")
        st.nnz = n0
        for g in gates:
            st.apply_gate(g, "top-k")
        return st.keys().copy(), st.alpha[: st.nnz].copy(), st.gamma

    saved = {name: [lib.bass_tune(knob, c, 0) for c in range(4)] for name, knob in T.KNOBS.items()}
    try:
        ref = {d: run(d) for d in ("table", "sort")}
        for name, knob in T.KNOBS.items():
            for v in T.VALUES[name]:
                for c in range(4):
                    lib.bass_tune(knob, c, v)
                for d in ("table", "sort"):
                    kx, ka, gm = run(d)
                    rk, ra, rg = ref[d]
                    assert np.array_equal(kx, rk) and np.array_equal(ka.view(np.float64), ra.view(np.float64)), \
                        (name, v, d)
                    assert gm == rg
                for c in range(4):
                    lib.bass_tune(knob, c, 0)
    finally:
        for name, knob in T.KNOBS.items():
            for c, v in enumerate(saved[name]):
                lib.bass_tune(knob, c, v)


def test_fresh_workspace_gate_then_truncate(gpu_ctx):
    """A 2-qubit gate whose tiles are smaller than the select's (emit_items_2q
    = 1: 256-entry tiles, twice the default count) followed by the truncate
    that reads its histogram, on a context whose workspace starts empty: the
    truncate finds the workspace the gate sized (it must not be replaced
    between the two) and every tile setting gives the default's keys and
    amplitudes bit for bit (the fixed-basis driver, N=20, k=10000, a 6-layer
    brickwork); its gamma multiplies kept/total ratios whose input totals are
    summed per emission tile, so the tile size may move gamma by a few ulps
    (1e-14 relative here)."""
    print("[case] fresh workspace: gate then truncate, every 2q tile setting", flush=True)
    L = pkg_module("_lib")
    T = pkg_module("tuning")
    D = pkg_module("_device")
    R = pkg_module("experimentation.runner")
    S = pkg_module("simulation.simulator")
    lib = L.load()
    circ = R.make_brickwork(20, 6, np.random.default_rng(20260520 * 10 ** 6))
    knob = T.KNOBS["emit_items_2q"]
    saved = [lib.bass_tune(knob, c, 0) for c in range(4)]
    key = str(torch.device("cuda:0"))
    shared = D._CONTEXTS.pop(key)
    try:
        out = []
        for v in (0,) + T.VALUES["emit_items_2q"]:
            for c in range(4):
                lib.bass_tune(knob, c, v)
            D._CONTEXTS.pop(key, None)  # a new context: its workspace grows from nothing
            st = S.FixedBasisSimulator(20, 10000, verbose=False, device="cuda:0").simulate(circ, seed=1)
            out.append((st.keys().copy(), st.alpha[: st.nnz].copy(), st.gamma))
        for v, (kx, ka, gm) in zip(T.VALUES["emit_items_2q"], out[1:]):
            assert np.array_equal(kx, out[0][0]), v
            assert np.array_equal(ka.view(np.float64), out[0][1].view(np.float64)), v
            assert abs(gm - out[0][2]) <= 1e-14 * out[0][2], v
    finally:
        D._CONTEXTS[key] = shared
        for c, v in enumerate(saved):
            lib.bass_tune(knob, c, v)


@pytest.mark.parametrize("n,ngates,count", [(2000, 40, None), (300001, 45, None), (70000, 56, 69877),
                                            (5000, 1, None), (100, 0, None)])
def test_diagonal_run(gpu_ctx, n, ngates, count):
    """bass_diag2q_run (a run of diagonal 2-qubit gates in one pass per 32
    gates) against one bass_diag2q launch per gate on the same state: bit for
    bit; with the count slot (counted) only its first `count` entries change.
    At n=2000 also against the oracle's per-gate kernel
    (src/core/sparse_state.py:179-190) applied gate by gate."""
    print(f"[case] diag run n={n} gates={ngates} count={count}", flush=True)
    D = pkg_module("_device")
    O = pkg_module("_ops")
    ctx = gpu_ctx
    rng = np.random.default_rng(n + ngates)
    x, a = random_state(n, 56, seed=n)
    gates = []
    for _ in range(ngates):
        q1, q2 = (int(v) for v in rng.choice(56, size=2, replace=False))
        d = rng.standard_normal(4) + 1j * rng.standard_normal(4)
        if rng.random() < 0.5:
            d = np.exp(1j * rng.uniform(-np.pi, np.pi, 4))
        gates.append((q1, q2, d))

    def load():
        b = D.Buffers(n, "cuda:0")
        b.x[:n] = torch.from_numpy(x.view(np.int64)).to("cuda:0")
        b.a[:n] = torch.from_numpy(a).to("cuda:0")
        return b

    m = n if count is None else count
    seq = load()
    for q1, q2, d in gates:
        O.diag2q(ctx, seq, m, q1, q2, d)
    run = load()
    if count is None:
        O.diag2q_run(ctx, run, n, gates)
    else:
        O.counted_begin(ctx, count, n)
        O.diag2q_run(ctx, run, n, gates, counted=True)
    ga = run.a[:n].cpu().numpy()
    sa = seq.a[:n].cpu().numpy()
    assert np.array_equal(ga.view(np.float64), sa.view(np.float64))
    assert np.array_equal(ga[m:].view(np.float64), a[m:].view(np.float64))
    if n <= 2000:
        oa = a.copy()
        for q1, q2, d in gates:
            OK.apply_2qubit_diagonal_kernel(x, oa, n, q1, q2, d)
        np.testing.assert_allclose(ga, oa, rtol=1e-12, atol=0)
        if exact():
            assert np.array_equal(ga.view(np.float64), oa.view(np.float64))
