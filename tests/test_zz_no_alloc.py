"""bass_truncate_noalloc (scratch in the library's static device array, no
workspace, no allocation) against bass_truncate (caller workspace), through
the C ABI: two calls in a row, the second on a smaller input, must give the
same kept entries bit for bit and the same stats."""

import ctypes

import numpy as np
import pytest
import torch

from conftest import pkg_module, random_state

pytestmark = pytest.mark.gpu

MAX = 1 << 20


def _noalloc(ctx, lib, src, n, k, km, L):
    dev = pkg_module("_device")
    dst = dev.Buffers(k, ctx.device)
    st = torch.zeros(8, dtype=torch.int64, device=ctx.device)  # one bass_stats (64 bytes)
    kp, kb = (None, 0) if km is None else (km.ctypes.data_as(ctypes.c_void_p), len(km))
    rc = lib.bass_truncate_noalloc(src.xp, src.ap, n, k, dst.xp, dst.ap, dst.cap, 1, kp, kb,
                                   ctypes.c_void_p(st.data_ptr()), ctx.stream)
    L.check(rc, "bass_truncate_noalloc")
    torch.cuda.synchronize()
    raw = st.cpu().numpy().tobytes()
    return dst, L.BassStats.from_buffer_copy(raw)


def _with_ws(ctx, src, n, k, km):
    ops, dev = pkg_module("_ops"), pkg_module("_device")
    dst = dev.Buffers(k, ctx.device)
    st = ops.truncate(ctx, src, n, k, dst, renorm=True, key_map=km)
    return dst, st


@pytest.mark.parametrize("ties", [False, True])
def test_truncate_noalloc_matches_workspace_call(gpu_ctx, ties):
    L = pkg_module("_lib")
    dev = pkg_module("_device")
    ops = pkg_module("_ops")
    lib = L.load()
    nbits = 30
    rng = np.random.default_rng(5 + ties)
    pos = rng.permutation(nbits)
    km = ops.key_map_arg(pos)
    for n, k in [(MAX, 300001), (70001, 9000)]:  # the second call smaller than the first
        print(f"[case] no-alloc truncate n={n} k={k} ties={ties}", flush=True)
        x, a = random_state(n, nbits, seed=n + 3, ties=ties)
        if ties:
            a[: n // 3] = a[0]  # a threshold bucket of > 65536 equal |a|^2 on the first call
        src = dev.Buffers(n, gpu_ctx.device)
        src.x[:n] = torch.from_numpy(x.view(np.int64)).to(gpu_ctx.device)
        src.a[:n] = torch.from_numpy(a).to(gpu_ctx.device)
        d0, s0 = _noalloc(gpu_ctx, lib, src, n, k, km, L)
        d1, s1 = _with_ws(gpu_ctx, src, n, k, km)
        assert s0.err == 0 and s0.nnz == k == s1.nnz and s0.n_in == n
        assert s0.kept == s1.kept and s0.total == s1.total and s0.norm == s1.norm
        assert s0.sum2 == s1.sum2 and s0.sum4 == s1.sum4 and s0.flags == s1.flags
        assert torch.equal(d0.x[:k], d1.x[:k])
        assert torch.equal(d0.a[:k].view(torch.float64), d1.a[:k].view(torch.float64))


def test_truncate_noalloc_rejects_oversize(gpu_ctx):
    L = pkg_module("_lib")
    lib = L.load()
    st = torch.zeros(8, dtype=torch.int64, device=gpu_ctx.device)
    rc = lib.bass_truncate_noalloc(None, None, MAX + 1, 10, None, None, 10, 1, None, 0,
                                   ctypes.c_void_p(st.data_ptr()), gpu_ctx.stream)
    assert rc == L.BASS_E_ARG and "exceeds the maximum" in L.last_error()
