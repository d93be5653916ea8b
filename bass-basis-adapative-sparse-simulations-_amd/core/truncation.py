"""Truncation kernels under the names of the reference's src/core/truncation.py."""

import numpy as np
import torch

from .. import _lib as L
from .. import _ops
from .._device import Buffers, context


def top_k_truncate_kernel_opt(x, alpha, nnz, k):
    """truncation.py:68-113: keep the k largest |a|^2 (ties: smaller key), compact
    them to the front (input order), return (nnz_new, gamma_sq) with gamma_sq the
    kept sum of |a|^2; amplitudes are not renormalised."""
    from .sparse_state import _pair

    ctx = context()
    buf, xs, as_ = _pair(x, alpha, ctx)
    n, k = int(nnz), int(k)
    if n <= k:
        s2, _ = _ops.moments(ctx, buf, n)
        return n, float(s2)
    tmp = Buffers(k, ctx.device)
    st = _ops.truncate(ctx, buf, n, k, tmp, renorm=False)
    buf.x[:k].copy_(tmp.x[:k])
    buf.a[:k].copy_(tmp.a[:k])
    xs.write_back(k)
    as_.write_back(k)
    return k, float(st.kept)


def lcg_next(rng):
    """(1103515245 * rng + 12345) & 0x7FFFFFFF (truncation.py:136, simulator.py:158)."""
    return (1103515245 * int(rng) + 12345) & 0x7FFFFFFF


def random_k_indices(nnz, k, seed):
    """Indices perm[:k] of the reference's partial Fisher-Yates shuffle
    (truncation.py:131-140).  The swap sequence depends only on (nnz, k, seed),
    so it is computed on the host with a sparse permutation; the gather and the
    sum run on the device."""
    perm = {}
    rng = int(seed)
    out = np.empty(k, dtype=np.int64)
    for i in range(k):
        rng = lcg_next(rng)
        r = i + (rng % (nnz - i))
        pi = perm.get(i, i)
        pr = perm.get(r, r)
        perm[i] = pr
        perm[r] = pi
        out[i] = pr
    return out


def _random_k_device(ctx, buf, n, k, seed, dst):
    idx = torch.from_numpy(random_k_indices(n, k, seed)).to(ctx.device)
    dst.ensure(k)
    rc = ctx.lib.bass_gather(buf.xp, buf.ap, ctypes_ptr(idx), k, dst.xp, dst.ap, ctx.stream)
    L.check(rc, "bass_gather")
    s2, _ = _ops.moments(ctx, dst, k)
    return s2


def ctypes_ptr(t):
    import ctypes

    return ctypes.c_void_p(t.data_ptr())


def random_k_truncate_kernel_opt(x, alpha, nnz, k, seed):
    """truncation.py:116-159: LCG partial Fisher-Yates selection of k entries,
    returns (k, gamma_sq) (or (nnz, sum |a|^2) when nnz <= k)."""
    from .sparse_state import _pair

    ctx = context()
    buf, xs, as_ = _pair(x, alpha, ctx)
    n, k = int(nnz), int(k)
    if n <= k:
        s2, _ = _ops.moments(ctx, buf, n)
        return n, float(s2)
    tmp = Buffers(k, ctx.device)
    gsq = _random_k_device(ctx, buf, n, k, seed, tmp)
    buf.x[:k].copy_(tmp.x[:k])
    buf.a[:k].copy_(tmp.a[:k])
    xs.write_back(k)
    as_.write_back(k)
    return k, float(gsq)
