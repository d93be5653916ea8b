"""Device operations on state buffers (thin wrappers over the C ABI).

Each function launches on the context's current stream.  Functions that
return scalars read them back through bass_read_stats (one synchronisation).
"""

import ctypes

import numpy as np
import torch

from . import _lib as L

SEG = 2048  # output slots per emission tile (segmented gate output)


def _tiles(n, nq):
    per = 1024 if nq == 1 else 512
    return (n + per - 1) // per


def gate(ctx, src, n, G, qubits, dst, mode=L.MODE_GATE, dedup=True, segmented=False):
    """Apply a 1- or 2-qubit gate (or rotation) to src[:n] into dst.

    Returns (n_out, hist_tiles).  With segmented=True the outputs sit in
    per-tile segments of dst (feed truncate(..., segmented=True) or densify())."""
    nq = len(qubits)
    M = 1 << nq
    need = _tiles(n, nq) * SEG if segmented else M * n
    # a truncate of the output reads this call's histogram from the workspace
    # and needs whole SEG-slot tiles of it: size for that now
    ctx.ensure_ws(max(-(-need // SEG) * SEG, n, 1))
    dst.ensure(max(need, 1))
    buf, gp = L.mat_arg(G)
    q0 = int(qubits[0])
    q1 = int(qubits[1]) if nq == 2 else 0
    ht = ctypes.c_int64(0)
    rc = ctx.lib.bass_gate(src.xp, src.ap, n, nq, q0, q1, gp, mode, 1 if dedup else 0, 1 if segmented else 0,
                           dst.xp, dst.ap, dst.cap, ctypes.byref(ht), ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_gate")
    st = ctx.read_stats("bass_gate")
    return st.nnz, int(ht.value)


def densify(ctx, src, seg_tiles, n, dst):
    """Dense copy of the preceding segmented gate output."""
    dst.ensure(max(n, 1))
    rc = ctx.lib.bass_densify(src.xp, src.ap, seg_tiles, dst.xp, dst.ap, dst.cap, ctx.wsp, ctx.ws_bytes,
                              ctx.stream)
    L.check(rc, "bass_densify")


def lift(ctx, src, n, j, dst, G=None, nq=1, mode=L.MODE_GATE, count=True, spec=False):
    """Key-ordered gate (mode GATE / ROT) or bit move (MODE_MOVE) on src[:n].

    Gates: returns (n_out, seg_tiles) with the output segmented in dst; with
    count=False n_out is None and nothing waits for the device (a following
    truncate(..., segmented=True) reports it as stats.n_in).  spec=True (with
    count=False): also the truncate's digit-1 pass against the last select's
    bucket; pass spec=True to the truncate_launch that follows.
    Move: returns (n, 0) with the output dense in dst."""
    tm = 1024 if mode == L.MODE_MOVE else SEG // (1 << nq)
    tiles = (n + tm - 1) // tm
    need = n if mode == L.MODE_MOVE else tiles * SEG
    ctx.ensure_ws(max(need, n, 1))
    dst.ensure(max(need, 1))
    if mode == L.MODE_MOVE:
        buf, gp = None, None
    else:
        buf, gp = L.mat_arg(G)
    st = ctypes.c_int64(0)
    # count=False: a segmented truncate reads the output next, which needs no
    # scan of its tile counts
    fl = mode | (L.LIFT_NO_SCAN if (not count and mode != L.MODE_MOVE) else 0)
    if spec and mode != L.MODE_MOVE:
        fl |= L.LIFT_SPEC
    rc = ctx.lib.bass_lift(src.xp, src.ap, n, nq, int(j), gp, fl, dst.xp, dst.ap, dst.cap, ctypes.byref(st),
                           ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_lift")
    if mode == L.MODE_MOVE:
        return n, 0
    if not count:
        return None, int(st.value)
    return ctx.read_stats("bass_lift").nnz, int(st.value)


def lift_counted(ctx, src, n_max, j, dst, G=None, nq=1, mode=L.MODE_GATE, spec=False):
    """lift() whose input count is the workspace's state-count slot (<= n_max),
    read on the device: nothing waits for the device.  Gates: returns the
    output's tile count (its segmented output goes to a counted truncate);
    a move (MODE_MOVE) copies the state densely and returns 0.  spec=True
    (gates): also the truncate's digit-1 pass against the last select's
    bucket; pass spec=True to the truncate_counted that follows."""
    tm = 1024 if mode == L.MODE_MOVE else SEG // (1 << nq)
    tiles = (n_max + tm - 1) // tm
    need = n_max if mode == L.MODE_MOVE else tiles * SEG
    ctx.ensure_ws(max(need, n_max, 1))
    dst.ensure(max(need, 1))
    buf, gp = (None, None) if mode == L.MODE_MOVE else L.mat_arg(G)
    st = ctypes.c_int64(0)
    # a counted truncate reads a gate's output next: no scan of its tile counts
    fl = mode if mode == L.MODE_MOVE else mode | L.LIFT_NO_SCAN | (L.LIFT_SPEC if spec else 0)
    rc = ctx.lib.bass_lift_counted(src.xp, src.ap, n_max, nq, int(j), gp, fl, dst.xp, dst.ap, dst.cap,
                                   ctypes.byref(st), ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_lift_counted")
    return int(st.value)


def truncate_counted(ctx, src, k, dst, hist_tiles, key_map=None, spec=False):
    """truncate(..., segmented=True) in counted mode: resolves every threshold
    itself, appends its stats to the op log and sets the state-count slot;
    nothing waits for the device.  spec=True: src is the output of a
    lift_counted(spec=True) issued just before."""
    ctx.ensure_ws(max(hist_tiles * SEG, 1))
    dst.ensure(k)
    if key_map is None:
        kp, kb = None, 0
    else:
        km = np.ascontiguousarray(key_map, dtype=np.int32)
        kp, kb = km.ctypes.data_as(ctypes.c_void_p), len(km)
    fl = 1 | 4 | (L.TRUNC_SPEC if spec else 0)
    rc = ctx.lib.bass_truncate(src.xp, src.ap, 0, k, dst.xp, dst.ap, dst.cap, 1, hist_tiles, fl, kp, kb,
                               ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_truncate")


def lift_truncate(ctx, src, n, j, seg, dst, k, G, nq=1, mode=L.MODE_GATE, renorm=True, counted=False, spec=False,
                  key_map=None, margin=1):
    """A gate or rotation lift of src[:n] into the segments seg, then the
    top-k select of them into dst (<= k entries) (bass_lift_truncate): bit for
    bit lift(..., count=False, spec) + truncate_launch(..., segmented=True,
    spec), or with counted=True (n bounds the count in the state-count slot)
    lift_counted + truncate_counted.  With spec the lift stores only the
    outputs that can matter to the predicted threshold (and stores them all
    again when the select finds it needs more); margin 1 keeps the binade
    below the predicted bucket too, margin 0 only the bucket and above
    (flags bit 2).  dst may be src; seg may
    share memory with neither.  Nothing waits for the device; a non-counted
    call's stats come from truncate_finish(ctx, handle)."""
    tm = SEG // (1 << nq)
    tiles = (n + tm - 1) // tm
    ctx.ensure_ws(max(tiles * SEG, n, 1))
    seg.ensure(max(tiles * SEG, 1))
    dst.ensure(k)
    buf, gp = L.mat_arg(G)
    if key_map is None:
        km, kp, kb = None, None, 0
    else:
        km = np.ascontiguousarray(key_map, dtype=np.int32)
        kp, kb = km.ctypes.data_as(ctypes.c_void_p), len(km)
    if margin not in (0, 1):
        raise ValueError("lift_truncate: margin must be 0 or 1")
    fl = (L.GT_COUNTED if counted else 0) | (L.GT_SPEC if spec else 0) | (4 if margin == 0 else 0)
    rc = ctx.lib.bass_lift_truncate(src.xp, src.ap, int(n), int(nq), int(j), gp, int(mode), int(k), seg.xp, seg.ap,
                                    seg.cap, dst.xp, dst.ap, dst.cap, 1 if renorm else 0, fl, kp, kb, ctx.wsp,
                                    ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_lift_truncate")
    return None, (km, buf)


def gate_truncate(ctx, src, n, j, dst, k, G, nq=1, mode=L.MODE_GATE, renorm=True, counted=False, spec=False,
                  key_map=None):
    """A gate or rotation lift of src[:n] followed by the top-k select of its
    output into dst (<= k entries), the output never stored
    (bass_gate_truncate): bit for bit lift(..., count=False, spec) +
    truncate_launch(..., segmented=True, spec), or with counted=True (n bounds
    the count in the state-count slot) lift_counted + truncate_counted.
    dst must not share memory with src.  Nothing waits for the device; a
    non-counted call's stats come from truncate_finish(ctx, handle)."""
    tm = SEG // (1 << nq)
    tiles = (n + tm - 1) // tm
    ctx.ensure_ws(max(tiles * SEG, n, 1))
    dst.ensure(k)
    buf, gp = L.mat_arg(G)
    if key_map is None:
        km, kp, kb = None, None, 0
    else:
        km = np.ascontiguousarray(key_map, dtype=np.int32)
        kp, kb = km.ctypes.data_as(ctypes.c_void_p), len(km)
    fl = (L.GT_COUNTED if counted else 0) | (L.GT_SPEC if spec else 0)
    rc = ctx.lib.bass_gate_truncate(src.xp, src.ap, int(n), int(nq), int(j), gp, int(mode), int(k), dst.xp, dst.ap,
                                    dst.cap, 1 if renorm else 0, fl, kp, kb, ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_gate_truncate")
    return None, (km, buf)


def diag2q_counted(ctx, buf, n_max, q1, q2, diag):
    """diag2q() on the state's count from the state-count slot (<= n_max)."""
    d = np.ascontiguousarray(np.asarray(diag, dtype=np.complex128)).view(np.float64)
    rc = ctx.lib.bass_diag2q_counted(buf.xp, buf.ap, int(n_max), int(q1), int(q2), d.ctypes.data_as(ctypes.c_void_p),
                                     ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_diag2q_counted")


DIAG_RUN_MAX = 32  # gates per launch of bass_diag2q_run


def diag2q_run(ctx, buf, n, gates, counted=False):
    """Diagonal 2-qubit gates [(q1, q2, diag[4]), ...] applied in order in one
    pass per DIAG_RUN_MAX gates (bit-identical to one diag2q() per gate).
    counted=True: n bounds the count read from the state-count slot."""
    if not gates:
        return
    qs = np.ascontiguousarray([[int(g[0]), int(g[1])] for g in gates], dtype=np.int32)
    ds = np.ascontiguousarray([np.asarray(g[2], dtype=np.complex128) for g in gates]).view(np.float64)
    rc = ctx.lib.bass_diag2q_run(buf.xp, buf.ap, int(n), len(gates), qs.ctypes.data_as(ctypes.c_void_p),
                                 ds.ctypes.data_as(ctypes.c_void_p), 1 if counted else 0, ctx.wsp, ctx.ws_bytes,
                                 ctx.stream)
    L.check(rc, "bass_diag2q_run")


def counted_begin(ctx, n, n_max):
    """Start a run of counted gates on a state of n entries (<= n_max): size the
    workspace for every operation of the run first (the largest is a 2-qubit
    gate's 4 n_max output slots; growing it later would drop the count slot
    and the op log), then write n into the state-count slot."""
    tm = SEG // 4
    ctx.ensure_ws(max(((n_max + tm - 1) // tm) * SEG, n_max, 1))
    L.check(ctx.lib.bass_set_count(ctx.wsp, ctx.ws_bytes, int(n), ctx.stream), "bass_set_count")


def counted_read(ctx, max_entries):
    """The op log of counted truncations since the last read, as a list of
    bass_stats records (synchronous); empties it."""
    out = (L.BassStats * max(1, int(max_entries)))()
    n = ctypes.c_int64(0)
    rc = ctx.lib.bass_counted_read(ctx.wsp, ctx.ws_bytes, ctypes.cast(out, ctypes.c_void_p), int(max_entries),
                                   ctypes.byref(n), ctx.stream)
    L.check(rc, "bass_counted_read")
    return list(out[: n.value])


def key_map_arg(pos):
    """Stored-bit -> qubit map (int32[N]) of an ordered state whose qubit q sits at bit pos[q]."""
    pos = np.asarray(pos)
    km = np.empty(len(pos), dtype=np.int32)
    km[pos] = np.arange(len(pos), dtype=np.int32)
    return km


def truncate_launch(ctx, src, n, k, dst, renorm=True, hist_tiles=0, segmented=False, key_map=None, spec=False):
    """The launches of truncate() without waiting for the device: returns the
    handle truncate_finish() takes.  Nothing may launch a stats-writing
    operation on the context in between.  spec=True: src is the output of a
    lift(..., spec=True) issued just before."""
    n = 0 if n is None else n
    # the select reads whole tiles of SEG slots, and hist_tiles per-tile sums
    # of the producing gate (segmented: SEG slots a tile; dense: emission tiles
    # of >= 512 slots, of which a workspace of S entries keeps S // 512 + 2)
    tiles = hist_tiles if segmented else -(-int(n) // SEG)
    psum = 0 if segmented else max(hist_tiles - 2, 0) * 512
    ctx.ensure_ws(max(n, tiles * SEG, psum), keep=hist_tiles > 0)
    dst.ensure(k)
    if key_map is None:
        km, kp, kb = None, None, 0
    else:
        km = np.ascontiguousarray(key_map, dtype=np.int32)
        kp, kb = km.ctypes.data_as(ctypes.c_void_p), len(km)
    flags = (1 if segmented else 0) | (L.TRUNC_SPEC if spec else 0)
    args = (src.xp, src.ap, n, k, dst.xp, dst.ap, dst.cap, 1 if renorm else 0, hist_tiles, flags, kp, kb)
    rc = ctx.lib.bass_truncate(*args, ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_truncate")
    return args, km


def truncate_finish(ctx, handle):
    """Stats of a truncate_launch() (waits for the device)."""
    args, _km = handle
    st = ctx.read_stats("bass_truncate")
    if st.flags & 2:
        # very many entries tie at the threshold: repeat with the multi-block path
        if args is None:  # a fused gate select resolves every threshold itself
            raise L.BassError("bass_gate_truncate: threshold left unresolved")
        a = list(args)
        a[9] = (a[9] | 2) & ~L.TRUNC_SPEC
        rc = ctx.lib.bass_truncate(*a, ctx.wsp, ctx.ws_bytes, ctx.stream)
        L.check(rc, "bass_truncate")
        st = ctx.read_stats("bass_truncate")
    return st


def truncate(ctx, src, n, k, dst, renorm=True, hist_tiles=0, segmented=False, key_map=None):
    """Top-k of src[:n] (n > k) into dst[:k].  Returns stats (kept, total, sum2, sum4, norm).

    Segmented input (a gate's output): n may be None and the input may hold
    <= k entries, which are then all kept, in order, without renormalisation
    (stats.flags bit 0 clear); stats.n_in is the input count.
    key_map: stored-bit -> qubit map (key_map_arg) when src holds stored words."""
    return truncate_finish(ctx, truncate_launch(ctx, src, n, k, dst, renorm, hist_tiles, segmented, key_map))


def diag2q(ctx, buf, n, q1, q2, diag):
    d = np.ascontiguousarray(np.asarray(diag, dtype=np.complex128).view(np.float64))
    rc = ctx.lib.bass_diag2q(buf.xp, buf.ap, n, int(q1), int(q2), d.ctypes.data_as(ctypes.c_void_p),
                             ctx.stream)
    L.check(rc, "bass_diag2q")


def moments(ctx, buf, n):
    """(sum |a|^2, sum |a|^4) of buf[:n]."""
    ctx.ensure_ws(n)
    rc = ctx.lib.bass_moments(buf.ap, n, ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_moments")
    st = ctx.read_stats("bass_moments")
    return st.sum2, st.sum4


def scale(ctx, buf, n, div):
    rc = ctx.lib.bass_scale(buf.ap, n, float(div), ctx.stream)
    L.check(rc, "bass_scale")


def rdm1q(ctx, buf, n, N, ordered=False):
    """All single-qubit RDMs -> complex128[N, 2, 2] (host).  ordered=True: buf
    is a key-ordered state and row p is the RDM of stored bit p."""
    ctx.ensure_ws(n)
    out = np.zeros((N, 4), dtype=np.float64)
    fn = ctx.lib.bass_rdm1q_ordered if ordered else ctx.lib.bass_rdm1q
    rc = fn(buf.xp, buf.ap, n, int(N), out.ctypes.data_as(ctypes.c_void_p), ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_rdm1q")
    rho = np.zeros((N, 2, 2), dtype=np.complex128)
    rho[:, 0, 0] = out[:, 0]
    rho[:, 1, 1] = out[:, 1]
    rho[:, 0, 1] = out[:, 2] + 1j * out[:, 3]
    rho[:, 1, 0] = out[:, 2] - 1j * out[:, 3]
    return rho


def rdm2q(ctx, buf, n, j):
    """4x4 RDM (host complex128) of the pair stored at bits (j, 0) of a
    key-ordered state, index (b_j << 1) | b_0."""
    ctx.ensure_ws(n)
    out = np.zeros(32, dtype=np.float64)
    rc = ctx.lib.bass_rdm2q_ordered(buf.xp, buf.ap, n, int(j), out.ctypes.data_as(ctypes.c_void_p), ctx.wsp,
                                    ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_rdm2q_ordered")
    return out.view(np.complex128).reshape(4, 4).copy()


# ------------------------------------------------------------------ dense --
def dense_state(N, device):
    """A zeroed complex128[2^N] device vector."""
    return torch.zeros(1 << int(N), dtype=torch.complex128, device=device)


def dense_scatter(ctx, buf, n, N, psi, key_map=None):
    """psi = 0, psi[key(x_i)] = a_i for the n entries of buf (key_map: stored
    bit -> qubit of a key-ordered state, None = keys stored as they are)."""
    if key_map is None:
        kp, kb = None, 0
    else:
        km = np.ascontiguousarray(key_map, dtype=np.int32)
        kp, kb = km.ctypes.data_as(ctypes.c_void_p), len(km)
    rc = ctx.lib.bass_dense_scatter(buf.xp, buf.ap, int(n), kp, kb, L.ptr(psi), int(N), ctx.stream)
    L.check(rc, "bass_dense_scatter")


def dense_gate(ctx, psi, N, G, qubits):
    """In place: gate G (2x2 or 4x4, index (b_q0 << 1) | b_q1) on qubits."""
    buf, gp = L.mat_arg(G)
    q0 = int(qubits[0])
    q1 = int(qubits[1]) if len(qubits) == 2 else 0
    rc = ctx.lib.bass_dense_gate(L.ptr(psi), int(N), len(qubits), q0, q1, gp, ctx.stream)
    L.check(rc, "bass_dense_gate")


def dense_vdot(ctx, u, v):
    """sum conj(u_i) v_i of two device vectors (host complex)."""
    if u.numel() != v.numel():
        raise ValueError("dense_vdot: lengths differ")
    ctx.ensure_ws(max(1, u.numel() >> 12))  # 16 B of partials per 4096 entries
    out = np.zeros(2, dtype=np.float64)
    rc = ctx.lib.bass_dense_vdot(L.ptr(u), L.ptr(v), int(u.numel()), out.ctypes.data_as(ctypes.c_void_p), ctx.wsp,
                                 ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_dense_vdot")
    return complex(out[0], out[1])


# ---------------------------------------------------------------- batches --
# A batch is a numpy array of L.JOB_DTYPE records (one per state).

STATS_DTYPE = np.dtype([("nnz", "<i8"), ("total", "<f8"), ("kept", "<f8"), ("norm", "<f8"), ("sum2", "<f8"),
                        ("sum4", "<f8"), ("n_in", "<i8"), ("err", "<i4"), ("flags", "<i4")])


def new_jobs(count):
    return np.zeros(count, dtype=L.JOB_DTYPE)


def batch_stats(ctx, njobs, what="bass_batch_stats"):
    out = np.zeros(njobs, dtype=STATS_DTYPE)
    rc = ctx.lib.bass_batch_stats(ctx.wsp, ctx.ws_bytes, int(njobs), out.ctypes.data_as(ctypes.c_void_p), ctx.stream)
    L.check(rc, what)
    return out


def batch_tiles(jobs, nq, mode=L.MODE_GATE):
    tm = 1024 if mode == L.MODE_MOVE else SEG // (1 << nq)
    return int(np.sum((jobs["n"] + tm - 1) // tm))


def batch_step(ctx, jobs, nq, mode, tiles):
    """Gate / rotation + truncate-or-keep for every job; returns per-job stats.
    tiles: Buffers for the gate outputs (grown as needed)."""
    T = batch_tiles(jobs, nq, mode)
    slots = max(T * SEG, 1)
    ctx.ensure_ws(max(slots, len(jobs) * 1024))
    tiles.ensure(slots)
    jp = jobs.ctypes.data_as(ctypes.c_void_p)
    rc = ctx.lib.bass_batch_step(jp, len(jobs), nq, mode, 0, tiles.xp, tiles.ap, tiles.cap, ctx.wsp, ctx.ws_bytes,
                                 ctx.stream)
    L.check(rc, "bass_batch_step")
    return batch_stats(ctx, len(jobs))


def batch_move(ctx, jobs):
    ctx.ensure_ws(max(batch_tiles(jobs, 1, L.MODE_MOVE) * 1024, len(jobs) * 1024))
    rc = ctx.lib.bass_batch_move(jobs.ctypes.data_as(ctypes.c_void_p), len(jobs), ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_batch_move")


def _batch_ew(ctx, fn, jobs, what):
    ctx.ensure_ws(len(jobs) * 1024 + int(np.sum(jobs["n"])) // 64)
    rc = fn(jobs.ctypes.data_as(ctypes.c_void_p), len(jobs), ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, what)


def batch_diag2q(ctx, jobs):
    _batch_ew(ctx, ctx.lib.bass_batch_diag2q, jobs, "bass_batch_diag2q")


def batch_scale(ctx, jobs):
    _batch_ew(ctx, ctx.lib.bass_batch_scale, jobs, "bass_batch_scale")


def batch_moments(ctx, jobs):
    """Per-job (sum |a|^2, sum |a|^4)."""
    _batch_ew(ctx, ctx.lib.bass_batch_moments, jobs, "bass_batch_moments")
    st = batch_stats(ctx, len(jobs))
    return st["sum2"], st["sum4"]


def batch_rdm1q(ctx, jobs):
    """RDMs of stored bits p < key_bits per job -> complex128[njobs, 64, 2, 2]."""
    n = len(jobs)
    ctx.ensure_ws(n * 2048 + int(np.sum(jobs["n"])) * 2)
    out = np.zeros((n, 64, 4), dtype=np.float64)
    rc = ctx.lib.bass_batch_rdm1q(jobs.ctypes.data_as(ctypes.c_void_p), n, out.ctypes.data_as(ctypes.c_void_p),
                                  ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_batch_rdm1q")
    rho = np.zeros((n, 64, 2, 2), dtype=np.complex128)
    rho[:, :, 0, 0] = out[:, :, 0]
    rho[:, :, 1, 1] = out[:, :, 1]
    rho[:, :, 0, 1] = out[:, :, 2] + 1j * out[:, :, 3]
    rho[:, :, 1, 0] = out[:, :, 2] - 1j * out[:, :, 3]
    return rho


class BatchPlan:
    """A batch plan (bass_bplan_*): its own workspace and gate-output tiles
    for up to jobs_cap states and tiles_cap tiles, fixed launch shapes, and
    with graphs=True each operation's launch sequence captured once in a HIP
    graph.  run() is asynchronous; its results land in the pinned host array
    `out` once the stream has passed the call."""

    def __init__(self, ctx, jobs_cap, tiles_cap, graphs=True):
        from ._device import Buffers

        self.ctx = ctx
        self.jobs_cap = int(jobs_cap)
        self.tiles_cap = int(tiles_cap)
        nb = int(ctx.lib.bass_bplan_ws_bytes(self.jobs_cap, self.tiles_cap))
        if nb <= 0:
            raise ValueError("bad batch plan capacities")
        self.ws = torch.zeros(nb, dtype=torch.uint8, device=ctx.device)
        self.tiles = Buffers(self.tiles_cap * SEG, ctx.device)
        h = ctypes.c_void_p()
        rc = ctx.lib.bass_bplan_create(self.jobs_cap, self.tiles_cap, self.tiles.xp, self.tiles.ap,
                                       L.ptr(self.ws), nb, 1 if graphs else 0, ctypes.byref(h))
        L.check(rc, "bass_bplan_create")
        self._h = h
        self._lib = ctx.lib

    def run(self, op, jobs, out=None):
        ptr = out.ctypes.data_as(ctypes.c_void_p) if out is not None else None
        rc = self._lib.bass_bplan_run(self._h, int(op), jobs.ctypes.data_as(ctypes.c_void_p), len(jobs), ptr,
                                      self.ctx.stream)
        L.check(rc, "bass_bplan_run")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bass_bplan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pinned_array(shape, dtype):
    """A numpy array over pinned host memory (asynchronous device copies)."""
    import torch

    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    t = torch.empty(max(n, 1), dtype=torch.uint8, pin_memory=True)
    arr = t.numpy()[:n].view(dt).reshape(shape)
    return arr, t
