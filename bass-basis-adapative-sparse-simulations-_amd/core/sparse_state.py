"""Device sparse state and the gate / sort / merge kernels, under the names of
the reference's src/core/sparse_state.py.

Kernel functions accept either device tensors (torch int64 keys viewed as
uint64, torch complex128 amplitudes) -- the fast path -- or NumPy arrays, which
are staged to the device and written back in place so reference call sites
keep working.  All arithmetic runs in libbass_hip.so.
"""

import numpy as np
import torch

from .. import _lib as L
from .. import _ops
from .._device import Buffers, context, to_host_amps, to_host_keys


# --------------------------------------------------------------------------
# key-ordered states: stored word <-> basis key


def lifted_positions(pos, j, t):
    """Qubit positions after moving stored bit j down to bit t (bits t..j-1 move up by one)."""
    pos = np.asarray(pos)
    new = pos.copy()
    new[(pos >= t) & (pos < j)] += 1
    new[pos == j] = t
    return new


def stored_to_keys(w, pos):
    """Basis keys of stored words w (uint64) of a state whose qubit q sits at bit pos[q]."""
    w = np.asarray(w, dtype=np.uint64)
    if pos is None or np.array_equal(pos, np.arange(len(pos))):
        return w.copy()
    # one 256-entry table per stored byte: the basis bits that byte carries
    pos = np.asarray(pos, dtype=np.int64)
    inv = np.full(64, -1, dtype=np.int64)
    inv[pos] = np.arange(len(pos))
    v = np.arange(256, dtype=np.uint64)
    out = np.zeros_like(w)
    for b in range((int(pos.max()) >> 3) + 1):
        table = np.zeros(256, dtype=np.uint64)
        for i in range(8):
            q = inv[8 * b + i]
            if q >= 0:
                table |= ((v >> np.uint64(i)) & np.uint64(1)) << np.uint64(q)
        out |= table[((w >> np.uint64(8 * b)) & np.uint64(255)).astype(np.intp)]
    return out


# --------------------------------------------------------------------------
# argument staging


class _Arr:
    """A keys or amplitude argument on the device, with write-back to the
    caller's NumPy array when it came from the host."""

    def __init__(self, arr, kind, ctx):
        self.host = None
        if isinstance(arr, torch.Tensor):
            if arr.device.type != "cuda":
                raise TypeError("device kernels need CUDA/ROCm tensors or NumPy arrays")
            self.t = arr
        else:
            self.host = arr
            if kind == "x":
                h = np.ascontiguousarray(np.asarray(arr, dtype=np.uint64)).view(np.int64)
            else:
                h = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128))
            self.t = torch.from_numpy(h).to(ctx.device)
        if kind == "x" and self.t.dtype != torch.int64:
            raise TypeError("keys must be int64 (uint64 bit pattern) tensors")
        if kind == "a" and self.t.dtype != torch.complex128:
            raise TypeError("amplitudes must be complex128 tensors")

    def write_back(self, n):
        if self.host is not None and n > 0:
            v = self.t[:n].cpu().numpy()
            if v.dtype == np.int64:
                v = v.view(np.uint64)
            self.host[:n] = v


def _pair(x, a, ctx):
    xs, as_ = _Arr(x, "x", ctx), _Arr(a, "a", ctx)
    b = Buffers.__new__(Buffers)
    b.x, b.a, b.device = xs.t, as_.t, ctx.device
    b._xt = b._at = None
    b.cap = min(xs.t.shape[0], as_.t.shape[0])
    return b, xs, as_


# --------------------------------------------------------------------------
# kernels (src/core/sparse_state.py)


def bitwise_sort_key(x_array, mask):
    """sparse_state.py:7-12."""
    if isinstance(x_array, torch.Tensor):
        return x_array & ~int(mask)
    return np.asarray(x_array, dtype=np.uint64) & ~np.uint64(mask)


def apply_1qubit_gate_kernel_opt(x_in, alpha_in, nnz_in, x_out, alpha_out, gate_matrix, q):
    """sparse_state.py:15-67: per entry emit (x & ~m, G[0,b] a) then (x | m, G[1,b] a),
    exact zeros dropped, input order.  Returns nnz_out."""
    ctx = context()
    src, _, _ = _pair(x_in, alpha_in, ctx)
    dst, xo, ao = _pair(x_out, alpha_out, ctx)
    n = int(nnz_in)
    if 2 * n > dst.cap:
        # the reference writes past a short buffer; refuse instead
        raise RuntimeError("Temporary buffer overflow")
    nout, _ = _ops.gate(ctx, src, n, gate_matrix, (q,), dst, dedup=False)
    xo.write_back(nout)
    ao.write_back(nout)
    return nout


def apply_2qubit_gate_kernel_opt(x_in, alpha_in, nnz_in, x_out, alpha_out, gate_matrix, q1, q2):
    """sparse_state.py:70-124: up to four emissions per entry; returns -1 when
    x_out is too short (:117-118)."""
    ctx = context()
    src, _, _ = _pair(x_in, alpha_in, ctx)
    dst, xo, ao = _pair(x_out, alpha_out, ctx)
    n = int(nnz_in)
    cap = dst.cap
    try:
        if 4 * n > cap:
            # emit into scratch first; overflow is decided on the real count
            tmp = Buffers(4 * n, ctx.device)
            nout, _ = _ops.gate(ctx, src, n, gate_matrix, (q1, q2), tmp, dedup=False)
            if nout > cap:
                return -1
            dst.x[:nout].copy_(tmp.x[:nout])
            dst.a[:nout].copy_(tmp.a[:nout])
        else:
            nout, _ = _ops.gate(ctx, src, n, gate_matrix, (q1, q2), dst, dedup=False)
    except RuntimeError:
        return -1
    xo.write_back(nout)
    ao.write_back(nout)
    return nout


def merge_sorted_kernel(x, alpha, nnz):
    """sparse_state.py:127-147: run-length sum of equal adjacent keys, in place,
    zero sums kept.  Returns the merged count."""
    ctx = context()
    buf, xs, as_ = _pair(x, alpha, ctx)
    n = int(nnz)
    if n == 0:
        return 0
    ctx.ensure_ws(n)
    tmp = Buffers(n, ctx.device)
    rc = ctx.lib.bass_merge_sorted(buf.xp, buf.ap, n, tmp.xp, tmp.ap, ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_merge_sorted")
    m = ctx.read_stats("bass_merge_sorted").nnz
    buf.x[:m].copy_(tmp.x[:m])
    buf.a[:m].copy_(tmp.a[:m])
    xs.write_back(m)
    as_.write_back(m)
    return m


def radix_sort_uint64(keys, values, n, key_bits=None):
    """sparse_state.py:150-172: stable LSD radix sort of keys[:n] carrying values.
    key_bits (extension) bounds the significant key bits; by default it is
    taken from the largest key."""
    ctx = context()
    buf, xs, as_ = _pair(keys, values, ctx)
    n = int(n)
    if n <= 1:
        return
    if key_bits is None:
        # keys live in an int64 tensor: any key with bit 63 set reads negative
        xs_ = buf.x[:n]
        if int(xs_.min().item()) < 0:
            key_bits = 64
        else:
            key_bits = max(1, int(xs_.max().item()).bit_length())
    ctx.ensure_ws(n)
    rc = ctx.lib.bass_radix_sort(buf.xp, buf.ap, n, int(key_bits), ctx.wsp, ctx.ws_bytes, ctx.stream)
    L.check(rc, "bass_radix_sort")
    ctx.read_stats("bass_radix_sort")
    xs.write_back(n)
    as_.write_back(n)


def apply_2qubit_diagonal_kernel(x, alpha, nnz, q1, q2, diag):
    """sparse_state.py:178-190: alpha[i] *= diag[(b1 << 1) | b2], in place."""
    ctx = context()
    buf, xs, as_ = _pair(x, alpha, ctx)
    _ops.diag2q(ctx, buf, int(nnz), q1, q2, diag)
    as_.write_back(int(nnz))


# --------------------------------------------------------------------------
# SparseState (src/core/sparse_state.py:193-415)


class SparseState:
    """Sparse state |phi> = sum_i alpha_i |x_i> held on the device.

    ``x`` / ``alpha`` return host copies (length = capacity) for reference
    call sites; ``x_dev`` / ``alpha_dev`` are the device tensors.  ``dedup``
    selects how gates merge the outputs of a group: "table" (group hash
    table, outputs in first-input order) or "sort" (emit, stable radix sort,
    merge: the reference's path and key-ascending order).  Both give the same
    entries and sums.
    """

    RADIX_THRESHOLD = 2000

    def __init__(self, num_qubits, k, init_state=None, device=None, dedup="table", capacity=None):
        self.N = num_qubits
        self.k = k
        self._ctx = context(device)
        # capacity: initial entries of the state buffer (default k); both
        # buffers grow on demand
        cap = k if capacity is None else min(k, capacity)
        self._buf = Buffers(max(cap, 1), self._ctx.device)
        self._tmp = Buffers(max(4 * cap, 4), self._ctx.device)
        self.gamma = 1.0
        if init_state is None:
            init_state = 0
        self._buf.x[0] = int(np.uint64(init_state).view(np.int64))
        self._buf.a[0] = 1.0 + 0.0j
        self.nnz = 1
        self._temp_nnz = 0
        self._hist_tiles = 0  # digit histogram of the current entries, if fresh
        self.dedup = dedup
        # key-ordered layout (set by the BASS driver): entries sorted by stored
        # word, qubit q at bit pos[q]; None = keys stored as is
        self.pos = None
        # optional operation counters (records produced, algorithmic bytes)
        self.counters = None

    def _count_gate(self, op, n_in, n_out):
        if self.counters is not None:
            c = self.counters
            c["amp_updates"] = c.get("amp_updates", 0) + n_out
            c["bytes_" + op] = c.get("bytes_" + op, 0) + 24 * (n_in + n_out)
            c[op] = c.get(op, 0) + 1

    def _count_select(self, n_in, k):
        if self.counters is not None:
            c = self.counters
            c["bytes_select"] = c.get("bytes_select", 0) + 40 * n_in + 24 * k
            c["truncate"] = c.get("truncate", 0) + 1

    # host views ------------------------------------------------------------
    @property
    def x(self):
        out = np.zeros(self._buf.cap, dtype=np.uint64)
        out[: self.nnz] = stored_to_keys(to_host_keys(self._buf.x, self.nnz), self.pos)
        return out

    @property
    def alpha(self):
        out = np.zeros(self._buf.cap, dtype=np.complex128)
        out[: self.nnz] = to_host_amps(self._buf.a, self.nnz)
        return out

    @property
    def x_dev(self):
        return self._buf.x

    @property
    def alpha_dev(self):
        return self._buf.a

    def __repr__(self):
        return f"SparseState(N={self.N}, k={self.k}, nnz={self.nnz}, γ²={self.gamma ** 2:.6f})"

    def copy(self):
        """sparse_state.py:232-238."""
        s = SparseState(self.N, self.k, device=self._ctx.device, dedup=self.dedup)
        s._buf.ensure(self.nnz)
        s._buf.x[: self.nnz].copy_(self._buf.x[: self.nnz])
        s._buf.a[: self.nnz].copy_(self._buf.a[: self.nnz])
        s.nnz = self.nnz
        s.gamma = self.gamma
        s.pos = None if self.pos is None else self.pos.copy()
        return s

    def keys(self):
        """Basis keys of the entries (host, uint64[nnz])."""
        return stored_to_keys(to_host_keys(self._buf.x, self.nnz), self.pos)

    def to_dense(self):
        """sparse_state.py:240-255 (stored amplitudes, no division by gamma)."""
        if self.N > 20:
            raise ValueError("Dense conversion only allowed for N <= 20")
        dense = np.zeros(2 ** self.N, dtype=np.complex128)
        dense[self.keys().astype(np.int64)] = to_host_amps(self._buf.a, self.nnz)
        return dense

    def probability_distribution(self):
        return np.abs(to_host_amps(self._buf.a, self.nnz)) ** 2 / (self.gamma ** 2)

    def sample(self, num_samples=1):
        p = self.probability_distribution()
        p /= p.sum()
        idx = np.random.choice(self.nnz, size=num_samples, p=p)
        return self.keys()[idx]

    def _swap(self):
        self._buf, self._tmp = self._tmp, self._buf

    # gates -----------------------------------------------------------------
    def apply_single_qubit_gate(self, gate, method, seed):
        """sparse_state.py:268-301: gate, deduplicate, truncate to k if needed."""
        q = gate.qubits[0]
        n_in = self.nnz
        if self.dedup == "sort":
            n, _ = _ops.gate(self._ctx, self._buf, self.nnz, gate.matrix, (q,), self._tmp, dedup=False)
            radix_sort_uint64(self._tmp.x, self._tmp.a, n, key_bits=max(1, self.N))
            n = merge_sorted_kernel(self._tmp.x, self._tmp.a, n)
            self._hist_tiles = 0
        else:
            n, ht = _ops.gate(self._ctx, self._buf, self.nnz, gate.matrix, (q,), self._tmp, dedup=True)
            self._hist_tiles = ht
        self._count_gate("gate1q", n_in, n)
        self._swap()
        self._tmp.ensure(4 * max(n, 1))
        self.nnz = n
        self._after_gate(method, seed)

    def apply_two_qubit_gate(self, gate, method, seed):
        """sparse_state.py:303-377: emit <= 4 outputs per entry, sort by key
        (stable radix sort), merge equal keys (dedup="sort", the reference's
        path); dedup="table" deduplicates through the group table instead
        (same entries and sums, no sort)."""
        n_in = self.nnz
        q1, q2 = gate.qubits
        if self.dedup == "sort":
            n, _ = _ops.gate(self._ctx, self._buf, self.nnz, gate.matrix, (q1, q2), self._tmp, dedup=False)
            radix_sort_uint64(self._tmp.x, self._tmp.a, n, key_bits=max(1, self.N))
            n = merge_sorted_kernel(self._tmp.x, self._tmp.a, n)
            ht = 0
        else:
            n, ht = _ops.gate(self._ctx, self._buf, self.nnz, gate.matrix, (q1, q2), self._tmp, dedup=True)
        self._count_gate("gate2q", n_in, n)
        self._hist_tiles = ht
        self._swap()
        self._tmp.ensure(4 * max(n, 1))
        self.nnz = n
        self._after_gate(method, seed)

    def _after_gate(self, method, seed):
        if self.nnz > self.k:
            if method == "top-k":
                self.truncate_top_k()
            else:
                if self.dedup == "table":
                    # random-k draws entries by array index: put them in the
                    # order the reference's sort-and-merge gate leaves them
                    # (ascending key, sparse_state.py:276-285) before the draw
                    radix_sort_uint64(self._buf.x, self._buf.a, self.nnz, key_bits=max(1, self.N))
                    self._hist_tiles = 0
                self.truncate_random_k(seed)

    def apply_gate(self, gate, method="top-k", seed="None"):
        """sparse_state.py:379-385."""
        if self.pos is not None:
            raise ValueError("state is in the BASS driver's key-ordered layout; apply gates through the driver")
        if gate.n_qubits == 1:
            self.apply_single_qubit_gate(gate, method, seed)
        elif gate.n_qubits == 2:
            self.apply_two_qubit_gate(gate, method, seed)
        else:
            raise NotImplementedError("Only 1- and 2-qubit gates supported")

    # truncation ------------------------------------------------------------
    def truncate_top_k(self):
        """sparse_state.py:388-400: keep k largest, divide by sqrt(gamma_sq)."""
        if self.nnz <= self.k:
            return
        st = _ops.truncate(self._ctx, self._buf, self.nnz, self.k, self._tmp, renorm=True,
                           hist_tiles=self._hist_tiles)
        self._swap()
        self._count_select(self.nnz, self.k)
        self.nnz = self.k
        self.gamma *= np.sqrt(st.kept)
        self._hist_tiles = 0

    def truncate_random_k(self, seed):
        """sparse_state.py:402-415."""
        if self.nnz <= self.k:
            return
        if seed is None:
            seed = int(np.random.randint(0, 2 ** 31))
        from .truncation import _random_k_device
        gsq = _random_k_device(self._ctx, self._buf, self.nnz, self.k, seed, self._tmp)
        self._swap()
        self.nnz = self.k
        gl = np.sqrt(gsq)
        _ops.scale(self._ctx, self._buf, self.nnz, gl)
        self.gamma *= gl
        self._hist_tiles = 0
