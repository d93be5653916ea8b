"""numpy restatement of the reference's Numba kernels (oracle; test infrastructure).

Each function keeps the reference's name, argument order and return value.  The
loops are vectorised, but every order-dependent result (output order,
first-appearance order, which duplicates are summed) follows the sequential
reference loop.  Citations are to files in the reference repository.
"""

import math

import numpy as np

EMPTY = np.uint64(0xFFFFFFFFFFFFFFFF)
_U1 = np.uint64(1)


def _u64(v):
    return np.uint64(int(v))


# ---------------------------------------------------------------------------
# src/core/sparse_state.py
# ---------------------------------------------------------------------------

def bitwise_sort_key(x_array, mask):
    """src/core/sparse_state.py:7-12 -- free bits of each key."""
    return x_array & ~_u64(mask)


def apply_1qubit_gate_kernel_opt(x_in, alpha_in, nnz_in, x_out, alpha_out, gate_matrix, q):
    """src/core/sparse_state.py:15-67.

    For each input entry (in input order) emit ``(x & ~m, G[0,b] a)`` then
    ``(x | m, G[1,b] a)``; an emission whose product is exactly zero (both parts)
    is dropped.  Returns nnz_out.
    """
    n = int(nnz_in)
    x = np.asarray(x_in[:n], dtype=np.uint64)
    a = np.asarray(alpha_in[:n], dtype=np.complex128)
    m = _U1 << _u64(q)
    bit = (x >> _u64(q)) & _U1
    base = x & ~m
    g = np.asarray(gate_matrix, dtype=np.complex128)
    out0 = np.where(bit == 0, g[0, 0] * a, g[0, 1] * a)
    out1 = np.where(bit == 0, g[1, 0] * a, g[1, 1] * a)
    keys = np.empty(2 * n, dtype=np.uint64)
    vals = np.empty(2 * n, dtype=np.complex128)
    keys[0::2] = base
    keys[1::2] = base | m
    vals[0::2] = out0
    vals[1::2] = out1
    keep = (vals.real != 0.0) | (vals.imag != 0.0)
    cnt = int(keep.sum())
    x_out[:cnt] = keys[keep]
    alpha_out[:cnt] = vals[keep]
    return cnt


def apply_2qubit_gate_kernel_opt(x_in, alpha_in, nnz_in, x_out, alpha_out, gate_matrix, q1, q2):
    """src/core/sparse_state.py:70-124.

    Up to four emissions per input entry, ``out_idx`` 0..3 in order, index
    ``(b_q1 << 1) | b_q2``; exact-zero products dropped; returns -1 when the
    output buffer would overflow (``:117-118``).
    """
    n = int(nnz_in)
    x = np.asarray(x_in[:n], dtype=np.uint64)
    a = np.asarray(alpha_in[:n], dtype=np.complex128)
    m1 = _U1 << _u64(q1)
    m2 = _U1 << _u64(q2)
    b1 = ((x >> _u64(q1)) & _U1).astype(np.int64)
    b2 = ((x >> _u64(q2)) & _U1).astype(np.int64)
    in_idx = (b1 << 1) | b2
    base = x & ~(m1 | m2)
    g = np.asarray(gate_matrix, dtype=np.complex128)
    keys = np.empty((n, 4), dtype=np.uint64)
    vals = np.empty((n, 4), dtype=np.complex128)
    for o in range(4):
        kx = base.copy()
        if (o >> 1) & 1:
            kx |= m1
        if o & 1:
            kx |= m2
        keys[:, o] = kx
        vals[:, o] = g[o, in_idx] * a
    keys = keys.ravel()
    vals = vals.ravel()
    keep = (vals.real != 0.0) | (vals.imag != 0.0)
    cnt = int(keep.sum())
    if cnt > x_out.shape[0]:
        return -1
    x_out[:cnt] = keys[keep]
    alpha_out[:cnt] = vals[keep]
    return cnt


def merge_sorted_kernel(x, alpha, nnz):
    """src/core/sparse_state.py:127-147.

    Run-length sum of equal adjacent keys, in place; zero sums are KEPT.
    """
    n = int(nnz)
    if n == 0:
        return 0
    xs = x[:n]
    starts = np.flatnonzero(np.concatenate(([True], xs[1:] != xs[:-1])))
    lens = np.diff(np.append(starts, n))
    # strictly left-to-right sums inside each run, as the reference loop adds
    # (np.add.reduceat does not keep that order for runs longer than 2)
    sums = alpha[starts].copy()
    for j in range(1, int(lens.max())):
        live = lens > j
        sums[live] = sums[live] + alpha[starts[live] + j]
    keys = xs[starts].copy()
    cnt = len(starts)
    x[:cnt] = keys
    alpha[:cnt] = sums
    return cnt


def radix_sort_uint64(keys, values, n):
    """src/core/sparse_state.py:150-172.

    4-pass 16-bit LSD radix sort of keys[:n] with values as payload.  LSD radix
    sort is stable, so the result equals a stable sort on the full 64-bit key.
    """
    n = int(n)
    order = np.argsort(keys[:n], kind="stable")
    keys[:n] = keys[:n][order]
    values[:n] = values[:n][order]


def apply_2qubit_diagonal_kernel(x, alpha, nnz, q1, q2, diag):
    """src/core/sparse_state.py:178-190 -- alpha[i] *= diag[(b1 << 1) | b2]."""
    n = int(nnz)
    xs = x[:n]
    b1 = ((xs >> _u64(q1)) & _U1).astype(np.int64)
    b2 = ((xs >> _u64(q2)) & _U1).astype(np.int64)
    d = np.asarray(diag, dtype=np.complex128)
    alpha[:n] = alpha[:n] * d[(b1 << 1) | b2]


# ---------------------------------------------------------------------------
# src/core/truncation.py
# ---------------------------------------------------------------------------

def top_k_truncate_kernel_opt(x, alpha, nnz, k, selector=None):
    """src/core/truncation.py:68-113.

    Keep the k largest |a|^2 (argpartition; tie order implementation-defined),
    compact them to the front in selection order and return ``(k, gamma_sq)``
    with gamma_sq the sum of the kept |a|^2 (not renormalised).  ``selector``
    (oracle extension) replaces argpartition by a caller-supplied rule.
    """
    n = int(nnz)
    k = int(k)
    a = alpha[:n]
    mags = a.real * a.real + a.imag * a.imag
    if n <= k:
        return n, float(_seq_sum(mags))
    if selector is None:
        idx = np.argpartition(mags, n - k)[n - k:]
    else:
        idx = selector(x[:n], alpha[:n], mags, k)
    gamma_sq = float(_seq_sum(mags[idx]))
    tx = x[idx].copy()
    ta = alpha[idx].copy()
    x[:k] = tx
    alpha[:k] = ta
    return k, gamma_sq


def lcg_next(rng):
    """The reference's LCG step, ``(1103515245 * rng + 12345) & 0x7FFFFFFF``
    (src/core/truncation.py:136, src/simulation/simulator.py:158)."""
    return (1103515245 * int(rng) + 12345) & 0x7FFFFFFF


def random_k_truncate_kernel_opt(x, alpha, nnz, k, seed):
    """src/core/truncation.py:116-159 -- LCG partial Fisher-Yates over the first k."""
    n = int(nnz)
    k = int(k)
    a = alpha[:n]
    if n <= k:
        return n, float(_seq_sum(a.real * a.real + a.imag * a.imag))
    perm = np.arange(n)
    rng = int(seed)
    for i in range(k):
        rng = lcg_next(rng)
        r = i + (rng % (n - i))
        perm[i], perm[r] = perm[r], perm[i]
    sel = perm[:k]
    tx = x[sel].copy()
    ta = alpha[sel].copy()
    gamma_sq = float(_seq_sum(ta.real * ta.real + ta.imag * ta.imag))
    x[:k] = tx
    alpha[:k] = ta
    return k, gamma_sq


def _seq_sum(v):
    """Left-to-right float sum (the Numba loops accumulate sequentially)."""
    s = 0.0
    for t in np.asarray(v, dtype=np.float64).tolist():
        s += t
    return s


# ---------------------------------------------------------------------------
# src/simulation/bass_simulator.py -- hash table
# ---------------------------------------------------------------------------

def ht_clear(ht_keys):
    """src/simulation/bass_simulator.py:100-105."""
    ht_keys[:] = EMPTY


def _ht_slot(ht_keys, key):
    cap = ht_keys.shape[0]
    mask = cap - 1
    h = int(key) & mask
    while ht_keys[h] != EMPTY:
        if ht_keys[h] == key:
            return h, True
        h = (h + 1) & mask
    return h, False


def ht_build(x, nnz, ht_keys, ht_vals):
    """src/simulation/bass_simulator.py:108-123 (h = key & (cap-1), linear probing)."""
    for i in range(int(nnz)):
        h, _ = _ht_slot(ht_keys, x[i])
        ht_keys[h] = x[i]
        ht_vals[h] = np.int32(i)


def ht_lookup(ht_keys, ht_vals, key):
    """src/simulation/bass_simulator.py:126-137 -- stored index or -1."""
    h, found = _ht_slot(ht_keys, np.uint64(int(key)))
    return int(ht_vals[h]) if found else -1


def ht_insert(ht_keys, ht_vals, key, val):
    """src/simulation/bass_simulator.py:140-153 -- insert or overwrite."""
    key = np.uint64(int(key))
    h, _ = _ht_slot(ht_keys, key)
    ht_keys[h] = key
    ht_vals[h] = np.int32(val)


# ---------------------------------------------------------------------------
# src/simulation/bass_simulator.py -- kernels
# ---------------------------------------------------------------------------

def apply_1qubit_rotation_sparse(x, alpha, nnz, qubit, U):
    """src/simulation/bass_simulator.py:156-199.

    Exact rotation of qubit ``qubit`` by U: entries grouped by the rest key,
    ``(a0', a1') = U (a0, a1)`` with an absent member read as 0; each output
    kept iff ``|a'| > 1e-30``; output ascending by rest (bit 0 before bit 1),
    i.e. ascending by key.  In place; x/alpha must hold 2*nnz entries.
    """
    n = int(nnz)
    xs = np.asarray(x[:n], dtype=np.uint64)
    a = np.asarray(alpha[:n], dtype=np.complex128)
    m = _U1 << _u64(qubit)
    rest = xs & ~m
    bits = ((xs >> _u64(qubit)) & _U1).astype(bool)
    uniq, inv = np.unique(rest, return_inverse=True)
    g = len(uniq)
    a0 = np.zeros(g, dtype=np.complex128)
    a1 = np.zeros(g, dtype=np.complex128)
    a0[inv[~bits]] = a[~bits]
    a1[inv[bits]] = a[bits]
    U = np.asarray(U, dtype=np.complex128)
    n0 = U[0, 0] * a0 + U[0, 1] * a1
    n1 = U[1, 0] * a0 + U[1, 1] * a1
    keys = np.empty(2 * g, dtype=np.uint64)
    vals = np.empty(2 * g, dtype=np.complex128)
    keys[0::2] = uniq
    keys[1::2] = uniq | m
    vals[0::2] = n0
    vals[1::2] = n1
    keep = np.abs(vals) > 1e-30
    cnt = int(keep.sum())
    x[:cnt] = keys[keep]
    alpha[:cnt] = vals[keep]
    return cnt


def compute_2qubit_rdm(x, alpha, nnz, q1, q2):
    """src/simulation/bass_simulator.py:202-234 -- 4x4 RDM, index (b1<<1)|b2."""
    n = int(nnz)
    xs = np.asarray(x[:n], dtype=np.uint64)
    a = np.asarray(alpha[:n], dtype=np.complex128)
    m12 = (_U1 << _u64(q1)) | (_U1 << _u64(q2))
    rest = xs & ~m12
    idx2 = (((xs >> _u64(q1)) & _U1).astype(np.int64) << 1) | ((xs >> _u64(q2)) & _U1).astype(np.int64)
    uniq, inv = np.unique(rest, return_inverse=True)
    A = np.zeros((len(uniq), 4), dtype=np.complex128)
    A[inv, idx2] = a
    return A.T @ A.conj()


def apply_2qubit_rotation_sparse(x, alpha, nnz, q1, q2, U4):
    """src/simulation/bass_simulator.py:237-289 -- 4x4 unitary on the pair,
    outputs with |a'| <= 1e-30 dropped, ascending by rest then out index."""
    n = int(nnz)
    xs = np.asarray(x[:n], dtype=np.uint64)
    a = np.asarray(alpha[:n], dtype=np.complex128)
    m1 = _U1 << _u64(q1)
    m2 = _U1 << _u64(q2)
    rest = xs & ~(m1 | m2)
    idx2 = (((xs >> _u64(q1)) & _U1).astype(np.int64) << 1) | ((xs >> _u64(q2)) & _U1).astype(np.int64)
    uniq, inv = np.unique(rest, return_inverse=True)
    A = np.zeros((len(uniq), 4), dtype=np.complex128)
    A[inv, idx2] = a
    U4 = np.asarray(U4, dtype=np.complex128)
    out = np.zeros_like(A)
    for o in range(4):
        acc = np.zeros(len(uniq), dtype=np.complex128)
        for i in range(4):
            acc = acc + U4[o, i] * A[:, i]
        out[:, o] = acc
    keys = np.empty((len(uniq), 4), dtype=np.uint64)
    for o in range(4):
        kx = uniq.copy()
        if (o >> 1) & 1:
            kx |= m1
        if o & 1:
            kx |= m2
        keys[:, o] = kx
    keys = keys.ravel()
    vals = out.ravel()
    keep = np.abs(vals) > 1e-30
    cnt = int(keep.sum())
    x[:cnt] = keys[keep]
    alpha[:cnt] = vals[keep]
    return cnt


def apply_2qubit_gate_ht(x_in, alpha_in, nnz_in, x_out, alpha_out, gate_matrix, q1, q2,
                         ht_keys=None, ht_vals=None):
    """src/simulation/bass_simulator.py:292-354.

    Up to four emissions per input entry (exact-zero products skipped); a key
    already emitted is accumulated into its first-seen slot, so the output order
    is first appearance and each sum runs in input order.  The reference resolves
    duplicates through its hash table; the result does not depend on the table,
    which is accepted and left untouched here.
    """
    n = int(nnz_in)
    x = np.asarray(x_in[:n], dtype=np.uint64)
    a = np.asarray(alpha_in[:n], dtype=np.complex128)
    m1 = _U1 << _u64(q1)
    m2 = _U1 << _u64(q2)
    b1 = ((x >> _u64(q1)) & _U1).astype(np.int64)
    b2 = ((x >> _u64(q2)) & _U1).astype(np.int64)
    in_idx = (b1 << 1) | b2
    base = x & ~(m1 | m2)
    g = np.asarray(gate_matrix, dtype=np.complex128)
    keys = np.empty((n, 4), dtype=np.uint64)
    vals = np.empty((n, 4), dtype=np.complex128)
    for o in range(4):
        kx = base.copy()
        if (o >> 1) & 1:
            kx |= m1
        if o & 1:
            kx |= m2
        keys[:, o] = kx
        vals[:, o] = g[o, in_idx] * a
    keys = keys.ravel()
    vals = vals.ravel()
    keep = (vals.real != 0.0) | (vals.imag != 0.0)
    keys = keys[keep]
    vals = vals[keep]
    if len(keys) == 0:
        return 0
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    acc = np.zeros(len(uniq), dtype=np.complex128)
    np.add.at(acc, inv, vals)  # unbuffered, in emission order
    order = np.argsort(first, kind="stable")
    cnt = len(uniq)
    x_out[:cnt] = uniq[order]
    alpha_out[:cnt] = acc[order]
    return cnt


def eigh_2x2(rho):
    """src/simulation/bass_simulator.py:387-437.

    Analytic 2x2 Hermitian eigensolver, eigenvalues descending, with the
    reference's eigenvector gauge ``v_max = [b, tau+disc]/n0``,
    ``v_min = [b, tau-disc]/n1`` and the ``|b|^2 < 1e-28`` identity/swap branch.
    """
    a = float(rho[0, 0].real)
    d = float(rho[1, 1].real)
    b = complex(rho[0, 1])
    # numba compiles b.real**2 to one rounded multiplication; Python's float
    # ** calls C pow, which differs from it by an ulp for some inputs
    b_sq = b.real * b.real + b.imag * b.imag
    tau = (d - a) * 0.5
    disc = math.sqrt(tau * tau + b_sq)
    lam0 = (a + d) * 0.5 + disc
    lam1 = (a + d) * 0.5 - disc
    evals = np.array([lam0, lam1])
    V = np.zeros((2, 2), dtype=np.complex128)
    if b_sq < 1e-28:
        if a >= d:
            V[0, 0] = 1.0
            V[1, 1] = 1.0
        else:
            V[0, 1] = 1.0
            V[1, 0] = 1.0
    else:
        td_p = tau + disc
        td_m = tau - disc
        n0 = math.sqrt(b_sq + td_p * td_p)
        n1 = math.sqrt(b_sq + td_m * td_m)
        V[0, 0] = b / n0
        V[1, 0] = complex(td_p) / n0
        V[0, 1] = b / n1
        V[1, 1] = complex(td_m) / n1
    return evals, V


def compute_all_1qubit_rdms(x, alpha, nnz, N, ht_keys=None, ht_vals=None):
    """src/simulation/bass_simulator.py:441-470.

    rho[q,b,b] = sum |a|^2 over entries with bit q = b;
    rho[q,0,1] = sum over bit-q-0 entries whose partner x|2^q is present of
    a_x conj(a_partner); rho[q,1,0] its conjugate sum.  The partner lookup here
    is a sorted search instead of the reference's table (same result).
    """
    n = int(nnz)
    xs = np.asarray(x[:n], dtype=np.uint64)
    a = np.asarray(alpha[:n], dtype=np.complex128)
    prob = a.real * a.real + a.imag * a.imag
    order = np.argsort(xs, kind="stable")
    sx = xs[order]
    rho = np.zeros((int(N), 2, 2), dtype=np.complex128)
    for q in range(int(N)):
        m = _U1 << _u64(q)
        bit = ((xs >> _u64(q)) & _U1).astype(bool)
        rho[q, 0, 0] = prob[~bit].sum()
        rho[q, 1, 1] = prob[bit].sum()
        sel = np.flatnonzero(~bit)
        if len(sel) == 0 or n == 0:
            continue
        partner = xs[sel] | m
        pos = np.searchsorted(sx, partner)
        pos_c = np.minimum(pos, n - 1)
        hit = sx[pos_c] == partner
        if not np.any(hit):
            continue
        j = order[pos_c[hit]]
        i = sel[hit]
        rho[q, 0, 1] = np.sum(a[i] * np.conj(a[j]))
        rho[q, 1, 0] = np.sum(a[j] * np.conj(a[i]))
    return rho


def is_gate_diagonal(G, tol=1e-12):
    """src/simulation/bass_simulator.py:473-475."""
    G = np.asarray(G)
    return np.max(np.abs(G - np.diag(np.diag(G)))) < tol


def manual_kron_2x2(A, B):
    """src/simulation/bass_simulator.py:357-383 -- A (x) B, elementwise products."""
    C = np.empty((4, 4), dtype=np.complex128)
    for i in range(2):
        for j in range(2):
            for k in range(2):
                for l in range(2):
                    C[2 * i + k, 2 * j + l] = A[i, j] * B[k, l]
    return C
