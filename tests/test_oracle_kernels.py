"""Pins of the CPU oracle's kernels against closed forms and dense brute force
(no reference code runs here).  A plausible mistake in the oracle -- a swapped
operand, a wrong bit, a dropped term -- changes a dense result below."""

import numpy as np
import pytest
from scipy.linalg import expm

from conftest import random_state, random_unitary
from oracle import circuits as OC
from oracle import drivers as OD
from oracle import exact as OE
from oracle import kernels as OK

X = np.array([[0, 1], [1, 0]], dtype=complex)
Z = np.diag([1.0, -1.0]).astype(complex)


def dense_apply(psi, N, G, qubits):
    """Full 2^N matrix of a 1/2-qubit gate with qubit q = bit q and 2-qubit
    index (b_q1 << 1) | b_q2, applied by explicit index arithmetic."""
    out = np.zeros_like(psi)
    idx = np.arange(2 ** N)
    if len(qubits) == 1:
        q = qubits[0]
        b = (idx >> q) & 1
        for o in range(2):
            tgt = (idx & ~(1 << q)) | (o << q)
            np.add.at(out, tgt, G[o, b] * psi)
    else:
        q1, q2 = qubits
        b = (((idx >> q1) & 1) << 1) | ((idx >> q2) & 1)
        for o in range(4):
            tgt = (idx & ~((1 << q1) | (1 << q2))) | (((o >> 1) & 1) << q1) | ((o & 1) << q2)
            np.add.at(out, tgt, G[o, b] * psi)
    return out


def sparse_to_dense(x, a, N):
    d = np.zeros(2 ** N, dtype=complex)
    d[np.asarray(x).astype(np.int64)] = a
    return d


def test_gate_matrices_closed_form():
    t = 0.37
    np.testing.assert_allclose(OC.RX(0, t).matrix, expm(1j * t * X), atol=1e-15)
    np.testing.assert_allclose(OC.RZZ(0, 1, t).matrix, expm(1j * t * np.kron(Z, Z)), atol=1e-15)
    np.testing.assert_allclose(OC.RZ(0, t).matrix, expm(-1j * t * Z / 2), atol=1e-15)
    Y = np.array([[0, -1j], [1j, 0]])
    np.testing.assert_allclose(OC.RY(0, t).matrix, expm(-1j * t * Y / 2), atol=1e-15)


@pytest.mark.parametrize("qubits", [(0,), (3,), (1, 4), (4, 1), (0, 5)])
def test_exact_simulator_vs_index_arithmetic(qubits):
    N = 6
    rng = np.random.default_rng(len(qubits) * 10 + qubits[0])
    psi = rng.standard_normal(2 ** N) + 1j * rng.standard_normal(2 ** N)
    G = random_unitary(2 ** len(qubits), rng)
    sim = OE.ExactSimulator(N)
    got = sim.apply_gate_exact(psi, OD._G(qubits, G))
    np.testing.assert_allclose(got, dense_apply(psi, N, G, qubits), atol=1e-13)


@pytest.mark.parametrize("q", [0, 2, 5])
def test_1q_emit_sort_merge_equals_dense(q):
    N = 7
    rng = np.random.default_rng(q)
    x, a = random_state(40, N, seed=q)
    G = random_unitary(2, rng)
    G[1, 0] = 0.0
    tx, ta = np.zeros(80, np.uint64), np.zeros(80, complex)
    m = OK.apply_1qubit_gate_kernel_opt(x, a, 40, tx, ta, G, q)
    OK.radix_sort_uint64(tx, ta, m)
    assert np.all(np.diff(tx[:m].astype(np.int64)) >= 0)
    m = OK.merge_sorted_kernel(tx, ta, m)
    assert len(np.unique(tx[:m])) == m
    np.testing.assert_allclose(sparse_to_dense(tx[:m], ta[:m], N),
                               dense_apply(sparse_to_dense(x, a, N), N, G, (q,)), atol=1e-14)


@pytest.mark.parametrize("qs", [(1, 4), (5, 0)])
def test_2q_ht_equals_dense_and_first_appearance(qs):
    N = 7
    rng = np.random.default_rng(sum(qs))
    x, a = random_state(50, N, seed=7)
    G = random_unitary(4, rng)
    tx, ta = np.zeros(200, np.uint64), np.zeros(200, complex)
    m = OK.apply_2qubit_gate_ht(x, a, 50, tx, ta, G, *qs)
    np.testing.assert_allclose(sparse_to_dense(tx[:m], ta[:m], N),
                               dense_apply(sparse_to_dense(x, a, N), N, G, qs), atol=1e-14)
    # first-appearance order: every key appears after the keys emitted earlier
    ex, ea = np.zeros(200, np.uint64), np.zeros(200, complex)
    me = OK.apply_2qubit_gate_kernel_opt(x, a, 50, ex, ea, G, *qs)
    _, first = np.unique(ex[:me], return_index=True)
    assert np.array_equal(tx[:m], ex[:me][np.sort(first)])


def test_diagonal_equals_dense():
    N = 6
    x, a = random_state(30, N, seed=3)
    d = np.exp(1j * np.array([0.2, -1.1, 0.5, 2.5]))
    b = a.copy()
    OK.apply_2qubit_diagonal_kernel(x, b, 30, 4, 1, d)
    np.testing.assert_allclose(sparse_to_dense(x, b, N),
                               dense_apply(sparse_to_dense(x, a, N), N, np.diag(d), (4, 1)), atol=1e-15)


@pytest.mark.parametrize("nq", [1, 2])
def test_rotation_equals_dense(nq):
    N = 6
    rng = np.random.default_rng(nq)
    x, a = random_state(25, N, seed=nq + 20)
    U = random_unitary(2 ** nq, rng)
    bx, ba = np.zeros(100, np.uint64), np.zeros(100, complex)
    bx[:25], ba[:25] = x, a
    if nq == 1:
        m = OK.apply_1qubit_rotation_sparse(bx, ba, 25, 2, U)
        ref = dense_apply(sparse_to_dense(x, a, N), N, U, (2,))
        gm = np.uint64(1 << 2)
    else:
        m = OK.apply_2qubit_rotation_sparse(bx, ba, 25, 3, 0, U)
        ref = dense_apply(sparse_to_dense(x, a, N), N, U, (3, 0))
        gm = np.uint64((1 << 3) | 1)
    rest = (bx[:m] & ~gm).astype(np.int64)
    assert np.all(np.diff(rest) >= 0), "rotation output ascends by the rest key (bass_simulator.py:165)"
    np.testing.assert_allclose(sparse_to_dense(bx[:m], ba[:m], N), ref, atol=1e-14)
    assert m == np.count_nonzero(np.abs(ref) > 1e-30)


def test_rdms_equal_partial_trace():
    N = 7
    x, a = random_state(60, N, seed=9)
    rho = OK.compute_all_1qubit_rdms(x, a, 60, N)
    psi = sparse_to_dense(x, a, N).reshape([2] * N)
    for q in range(N):
        # axis of qubit q in the C-order reshape is N-1-q
        m = np.moveaxis(psi, N - 1 - q, 0).reshape(2, -1)
        np.testing.assert_allclose(rho[q], m @ m.conj().T, atol=1e-14)


@pytest.mark.parametrize("q1,q2", [(0, 1), (5, 2), (3, 6)])
def test_2qubit_rdm_equals_partial_trace(q1, q2):
    """rho[p][q] = sum_rest psi[p, rest] conj(psi[q, rest]), index (b_q1 << 1) | b_q2."""
    N = 7
    x, a = random_state(70, N, seed=q1 * 7 + q2)
    rho = OK.compute_2qubit_rdm(x, a, 70, q1, q2)
    psi = sparse_to_dense(x, a, N).reshape([2] * N)
    m = np.moveaxis(psi, (N - 1 - q1, N - 1 - q2), (0, 1)).reshape(4, -1)
    np.testing.assert_allclose(rho, m @ m.conj().T, atol=1e-14)
    assert abs(np.trace(rho) - np.sum(np.abs(a) ** 2)) < 1e-13


def test_eigh_2x2_against_numpy_and_gauge():
    rng = np.random.default_rng(4)
    for _ in range(50):
        A = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        H = A @ A.conj().T
        ev, V = OK.eigh_2x2(H)
        np.testing.assert_allclose(np.sort(ev), np.linalg.eigvalsh(H), atol=1e-13)
        assert ev[0] >= ev[1]
        np.testing.assert_allclose(H @ V, V * ev, atol=1e-12)
        np.testing.assert_allclose(V.conj().T @ V, np.eye(2), atol=1e-13)
        # gauge: first components equal b / n (same phase as rho01), second real
        assert abs(V[1, 0].imag) == 0 and abs(V[1, 1].imag) == 0
        assert abs(np.angle(V[0, 0]) - np.angle(H[0, 1])) < 1e-12
    ev, V = OK.eigh_2x2(np.diag([0.2, 0.8]).astype(complex))
    np.testing.assert_array_equal(V, np.array([[0, 1], [1, 0]]))
    ev, V = OK.eigh_2x2(np.diag([0.8, 0.2]).astype(complex))
    np.testing.assert_array_equal(V, np.eye(2))


def test_topk_brute_force_and_gamma():
    x, a = random_state(500, 16, seed=2)
    bx, ba = x.copy(), a.copy()
    k, gsq = OK.top_k_truncate_kernel_opt(bx, ba, 500, 123)
    p = np.abs(a) ** 2
    top = set(x[np.argsort(-p)[:123]].tolist())
    assert set(bx[:k].tolist()) == top
    assert abs(gsq - np.sum(np.sort(p)[::-1][:123])) < 1e-14
    sel = OD.key_tiebreak_selector(x, a, p, 123)
    assert set(x[sel].tolist()) == top


def test_key_tiebreak_prefers_smaller_keys():
    x = np.array([9, 3, 7, 1, 5], dtype=np.uint64)
    a = np.array([0.5, 0.5, 0.5, 0.1, 0.5], dtype=complex)
    sel = OD.key_tiebreak_selector(x, a, None, 2)
    assert sorted(x[sel].tolist()) == [3, 5]


def test_lcg_fisher_yates_known_answer():
    # LCG from seed 1: 1103515245*1+12345 = 1103527590 (< 2^31)
    assert OK.lcg_next(1) == 1103527590
    assert OK.lcg_next(1103527590) == (1103515245 * 1103527590 + 12345) & 0x7FFFFFFF
    x = np.arange(10, dtype=np.uint64)
    a = (np.arange(10) + 1.0).astype(complex)
    k, gsq = OK.random_k_truncate_kernel_opt(x, a, 10, 3, 1)
    r = 1
    perm = list(range(10))
    for i in range(3):
        r = (1103515245 * r + 12345) & 0x7FFFFFFF
        j = i + r % (10 - i)
        perm[i], perm[j] = perm[j], perm[i]
    assert x[:3].tolist() == perm[:3]
    assert gsq == sum((p + 1.0) ** 2 for p in perm[:3])


def test_radix_and_merge_keep_zero_sums():
    x = np.array([4, 1, 4, 2, 1], dtype=np.uint64)
    a = np.array([1.0, 2.0, -1.0, 3.0, 0.5], dtype=complex)
    OK.radix_sort_uint64(x, a, 5)
    assert x.tolist() == [1, 1, 2, 4, 4]
    assert a.tolist() == [2.0, 0.5, 3.0, 1.0, -1.0]
    m = OK.merge_sorted_kernel(x, a, 5)
    assert m == 3 and x[:3].tolist() == [1, 2, 4] and a[:3].tolist() == [2.5, 3.0, 0.0]
