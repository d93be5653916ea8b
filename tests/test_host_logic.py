"""Host-side pieces of the product (no GPU): gate matrices, circuit
generators, frame algebra, random-k index sequence."""

import numpy as np

from conftest import pkg_module
from oracle import circuits as OC
from oracle import kernels as OK


def _same_circuit(a, b):
    assert len(a) == len(b)
    for g, h in zip(a, b):
        assert tuple(g.qubits) == tuple(h.qubits)
        assert g.n_qubits == h.n_qubits
        assert np.array_equal(g.matrix, h.matrix)


def test_tfim_generator_matches_oracle():
    rc = pkg_module("utils.random_circuits")
    for seed, kw in [(1, {}), (2, dict(disorder_strength_J=0.3, disorder_strength_h=0.5))]:
        a = rc.generate_tfim_circuit(9, 4, J=1.0, h=0.7, dt=0.1, rng=np.random.default_rng(seed), **kw)
        b = OC.tfim(9, 4, J=1.0, h=0.7, dt=0.1, rng=np.random.default_rng(seed),
                    dJ=kw.get("disorder_strength_J", 0.0), dh=kw.get("disorder_strength_h", 0.0))
        _same_circuit(a, b)


def test_rfim_generator_matches_oracle():
    rc = pkg_module("utils.random_circuits")
    _same_circuit(rc.generate_rfim_circuit(11, 3, W=2.0, dt=0.2, rng=np.random.default_rng(5)),
                  OC.rfim(11, 3, W=2.0, dt=0.2, rng=np.random.default_rng(5)))


def test_brickwork_and_kicked_ising_structure():
    rc = pkg_module("utils.random_circuits")
    c, ends = rc.brickwork_haar_circuit(6, 3, np.random.default_rng(0))
    assert [g.n_qubits for g in c[:9]] == [1] * 6 + [2] * 3
    assert [g.qubits for g in c[6:9]] == [(0, 1), (2, 3), (4, 5)]
    assert [g.qubits for g in c[15:17]] == [(1, 2), (3, 4)]
    assert ends == [8, 16, 25]
    for g in c:
        if g.n_qubits == 1:
            np.testing.assert_allclose(g.matrix @ g.matrix.conj().T, np.eye(2), atol=1e-14)
    k, kends = rc.kicked_ising_circuit(5, 2)
    assert [g.qubits for g in k[:5]] == [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]
    assert np.allclose(k[5].matrix, OC.RX(0, 0.3).matrix) and np.allclose(k[15].matrix, OC.RX(0, 0.31).matrix)
    assert kends == [9, 19]


def test_eigh_and_kron_match_oracle_bitwise():
    B = pkg_module("simulation.bass_simulator")
    rng = np.random.default_rng(3)
    for _ in range(20):
        A = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        H = A @ A.conj().T
        e1, V1 = B.eigh_2x2(H)
        e2, V2 = OK.eigh_2x2(H)
        assert np.array_equal(e1, e2) and np.array_equal(V1, V2)
        U = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        assert np.array_equal(B._manual_kron_2x2(A, U), OK.manual_kron_2x2(A, U))
        np.testing.assert_allclose(B._manual_kron_2x2(A, U), np.kron(A, U), atol=1e-15)


def test_eigh_single_batch_oracle_bitwise_where_pow_rounds_differently():
    """The single-circuit driver's eigh_2x2, the batch engine's row-wise form
    and the oracle's agree bit for bit, also on off-diagonals whose square by
    C pow (Python's float **) is an ulp off the correctly rounded product --
    the reference's compiled b.real**2 -- which once made the two drivers take
    different rotation decisions."""
    from fractions import Fraction

    B = pkg_module("simulation.bass_simulator")
    Bm = pkg_module("simulation.batched")
    rng = np.random.default_rng(0)
    rhos = []
    while len(rhos) < 12:
        a = rng.random()
        b = complex(rng.standard_normal(), rng.standard_normal()) * rng.random() * 0.5 * 10 ** rng.uniform(-6, 0)
        if b.real ** 2 != b.real * b.real or b.imag ** 2 != b.imag * b.imag:
            for v in (b.real, b.imag):
                assert v * v == float(Fraction(v) * Fraction(v))  # the product rounds once, correctly
            rhos.append(np.array([[a, b], [b.conjugate(), 1 - a]]))
    Vb, _ = Bm._eigh2x2_rows(np.array(rhos))
    for r, vb in zip(rhos, Vb):
        e1, V1 = B.eigh_2x2(r)
        e2, V2 = OK.eigh_2x2(r)
        assert np.array_equal(e1.view(np.uint64), e2.view(np.uint64))
        assert np.array_equal(V1.view(np.uint64), V2.view(np.uint64))
        assert np.array_equal(V1.view(np.uint64), vb.view(np.uint64))


def test_random_k_indices_match_oracle_lcg():
    T = pkg_module("core.truncation")
    for n, k, seed in [(10, 3, 1), (1000, 400, 12345), (50, 50 - 1, 7)]:
        idx = T.random_k_indices(n, k, seed)
        x = np.arange(n, dtype=np.uint64)
        a = np.ones(n, dtype=complex)
        OK.random_k_truncate_kernel_opt(x, a, n, k, seed)
        assert np.array_equal(idx, x[:k].astype(np.int64))


def test_gate_classes_match_reference_conventions():
    G = pkg_module("core.gates")
    np.testing.assert_array_equal(G.RXGate(0, 0.3).matrix, OC.RX(0, 0.3).matrix)
    np.testing.assert_array_equal(G.RZZGate(0, 1, 0.3).matrix, OC.RZZ(0, 1, 0.3).matrix)
    np.testing.assert_array_equal(G.RZGate(0, 0.3).matrix, OC.RZ(0, 0.3).matrix)
    np.testing.assert_array_equal(G.CZGate(0, 1).matrix, OC.CZ(0, 1).matrix)
    assert G.CNOTGate(3, 1).qubits == (3, 1)


def test_stored_to_keys_inverts_the_bit_placement():
    """Stored word: basis bit q placed at bit pos[q]; stored_to_keys must give
    the basis key back (placement done here one key at a time)."""
    ss = pkg_module("core.sparse_state")
    rng = np.random.default_rng(3)
    for N in (1, 7, 8, 9, 33, 63):
        pos = rng.permutation(N)
        keys = rng.integers(0, 1 << min(N, 62), size=200, dtype=np.uint64)
        if N == 63:
            keys |= np.uint64(1) << np.uint64(62)
        w = np.array([sum(((int(k) >> q) & 1) << int(pos[q]) for q in range(N)) for k in keys], dtype=np.uint64)
        assert np.array_equal(ss.stored_to_keys(w, pos), keys)
    assert np.array_equal(ss.stored_to_keys(keys, None), keys)


def test_batched_host_helpers():
    """Row-wise position lifts and key maps of the batched driver agree with
    the single-circuit helpers; the slab allocator is first fit and merges
    neighbours on release."""
    Bm = pkg_module("simulation.batched")
    ss = pkg_module("core.sparse_state")
    rng = np.random.default_rng(4)
    P = np.stack([np.concatenate([rng.permutation(20), Bm._BAD + np.arange(20, Bm.NMAX)]) for _ in range(6)])
    jb = rng.integers(0, 20, size=6)
    for t in (0, 1):
        jj = np.maximum(jb, t)
        L = Bm._lift_rows(P, jj, t)
        for r in range(6):
            assert np.array_equal(L[r, :20], ss.lifted_positions(P[r, :20], int(jj[r]), t))
            assert np.array_equal(L[r, 20:], P[r, 20:])
        km = Bm._key_maps(L)
        for r in range(6):
            assert all(km[r, L[r, q]] == q for q in range(20))
    A = rng.normal(size=(5, 2, 2)) + 1j * rng.normal(size=(5, 2, 2))
    B = rng.normal(size=(5, 2, 2)) + 1j * rng.normal(size=(5, 2, 2))
    K = Bm._kron2(A, B)
    for r in range(5):
        assert np.allclose(K[r], np.kron(A[r], B[r]), rtol=0, atol=1e-14)
    eng = Bm.BatchEngine.__new__(Bm.BatchEngine)
    eng.free = [(0, 100)]
    a, b, c = eng._alloc(30), eng._alloc(50), eng._alloc(20)
    assert (a, b, c) == (0, 30, 80) and eng._alloc(1) is None
    eng._release(30, 50)
    assert eng._alloc(60) is None and eng._alloc(40) == 30
    eng._release(0, 30)
    eng._release(30, 40)
    eng._release(80, 20)
    assert eng.free == [(0, 100)]
