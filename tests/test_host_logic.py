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
