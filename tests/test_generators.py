"""Host-side circuit generators and metrics of the drop-in surface
(src/core/gates.py, src/utils/random_circuits.py, src/experimentation/runner.py,
src/benchmarking/fidelity.py), checked against what the mathematics fixes:
unitarity, the pairing structure, the random draws they are defined by, a
hand-computed depth and the closed form of the cross-entropy estimate."""

import numpy as np
from scipy.stats import unitary_group

from conftest import pkg_module


def _gates():
    return pkg_module("core.gates")


def _rc():
    return pkg_module("utils.random_circuits")


def test_haar_random_unitary_seeded_draw():
    G = _gates()
    u = G.haar_random_unitary(4, seed=11)
    assert np.allclose(u.conj().T @ u, np.eye(4), atol=1e-13)
    # reseeding the legacy global generator is the same stream as a RandomState
    # of that seed (the form the runner's per-gate draws use)
    assert np.array_equal(u, unitary_group.rvs(4, random_state=np.random.RandomState(11)))
    assert np.array_equal(G.haar_random_unitary(4, seed=11), u)
    assert not np.array_equal(G.haar_random_unitary(4, seed=12), u)
    g = G.RandomTwoQubitGate(3, 1, seed=11)
    assert g.qubits == (3, 1) and np.array_equal(g.matrix, u)


def test_generate_random_circuit_pairings():
    rc = _rc()
    np.random.seed(5)
    N, L = 7, 4
    c = rc.generate_random_circuit(N, L)
    assert len(c) == L * (N // 2)
    for layer in range(L):
        qs = [q for g in c[layer * (N // 2):(layer + 1) * (N // 2)] for q in g.qubits]
        assert len(set(qs)) == len(qs) == 2 * (N // 2)  # a matching: no qubit twice in a layer
    for g in c:
        assert np.allclose(g.matrix.conj().T @ g.matrix, np.eye(4), atol=1e-12)
    # the layer's pairing is the global generator's shuffle, then one draw per gate
    np.random.seed(5)
    order = list(range(N))
    np.random.shuffle(order)
    assert c[0].qubits == (order[0], order[1])
    np.random.seed(5)
    again = rc.generate_random_circuit(N, L)
    assert all(a.qubits == b.qubits and np.array_equal(a.matrix, b.matrix) for a, b in zip(c, again))


def test_make_haar_reseeds_from_rng():
    R = pkg_module("experimentation.runner")
    rc = _rc()
    c = R.make_haar(6, 3, np.random.default_rng(42))
    seed = int(np.random.default_rng(42).integers(0, 2 ** 31))
    np.random.seed(seed)
    ref = rc.generate_random_circuit(6, 3)
    assert len(c) == 9
    assert all(a.qubits == b.qubits and np.array_equal(a.matrix, b.matrix) for a, b in zip(c, ref))


def test_generate_rfitm_circuit_fields():
    rc = _rc()
    N, L, J, h, dt, sh = 5, 2, 0.7, 1.3, 0.05, 0.2
    c = rc.generate_rfitm_circuit(N, L, J=J, h=h, dt=dt, sigma_h=sh, rng=np.random.default_rng(9))
    fields = np.random.default_rng(9).normal(loc=h, scale=sh, size=N)
    assert len(c) == L * (2 * N - 1)
    layer = c[: 2 * N - 1]
    for q in range(N - 1):
        g = layer[q]
        assert g.qubits == (q, q + 1)
        assert np.allclose(np.diag(g.matrix), np.exp(1j * J * dt * np.array([1, -1, -1, 1])), atol=1e-15)
    for q in range(N):
        g = layer[N - 1 + q]
        t = fields[q] * dt
        assert g.qubits == (q,)
        assert np.allclose(g.matrix, [[np.cos(t), 1j * np.sin(t)], [1j * np.sin(t), np.cos(t)]], atol=1e-15)
    # no rng: the reproducible default generator
    d0 = rc.generate_rfitm_circuit(N, 1)
    f0 = np.random.default_rng(0).normal(loc=1.0, scale=0.1, size=N)
    assert np.allclose(d0[N - 1].matrix[0, 0], np.cos(f0[0] * 0.1))


def test_circuit_depth_and_info():
    rc = _rc()
    G = _gates()
    # one TFIM layer on 3 qubits: RZZ(0,1) -> 1, RZZ(1,2) -> 2, RX(0) -> 2,
    # RX(1) -> 3, RX(2) -> 3
    c = [G.RZZGate(0, 1, 0.1), G.RZZGate(1, 2, 0.1), G.RXGate(0, 0.2), G.RXGate(1, 0.2), G.RXGate(2, 0.2)]
    assert rc.estimate_circuit_depth(c, 3) == 3
    assert rc.circuit_info(c) == {"num_gates": 5, "num_qubits": 3, "depth": 3}
    assert rc.circuit_info([]) == {"num_gates": 0, "num_qubits": 0, "depth": 0}
    # a brickwork layer pair on 4 qubits is depth 2
    b = [G.CZGate(0, 1), G.CZGate(2, 3), G.CZGate(1, 2)]
    assert rc.estimate_circuit_depth(b, 4) == 2


def test_cross_entropy_fidelity_closed_form():
    F = pkg_module("benchmarking.fidelity")
    N = 3
    exact = np.zeros(2 ** N, dtype=np.complex128)
    exact[[1, 6]] = np.sqrt([0.75, 0.25])

    class Sampled:
        def __init__(self, keys):
            self.N, self._k = N, np.asarray(keys, dtype=np.uint64)

        def sample(self, num_samples=1):
            return self._k[:num_samples]

    # 2^N <P> - 1 over the samples: keys 1, 1, 6, 0 -> mean P = (0.75 + 0.75 + 0.25 + 0) / 4
    xeb = F.estimate_cross_entropy_fidelity(Sampled([1, 1, 6, 0]), exact, num_samples=4)
    assert abs(xeb - (8 * 1.75 / 4 - 1)) < 1e-15
