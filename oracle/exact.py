"""Dense FP64 oracle: exact simulator and metrics (test infrastructure).

``ExactSimulator``            -- src/simulation/exact_simulator.py:66-144
``compute_fidelity``          -- src/benchmarking/fidelity.py:48-65
``compute_participation_ratio`` -- src/benchmarking/fidelity.py:79-89
``fidelity_bass``             -- src/experimentation/runner.py:620-626
"""

import numpy as np


class ExactSimulator:
    def __init__(self, num_qubits, verbose=False):
        if num_qubits > 24:
            raise ValueError(f"N={num_qubits} too large for exact simulation")
        self.N = num_qubits

    def simulate(self, circuit, initial_state=None):
        """exact_simulator.py:78-97."""
        psi = np.zeros(2 ** self.N, dtype=np.complex128)
        psi[0 if initial_state is None else initial_state] = 1.0
        for g in circuit:
            psi = self.apply_gate_exact(psi, g)
        return psi

    def apply_gate_exact(self, psi, gate):
        if gate.n_qubits == 1:
            return self._apply_1qubit_dense(psi, gate.matrix, gate.qubits[0])
        if gate.n_qubits == 2:
            return self._apply_2qubit_dense(psi, gate.matrix, gate.qubits[0], gate.qubits[1])
        raise NotImplementedError()

    def _apply_1qubit_dense(self, psi, matrix, q):
        """exact_simulator.py:114-119 -- axis 1 of (2^(N-q-1), 2, 2^q) is bit q."""
        psi3d = psi.reshape(-1, 2, 1 << q)
        return np.einsum("ij,kjl->kil", matrix, psi3d).reshape(-1)

    def _apply_2qubit_dense(self, psi, matrix, q1, q2):
        """exact_simulator.py:121-144 -- index (b1 << 1) | b2, q1 < q2 enforced by
        permuting the matrix."""
        N = self.N
        if q1 > q2:
            q1, q2 = q2, q1
            perm = [0, 2, 1, 3]
            matrix = matrix[np.ix_(perm, perm)]
        A, B, C = 1 << (N - q2 - 1), 1 << (q2 - q1 - 1), 1 << q1
        M = matrix.reshape(2, 2, 2, 2)
        psi5 = psi.reshape(A, 2, B, 2, C)
        return np.einsum("pqrs,asbrc->aqbpc", M, psi5).reshape(-1)


def exact_statevector(N, circuit):
    """runner.py:615-617."""
    return ExactSimulator(N).simulate(circuit)


def compute_fidelity(x, alpha, exact_state):
    """fidelity.py:48-65 -- |<psi_exact|phi>|^2 over the stored keys."""
    ov = np.dot(exact_state[np.asarray(x).astype(np.int64)].conj(), alpha)
    return float(np.abs(ov) ** 2)


def compute_participation_ratio(alpha):
    """fidelity.py:79-89 -- 1 / sum |a|^4."""
    d = float(np.sum(np.abs(alpha) ** 4))
    return 1.0 / d if d > 1e-60 else float("inf")


def to_statevector(N, x, alpha, U):
    """bass_simulator.py:608-621 -- scatter, then apply U_j on every qubit j."""
    psi = np.zeros(2 ** N, dtype=np.complex128)
    psi[np.asarray(x).astype(np.int64)] = alpha
    for j in range(N):
        psi = np.einsum("ij,kjl->kil", U[j], psi.reshape(-1, 2, 1 << j)).reshape(-1)
    return psi


def fidelity_bass(N, x, alpha, U, exact_sv):
    """runner.py:620-626."""
    if N > 24:
        return 0.0
    return float(np.abs(np.vdot(exact_sv, to_statevector(N, x, alpha, U))) ** 2)
