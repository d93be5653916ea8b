"""Gate definitions (host side), same classes and matrices as the reference's
src/core/gates.py:6-164.

Conventions kept from the reference: qubits are stored unsorted; the 2-qubit
matrix index is (b_q1 << 1) | b_q2; RXGate(t) = exp(+i t X),
RZZGate(t) = exp(+i t Z(x)Z), RZGate(t) = exp(-i t Z / 2), RYGate(t) = exp(-i t Y / 2).
"""

import numpy as np


class QuantumGate:
    """gates.py:6-24."""

    def __init__(self, qubits, matrix):
        self.qubits = tuple(qubits)
        self.matrix = np.array(matrix, dtype=np.complex128)
        self.n_qubits = len(qubits)
        dim = 2 ** self.n_qubits
        assert self.matrix.shape == (dim, dim), "Matrix dimension mismatch"

    def __repr__(self):
        return f"{self.__class__.__name__}(qubits={self.qubits})"


class SingleQubitGate(QuantumGate):
    def __init__(self, qubit, matrix):
        super().__init__((qubit,), matrix)


class TwoQubitGate(QuantumGate):
    def __init__(self, qubit1, qubit2, matrix):
        super().__init__((qubit1, qubit2), matrix)


class HGate(SingleQubitGate):
    def __init__(self, qubit):
        super().__init__(qubit, np.array([[1, 1], [1, -1]]) / np.sqrt(2))


class XGate(SingleQubitGate):
    def __init__(self, qubit):
        super().__init__(qubit, np.array([[0, 1], [1, 0]]))


class YGate(SingleQubitGate):
    def __init__(self, qubit):
        super().__init__(qubit, np.array([[0, -1j], [1j, 0]]))


class ZGate(SingleQubitGate):
    def __init__(self, qubit):
        super().__init__(qubit, np.array([[1, 0], [0, -1]]))


class SGate(SingleQubitGate):
    def __init__(self, qubit):
        super().__init__(qubit, np.array([[1, 0], [0, 1j]]))


class TGate(SingleQubitGate):
    def __init__(self, qubit):
        super().__init__(qubit, np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]]))


class CNOTGate(TwoQubitGate):
    def __init__(self, control, target):
        super().__init__(control, target,
                         np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]))


class CZGate(TwoQubitGate):
    def __init__(self, qubit1, qubit2):
        super().__init__(qubit1, qubit2, np.diag([1, 1, 1, -1]))


class RXGate(SingleQubitGate):
    """exp(i theta X) (gates.py:106-113)."""

    def __init__(self, qubit, theta):
        c, s = np.cos(theta), np.sin(theta)
        super().__init__(qubit, np.array([[c, 1j * s], [1j * s, c]], dtype=np.complex128))


class RZZGate(TwoQubitGate):
    """exp(i theta Z(x)Z) (gates.py:115-121)."""

    def __init__(self, qubit1, qubit2, theta):
        p, m = np.exp(1j * theta), np.exp(-1j * theta)
        super().__init__(qubit1, qubit2, np.diag([p, m, m, p]).astype(np.complex128))


class RZGate(SingleQubitGate):
    """exp(-i theta Z / 2) (gates.py:124-130)."""

    def __init__(self, qubit, theta):
        p, m = np.exp(-1j * theta / 2), np.exp(1j * theta / 2)
        super().__init__(qubit, np.array([[p, 0], [0, m]], dtype=np.complex128))


class RYGate(SingleQubitGate):
    """exp(-i theta Y / 2) (gates.py:133-139)."""

    def __init__(self, qubit, theta):
        c, s = np.cos(theta / 2), np.sin(theta / 2)
        super().__init__(qubit, np.array([[c, -s], [s, c]], dtype=np.complex128))


def haar_unitary_2x2(rng):
    """Haar-random U(2) from a numpy Generator: QR of a complex Ginibre matrix
    with the phases of R's diagonal moved into Q (Mezzadri 2007)."""
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))


class HaarSingleQubitGate(SingleQubitGate):
    """Seeded Haar-random U(2) on one qubit (the reference draws Haar matrices
    with scipy's unitary_group, gates.py:142-164; this draws from a numpy
    Generator so circuits are reproducible without scipy)."""

    def __init__(self, qubit, rng):
        super().__init__(qubit, haar_unitary_2x2(rng))


def haar_random_unitary(dimension, seed=None):
    """gates.py:152-164: Haar-random unitary of the given dimension from scipy's
    unitary_group on numpy's legacy global generator, which a seed reseeds
    first (as the reference does, so seeded draws match it)."""
    from scipy.stats import unitary_group

    if seed is not None:
        np.random.seed(seed)
    return unitary_group.rvs(dimension)


class RandomTwoQubitGate(TwoQubitGate):
    """gates.py:142-149: Haar-random 2-qubit gate, haar_random_unitary(4, seed)."""

    def __init__(self, qubit1, qubit2, seed=None):
        super().__init__(qubit1, qubit2, haar_random_unitary(4, seed))
