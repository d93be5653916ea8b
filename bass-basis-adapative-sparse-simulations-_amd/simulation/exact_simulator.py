"""Exact state-vector simulation on the device (reference:
src/simulation/exact_simulator.py:66-144).

The dense vector (complex128[2^N], index = basis key) lives in HBM and every
gate is one in-place streaming kernel (bass_dense_gate).  simulate() returns a
host array, as the reference does; simulate_device() keeps the vector on the
device for device-side fidelities.
"""

import numpy as np

from .. import _ops
from .._device import context


class ExactSimulator:
    """exact_simulator.py:66-144.  max_qubits: the reference refuses N > 24
    (host memory); the device vector can be larger when asked for."""

    def __init__(self, num_qubits, verbose=True, device=None, max_qubits=24):
        if num_qubits > max_qubits:
            raise ValueError(f"N={num_qubits} too large for exact simulation")
        self.N = num_qubits
        self.verbose = verbose
        self._ctx = context(device)

    def simulate_device(self, circuit, initial_state=None):
        psi = _ops.dense_state(self.N, self._ctx.device)
        psi[0 if initial_state is None else int(initial_state)] = 1.0
        for gate in circuit:
            self.apply_gate_exact(psi, gate)
        return psi

    def simulate(self, circuit, initial_state=None):
        """Dense state vector (host complex128[2^N])."""
        return self.simulate_device(circuit, initial_state).cpu().numpy()

    def apply_gate_exact(self, psi, gate):
        """In place on the device vector psi (the reference returns a new array)."""
        if gate.n_qubits not in (1, 2):
            raise NotImplementedError()
        _ops.dense_gate(self._ctx, psi, self.N, np.asarray(gate.matrix, dtype=np.complex128), gate.qubits)
        return psi


def exact_statevector(N, circuit, device=None):
    """src/experimentation/runner.py:615-617 (host array)."""
    return ExactSimulator(N, verbose=False, device=device).simulate(circuit)


def exact_statevector_device(N, circuit, device=None, max_qubits=24):
    return ExactSimulator(N, verbose=False, device=device, max_qubits=max_qubits).simulate_device(circuit)
