"""Circuit generators.

``generate_tfim_circuit`` and ``generate_rfim_circuit`` follow
src/utils/random_circuits.py:100-165 and :212-331 of the reference (same
random draws, same gate order).  The brickwork-Haar and kicked-Ising
workloads have no generator in the reference; they are built here from the
reference's gate classes.
"""

import numpy as np

from ..core.gates import CZGate, HaarSingleQubitGate, RandomTwoQubitGate, RXGate, RZGate, RZZGate

SEED_BASE = 20260520


def generate_tfim_circuit(num_qubits, num_layers, J=1.0, h=3.0, dt=0.3, rng=None,
                          disorder_strength_J=0.0, disorder_strength_h=0.0):
    """random_circuits.py:100-165: per layer draw J_i (N-1 values) then h_i (N
    values) from rng.uniform, append RZZ(i, i+1, J_i dt) on the open chain, then
    RX(i, h_i dt) on every qubit."""
    if rng is None:
        rng = np.random.default_rng(0)
    circuit = []
    for _ in range(num_layers):
        J_vals = rng.uniform(J - disorder_strength_J, J + disorder_strength_J, size=num_qubits - 1)
        h_vals = rng.uniform(h - disorder_strength_h, h + disorder_strength_h, size=num_qubits)
        for i in range(num_qubits - 1):
            circuit.append(RZZGate(i, i + 1, J_vals[i] * dt))
        for i in range(num_qubits):
            circuit.append(RXGate(i, h_vals[i] * dt))
    return circuit


def generate_rfim_circuit(num_qubits, num_layers, J=1.0, h0=1.0, W=2.0, dt=0.2, rng=None):
    """random_circuits.py:212-331: Delta_i ~ U[-W, W] once per circuit; per
    layer RZZ(J dt) on the open chain, RX(h0 dt), RZ(-2 Delta_i dt)."""
    if rng is None:
        rng = np.random.default_rng(0)
    Delta = rng.uniform(-W, W, size=num_qubits)
    circuit = []
    for _ in range(num_layers):
        for i in range(num_qubits - 1):
            circuit.append(RZZGate(i, i + 1, J * dt))
        for i in range(num_qubits):
            circuit.append(RXGate(i, h0 * dt))
        for i in range(num_qubits):
            circuit.append(RZGate(i, -2.0 * float(Delta[i]) * dt))
    return circuit


def brickwork_haar_circuit(num_qubits, num_layers, rng):
    """Per layer: a Haar-random U(2) on every qubit, then CZ on the even bonds
    (0,1),(2,3),... in even layers and the odd bonds (1,2),(3,4),... in odd
    layers (open chain).  Returns (circuit, layer_end) with layer_end[l] the
    index of the last gate of layer l."""
    circuit = []
    ends = []
    for layer in range(num_layers):
        for q in range(num_qubits):
            circuit.append(HaarSingleQubitGate(q, rng))
        for i in range(layer % 2, num_qubits - 1, 2):
            circuit.append(CZGate(i, i + 1))
        ends.append(len(circuit) - 1)
    return circuit, ends


def kicked_ising_circuit(num_qubits, num_steps, zz_angle=np.pi / 4, theta0=0.3, dtheta=0.01):
    """Periodic kicked Ising chain: per step RZZ(zz_angle) on every bond
    (i, i+1 mod N) then RX(theta0 + dtheta * step) on every qubit.  Returns
    (circuit, step_end)."""
    circuit = []
    ends = []
    for s in range(num_steps):
        for i in range(num_qubits):
            circuit.append(RZZGate(i, (i + 1) % num_qubits, zz_angle))
        th = theta0 + dtheta * s
        for i in range(num_qubits):
            circuit.append(RXGate(i, th))
        ends.append(len(circuit) - 1)
    return circuit, ends


def tfim_config_circuit(num_qubits=20, steps=40, J=1.0, h=1.0, dt=0.05):
    """Uniform open-chain TFIM Trotter circuit (reference gate conventions):
    per step N-1 RZZ(J dt) then N RX(h dt)."""
    return generate_tfim_circuit(num_qubits, steps, J=J, h=h, dt=dt, rng=np.random.default_rng(0))


def generate_random_circuit(num_qubits, num_layers, gate_type="haar"):
    """random_circuits.py:60-97: per layer, a random pairing of the qubits
    (numpy's legacy global shuffle) with a Haar-random 2-qubit gate on each of
    the num_qubits // 2 pairs (drawn from the same global generator)."""
    if gate_type != "haar":
        raise NotImplementedError(f"Gate type {gate_type} not implemented")
    circuit = []
    for _ in range(num_layers):
        order = list(range(num_qubits))
        np.random.shuffle(order)
        for p in range(num_qubits // 2):
            circuit.append(RandomTwoQubitGate(order[2 * p], order[2 * p + 1]))
    return circuit


def generate_rfitm_circuit(num_qubits, num_layers, J=1.0, h=1.0, dt=0.1, sigma_h=0.1, rng=None):
    """random_circuits.py:168-209: random-field transverse Ising Trotter circuit.
    Site fields h_j ~ N(h, sigma_h^2), one draw per qubit per circuit from rng
    (default_rng(0) when None); per layer RZZ(J dt) on the open chain, then
    RX(h_j dt) on every qubit."""
    if rng is None:
        rng = np.random.default_rng(0)
    fields = rng.normal(loc=h, scale=sigma_h, size=num_qubits)
    circuit = []
    for _ in range(num_layers):
        circuit.extend(RZZGate(q, q + 1, J * dt) for q in range(num_qubits - 1))
        circuit.extend(RXGate(q, float(fields[q]) * dt) for q in range(num_qubits))
    return circuit


def estimate_circuit_depth(circuit, num_qubits):
    """random_circuits.py:352-364: longest chain of gates sharing qubits (each
    gate starts after the latest gate on any of its qubits)."""
    t = [0] * num_qubits
    for g in circuit:
        start = max(t[q] for q in g.qubits)
        for q in g.qubits:
            t[q] = start + 1
    return max(t) if t else 0


def circuit_info(circuit):
    """random_circuits.py:334-349: gate count, qubit count (highest qubit + 1)
    and depth."""
    used = set()
    for g in circuit:
        used.update(g.qubits)
    n = max(used) + 1 if used else 0
    return {"num_gates": len(circuit), "num_qubits": n, "depth": estimate_circuit_depth(circuit, n)}
