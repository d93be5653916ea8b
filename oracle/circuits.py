"""Gate matrices and circuit generators restated from the reference (oracle).

Matrices follow src/core/gates.py (note the reference conventions
RX(t) = exp(+i t X), RZZ(t) = exp(+i t Z(x)Z), RZ(t) = exp(-i t Z / 2)).
Generators follow src/utils/random_circuits.py and src/experimentation/runner.py
so that the reference-held fixtures (seeded with numpy PCG64) can be replayed.
"""

import numpy as np

from .drivers import _G


def RX(q, theta):
    """gates.py:106-113."""
    c, s = np.cos(theta), np.sin(theta)
    return _G((q,), np.array([[c, 1j * s], [1j * s, c]], dtype=np.complex128))


def RZZ(q1, q2, theta):
    """gates.py:115-121."""
    p, m = np.exp(1j * theta), np.exp(-1j * theta)
    return _G((q1, q2), np.diag([p, m, m, p]).astype(np.complex128))


def RZ(q, theta):
    """gates.py:124-130."""
    p, m = np.exp(-1j * theta / 2), np.exp(1j * theta / 2)
    return _G((q,), np.array([[p, 0], [0, m]], dtype=np.complex128))


def RY(q, theta):
    """gates.py:133-139."""
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return _G((q,), np.array([[c, -s], [s, c]], dtype=np.complex128))


def H(q):
    """gates.py:42-47."""
    return _G((q,), np.array([[1, 1], [1, -1]]) / np.sqrt(2))


def CZ(q1, q2):
    """gates.py:98-103."""
    return _G((q1, q2), np.diag([1, 1, 1, -1]))


def CNOT(c, t):
    """gates.py:90-95."""
    return _G((c, t), np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]))


def tfim(N, layers, J=1.0, h=3.0, dt=0.3, rng=None, dJ=0.0, dh=0.0):
    """random_circuits.py:100-165 -- per layer draw J_i (N-1) then h_i (N), then
    N-1 RZZ(i, i+1, J_i dt) and N RX(i, h_i dt)."""
    if rng is None:
        rng = np.random.default_rng(0)
    c = []
    for _ in range(layers):
        Jv = rng.uniform(J - dJ, J + dJ, size=N - 1)
        hv = rng.uniform(h - dh, h + dh, size=N)
        for i in range(N - 1):
            c.append(RZZ(i, i + 1, Jv[i] * dt))
        for i in range(N):
            c.append(RX(i, hv[i] * dt))
    return c


def rfim(N, layers, J=1.0, h0=1.0, W=2.0, dt=0.2, rng=None):
    """random_circuits.py:212-331 -- Delta_i ~ U[-W, W] once; per layer RZZ(J dt)
    on the open chain, RX(h0 dt), RZ(-2 Delta_i dt)."""
    if rng is None:
        rng = np.random.default_rng(0)
    D = rng.uniform(-W, W, size=N)
    c = []
    for _ in range(layers):
        for i in range(N - 1):
            c.append(RZZ(i, i + 1, J * dt))
        for i in range(N):
            c.append(RX(i, h0 * dt))
        for i in range(N):
            c.append(RZ(i, -2.0 * float(D[i]) * dt))
    return c


def random_3regular_graph(N, rng, max_tries=500):
    """runner.py:466-497 -- pairing model; edges returned in Python set order."""
    for _ in range(max_tries):
        stubs = list(range(N)) * 3
        rng.shuffle(stubs)
        edges = set()
        ok = True
        for k in range(0, len(stubs), 2):
            u, v = stubs[k], stubs[k + 1]
            if u == v or (min(u, v), max(u, v)) in edges:
                ok = False
                break
            edges.add((min(u, v), max(u, v)))
        if ok:
            return list(edges)
    raise RuntimeError("no 3-regular graph")


def qaoa(N, rounds, rng):
    """runner.py:500-609 -- H on all qubits, then per round RZZ(gamma/2) on every
    edge and RX(-beta) on every qubit; Fourier angles plus U(-0.05, 0.05)."""
    p = rounds
    edges = random_3regular_graph(N, rng)
    A = np.pi / 4.0
    gammas = [A * np.sin((2 * k - 1) * np.pi / (4 * p)) + rng.uniform(-0.05, 0.05)
              for k in range(1, p + 1)]
    betas = [A * np.cos((2 * k - 1) * np.pi / (4 * p)) + rng.uniform(-0.05, 0.05)
             for k in range(1, p + 1)]
    c = [H(q) for q in range(N)]
    for k in range(p):
        for i, j in edges:
            c.append(RZZ(i, j, gammas[k] / 2.0))
        for q in range(N):
            c.append(RX(q, -betas[k]))
    return c
