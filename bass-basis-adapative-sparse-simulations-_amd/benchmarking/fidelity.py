"""Metrics on device states (reference: src/benchmarking/fidelity.py:48-89)."""

import numpy as np

from .. import _ops
from .._device import to_host_amps, to_host_keys


def compute_fidelity(sparse_state, exact_state):
    """fidelity.py:48-65: |<psi_exact | phi>|^2 over the stored keys (stored
    amplitudes are already normalised)."""
    xs = to_host_keys(sparse_state._buf.x, sparse_state.nnz).astype(np.int64)
    al = to_host_amps(sparse_state._buf.a, sparse_state.nnz)
    return float(np.abs(np.dot(exact_state[xs].conj(), al)) ** 2)


def compute_fidelity_bounds(sparse_state):
    """fidelity.py:68-76."""
    return sparse_state.gamma ** 2, 1.0


def compute_participation_ratio(sparse_state):
    """fidelity.py:79-89: 1 / sum |alpha|^4 (device reduction)."""
    _, s4 = _ops.moments(sparse_state._ctx, sparse_state._buf, sparse_state.nnz)
    return 1.0 / s4 if s4 > 1e-60 else float("inf")


def fidelity_bass(sim, state, exact_sv):
    """src/experimentation/runner.py:620-626."""
    if sim.N > 24:
        return 0.0
    return float(np.abs(np.vdot(exact_sv, sim.to_statevector(state))) ** 2)
