"""Metrics on device states (reference: src/benchmarking/fidelity.py:48-89)."""

import numpy as np
import torch

from .. import _ops
from .._device import to_host_amps


def compute_fidelity(sparse_state, exact_state):
    """fidelity.py:48-65: |<psi_exact | phi>|^2 over the stored entries (the
    stored amplitudes are already normalised).  Keys of a key-ordered state
    are mapped back from the stored words."""
    xs = sparse_state.keys().astype(np.int64)
    al = to_host_amps(sparse_state._buf.a, sparse_state.nnz)
    return float(np.abs(np.dot(np.asarray(exact_state)[xs].conj(), al)) ** 2)


def compute_fidelity_bounds(sparse_state):
    """fidelity.py:68-76."""
    return sparse_state.gamma ** 2, 1.0


def compute_participation_ratio(sparse_state):
    """fidelity.py:79-89: 1 / sum |alpha|^4 (device reduction)."""
    _, s4 = _ops.moments(sparse_state._ctx, sparse_state._buf, sparse_state.nnz)
    return 1.0 / s4 if s4 > 1e-60 else float("inf")


def fidelity_bass(sim, state, exact_sv):
    """src/experimentation/runner.py:620-626: |<psi_exact | (U_0 x ... x
    U_{N-1}) phi>|^2, reconstruction and overlap on the device.  exact_sv: a
    device tensor (kept there) or a host array (uploaded)."""
    if sim.N > 24:
        return 0.0
    psi = sim.to_statevector_device(state)
    ex = exact_sv if isinstance(exact_sv, torch.Tensor) else torch.from_numpy(
        np.ascontiguousarray(exact_sv, dtype=np.complex128))
    ex = ex.to(device=psi.device, dtype=torch.complex128)
    return float(abs(_ops.dense_vdot(sim._ctx, ex, psi)) ** 2)


def estimate_cross_entropy_fidelity(sparse_state, exact_state, num_samples=1000):
    """fidelity.py:92-102: cross-entropy benchmark XEB = 2^N <P_exact(x)> - 1
    over num_samples keys x sampled from the sparse state (its sample())."""
    keys = np.asarray(sparse_state.sample(num_samples), dtype=np.uint64).astype(np.int64)
    p = np.abs(np.asarray(exact_state)[keys]) ** 2
    return float(2 ** sparse_state.N * np.mean(p) - 1)
