"""Fixed-basis sparse simulator on the device, under the names of the
reference's src/simulation/simulator.py:61-253."""

from time import perf_counter

import numpy as np

from .. import _ops
from .._device import context
from ..core.sparse_state import SparseState
from ..core.truncation import lcg_next


class FixedBasisSimulator:
    """simulator.py:61-253.  2-qubit gates: deduplicated gate + top-k with
    gamma *= sqrt(kept / total) and renormalisation (:161-201); 1-qubit gates:
    SparseState.apply_gate, which truncates to k inside the state (:204-208).
    ``dedup1q`` = "table" or "sort" selects the 1-qubit deduplication path."""

    def __init__(self, num_qubits, k, truncation="top-k", verbose=True, log_every=1, max_nnz_factor=8,
                 device=None, dedup1q="table"):
        self.N = num_qubits
        self.k = k
        self.verbose = verbose
        self.log_every = log_every
        if truncation not in ["top-k", "random-k"]:
            raise ValueError("Unknown truncation")
        self._truncate_name = truncation
        self._ctx = context(device)
        self.dedup1q = dedup1q
        self.counters = {"amp_updates": 0}  # records produced, algorithmic bytes per operation

    def simulate(self, circuit, initial_state=None, seed=None):
        """simulator.py:134-244."""
        self.begin(len(circuit), initial_state, seed)
        for i, gate in enumerate(circuit):
            self.step(i, gate)
        return self.end()

    # The loop body of simulate(), exposed so callers can time it gate by gate.
    def begin(self, num_gates, initial_state=None, seed=None):
        state = SparseState(self.N, self.k, initial_state, device=self._ctx.device, dedup=self.dedup1q)
        state.counters = self.counters = {"amp_updates": 0}
        self.state = state
        self.runtime_per_gate = np.empty(num_gates)
        self.gamma_history = np.empty(num_gates + 1)
        self.nnz_history = np.empty(num_gates + 1)
        self.gamma_history[0] = 1.0
        self.nnz_history[0] = state.nnz
        if seed is None:
            seed = int(np.random.randint(0, 2 ** 31))
        self._cur = seed
        return state

    def step(self, i, gate):
        state, ctx = self.state, self._ctx
        t0 = perf_counter()
        self._cur = lcg_next(self._cur)
        if gate.n_qubits == 2:
            q1, q2 = gate.qubits
            n_in = state.nnz
            n, ht = _ops.gate(ctx, state._buf, state.nnz, gate.matrix, (q1, q2), state._tmp, dedup=True)
            state._buf, state._tmp = state._tmp, state._buf
            state.nnz, state._hist_tiles = n, ht
            state._count_gate("gate2q", n_in, n)
            if state.nnz > self.k:
                if self._truncate_name == "top-k":
                    st = _ops.truncate(ctx, state._buf, state.nnz, self.k, state._tmp, renorm=True,
                                       hist_tiles=state._hist_tiles)
                    state._buf, state._tmp = state._tmp, state._buf
                    state._count_select(state.nnz, self.k)
                    total, kept = st.total, st.kept
                    state.nnz = self.k
                    state.gamma *= np.sqrt(kept / total) if total > 1e-30 else 0
                    state._hist_tiles = 0
                else:
                    raise NotImplementedError(
                        "random-k 2-qubit truncation draws with the legacy np.random.choice; "
                        "not supported on the device")
        else:
            state.apply_gate(gate, self._truncate_name, self._cur)
        self.runtime_per_gate[i] = perf_counter() - t0
        self.gamma_history[i + 1] = state.gamma
        self.nnz_history[i + 1] = state.nnz

    def end(self):
        return self.state

    def get_metrics(self):
        """simulator.py:246-253."""
        return {
            "runtime_per_gate": self.runtime_per_gate,
            "gamma_squared": self.gamma_history ** 2,
            "nnz": self.nnz_history,
            "total_runtime": self.runtime_per_gate.sum(),
            "avg_runtime_per_gate": self.runtime_per_gate.mean(),
        }
