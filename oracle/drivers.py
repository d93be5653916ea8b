"""numpy restatement of the reference drivers (oracle; test infrastructure).

``SparseState``  -- src/core/sparse_state.py:193-415
``BASS``         -- src/simulation/bass_simulator.py:478-929
``FixedBasisSimulator`` -- src/simulation/simulator.py:61-253

Oracle extension: every top-k selection goes through ``select_topk``, whose
default is the reference's ``np.argpartition(probs, -k)[-k:]``.  A ``selector``
callable ``(keys, probs, k) -> indices`` may replace it (for instance the
"largest |a|^2, ties to the smaller key" rule).  ``BASS(accept_rtol=r)``
accepts a basis rotation only when it lowers PR by more than r relative
(default 0: the reference's ``not (pr_new >= pr)``); r > 0 makes rotations that
leave PR unchanged up to rounding -- relabelings -- a deterministic reject
instead of a summation-order coin flip.  Nothing else differs.
"""

import numpy as np

from . import kernels as K


def argpartition_selector(keys, alpha, probs, k):
    """The reference rule (src/simulation/bass_simulator.py:909)."""
    return np.argpartition(probs, -k)[-k:]


def key_tiebreak_selector(keys, alpha, probs, k):
    """The rule the device implements: largest |a|^2 first, |a|^2 computed as
    re*re + im*im; equal values ordered by ascending key.  Returns the selected
    indices in input order."""
    p = alpha.real * alpha.real + alpha.imag * alpha.imag
    order = np.lexsort((keys, -p))
    return np.sort(order[:k])


class SparseState:
    """src/core/sparse_state.py:193-227 (SoA keys/amplitudes, nnz, gamma)."""

    RADIX_THRESHOLD = 2000

    def __init__(self, num_qubits, k, init_state=None, selector=None):
        self.N = num_qubits
        self.k = k
        self.x = np.zeros(k, dtype=np.uint64)
        self.alpha = np.zeros(k, dtype=np.complex128)
        self.gamma = 1.0
        if init_state is None:
            init_state = 0
        self.x[0] = np.uint64(init_state)
        self.alpha[0] = 1.0 + 0.0j
        self.nnz = 1
        self._temp_x = np.zeros(4 * k, dtype=np.uint64)
        self._temp_alpha = np.zeros(4 * k, dtype=np.complex128)
        self._temp_nnz = 0
        self._selector = selector

    def copy(self):
        """src/core/sparse_state.py:232-238."""
        s = SparseState(self.N, self.k, selector=self._selector)
        s.x[: self.nnz] = self.x[: self.nnz]
        s.alpha[: self.nnz] = self.alpha[: self.nnz]
        s.nnz = self.nnz
        s.gamma = self.gamma
        return s

    def to_dense(self):
        """src/core/sparse_state.py:240-255."""
        if self.N > 20:
            raise ValueError("Dense conversion only allowed for N <= 20")
        dense = np.zeros(2 ** self.N, dtype=np.complex128)
        dense[self.x[: self.nnz].astype(np.int64)] = self.alpha[: self.nnz]
        return dense

    def _ensure_temp(self, need):
        if need > self._temp_x.shape[0]:
            self._temp_x = np.zeros(need, dtype=np.uint64)
            self._temp_alpha = np.zeros(need, dtype=np.complex128)

    def _sort_merge_copy(self):
        n = self._temp_nnz
        # np.argsort below RADIX_THRESHOLD, radix sort above (:276-281); both
        # leave the keys ascending, and duplicate pairs sum commutatively.
        K.radix_sort_uint64(self._temp_x, self._temp_alpha, n)
        self._temp_nnz = K.merge_sorted_kernel(self._temp_x, self._temp_alpha, n)
        if self._temp_nnz > self.x.shape[0]:
            new_cap = max(self._temp_nnz, 2 * self.x.shape[0])
            self.x = np.zeros(new_cap, dtype=np.uint64)
            self.alpha = np.zeros(new_cap, dtype=np.complex128)
        self.x[: self._temp_nnz] = self._temp_x[: self._temp_nnz]
        self.alpha[: self._temp_nnz] = self._temp_alpha[: self._temp_nnz]
        self.nnz = self._temp_nnz

    def apply_single_qubit_gate(self, gate, method, seed):
        """src/core/sparse_state.py:268-301: emit -> sort -> merge -> copy -> truncate."""
        q = gate.qubits[0]
        self._ensure_temp(2 * self.nnz)
        self._temp_nnz = K.apply_1qubit_gate_kernel_opt(
            self.x, self.alpha, self.nnz, self._temp_x, self._temp_alpha, gate.matrix, q)
        self._sort_merge_copy()
        if self.nnz > self.k:
            if method == "top-k":
                self.truncate_top_k()
            else:
                self.truncate_random_k(seed)

    def apply_two_qubit_gate(self, gate, method, seed):
        """src/core/sparse_state.py:303-377."""
        q1, q2 = gate.qubits
        self._ensure_temp(4 * self.nnz)
        self._temp_nnz = K.apply_2qubit_gate_kernel_opt(
            self.x, self.alpha, self.nnz, self._temp_x, self._temp_alpha, gate.matrix, q1, q2)
        if self._temp_nnz == -1:
            raise RuntimeError("Temporary buffer overflow")
        self._sort_merge_copy()
        if self.nnz > self.k:
            if method == "top-k":
                self.truncate_top_k()
            else:
                self.truncate_random_k(seed)

    def apply_gate(self, gate, method="top-k", seed="None"):
        """src/core/sparse_state.py:379-385."""
        if gate.n_qubits == 1:
            self.apply_single_qubit_gate(gate, method, seed)
        elif gate.n_qubits == 2:
            self.apply_two_qubit_gate(gate, method, seed)
        else:
            raise NotImplementedError("Only 1- and 2-qubit gates supported")

    def truncate_top_k(self):
        """src/core/sparse_state.py:388-400."""
        if self.nnz <= self.k:
            return
        self.nnz, gsq = K.top_k_truncate_kernel_opt(
            self.x, self.alpha, self.nnz, self.k, selector=self._selector)
        gl = np.sqrt(gsq)
        self.alpha[: self.nnz] /= gl
        self.gamma *= gl

    def truncate_random_k(self, seed):
        """src/core/sparse_state.py:402-415."""
        if self.nnz <= self.k:
            return
        if seed is None:
            seed = int(np.random.randint(0, 2 ** 31))
        self.nnz, gsq = K.random_k_truncate_kernel_opt(self.x, self.alpha, self.nnz, self.k, seed)
        gl = np.sqrt(gsq)
        self.alpha[: self.nnz] /= gl
        self.gamma *= gl


class _G:
    """Minimal gate record (qubits, matrix, n_qubits)."""

    def __init__(self, qubits, matrix):
        self.qubits = tuple(qubits)
        self.matrix = np.array(matrix, dtype=np.complex128)
        self.n_qubits = len(self.qubits)


class BASS:
    """src/simulation/bass_simulator.py:478-929 (1-qubit coordinate descent; the
    2-qubit brick-wall pass of :679-728, :820-824 is included for completeness)."""

    def __init__(self, num_qubits, k, optimize_every=5, buffer_factor=2, truncate_every=1,
                 max_nnz_factor=8, use_2qubit_rotations=False, verbose=False,
                 use_fast_eigh=True, use_diag_gate=True, use_kron_cache=True,
                 pr_opt_threshold=0.90, selector=None, optimize_at=None, trace=None, accept_rtol=0.0):
        self.N = num_qubits
        self.k = k
        self.optimize_every = optimize_every
        self.buffer_factor = buffer_factor
        self.verbose = verbose
        self.use_fast_eigh = use_fast_eigh
        self.use_diag_gate = use_diag_gate
        self.use_kron_cache = use_kron_cache
        self._kron_cache = {}
        self.U = None
        self._gamma_running = 1.0
        self.truncate_every = truncate_every
        self.max_nnz_factor = max_nnz_factor
        self.use_2qubit_rotations = use_2qubit_rotations
        self.pr_opt_threshold = pr_opt_threshold
        self._pr_at_last_opt = 0.0
        self.selector = selector or argpartition_selector
        # oracle extension: explicit gate indices (0-based, "after gate i") at
        # which the optimiser check runs, instead of (i+1) % optimize_every == 0
        self.optimize_at = None if optimize_at is None else set(int(i) for i in optimize_at)
        self.trace = trace  # optional list: decision log
        self.accept_rtol = float(accept_rtol)

    # -- helpers (:551-593) -------------------------------------------------
    def _renorm(self, state):
        norm = np.sqrt(np.sum(np.abs(state.alpha[: state.nnz]) ** 2))
        if norm > 1e-30:
            state.alpha[: state.nnz] /= norm

    def _select(self, state, probs):
        return self.selector(state.x[: state.nnz], state.alpha[: state.nnz], probs, self.k)

    def _truncate_state(self, state):
        """:565-572."""
        if state.nnz <= self.k:
            return
        probs = np.abs(state.alpha[: state.nnz]) ** 2
        top_idx = self._select(state, probs)
        state.x[: self.k] = state.x[top_idx].copy()
        state.alpha[: self.k] = state.alpha[top_idx].copy()
        state.nnz = self.k

    def _pr(self, state):
        """:574-578."""
        if state.nnz == 0:
            return float("inf")
        denom = float(np.sum(np.abs(state.alpha[: state.nnz]) ** 4))
        return 1.0 / denom if denom > 1e-60 else float("inf")

    def _get_kron(self, q1, q2):
        if not self.use_kron_cache:
            return K.manual_kron_2x2(self.U[q1], self.U[q2])
        key = (q1, q2)
        if key not in self._kron_cache:
            self._kron_cache[key] = K.manual_kron_2x2(self.U[q1], self.U[q2])
        return self._kron_cache[key]

    def _invalidate_kron(self, qubit):
        for kk in [kk for kk in self._kron_cache if qubit in kk]:
            del self._kron_cache[kk]

    def _rotate_qubit(self, state, qubit, U):
        """:595-606."""
        needed = 2 * state.nnz
        if needed > len(state.x):
            ox = state.x[: state.nnz].copy()
            oa = state.alpha[: state.nnz].copy()
            state.x = np.empty(needed, dtype=np.uint64)
            state.alpha = np.empty(needed, dtype=np.complex128)
            state.x[: state.nnz] = ox
            state.alpha[: state.nnz] = oa
        state.nnz = K.apply_1qubit_rotation_sparse(state.x, state.alpha, state.nnz, qubit, U)

    def to_statevector(self, state):
        """:608-621."""
        if self.N > 24:
            raise ValueError(f"N={self.N} too large for statevector")
        psi = np.zeros(2 ** self.N, dtype=np.complex128)
        psi[state.x[: state.nnz].astype(np.int64)] = state.alpha[: state.nnz]
        for j in range(self.N):
            psi3d = psi.reshape(-1, 2, 1 << j)
            psi3d = np.einsum("ij,kjl->kil", self.U[j], psi3d)
            psi = psi3d.reshape(-1)
        return psi

    def _apply_2q_rotation(self, state, q1, q2, V4):
        needed = 4 * state.nnz
        if needed > len(state.x):
            nx = np.empty(needed, dtype=np.uint64)
            na = np.empty(needed, dtype=np.complex128)
            nx[: state.nnz] = state.x[: state.nnz]
            na[: state.nnz] = state.alpha[: state.nnz]
            state.x, state.alpha = nx, na
        state.nnz = K.apply_2qubit_rotation_sparse(state.x, state.alpha, state.nnz, q1, q2, V4)

    def _apply_2q_gate_ht(self, state, q1, q2, G):
        """:644-675."""
        needed = 4 * state.nnz
        if needed > len(state._temp_x):
            state._temp_x = np.empty(needed, dtype=np.uint64)
            state._temp_alpha = np.empty(needed, dtype=np.complex128)
        state._temp_nnz = K.apply_2qubit_gate_ht(state.x, state.alpha, state.nnz,
                                                 state._temp_x, state._temp_alpha, G, q1, q2)
        if state._temp_nnz > state.x.shape[0]:
            new_cap = max(state._temp_nnz, 2 * state.x.shape[0])
            state.x = np.empty(new_cap, dtype=np.uint64)
            state.alpha = np.empty(new_cap, dtype=np.complex128)
        state.x[: state._temp_nnz] = state._temp_x[: state._temp_nnz]
        state.alpha[: state._temp_nnz] = state._temp_alpha[: state._temp_nnz]
        state.nnz = state._temp_nnz

    def _optimize_2q_pass(self, state, pairs):
        """:679-728."""
        for q1, q2 in pairs:
            rho = K.compute_2qubit_rdm(state.x, state.alpha, state.nnz, q1, q2)
            evals, evecs = np.linalg.eigh(rho)
            order = np.argsort(-evals)
            V = evecs[:, order]
            V_rot = V.conj().T
            x_save = state.x[: state.nnz].copy()
            a_save = state.alpha[: state.nnz].copy()
            nnz_save = state.nnz
            pr_before = self._pr(state)
            self._apply_2q_rotation(state, q1, q2, V_rot)
            self._truncate_state(state)
            frac1 = float(np.sum(np.abs(state.alpha[: state.nnz]) ** 2))
            self._renorm(state)
            self._apply_2q_rotation(state, q1, q2, V)
            self._truncate_state(state)
            frac2 = float(np.sum(np.abs(state.alpha[: state.nnz]) ** 2))
            self._renorm(state)
            pr_after = self._pr(state)
            # accept_rtol = 0: the reference's "revert iff pr_after > pr_before"
            accept = not (pr_after > pr_before * (1.0 - self.accept_rtol))
            if self.trace is not None:
                self.trace.append(("rot2", (q1, q2), bool(accept), pr_before, pr_after, frac1 * frac2))
            if not accept:
                state.x[:nnz_save] = x_save
                state.alpha[:nnz_save] = a_save
                state.nnz = nnz_save
            else:
                self._gamma_running *= np.sqrt(frac1 * frac2)

    def _optimize_basis(self, state):
        """:730-827 -- coordinate descent over single-qubit frames."""
        if self._pr(state) < self.k / 4:
            return
        for _pass in range(3):
            all_rho = K.compute_all_1qubit_rdms(state.x, state.alpha, state.nnz, self.N)
            pr = self._pr(state)
            improved = False
            for j in range(self.N):
                rho = all_rho[j]
                if self.use_fast_eigh:
                    _, V = K.eigh_2x2(rho)
                else:
                    evals, V = np.linalg.eigh(rho)
                    V = V[:, np.argsort(-evals)]
                    if np.linalg.det(V).real < 0:
                        V[:, 1] *= -1
                if abs(V[0, 1]) + abs(V[1, 0]) < 1e-10:
                    continue
                x_save = state.x[: state.nnz].copy()
                a_save = state.alpha[: state.nnz].copy()
                nnz_save = state.nnz
                U_save = self.U[j].copy()
                self._rotate_qubit(state, j, V.conj().T)
                self._truncate_state(state)
                frac = float(np.sum(np.abs(state.alpha[: state.nnz]) ** 2))
                self._renorm(state)
                pr_new = self._pr(state)
                accept = not (pr_new >= pr * (1.0 - self.accept_rtol))
                if self.trace is not None:
                    self.trace.append(("rot", j, bool(accept), pr, pr_new, frac))
                if not accept:
                    state.x[:nnz_save] = x_save
                    state.alpha[:nnz_save] = a_save
                    state.nnz = nnz_save
                    self.U[j] = U_save
                else:
                    self.U[j] = self.U[j] @ V
                    pr = pr_new
                    self._gamma_running *= np.sqrt(frac if frac > 1e-30 else 0)
                    self._invalidate_kron(j)
                    improved = True
            if not improved:
                break
        if self.use_2qubit_rotations and self.N >= 2:
            self._optimize_2q_pass(state, [(i, i + 1) for i in range(0, self.N - 1, 2)])
            self._optimize_2q_pass(state, [(i, i + 1) for i in range(1, self.N - 1, 2)])
        self._truncate_state(state)
        self._renorm(state)

    def optimize_trigger(self, state):
        """:829-845."""
        if self.pr_opt_threshold >= 1.0:
            return True
        return self._pr(state) > self._pr_at_last_opt / self.pr_opt_threshold

    def _check_point(self, i, M):
        if self.optimize_at is not None:
            return i in self.optimize_at and i + 1 < M
        return self.optimize_every > 0 and (i + 1) % self.optimize_every == 0 and i + 1 < M

    def simulate(self, circuit, initial_state=0):
        """:847-929."""
        self.U = [np.eye(2, dtype=np.complex128) for _ in range(self.N)]
        self._kron_cache = {}
        self._pr_at_last_opt = 0.0
        hard_cap = min(self.max_nnz_factor * self.k, 2 ** self.N)
        internal_k = min(max(max(self.buffer_factor, 2) * self.k, hard_cap), 2 ** self.N)
        state = SparseState(self.N, internal_k, initial_state, selector=self.selector)
        M = len(circuit)
        self._gamma_running = 1.0
        truncation_active = False
        gates_since_trunc = 0
        for i, gate in enumerate(circuit):
            if gate.n_qubits == 1:
                q = gate.qubits[0]
                G = self.U[q].conj().T @ gate.matrix @ self.U[q]
                state.apply_gate(_G((q,), G), "top-k")
                gates_since_trunc += 1
            elif gate.n_qubits == 2:
                q1, q2 = gate.qubits
                Ut = self._get_kron(q1, q2)
                G = Ut.conj().T @ gate.matrix @ Ut
                if self.use_diag_gate and K.is_gate_diagonal(G):
                    K.apply_2qubit_diagonal_kernel(state.x, state.alpha, state.nnz, q1, q2,
                                                   np.diag(G).astype(np.complex128))
                    if self._check_point(i, M):
                        if truncation_active and self.optimize_trigger(state):
                            self._optimize_basis(state)
                            self._pr_at_last_opt = self._pr(state)
                        truncation_active = False
                    continue
                self._apply_2q_gate_ht(state, q1, q2, G)
                gates_since_trunc += 1
            else:
                raise ValueError(f"{gate.n_qubits}-qubit gates unsupported")

            over_hard = state.nnz > hard_cap
            over_budget = state.nnz > self.k
            if over_hard or (over_budget and gates_since_trunc >= self.truncate_every):
                truncation_active = True
                probs = np.abs(state.alpha[: state.nnz]) ** 2
                total = np.sum(probs)
                top_idx = self._select(state, probs)
                kept = np.sum(probs[top_idx])
                self._gamma_running *= kept / total if total > 1e-30 else 0
                state.x[: self.k] = state.x[top_idx].copy()
                state.alpha[: self.k] = state.alpha[top_idx].copy()
                state.nnz = self.k
                self._renorm(state)
                gates_since_trunc = 0
                if self.trace is not None:
                    self.trace.append(("trunc", i, float(kept), float(total)))

            if self._check_point(i, M):
                if truncation_active and self.optimize_trigger(state):
                    self._optimize_basis(state)
                    self._pr_at_last_opt = self._pr(state)
                truncation_active = False

        state.gamma = np.sqrt(max(self._gamma_running, 0))
        return state


class FixedBasisSimulator:
    """src/simulation/simulator.py:61-253 (top-k branch; random-k keeps the
    reference's legacy np.random.choice draw)."""

    def __init__(self, num_qubits, k, truncation="top-k", verbose=False, log_every=1,
                 max_nnz_factor=8, selector=None):
        self.N = num_qubits
        self.k = k
        if truncation not in ["top-k", "random-k"]:
            raise ValueError("Unknown truncation")
        self._truncate_name = truncation
        self.selector = selector or argpartition_selector

    def simulate(self, circuit, initial_state=None, seed=None):
        """:134-244 (no per-gate timing/logging)."""
        state = SparseState(self.N, self.k, initial_state, selector=(
            None if self.selector is argpartition_selector else self.selector))
        L = len(circuit)
        gamma_hist = np.empty(L + 1)
        nnz_hist = np.empty(L + 1)
        gamma_hist[0] = 1.0
        nnz_hist[0] = state.nnz
        if seed is None:
            seed = int(np.random.randint(0, 2 ** 31))
        cur = seed
        for i, gate in enumerate(circuit):
            cur = K.lcg_next(cur)
            if gate.n_qubits == 2:
                q1, q2 = gate.qubits
                need = 4 * state.nnz
                tx = np.empty(need, dtype=np.uint64)
                ta = np.empty(need, dtype=np.complex128)
                n = K.apply_2qubit_gate_ht(state.x, state.alpha, state.nnz, tx, ta, gate.matrix, q1, q2)
                if n > state.x.shape[0]:
                    new_cap = max(n, 2 * state.x.shape[0])
                    state.x = np.empty(new_cap, dtype=np.uint64)
                    state.alpha = np.empty(new_cap, dtype=np.complex128)
                state.x[:n] = tx[:n]
                state.alpha[:n] = ta[:n]
                state.nnz = n
                if state.nnz > self.k:
                    probs = np.abs(state.alpha[: state.nnz]) ** 2
                    total = np.sum(probs)
                    if self._truncate_name == "top-k":
                        top_idx = self.selector(state.x[: state.nnz], state.alpha[: state.nnz], probs,
                                                self.k)
                    else:
                        np.random.seed(cur)
                        pn = probs / total if total > 1e-30 else np.ones(state.nnz) / state.nnz
                        top_idx = np.random.choice(state.nnz, size=self.k, replace=False, p=pn)
                    kept = np.sum(probs[top_idx])
                    state.x[: self.k] = state.x[top_idx].copy()
                    state.alpha[: self.k] = state.alpha[top_idx].copy()
                    state.nnz = self.k
                    state.gamma *= np.sqrt(kept / total) if total > 1e-30 else 0
                    norm = np.sqrt(np.sum(np.abs(state.alpha[: state.nnz]) ** 2))
                    if norm > 1e-30:
                        state.alpha[: state.nnz] /= norm
            else:
                state.apply_gate(gate, self._truncate_name, cur)
            gamma_hist[i + 1] = state.gamma
            nnz_hist[i + 1] = state.nnz
        self.gamma_history = gamma_hist
        self.nnz_history = nnz_hist
        return state
