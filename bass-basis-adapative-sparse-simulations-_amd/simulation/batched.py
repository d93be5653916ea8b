"""BASS over many independent circuits in lock-step on one GPU.

Every circuit follows the BASS driver's rules (bass_simulator.BASS with
truncate_every = 1: truncate iff nnz > k, optimiser checks at the same gate
indices, the same accept rule); the batch advances one gate index ("slot") at a
time and runs each kind of operation for all circuits of the slot as one
launch sequence over a job table (bass_batch_*), so the launch and
synchronisation cost of a slot is shared by the whole batch.  The optimiser
is batched the same way: one RDM launch per pass for every circuit in it, and
one rotation launch per qubit index for every circuit that tries it.

States live in two slabs; circuit c owns entries [off_c, off_c + k_c) of both
and a flag telling which slab holds its state (moves and tentative rotations
write the other one, accepting flips the flag).  Per circuit the device
operations and their order are those of the single-circuit driver, so a
circuit's result does not depend on the batch it ran in.
"""

import math

import numpy as np

from .. import _lib as L
from .. import _ops
from .._device import Buffers, context, to_host_amps, to_host_keys
from ..core.sparse_state import lifted_positions, stored_to_keys
from .bass_simulator import _SWAP_IDX, _is_gate_diagonal, _manual_kron_2x2, eigh_2x2


def _inverse(pos):
    km = np.empty(len(pos), dtype=np.int32)
    km[pos] = np.arange(len(pos), dtype=np.int32)
    return km


class BatchedBASS:
    """specs: list of (num_qubits, k, circuit, optimize_at or None).

    Keywords as bass_simulator.BASS (optimize_every, pr_opt_threshold,
    use_diag_gate, use_fast_eigh, accept_rtol, trace); truncate_every must be 1.
    """

    def __init__(self, specs, device=None, optimize_every=5, pr_opt_threshold=0.90, use_diag_gate=True,
                 use_fast_eigh=True, accept_rtol=1e-10, truncate_every=1, trace=False):
        if truncate_every != 1:
            raise NotImplementedError("the batched driver implements truncate_every = 1")
        self._ctx = context(device)
        self.B = len(specs)
        self.N = np.array([int(s[0]) for s in specs])
        self.k = np.array([int(s[1]) for s in specs], dtype=np.int64)
        self.circuits = [list(s[2]) for s in specs]
        self.optimize_at = [None if s[3] is None else frozenset(int(i) for i in s[3]) for s in specs]
        self.optimize_every = optimize_every
        self.pr_opt_threshold = pr_opt_threshold
        self.use_diag_gate = use_diag_gate
        self.use_fast_eigh = use_fast_eigh
        self.accept_rtol = float(accept_rtol)
        self.traces = [[] for _ in range(self.B)] if trace else None
        self.counters = {"amp_updates": 0}

    # --------------------------------------------------------------- state --
    def _setup(self):
        B = self.B
        self.off = np.concatenate([[0], np.cumsum(self.k)[:-1]]).astype(np.int64)
        total = int(np.sum(self.k))
        self.slabs = [Buffers(total, self._ctx.device), Buffers(total, self._ctx.device)]
        self.tiles = Buffers(1 << 16, self._ctx.device)
        self.cur = np.zeros(B, dtype=np.int64)
        self.nnz = np.ones(B, dtype=np.int64)
        for s in self.slabs:
            s.x[self.off] = 0
            s.a[self.off] = 1.0 + 0.0j
        self.pos = [np.arange(n) for n in self.N]
        self.U = [[np.eye(2, dtype=np.complex128) for _ in range(n)] for n in self.N]
        self.gamma_run = np.ones(B)
        self.trunc_active = np.zeros(B, dtype=bool)
        self.pr_last = np.zeros(B)
        self.sum2 = np.full(B, np.nan)
        self.sum4 = np.full(B, np.nan)
        self.slab_ptr = [(int(s.x.data_ptr()), int(s.a.data_ptr())) for s in self.slabs]

    def _region(self, c, which):
        xp, ap = self.slab_ptr[which]
        return xp + 8 * int(self.off[c]), ap + 16 * int(self.off[c])

    def _jobs(self, cs, out_other=False):
        """Job table for circuits cs: input = current slab, output = the same
        slab (truncations) or the other one (moves, tentative rotations)."""
        cs = np.asarray(cs, dtype=np.int64)
        J = _ops.new_jobs(len(cs))
        for i, c in enumerate(cs):
            J["x"][i], J["a"][i] = self._region(c, int(self.cur[c]))
            J["xo"][i], J["ao"][i] = self._region(c, int(self.cur[c]) ^ (1 if out_other else 0))
        J["n"] = self.nnz[cs]
        J["k"] = self.k[cs]
        J["cap_out"] = self.k[cs]
        J["renorm"] = 1
        return J

    def _count(self, key, v=1):
        self.counters[key] = self.counters.get(key, 0) + int(v)

    # ------------------------------------------------------------- driver --
    def _check_point(self, c, i):
        M = len(self.circuits[c])
        if self.optimize_at[c] is not None:
            return i in self.optimize_at[c] and i + 1 < M
        return self.optimize_every > 0 and (i + 1) % self.optimize_every == 0 and i + 1 < M

    def _apply_trunc_stats(self, i, cs, st):
        """The fast-truncation bookkeeping of BASS._fast_truncate per circuit."""
        for r, c in enumerate(cs):
            n = int(st["n_in"][r])
            self._count("amp_updates", n)
            self.sum2[c], self.sum4[c] = st["sum2"][r], st["sum4"][r]
            if st["flags"][r] & 1:
                self.trunc_active[c] = True
                total, kept = float(st["total"][r]), float(st["kept"][r])
                self.gamma_run[c] *= kept / total if total > 1e-30 else 0
                if self.traces is not None:
                    self.traces[c].append(("trunc", i, kept, total))
                self.nnz[c] = self.k[c]
                self._count("truncate")
            else:
                self.nnz[c] = n

    def _gates_1q(self, i, cs):
        J = self._jobs(cs)
        for r, c in enumerate(cs):
            gate = self.circuits[c][i]
            q = gate.qubits[0]
            G = self.U[c][q].conj().T @ gate.matrix @ self.U[c][q]
            jb = int(self.pos[c][q])
            J["j"][r] = jb
            J["g"][r, :8] = np.ascontiguousarray(G, dtype=np.complex128).view(np.float64).ravel()
            self.pos[c] = lifted_positions(self.pos[c], jb, 0)
            J["key_bits"][r] = self.N[c]
            J["key_map"][r, : self.N[c]] = _inverse(self.pos[c])
        st = _ops.batch_step(self._ctx, J, 1, L.MODE_GATE, self.tiles)
        self._count("gate1q", len(cs))
        self._apply_trunc_stats(i, cs, st)

    def _gates_2q(self, i, cs, mats):
        # bring the second qubit to bit 0 (a move into the other slab) unless
        # it is there, or the first one is (then the qubits swap roles)
        plan = {}
        movers = []
        for c in cs:
            gate = self.circuits[c][i]
            q1, q2 = gate.qubits
            G = mats[c]
            if self.pos[c][q1] == 0 and self.pos[c][q2] != 0:
                G = G[np.ix_(_SWAP_IDX, _SWAP_IDX)]
                q1, q2 = q2, q1
            plan[c] = (q1, q2, G)
            if self.pos[c][q2] != 0:
                movers.append(c)
        if movers:
            J = self._jobs(movers, out_other=True)
            for r, c in enumerate(movers):
                J["j"][r] = int(self.pos[c][plan[c][1]])
            _ops.batch_move(self._ctx, J)
            for r, c in enumerate(movers):
                self.pos[c] = lifted_positions(self.pos[c], int(J["j"][r]), 0)
                self.cur[c] ^= 1
        J = self._jobs(cs)
        for r, c in enumerate(cs):
            q1, q2, G = plan[c]
            jb = int(self.pos[c][q1])
            J["j"][r] = jb
            J["g"][r, :32] = np.ascontiguousarray(G, dtype=np.complex128).view(np.float64).ravel()
            self.pos[c] = lifted_positions(self.pos[c], jb, 1)
            J["key_bits"][r] = self.N[c]
            J["key_map"][r, : self.N[c]] = _inverse(self.pos[c])
        st = _ops.batch_step(self._ctx, J, 2, L.MODE_GATE, self.tiles)
        self._count("gate2q", len(cs))
        self._apply_trunc_stats(i, cs, st)

    def _diag(self, cs, mats):
        J = self._jobs(cs)
        for r, c in enumerate(cs):
            q1, q2 = self._cur_gate[c].qubits
            J["j"][r] = int(self.pos[c][q1])
            J["j2"][r] = int(self.pos[c][q2])
            J["g"][r, :8] = np.diag(mats[c]).astype(np.complex128).view(np.float64)
            self.sum4[c] = np.nan
            self._count("amp_updates", self.nnz[c])
        _ops.batch_diag2q(self._ctx, J)
        self._count("diag2q", len(cs))

    def _ensure_sums(self, cs):
        need = [c for c in cs if np.isnan(self.sum4[c]) or np.isnan(self.sum2[c])]
        if need:
            s2, s4 = _ops.batch_moments(self._ctx, self._jobs(need))
            self.sum2[need], self.sum4[need] = s2, s4

    def _pr(self, c):
        if self.nnz[c] == 0:
            return float("inf")
        s4 = self.sum4[c]
        return 1.0 / s4 if s4 > 1e-60 else float("inf")

    def _trigger(self, c):
        if self.pr_opt_threshold >= 1.0:
            return True
        return self._pr(c) > self.pr_last[c] / self.pr_opt_threshold

    def run(self):
        """Simulate every circuit; returns a list of per-circuit results
        (keys, amplitudes, gamma, U, nnz)."""
        self._setup()
        lens = np.array([len(c) for c in self.circuits])
        for i in range(int(lens.max()) if self.B else 0):
            act = np.flatnonzero(lens > i)
            one, two, diag, mats = [], [], [], {}
            self._cur_gate = {}
            for c in act:
                gate = self.circuits[c][i]
                self._cur_gate[c] = gate
                if gate.n_qubits == 1:
                    one.append(c)
                elif gate.n_qubits == 2:
                    q1, q2 = gate.qubits
                    Ut = _manual_kron_2x2(self.U[c][q1], self.U[c][q2])
                    G = Ut.conj().T @ gate.matrix @ Ut
                    mats[c] = G
                    (diag if self.use_diag_gate and _is_gate_diagonal(G) else two).append(c)
                else:
                    raise ValueError(f"{gate.n_qubits}-qubit gates unsupported")
            if diag:
                self._diag(diag, mats)
            if one:
                self._gates_1q(i, one)
            if two:
                self._gates_2q(i, two, mats)
            chk = [c for c in act if self._check_point(c, i)]
            if chk:
                self._ensure_sums([c for c in chk if self.trunc_active[c]])
                opt = [c for c in chk if self.trunc_active[c] and self._trigger(c)]
                if opt:
                    self._optimize(opt)
                    self._ensure_sums(opt)
                    for c in opt:
                        self.pr_last[c] = self._pr(c)
                self.trunc_active[chk] = False
        return self._results()

    # ------------------------------------------------------------ optimiser --
    def _optimize(self, cs):
        """BASS._optimize_basis (:730-827) for every circuit of cs at once."""
        ctx = self._ctx
        self._ensure_sums(cs)
        cs = [c for c in cs if not (self._pr(c) < self.k[c] / 4)]
        entered = list(cs)
        active = list(cs)
        for _pass in range(3):
            if not active:
                break
            J = self._jobs(active)
            J["key_bits"] = self.N[active]
            rho = _ops.batch_rdm1q(ctx, J)
            self._count("rdm", len(active))
            pr = {c: self._pr(c) for c in active}
            improved = {c: False for c in active}
            Vs = {}
            for r, c in enumerate(active):
                rl = rho[r][self.pos[c]]
                vv = []
                for j in range(self.N[c]):
                    if self.use_fast_eigh:
                        _, V = eigh_2x2(rl[j])
                    else:
                        evals, V = np.linalg.eigh(rl[j])
                        V = V[:, np.argsort(-evals)]
                        if np.linalg.det(V).real < 0:
                            V[:, 1] *= -1
                    vv.append(None if abs(V[0, 1]) + abs(V[1, 0]) < 1e-10 else V)
                Vs[c] = vv
            for j in range(int(max(self.N[c] for c in active))):
                trial = [c for c in active if j < self.N[c] and Vs[c][j] is not None]
                if not trial:
                    continue
                J = self._jobs(trial, out_other=True)
                tpos = {}
                for r, c in enumerate(trial):
                    jb = int(self.pos[c][j])
                    J["j"][r] = jb
                    J["g"][r, :8] = np.ascontiguousarray(Vs[c][j].conj().T).view(np.float64).ravel()
                    tpos[c] = lifted_positions(self.pos[c], jb, 0)
                    J["key_bits"][r] = self.N[c]
                    J["key_map"][r, : self.N[c]] = _inverse(tpos[c])
                st = _ops.batch_step(ctx, J, 1, L.MODE_ROT, self.tiles)
                self._count("rotate", len(trial))
                # inputs of <= k entries were kept whole: renormalise them and
                # take their moments (the candidates sit in the other slab)
                keep = [r for r in range(len(trial)) if not st["flags"][r] & 1]
                s4k = {}
                if keep:
                    sub = _ops.new_jobs(len(keep))
                    for t, r in enumerate(keep):
                        sub[t] = J[r]
                        sub["x"][t], sub["a"][t] = J["xo"][r], J["ao"][r]
                        sub["n"][t] = st["n_in"][r]
                        norm = math.sqrt(float(st["sum2"][r]))
                        sub["scale"][t] = norm if norm > 1e-30 else 1.0
                    _ops.batch_scale(ctx, sub)
                    _, s4 = _ops.batch_moments(ctx, sub)
                    for t, r in enumerate(keep):
                        s4k[r] = float(s4[t])
                for r, c in enumerate(trial):
                    m = int(st["n_in"][r])
                    self._count("amp_updates", m)
                    if st["flags"][r] & 1:
                        nn, frac, s4 = int(self.k[c]), float(st["kept"][r]), float(st["sum4"][r])
                    else:
                        nn, frac, s4 = m, float(st["sum2"][r]), s4k[r]
                    pr_new = 1.0 / s4 if s4 > 1e-60 else float("inf")
                    reject = pr_new >= pr[c] * (1.0 - self.accept_rtol)
                    if self.traces is not None:
                        self.traces[c].append(("rot", j, not reject, pr[c], pr_new, frac))
                    if reject:
                        self._count("reject")
                        continue
                    self.cur[c] ^= 1
                    self.nnz[c] = nn
                    self.pos[c] = tpos[c]
                    self.sum2[c], self.sum4[c] = np.nan, s4
                    self.U[c][j] = self.U[c][j] @ Vs[c][j]
                    pr[c] = pr_new
                    self.gamma_run[c] *= np.sqrt(frac if frac > 1e-30 else 0)
                    improved[c] = True
                    self._count("accept")
            active = [c for c in active if improved[c]]
        # final renormalisation (:551-554; nothing to truncate when nnz <= k)
        if entered:
            self._ensure_sums(entered)
            J = self._jobs(entered)
            scale = []
            for r, c in enumerate(entered):
                norm = math.sqrt(float(self.sum2[c]))
                J["scale"][r] = norm if norm > 1e-30 else 1.0
                scale.append(norm > 1e-30)
                self.sum4[c] = np.nan
            sel = [r for r in range(len(entered)) if scale[r]]
            if sel:
                _ops.batch_scale(ctx, np.ascontiguousarray(J[sel]))

    # -------------------------------------------------------------- results --
    def _results(self):
        out = []
        for c in range(self.B):
            n = int(self.nnz[c])
            buf = self.slabs[int(self.cur[c])]
            o = int(self.off[c])
            w = to_host_keys(buf.x[o: o + n], n)
            out.append({"keys": stored_to_keys(w, self.pos[c]), "amps": to_host_amps(buf.a[o: o + n], n),
                        "nnz": n, "gamma": float(np.sqrt(max(self.gamma_run[c], 0))), "U": self.U[c]})
        return out
