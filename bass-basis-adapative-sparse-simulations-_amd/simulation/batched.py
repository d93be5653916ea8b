"""BASS over many independent circuits in lock-step on one GPU.

Every circuit follows the BASS driver's rules (bass_simulator.BASS with
truncate_every = 1: truncate iff nnz > k, optimiser checks at the same gate
indices, the same accept rule); the batch advances one gate index ("slot") at a
time and runs each kind of operation for all circuits of the slot as one
launch sequence over a job table (bass_batch_*), so the launch and
synchronisation cost of a slot is shared by the whole batch.  The optimiser
is batched the same way: one RDM launch per pass for every circuit in it, and
one rotation launch per qubit index for every circuit that tries it.  The host
bookkeeping (frames, qubit positions, key maps, decisions) is vectorised over
the circuits of a slot.

States live in two slabs; circuit c owns entries [off_c, off_c + k_c) of both
and a flag telling which slab holds its state (moves and tentative rotations
write the other one, accepting flips the flag).  Per circuit the device
operations and their order are those of the single-circuit driver.
"""

import numpy as np

from .. import _lib as L
from .. import _ops
from .._device import Buffers, context, to_host_amps, to_host_keys
from ..core.sparse_state import stored_to_keys
from .bass_simulator import _SWAP_IDX

NMAX = 64
_BAD = 1000  # position of qubits a circuit does not have: never touched by a lift


def _lift_rows(P, jb, t):
    """lifted_positions for every row of P (row r lifts bit jb[r] to bit t)."""
    J = jb[:, None]
    out = P + ((P >= t) & (P < J))
    return np.where(P == J, t, out)


def _key_maps(P):
    """Stored-bit -> qubit maps (int32[m, 64]) of position rows P."""
    km = np.zeros((len(P), NMAX), dtype=np.int32)
    r, q = np.nonzero(P < NMAX)
    km[r, P[r, q]] = q
    return km


def _dagger(U):
    return np.conj(np.swapaxes(U, -1, -2))


def _kron2(A, B):
    """A (x) B for stacks of 2x2 matrices, index (2i+k, 2j+l) = A[i,j] B[k,l].
    Products in real arithmetic, (ac - bd, ad + bc) without fused operations:
    the rounding of the scalar products in bass_simulator._manual_kron_2x2."""
    a = A[:, :, None, :, None]
    b = B[:, None, :, None, :]
    re = a.real * b.real - a.imag * b.imag
    im = a.real * b.imag + a.imag * b.real
    out = np.empty(re.shape, dtype=np.complex128)
    out.real, out.imag = re, im
    return out.reshape(len(A), 4, 4)


def _eigh2x2_rows(rho):
    """eigh_2x2 (bass_simulator.py:387-437) for a stack of 2x2 Hermitian
    matrices: eigenvectors V (same gauge) and the 'already diagonal' mask."""
    a = rho[..., 0, 0].real
    d = rho[..., 1, 1].real
    b = rho[..., 0, 1]
    b_sq = b.real ** 2 + b.imag ** 2
    tau = (d - a) * 0.5
    disc = np.sqrt(tau * tau + b_sq)
    td_p = tau + disc
    td_m = tau - disc
    small = b_sq < 1e-28
    with np.errstate(divide="ignore", invalid="ignore"):
        n0 = np.sqrt(b_sq + td_p * td_p)
        n1 = np.sqrt(b_sq + td_m * td_m)
        # divide re and im separately: the rounding of a Python complex / float
        V = np.empty(rho.shape, dtype=np.complex128)
        V[..., 0, 0].real, V[..., 0, 0].imag = b.real / n0, b.imag / n0
        V[..., 1, 0] = td_p / n0
        V[..., 0, 1].real, V[..., 0, 1].imag = b.real / n1, b.imag / n1
        V[..., 1, 1] = td_m / n1
    ident = small & (a >= d)
    swap = small & ~(a >= d)
    V[ident] = np.eye(2)
    V[swap] = np.array([[0.0, 1.0], [1.0, 0.0]])
    trivial = (np.abs(V[..., 0, 1]) + np.abs(V[..., 1, 0])) < 1e-10
    return V, trivial


class BatchedBASS:
    """specs: list of (num_qubits, k, circuit, optimize_at or None).

    Keywords as bass_simulator.BASS (optimize_every, pr_opt_threshold,
    use_diag_gate, accept_rtol, trace); truncate_every must be 1.
    """

    def __init__(self, specs, device=None, optimize_every=5, pr_opt_threshold=0.90, use_diag_gate=True,
                 accept_rtol=1e-10, truncate_every=1, trace=False):
        if truncate_every != 1:
            raise NotImplementedError("the batched driver implements truncate_every = 1")
        self._ctx = context(device)
        self.B = len(specs)
        self.N = np.array([int(s[0]) for s in specs], dtype=np.int64)
        if self.B and self.N.max() > 63:
            raise ValueError("at most 63 qubits")
        self.k = np.array([int(s[1]) for s in specs], dtype=np.int64)
        self.circuits = [list(s[2]) for s in specs]
        self.optimize_at = [None if s[3] is None else frozenset(int(i) for i in s[3]) for s in specs]
        self.optimize_every = optimize_every
        self.pr_opt_threshold = pr_opt_threshold
        self.use_diag_gate = use_diag_gate
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
        for s in self.slabs:
            s.x[self.off] = 0
            s.a[self.off] = 1.0 + 0.0j
        self.xbase = np.array([s.x.data_ptr() for s in self.slabs], dtype=np.uint64)
        self.abase = np.array([s.a.data_ptr() for s in self.slabs], dtype=np.uint64)
        self.cur = np.zeros(B, dtype=np.int64)
        self.nnz = np.ones(B, dtype=np.int64)
        q = np.arange(NMAX)
        self.pos = np.where(q[None, :] < self.N[:, None], q[None, :], _BAD + q[None, :])
        self.U = np.zeros((B, NMAX, 2, 2), dtype=np.complex128)
        self.U[:, :, 0, 0] = 1.0
        self.U[:, :, 1, 1] = 1.0
        self.gamma_run = np.ones(B)
        self.trunc_active = np.zeros(B, dtype=bool)
        self.pr_last = np.zeros(B)
        self.sum2 = np.full(B, np.nan)
        self.sum4 = np.full(B, np.nan)
        self.lens = np.array([len(c) for c in self.circuits], dtype=np.int64)
        # check-point slots per circuit
        self.chk = [set() for _ in range(B)]
        for c in range(B):
            M = int(self.lens[c])
            if self.optimize_at[c] is not None:
                self.chk[c] = {i for i in self.optimize_at[c] if i + 1 < M}
            elif self.optimize_every > 0:
                self.chk[c] = set(range(self.optimize_every - 1, M - 1, self.optimize_every))

    def _jobs(self, cs, out_other=False):
        """Job table for circuits cs: input = current slab, output = the same
        slab (truncations) or the other one (moves, tentative rotations)."""
        J = _ops.new_jobs(len(cs))
        cur = self.cur[cs]
        oth = cur ^ 1 if out_other else cur
        off = self.off[cs].astype(np.uint64)
        J["x"] = self.xbase[cur] + np.uint64(8) * off
        J["a"] = self.abase[cur] + np.uint64(16) * off
        J["xo"] = self.xbase[oth] + np.uint64(8) * off
        J["ao"] = self.abase[oth] + np.uint64(16) * off
        J["n"] = self.nnz[cs]
        J["k"] = self.k[cs]
        J["cap_out"] = self.k[cs]
        J["renorm"] = 1
        J["key_bits"] = self.N[cs]
        return J

    def _count(self, key, v=1):
        self.counters[key] = self.counters.get(key, 0) + int(v)

    def _apply_trunc_stats(self, i, cs, st):
        """The fast-truncation bookkeeping of BASS._fast_truncate, per circuit."""
        n_in = st["n_in"]
        tr = (st["flags"] & 1) != 0
        self._count("amp_updates", n_in.sum())
        self.sum2[cs], self.sum4[cs] = st["sum2"], st["sum4"]
        if tr.any():
            t = cs[tr]
            total, kept = st["total"][tr], st["kept"][tr]
            self.trunc_active[t] = True
            self.gamma_run[t] *= np.where(total > 1e-30, kept / np.where(total > 1e-30, total, 1.0), 0.0)
            if self.traces is not None:
                for c, kp, to in zip(t, kept, total):
                    self.traces[c].append(("trunc", i, float(kp), float(to)))
            self._count("truncate", tr.sum())
        self.nnz[cs] = np.where(tr, self.k[cs], n_in)

    # ------------------------------------------------------------- gates ----
    def _gates_1q(self, i, cs, gates):
        q = np.fromiter((g.qubits[0] for g in gates), dtype=np.int64, count=len(gates))
        gm = np.stack([g.matrix for g in gates])
        U = self.U[cs, q]
        G = _dagger(U) @ gm @ U
        jb = self.pos[cs, q]
        self.pos[cs] = _lift_rows(self.pos[cs], jb, 0)
        J = self._jobs(cs)
        J["j"] = jb
        J["g"][:, :8] = np.ascontiguousarray(G).reshape(len(cs), 4).view(np.float64)
        J["key_map"] = _key_maps(self.pos[cs])
        st = _ops.batch_step(self._ctx, J, 1, L.MODE_GATE, self.tiles)
        self._count("gate1q", len(cs))
        self._apply_trunc_stats(i, cs, st)

    def _gates_2q(self, i, cs, gates):
        q1 = np.fromiter((g.qubits[0] for g in gates), dtype=np.int64, count=len(gates))
        q2 = np.fromiter((g.qubits[1] for g in gates), dtype=np.int64, count=len(gates))
        gm = np.stack([g.matrix for g in gates])
        Ut = _kron2(self.U[cs, q1], self.U[cs, q2])
        G = _dagger(Ut) @ gm @ Ut
        # in the frame the gate may be diagonal: in-place phases
        off = G.copy()
        off[:, np.arange(4), np.arange(4)] = 0
        diag = self.use_diag_gate & (np.abs(off).max(axis=(1, 2)) < 1e-12)
        if diag.any():
            d = cs[diag]
            J = self._jobs(d)
            J["j"] = self.pos[d, q1[diag]]
            J["j2"] = self.pos[d, q2[diag]]
            J["g"][:, :8] = np.ascontiguousarray(np.diagonal(G[diag], axis1=1, axis2=2)).view(np.float64)
            _ops.batch_diag2q(self._ctx, J)
            self.sum4[d] = np.nan
            self._count("amp_updates", self.nnz[d].sum())
            self._count("diag2q", len(d))
        gen = ~diag
        if not gen.any():
            return
        cs, q1, q2, G = cs[gen], q1[gen], q2[gen], G[gen]
        # the second qubit goes to bit 0 (a move) unless it is there, or the
        # first one is (then the qubits swap roles)
        sw = (self.pos[cs, q1] == 0) & (self.pos[cs, q2] != 0)
        if sw.any():
            G[sw] = G[sw][:, _SWAP_IDX][:, :, _SWAP_IDX]
            q1[sw], q2[sw] = q2[sw], q1[sw].copy()
        mv = self.pos[cs, q2] != 0
        if mv.any():
            m = cs[mv]
            J = self._jobs(m, out_other=True)
            jb = self.pos[m, q2[mv]]
            J["j"] = jb
            _ops.batch_move(self._ctx, J)
            self.pos[m] = _lift_rows(self.pos[m], jb, 0)
            self.cur[m] ^= 1
        jb = self.pos[cs, q1]
        self.pos[cs] = _lift_rows(self.pos[cs], jb, 1)
        J = self._jobs(cs)
        J["j"] = jb
        J["g"][:, :32] = np.ascontiguousarray(G).reshape(len(cs), 16).view(np.float64)
        J["key_map"] = _key_maps(self.pos[cs])
        st = _ops.batch_step(self._ctx, J, 2, L.MODE_GATE, self.tiles)
        self._count("gate2q", len(cs))
        self._apply_trunc_stats(i, cs, st)

    def _ensure_sums(self, cs):
        need = cs[np.isnan(self.sum4[cs]) | np.isnan(self.sum2[cs])]
        if len(need):
            s2, s4 = _ops.batch_moments(self._ctx, self._jobs(need))
            self.sum2[need], self.sum4[need] = s2, s4

    def _pr(self, cs):
        s4 = self.sum4[cs]
        pr = np.where(s4 > 1e-60, 1.0 / np.where(s4 > 1e-60, s4, 1.0), np.inf)
        return np.where(self.nnz[cs] == 0, np.inf, pr)

    def run(self):
        """Simulate every circuit; returns per-circuit results (keys, amps,
        nnz, gamma, U)."""
        self._setup()
        for i in range(int(self.lens.max()) if self.B else 0):
            act = np.flatnonzero(self.lens > i)
            gates = [self.circuits[c][i] for c in act]
            nq = np.fromiter((g.n_qubits for g in gates), dtype=np.int64, count=len(gates))
            if np.any((nq != 1) & (nq != 2)):
                raise ValueError("only 1- and 2-qubit gates are supported")
            one = np.flatnonzero(nq == 1)
            two = np.flatnonzero(nq == 2)
            if len(two):
                self._gates_2q(i, act[two], [gates[r] for r in two])
            if len(one):
                self._gates_1q(i, act[one], [gates[r] for r in one])
            chk = np.array([c for c in act if i in self.chk[c]], dtype=np.int64)
            if len(chk):
                ta = chk[self.trunc_active[chk]]
                if len(ta):
                    self._ensure_sums(ta)
                    pr = self._pr(ta)
                    trig = np.ones(len(ta), bool) if self.pr_opt_threshold >= 1.0 else \
                        pr > self.pr_last[ta] / self.pr_opt_threshold
                    opt = ta[trig]
                    if len(opt):
                        self._optimize(opt)
                        self._ensure_sums(opt)
                        self.pr_last[opt] = self._pr(opt)
                self.trunc_active[chk] = False
        return self._results()

    # ------------------------------------------------------------ optimiser --
    def _optimize(self, cs):
        """BASS._optimize_basis (:730-827) for every circuit of cs at once."""
        ctx = self._ctx
        self._ensure_sums(cs)
        cs = cs[~(self._pr(cs) < self.k[cs] / 4)]
        entered = cs.copy()
        active = cs
        for _pass in range(3):
            if not len(active):
                break
            rho_st = _ops.batch_rdm1q(ctx, self._jobs(active))
            self._count("rdm", len(active))
            P = np.minimum(self.pos[active], NMAX - 1)
            rho = np.take_along_axis(rho_st, P[:, :, None, None], axis=1)
            V, trivial = _eigh2x2_rows(rho)
            trivial |= np.arange(NMAX)[None, :] >= self.N[active][:, None]
            pr = self._pr(active)
            improved = np.zeros(len(active), dtype=bool)
            for j in range(int(self.N[active].max())):
                rows = np.flatnonzero(~trivial[:, j])
                if not len(rows):
                    continue
                trial = active[rows]
                Vj = V[rows, j]
                jb = self.pos[trial, j]
                tpos = _lift_rows(self.pos[trial], jb, 0)
                J = self._jobs(trial, out_other=True)
                J["j"] = jb
                J["g"][:, :8] = np.ascontiguousarray(_dagger(Vj)).reshape(len(rows), 4).view(np.float64)
                J["key_map"] = _key_maps(tpos)
                st = _ops.batch_step(ctx, J, 1, L.MODE_ROT, self.tiles)
                self._count("rotate", len(rows))
                m = st["n_in"]
                tr = (st["flags"] & 1) != 0
                self._count("amp_updates", m.sum())
                frac = np.where(tr, st["kept"], st["sum2"])
                s4 = st["sum4"].copy()
                keep = np.flatnonzero(~tr)
                if len(keep):
                    # kept whole: renormalise the candidates, then their moments
                    sub = np.ascontiguousarray(J[keep])
                    sub["x"], sub["a"] = sub["xo"], sub["ao"]
                    sub["n"] = m[keep]
                    norm = np.sqrt(st["sum2"][keep])
                    sub["scale"] = np.where(norm > 1e-30, norm, 1.0)
                    _ops.batch_scale(ctx, sub)
                    _, s4k = _ops.batch_moments(ctx, sub)
                    s4[keep] = s4k
                pr_new = np.where(s4 > 1e-60, 1.0 / np.where(s4 > 1e-60, s4, 1.0), np.inf)
                reject = pr_new >= pr[rows] * (1.0 - self.accept_rtol)
                if self.traces is not None:
                    for r, c in enumerate(trial):
                        self.traces[c].append(("rot", j, bool(not reject[r]), float(pr[rows][r]),
                                               float(pr_new[r]), float(frac[r])))
                acc = np.flatnonzero(~reject)
                self._count("reject", int(reject.sum()))
                self._count("accept", len(acc))
                if not len(acc):
                    continue
                a = trial[acc]
                self.cur[a] ^= 1
                self.nnz[a] = np.where(tr[acc], self.k[a], m[acc])
                self.pos[a] = tpos[acc]
                self.sum2[a], self.sum4[a] = np.nan, s4[acc]
                self.U[a, j] = self.U[a, j] @ Vj[acc]
                pr[rows[acc]] = pr_new[acc]
                f = frac[acc]
                self.gamma_run[a] *= np.sqrt(np.where(f > 1e-30, f, 0.0))
                improved[rows[acc]] = True
            active = active[improved]
        # final renormalisation (:551-554; nothing to truncate when nnz <= k)
        if len(entered):
            self._ensure_sums(entered)
            norm = np.sqrt(self.sum2[entered])
            do = norm > 1e-30
            if do.any():
                J = self._jobs(entered[do])
                J["scale"] = norm[do]
                _ops.batch_scale(ctx, J)
            self.sum4[entered] = np.nan

    # -------------------------------------------------------------- results --
    def _results(self):
        out = []
        for c in range(self.B):
            n = int(self.nnz[c])
            buf = self.slabs[int(self.cur[c])]
            o = int(self.off[c])
            w = to_host_keys(buf.x[o: o + n], n)
            pos = self.pos[c, : self.N[c]]
            out.append({"keys": stored_to_keys(w, pos), "amps": to_host_amps(buf.a[o: o + n], n), "nnz": n,
                        "gamma": float(np.sqrt(max(self.gamma_run[c], 0))),
                        "U": [self.U[c, j].copy() for j in range(self.N[c])]})
        return out
