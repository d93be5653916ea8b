"""BASS over many independent circuits at once on one GPU.

Every circuit follows the BASS driver's rules (bass_simulator.BASS with
truncate_every = 1: truncate iff nnz > k, optimiser checks at the same gate
indices, the same accept rule).  The engine advances every running circuit by
one gate per tick and runs each kind of operation for all of them as one
launch sequence over a job table (bass_batch_*), so the launch and
synchronisation cost of a tick is shared by the whole batch.  The optimiser is
batched the same way: one RDM launch per pass for every circuit in it, and one
rotation launch per qubit index for every circuit that tries it.  The host
bookkeeping (frames, qubit positions, key maps, decisions) is vectorised over
the circuits of a tick.

Circuits keep their own gate counters, so a circuit can join between ticks
(continuous batching, BatchEngine.admit) and leaves as soon as its last gate
is done.  States live in two pooled slabs; circuit c owns entries
[off_c, off_c + k_c) of both and a flag telling which slab holds its state
(moves and tentative rotations write the other one, accepting flips the flag).
Per circuit the device operations and their order are those of the
single-circuit driver, so results do not depend on what else is in the batch.
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
    b_sq = b.real * b.real + b.imag * b.imag
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


class BatchEngine:
    """Continuous batch of BASS circuits on one GPU.

    Up to max_circuits circuits run at once; each keeps its own gate counter,
    and one tick() applies the next gate of every running circuit (and the
    optimiser of those at a check point) with one launch sequence per
    operation kind.  admit() places a circuit in a free slot and a free region
    of the two pooled state slabs (first fit, slab_entries entries each);
    tick() returns the circuits that finished, whose regions are released.

    Keywords as bass_simulator.BASS (optimize_every, pr_opt_threshold,
    use_diag_gate, accept_rtol); truncate_every = 1 (truncate iff nnz > k).
    """

    def __init__(self, max_circuits, slab_entries, device=None, optimize_every=5, pr_opt_threshold=0.90,
                 use_diag_gate=True, accept_rtol=0.0, trace=False):
        self._ctx = context(device)
        C = int(max_circuits)
        self.C = C
        self.optimize_every = optimize_every
        self.pr_opt_threshold = pr_opt_threshold
        self.use_diag_gate = use_diag_gate
        self.accept_rtol = float(accept_rtol)
        self.trace = trace
        self.counters = {"amp_updates": 0}
        self.slab_entries = int(slab_entries)
        self.slabs = [Buffers(self.slab_entries, self._ctx.device), Buffers(self.slab_entries, self._ctx.device)]
        self.tiles = Buffers(1 << 16, self._ctx.device)
        self.xbase = np.array([s.x.data_ptr() for s in self.slabs], dtype=np.uint64)
        self.abase = np.array([s.a.data_ptr() for s in self.slabs], dtype=np.uint64)
        self.free = [(0, self.slab_entries)]  # free regions (start, length), sorted
        self.alive = np.zeros(C, dtype=bool)
        self.N = np.ones(C, dtype=np.int64)
        self.k = np.ones(C, dtype=np.int64)
        self.off = np.zeros(C, dtype=np.int64)
        self.cur = np.zeros(C, dtype=np.int64)
        self.nnz = np.ones(C, dtype=np.int64)
        self.pos = np.zeros((C, NMAX), dtype=np.int64)
        self.U = np.zeros((C, NMAX, 2, 2), dtype=np.complex128)
        self.gamma_run = np.ones(C)
        self.trunc_active = np.zeros(C, dtype=bool)
        self.pr_last = np.zeros(C)
        self.sum2 = np.full(C, np.nan)
        self.sum4 = np.full(C, np.nan)
        self.gidx = np.zeros(C, dtype=np.int64)
        self.lens = np.zeros(C, dtype=np.int64)
        self.circuits = [None] * C
        self.chk = [set() for _ in range(C)]
        self.tags = [None] * C
        self.traces = [[] for _ in range(C)] if trace else None
        # optimiser state machine (phase 1): stage 0 = RDM due (pass start),
        # 1 = rotation trials along the pass's chain, 2 = final renormalisation
        self.phase = np.zeros(C, dtype=np.int64)
        self.ostage = np.zeros(C, dtype=np.int64)
        self.opass = np.zeros(C, dtype=np.int64)
        self.oentered = np.zeros(C, dtype=bool)
        self.oimproved = np.zeros(C, dtype=bool)
        self.opr = np.zeros(C)
        self.ochain = np.zeros((C, NMAX), dtype=np.int64)
        self.oclen = np.zeros(C, dtype=np.int64)
        self.optr = np.zeros(C, dtype=np.int64)
        self.oV = np.zeros((C, NMAX, 2, 2), dtype=np.complex128)

    @property
    def active(self):
        return int(self.alive.sum())

    def _alloc(self, n):
        for r, (st, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(r)
                else:
                    self.free[r] = (st + n, ln - n)
                return st
        return None

    def _release(self, st, n):
        self.free.append((st, n))
        self.free.sort()
        merged = []
        for a, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((a, ln))
        self.free = merged

    def admit(self, num_qubits, k, circuit, optimize_at=None, tag=None):
        """Start a circuit from |0...0>; False if no slot or slab region is free."""
        if not (1 <= num_qubits <= 63):
            raise ValueError("1 to 63 qubits")
        slots = np.flatnonzero(~self.alive)
        if not len(slots):
            return False
        st = self._alloc(int(k))
        if st is None:
            return False
        c = int(slots[0])
        self._last_admitted = c
        M = len(circuit)
        self.alive[c] = True
        self.N[c], self.k[c], self.off[c] = num_qubits, k, st
        self.cur[c], self.nnz[c], self.gidx[c], self.lens[c] = 0, 1, 0, M
        q = np.arange(NMAX)
        self.pos[c] = np.where(q < num_qubits, q, _BAD + q)
        self.U[c] = 0
        self.U[c, :, 0, 0] = 1.0
        self.U[c, :, 1, 1] = 1.0
        self.gamma_run[c], self.trunc_active[c], self.pr_last[c] = 1.0, False, 0.0
        self.phase[c] = 0
        self.sum2[c] = self.sum4[c] = np.nan
        self.circuits[c] = list(circuit)
        if optimize_at is not None:
            self.chk[c] = {int(i) for i in optimize_at if int(i) + 1 < M}
        elif self.optimize_every > 0:
            self.chk[c] = set(range(self.optimize_every - 1, M - 1, self.optimize_every))
        else:
            self.chk[c] = set()
        self.tags[c] = tag
        if self.traces is not None:
            self.traces[c] = []
        self.slabs[0].x[st] = 0
        self.slabs[0].a[st] = 1.0 + 0.0j
        return True

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

    def _apply_trunc_stats(self, ii, cs, st):
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
                for c, g, kp, to in zip(t, ii[tr], kept, total):
                    self.traces[c].append(("trunc", int(g), float(kp), float(to)))
            self._count("truncate", tr.sum())
        self.nnz[cs] = np.where(tr, self.k[cs], n_in)

    # ------------------------------------------------------------- gates ----
    def _gates_1q(self, ii, cs, gates):
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
        self._apply_trunc_stats(ii, cs, st)

    def _gates_2q(self, ii, cs, gates):
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
        cs, q1, q2, G, ii = cs[gen], q1[gen], q2[gen], G[gen], ii[gen]
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
        self._apply_trunc_stats(ii, cs, st)

    def _ensure_sums(self, cs):
        need = cs[np.isnan(self.sum4[cs]) | np.isnan(self.sum2[cs])]
        if len(need):
            s2, s4 = _ops.batch_moments(self._ctx, self._jobs(need))
            self.sum2[need], self.sum4[need] = s2, s4

    def _pr(self, cs):
        s4 = self.sum4[cs]
        pr = np.where(s4 > 1e-60, 1.0 / np.where(s4 > 1e-60, s4, 1.0), np.inf)
        return np.where(self.nnz[cs] == 0, np.inf, pr)

    def tick(self):
        """One operation of every running circuit: its next gate (with the
        check-point test after it), or its next optimiser step (an RDM at a
        pass start, one rotation trial, or the final renormalisation).
        Returns [(tag, result)] of the circuits that finished."""
        g = np.flatnonzero(self.alive & (self.phase == 0) & (self.gidx < self.lens))
        if len(g):
            self._gate_step(g)
        if (self.alive & (self.phase == 1)).any():
            self._opt_step()
        done = np.flatnonzero(self.alive & (self.phase == 0) & (self.gidx >= self.lens))
        out = []
        for c in done:
            out.append((self.tags[c], self._result(c)))
            self.alive[c] = False
            self._release(int(self.off[c]), int(self.k[c]))
            self.circuits[c] = None
        return out

    def _gate_step(self, act):
        ii = self.gidx[act]
        gates = [self.circuits[c][i] for c, i in zip(act, ii)]
        nq = np.fromiter((g.n_qubits for g in gates), dtype=np.int64, count=len(gates))
        if np.any((nq != 1) & (nq != 2)):
            raise ValueError("only 1- and 2-qubit gates are supported")
        one = np.flatnonzero(nq == 1)
        two = np.flatnonzero(nq == 2)
        if len(two):
            self._gates_2q(ii[two], act[two], [gates[r] for r in two])
        if len(one):
            self._gates_1q(ii[one], act[one], [gates[r] for r in one])
        self.gidx[act] += 1
        # check point after gate i: optimise if truncation happened since the
        # last one and the participation ratio grew past the trigger
        chk = np.array([c for c, i in zip(act, ii) if i in self.chk[c]], dtype=np.int64)
        if not len(chk):
            return
        ta = chk[self.trunc_active[chk]]
        self.trunc_active[chk] = False
        if not len(ta):
            return
        self._ensure_sums(ta)
        pr = self._pr(ta)
        trig = np.ones(len(ta), bool) if self.pr_opt_threshold >= 1.0 else pr > self.pr_last[ta] / self.pr_opt_threshold
        opt = ta[trig]
        if not len(opt):
            return
        # BASS._optimize_basis (:730-827): skipped while PR < k/4
        enter = ~(self._pr(opt) < self.k[opt] / 4)
        e = opt[enter]
        self.phase[e] = 1
        self.ostage[e] = 0
        self.opass[e] = 0
        self.oentered[opt] = enter
        skip = opt[~enter]
        if len(skip):
            self._opt_leave(skip)

    # ------------------------------------------------------------ optimiser --
    def _opt_step(self):
        """Advance every circuit in the optimiser by one step (RDM pass starts
        first, so a circuit starting a pass also makes its first trial)."""
        ph = self.alive & (self.phase == 1)
        s0 = np.flatnonzero(ph & (self.ostage == 0))
        if len(s0):
            self._opt_rdm(s0)
        s1 = np.flatnonzero(ph & (self.ostage == 1) & (self.optr < self.oclen))
        if len(s1):
            self._opt_trial(s1)
        # pass ends: another pass after an improvement (at most 3), else finish
        end = np.flatnonzero(ph & (self.ostage == 1) & (self.optr >= self.oclen))
        if len(end):
            again = self.oimproved[end] & (self.opass[end] < 2)
            a = end[again]
            self.opass[a] += 1
            self.ostage[a] = 0
            self.ostage[end[~again]] = 2
        s2 = np.flatnonzero(ph & (self.ostage == 2))
        if len(s2):
            self._opt_finish(s2)

    def _opt_rdm(self, cs):
        rho_st = _ops.batch_rdm1q(self._ctx, self._jobs(cs))
        self._count("rdm", len(cs))
        P = np.minimum(self.pos[cs], NMAX - 1)
        rho = np.take_along_axis(rho_st, P[:, :, None, None], axis=1)
        V, trivial = _eigh2x2_rows(rho)
        trivial |= np.arange(NMAX)[None, :] >= self.N[cs][:, None]
        self.oV[cs] = V
        # the pass tries the qubits with a non-trivial eigenbasis in index order
        nt = ~trivial
        self.oclen[cs] = nt.sum(axis=1)
        order = np.argsort(trivial, axis=1, kind="stable")
        self.ochain[cs] = order
        self.optr[cs] = 0
        # PR at the pass start from the state's moments, recomputed after an
        # accepted rotation (as the single-circuit driver's _pr does: its sum
        # of |a|^2 is then unknown), so the accept rule compares the same value
        self._ensure_sums(cs)
        self.opr[cs] = self._pr(cs)
        self.oimproved[cs] = False
        self.ostage[cs] = 1

    def _opt_trial(self, trial):
        """One rotation trial per circuit (its next qubit), accept / reject."""
        ctx = self._ctx
        j = self.ochain[trial, self.optr[trial]]
        self.optr[trial] += 1
        Vj = self.oV[trial, j]
        jb = self.pos[trial, j]
        tpos = _lift_rows(self.pos[trial], jb, 0)
        J = self._jobs(trial, out_other=True)
        J["j"] = jb
        J["g"][:, :8] = np.ascontiguousarray(_dagger(Vj)).reshape(len(trial), 4).view(np.float64)
        J["key_map"] = _key_maps(tpos)
        st = _ops.batch_step(ctx, J, 1, L.MODE_ROT, self.tiles)
        self._count("rotate", len(trial))
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
        pr = self.opr[trial]
        reject = pr_new >= pr * (1.0 - self.accept_rtol)
        if self.traces is not None:
            for r, c in enumerate(trial):
                self.traces[c].append(("rot", int(j[r]), bool(not reject[r]), float(pr[r]),
                                       float(pr_new[r]), float(frac[r])))
        acc = np.flatnonzero(~reject)
        self._count("reject", int(reject.sum()))
        self._count("accept", len(acc))
        if not len(acc):
            return
        a = trial[acc]
        ja = j[acc]
        self.cur[a] ^= 1
        self.nnz[a] = np.where(tr[acc], self.k[a], m[acc])
        self.pos[a] = tpos[acc]
        self.sum2[a], self.sum4[a] = np.nan, s4[acc]
        self.U[a, ja] = self.U[a, ja] @ Vj[acc]
        self.opr[a] = pr_new[acc]
        f = frac[acc]
        self.gamma_run[a] *= np.sqrt(np.where(f > 1e-30, f, 0.0))
        self.oimproved[a] = True

    def _opt_finish(self, cs):
        """Final renormalisation (:551-554; nothing to truncate when nnz <= k)."""
        self._ensure_sums(cs)
        norm = np.sqrt(self.sum2[cs])
        do = norm > 1e-30
        if do.any():
            J = self._jobs(cs[do])
            J["scale"] = norm[do]
            _ops.batch_scale(self._ctx, J)
        self.sum4[cs] = np.nan
        self._opt_leave(cs)

    def _opt_leave(self, cs):
        self._ensure_sums(cs)
        self.pr_last[cs] = self._pr(cs)
        self.phase[cs] = 0

    # -------------------------------------------------------------- results --
    def _result(self, c):
        n = int(self.nnz[c])
        buf = self.slabs[int(self.cur[c])]
        o = int(self.off[c])
        w = to_host_keys(buf.x[o: o + n], n)
        pos = self.pos[c, : self.N[c]]
        res = {"keys": stored_to_keys(w, pos), "amps": to_host_amps(buf.a[o: o + n], n), "nnz": n,
               "gamma": float(np.sqrt(max(self.gamma_run[c], 0))),
               "U": [self.U[c, j].copy() for j in range(self.N[c])]}
        if self.traces is not None:
            res["trace"] = self.traces[c]
        return res


class BatchedBASS:
    """A fixed batch: every circuit starts at once (specs: list of
    (num_qubits, k, circuit, optimize_at or None)); run() returns the results
    in spec order.  pipelined: PipelinedBatchEngine (graphs: its HIP graphs)
    or BatchEngine; other keywords as BatchEngine."""

    def __init__(self, specs, device=None, trace=False, pipelined=True, graphs=True, **kw):
        if kw.pop("truncate_every", 1) != 1:
            raise NotImplementedError("the batched driver implements truncate_every = 1")
        self.specs = list(specs)
        C, slab = max(1, len(self.specs)), max(1, sum(int(s[1]) for s in self.specs))
        if pipelined:
            self.engine = PipelinedBatchEngine(C, slab, device=device, trace=trace, graphs=graphs, **kw)
        else:
            self.engine = BatchEngine(C, slab, device=device, trace=trace, **kw)
        self.traces = None

    @property
    def counters(self):
        return self.engine.counters

    def run(self):
        for r, (n, k, circ, oat) in enumerate(self.specs):
            if not self.engine.admit(n, k, circ, oat, tag=r):
                raise RuntimeError("batch does not fit")
        out = [None] * len(self.specs)
        while self.engine.active:
            for tag, res in self.engine.tick():
                out[tag] = res
        if self.engine.trace:
            self.traces = [r.pop("trace") for r in out]
        return out


# next action of a circuit in the pipelined engine
_A_GATE, _A_WAIT, _A_RDM, _A_TRIAL, _A_KEEP, _A_SUMS, _A_FIN, _A_DONE = range(8)
# what a moments result leads to
_C_CHK, _C_RDM, _C_FIN = range(1, 4)


class PipelinedBatchEngine(BatchEngine):
    """BatchEngine whose ticks overlap with the device.

    State sizes live in device count cells (one per circuit and slab): a run
    of gates reads and writes them on the device, so it needs no host read.
    tick() launches the next operation of every circuit that can go on, then
    waits for the previous tick's device work and applies its results (stats
    copied asynchronously into pinned host memory).  A circuit whose next
    operation depends on a result (a check point, the RDMs of a pass, a
    rotation trial's decision, moments) sits out the ticks until that result
    is in.  Per circuit the device operations, their order and every decision
    are those of BatchEngine, so the results are identical.

    The operations run through a batch plan (_ops.BatchPlan): fixed launch
    shapes whose geometry is derived on the device, and with graphs=True each
    operation's launch sequence is captured once in a HIP graph and replayed.
    """

    def __init__(self, max_circuits, slab_entries, device=None, graphs=True, **kw):
        super().__init__(max_circuits, slab_entries, device=device, **kw)
        import torch

        C = self.C
        self.graphs = bool(graphs)
        self.plan = _ops.BatchPlan(self._ctx, C, self.slab_entries // 512 + C + 1, graphs=self.graphs)
        self.tiles = None
        self._cnt = torch.zeros((2, C), dtype=torch.int64, device=self._ctx.device)
        self._cbase = np.uint64(self._cnt.data_ptr())
        self.act = np.full(C, _A_DONE, dtype=np.int64)
        self.after = np.zeros(C, dtype=np.int64)
        # the pending rotation trial of each circuit
        self.tj = np.zeros(C, dtype=np.int64)
        self.tV = np.zeros((C, 2, 2), dtype=np.complex128)
        self.tpos = np.zeros((C, NMAX), dtype=np.int64)
        self.tm = np.zeros(C, dtype=np.int64)
        self.ttr = np.zeros(C, dtype=bool)
        self.tfrac = np.zeros(C)
        self._keep_sum2 = np.zeros(C)
        # gates of the admitted circuits, flattened: [gbase_c, gbase_c + len_c)
        self._gcap = 0
        self._gtop = 0
        self.gbase = np.zeros(C, dtype=np.int64)
        self._g_nq = np.zeros(0, dtype=np.int8)
        self._g_q = np.zeros((0, 2), dtype=np.int64)
        self._g_m = np.zeros((0, 4, 4), dtype=np.complex128)
        self._g_chk = np.zeros(0, dtype=bool)  # an optimiser check point follows the gate
        # two tick slots of pinned result buffers (stats rows, RDM rows)
        self._stats = [_ops.pinned_array((8 * C,), _ops.STATS_DTYPE) for _ in range(2)]
        self._rho = [_ops.pinned_array((C, 64, 4), np.float64) for _ in range(2)]
        self._srow = [0, 0]
        self._records = [[], []]
        self._events = [None, None]
        self._tickno = 0
        self._finishing = []  # (event, [(c, tag, meta, x_pinned, a_pinned)])

    # ---------------------------------------------------------- admission --
    def _store_gates(self, circuit, new):
        """Append circuit's gates (circuit slot `new`, not stored yet) and its
        check points."""
        M = len(circuit)
        nqs = np.fromiter((g.n_qubits for g in circuit), dtype=np.int8, count=M)
        if np.any((nqs != 1) & (nqs != 2)):
            raise ValueError("only 1- and 2-qubit gates are supported")
        if self._gtop + M > self._gcap:
            # compact the stored circuits' gates, then grow
            live = np.flatnonzero(self.alive)
            live = live[live != new]
            keep = [(c, int(self.gbase[c]), int(self.lens[c])) for c in live]
            need = sum(m for _, _, m in keep) + M
            cap = max(1 << 16, 2 * need)
            nq = np.zeros(cap, dtype=np.int8)
            q = np.zeros((cap, 2), dtype=np.int64)
            m = np.zeros((cap, 4, 4), dtype=np.complex128)
            ck = np.zeros(cap, dtype=bool)
            top = 0
            for c, b, ln in keep:
                nq[top: top + ln] = self._g_nq[b: b + ln]
                q[top: top + ln] = self._g_q[b: b + ln]
                m[top: top + ln] = self._g_m[b: b + ln]
                ck[top: top + ln] = self._g_chk[b: b + ln]
                self.gbase[c] = top
                top += ln
            self._g_nq, self._g_q, self._g_m, self._g_chk = nq, q, m, ck
            self._gcap, self._gtop = cap, top
        b = self._gtop
        self._g_nq[b: b + M] = nqs
        i1 = np.flatnonzero(nqs == 1)
        i2 = np.flatnonzero(nqs == 2)
        if len(i1):
            self._g_q[b + i1, 0] = np.fromiter((circuit[i].qubits[0] for i in i1), dtype=np.int64, count=len(i1))
            self._g_m[b + i1, :2, :2] = np.array([circuit[i].matrix for i in i1], dtype=np.complex128)
        if len(i2):
            self._g_q[b + i2] = np.array([circuit[i].qubits[:2] for i in i2], dtype=np.int64)
            self._g_m[b + i2] = np.array([circuit[i].matrix for i in i2], dtype=np.complex128)
        self._g_chk[b: b + M] = False
        pts = np.fromiter(self.chk[new], dtype=np.int64, count=len(self.chk[new]))
        self._g_chk[b + pts] = True
        self._gtop = b + M
        return b

    def admit(self, num_qubits, k, circuit, optimize_at=None, tag=None):
        circuit = list(circuit)
        if not super().admit(num_qubits, k, circuit, optimize_at, tag):
            return False
        c = self._last_admitted
        self.gbase[c] = self._store_gates(circuit, c)
        self.circuits[c] = None  # the flattened copy is used
        self.act[c] = _A_GATE if self.lens[c] > 0 else _A_DONE
        self._cnt[0, c] = 1
        if self.lens[c] == 0:
            self._finish(np.array([c]))
        return True

    # ------------------------------------------------------------ helpers --
    def _jobs_dev(self, cs, out_other=False):
        """Jobs whose counts are the device cells (n = the bound k)."""
        J = self._jobs(cs, out_other)
        cur = self.cur[cs].astype(np.uint64)
        oth = cur ^ np.uint64(1) if out_other else cur
        C = np.uint64(self.C)
        csu = cs.astype(np.uint64)
        J["cnt_in"] = self._cbase + np.uint64(8) * (cur * C + csu)
        J["cnt_out"] = self._cbase + np.uint64(8) * (oth * C + csu)
        J["n"] = self.k[cs]
        return J

    def _cells(self, J, cs, out_other):
        cur = self.cur[cs].astype(np.uint64)
        oth = cur ^ np.uint64(1) if out_other else cur
        C = np.uint64(self.C)
        csu = cs.astype(np.uint64)
        J["cnt_out"] = self._cbase + np.uint64(8) * (oth * C + csu)

    def _run(self, op, J, kind, cs, rows=True, extra=None):
        slot = self._tickno & 1
        out = None
        if kind == "rdm":
            out = self._rho[slot][0][: len(cs)]
        elif rows:
            r0 = self._srow[slot]
            out = self._stats[slot][0][r0: r0 + len(cs)]
            self._srow[slot] = r0 + len(cs)
        self.plan.run(op, J, out)
        if kind is not None:
            self._records[slot].append((kind, cs, out, extra))

    # ------------------------------------------------------------- issue ----
    def _issue_gates(self, g):
        ii = self.gidx[g]
        idx = self.gbase[g] + ii
        nq = self._g_nq[idx]
        # after this gate: 2 = the circuit's last, 1 = a check point, 0 = go on
        last = ii + 1 >= self.lens[g]
        post = np.where(last, 2, np.where(self._g_chk[idx], 1, 0))
        one = nq == 1
        # the general gates of both widths go into one step (one lift pass per
        # width, one select over all of them)
        parts = []
        if (~one).any():
            p2 = self._issue_2q(ii[~one], g[~one], idx[~one], post[~one])
            if p2 is not None:
                parts.append(p2)
        if one.any():
            cs, gi = g[one], idx[one]
            q = self._g_q[gi, 0]
            gm = self._g_m[gi, :2, :2]
            U = self.U[cs, q]
            G = _dagger(U) @ gm @ U
            jb = self.pos[cs, q]
            self.pos[cs] = _lift_rows(self.pos[cs], jb, 0)
            J = self._jobs_dev(cs)
            J["j"] = jb
            J["nq"] = 1
            J["g"][:, :8] = np.ascontiguousarray(G).reshape(len(cs), 4).view(np.float64)
            J["key_map"] = _key_maps(self.pos[cs])
            parts.append((cs, J, ii[one], post[one]))
            self._count("gate1q", len(cs))
        if len(parts) == 1:
            cs, J, pi, pp = parts[0]
            self._run(L.BOP_GATE1 if J["nq"][0] == 1 else L.BOP_GATE2, J, "gate", cs, extra=(pi, pp))
        elif parts:
            cs = np.concatenate([p[0] for p in parts])
            J = np.concatenate([p[1] for p in parts])
            self._run(L.BOP_GATE12, J, "gate", cs, extra=(np.concatenate([p[2] for p in parts]),
                                                          np.concatenate([p[3] for p in parts])))
        self.gidx[g] += 1
        self.act[g] = np.where(post > 0, _A_WAIT, _A_GATE)

    def _issue_2q(self, ii, cs, gi, post):
        """Diagonal gates (launched) and moves (launched); returns the general
        gates' step table (cs, jobs, gate indices, post) or None."""
        q1 = self._g_q[gi, 0].copy()
        q2 = self._g_q[gi, 1].copy()
        gm = self._g_m[gi]
        Ut = _kron2(self.U[cs, q1], self.U[cs, q2])
        G = _dagger(Ut) @ gm @ Ut
        off = G.copy()
        off[:, np.arange(4), np.arange(4)] = 0
        diag = self.use_diag_gate & (np.abs(off).max(axis=(1, 2)) < 1e-12)
        if diag.any():
            d = cs[diag]
            J = self._jobs_dev(d)
            J["j"] = self.pos[d, q1[diag]]
            J["j2"] = self.pos[d, q2[diag]]
            J["g"][:, :8] = np.ascontiguousarray(np.diagonal(G[diag], axis1=1, axis2=2)).view(np.float64)
            self._run(L.BOP_DIAG2Q, J, "diag", d, rows=False, extra=(ii[diag], post[diag]))
            self._count("diag2q", len(d))
        gen = ~diag
        if not gen.any():
            return None
        cs, q1, q2, G, ii, post = cs[gen], q1[gen], q2[gen], G[gen], ii[gen], post[gen]
        sw = (self.pos[cs, q1] == 0) & (self.pos[cs, q2] != 0)
        if sw.any():
            G[sw] = G[sw][:, _SWAP_IDX][:, :, _SWAP_IDX]
            q1[sw], q2[sw] = q2[sw], q1[sw].copy()
        mv = self.pos[cs, q2] != 0
        if mv.any():
            m = cs[mv]
            J = self._jobs_dev(m, out_other=True)
            jb = self.pos[m, q2[mv]]
            J["j"] = jb
            self._run(L.BOP_MOVE, J, None, m, rows=False)
            self.pos[m] = _lift_rows(self.pos[m], jb, 0)
            self.cur[m] ^= 1
        jb = self.pos[cs, q1]
        self.pos[cs] = _lift_rows(self.pos[cs], jb, 1)
        J = self._jobs_dev(cs)
        J["j"] = jb
        J["nq"] = 2
        J["g"][:, :32] = np.ascontiguousarray(G).reshape(len(cs), 16).view(np.float64)
        J["key_map"] = _key_maps(self.pos[cs])
        self._count("gate2q", len(cs))
        return cs, J, ii, post

    def _issue_sums(self, cs):
        self._run(L.BOP_MOMENTS, self._jobs(cs), "sums", cs)

    def _issue_rdm(self, cs):
        self._run(L.BOP_RDM1Q, self._jobs(cs), "rdm", cs)
        self._count("rdm", len(cs))

    def _issue_trials(self, trial):
        j = self.ochain[trial, self.optr[trial]]
        self.optr[trial] += 1
        Vj = self.oV[trial, j]
        jb = self.pos[trial, j]
        tpos = _lift_rows(self.pos[trial], jb, 0)
        self.tj[trial], self.tV[trial], self.tpos[trial] = j, Vj, tpos
        J = self._jobs(trial, out_other=True)
        self._cells(J, trial, True)
        J["j"] = jb
        J["g"][:, :8] = np.ascontiguousarray(_dagger(Vj)).reshape(len(trial), 4).view(np.float64)
        J["key_map"] = _key_maps(tpos)
        self._run(L.BOP_ROT1, J, "trial", trial)
        self._count("rotate", len(trial))

    def _issue_keep(self, cs):
        # kept whole: renormalise the candidates, then their moments
        J = self._jobs(cs, out_other=True)
        J["x"], J["a"] = J["xo"], J["ao"]
        J["n"] = self.tm[cs]
        norm = np.sqrt(self._keep_sum2[cs])
        J["scale"] = np.where(norm > 1e-30, norm, 1.0)
        self._run(L.BOP_SCALE, J, None, cs, rows=False)
        self._run(L.BOP_MOMENTS, J, "keep", cs)

    def _issue_fin(self, cs):
        """Final renormalisation (:551-554), then the moments _opt_leave reads."""
        norm = np.sqrt(self.sum2[cs])
        do = norm > 1e-30
        if do.any():
            J = self._jobs(cs[do])
            J["scale"] = norm[do]
            self._run(L.BOP_SCALE, J, None, cs[do], rows=False)
        self.sum4[cs] = np.nan
        self._run(L.BOP_MOMENTS, self._jobs(cs), "leave", cs)

    # ------------------------------------------------------------ results --
    def _process(self, slot):
        chk, done = [], []
        for kind, cs, out, extra in self._records[slot]:
            if kind in ("gate", "diag"):
                chk.append(cs[extra[1] == 1])
                done.append(cs[extra[1] == 2])
            if out is not None and kind != "rdm":
                bad = np.flatnonzero(out["err"] != 0)
                if len(bad):
                    raise L.BassError(f"batch {kind}: circuit {int(cs[bad[0]])}: device error bits "
                                      f"0x{int(out['err'][bad[0]]):x}")
            if kind == "gate":
                self._apply_trunc_stats(extra[0], cs, out)
            elif kind == "diag":
                self.sum4[cs] = np.nan
                self._count("amp_updates", self.nnz[cs].sum())
            elif kind == "sums":
                self.sum2[cs], self.sum4[cs] = out["sum2"], out["sum4"]
                a = self.after[cs]
                self._checkpoint2(cs[a == _C_CHK])
                self._rdm2(cs[a == _C_RDM])
                self.act[cs[a == _C_FIN]] = _A_FIN
            elif kind == "rdm":
                self._rdm_result(cs, out)
            elif kind == "trial":
                self._trial_result(cs, out)
            elif kind == "keep":
                self._decide(cs, out["sum4"].copy())
            elif kind == "leave":
                self.sum2[cs], self.sum4[cs] = out["sum2"], out["sum4"]
                self.pr_last[cs] = self._pr(cs)
                self.phase[cs] = 0
                self.act[cs] = _A_GATE
        self._records[slot] = []
        self._srow[slot] = 0
        if chk:
            self._checkpoint(np.concatenate(chk))
            self._finish(np.concatenate(done))

    def _checkpoint(self, ch):
        """Check point after a gate (BatchEngine._gate_step)."""
        if not len(ch):
            return
        ta_mask = self.trunc_active[ch]
        self.trunc_active[ch] = False
        self.act[ch[~ta_mask]] = _A_GATE
        ta = ch[ta_mask]
        if not len(ta):
            return
        need = np.isnan(self.sum4[ta]) | np.isnan(self.sum2[ta])
        self.act[ta[need]] = _A_SUMS
        self.after[ta[need]] = _C_CHK
        self._checkpoint2(ta[~need])

    def _checkpoint2(self, ta):
        if not len(ta):
            return
        pr = self._pr(ta)
        trig = np.ones(len(ta), bool) if self.pr_opt_threshold >= 1.0 else pr > self.pr_last[ta] / self.pr_opt_threshold
        self.act[ta[~trig]] = _A_GATE
        opt = ta[trig]
        if not len(opt):
            return
        enter = ~(self._pr(opt) < self.k[opt] / 4)
        e = opt[enter]
        self.phase[e] = 1
        self.ostage[e] = 0
        self.opass[e] = 0
        self.oentered[opt] = enter
        self.act[e] = _A_RDM
        skip = opt[~enter]
        if len(skip):
            # _opt_leave: the sums are known here
            self.pr_last[skip] = self._pr(skip)
            self.phase[skip] = 0
            self.act[skip] = _A_GATE

    def _rdm_result(self, cs, out):
        rho_st = np.zeros((len(cs), 64, 2, 2), dtype=np.complex128)
        rho_st[:, :, 0, 0] = out[:, :, 0]
        rho_st[:, :, 1, 1] = out[:, :, 1]
        rho_st[:, :, 0, 1] = out[:, :, 2] + 1j * out[:, :, 3]
        rho_st[:, :, 1, 0] = out[:, :, 2] - 1j * out[:, :, 3]
        P = np.minimum(self.pos[cs], NMAX - 1)
        rho = np.take_along_axis(rho_st, P[:, :, None, None], axis=1)
        V, trivial = _eigh2x2_rows(rho)
        trivial |= np.arange(NMAX)[None, :] >= self.N[cs][:, None]
        self.oV[cs] = V
        self.oclen[cs] = (~trivial).sum(axis=1)
        self.ochain[cs] = np.argsort(trivial, axis=1, kind="stable")
        self.optr[cs] = 0
        need = np.isnan(self.sum4[cs]) | np.isnan(self.sum2[cs])
        self.act[cs[need]] = _A_SUMS
        self.after[cs[need]] = _C_RDM
        self._rdm2(cs[~need])

    def _rdm2(self, cs):
        if not len(cs):
            return
        self.opr[cs] = self._pr(cs)
        self.oimproved[cs] = False
        self.ostage[cs] = 1
        self.act[cs] = _A_TRIAL

    def _trial_result(self, cs, st):
        m = st["n_in"].copy()
        tr = (st["flags"] & 1) != 0
        self._count("amp_updates", m.sum())
        self.tm[cs] = m
        self.ttr[cs] = tr
        self.tfrac[cs] = np.where(tr, st["kept"], st["sum2"])
        self._keep_sum2[cs] = st["sum2"]
        keep = cs[~tr]
        self.act[keep] = _A_KEEP
        if tr.any():
            self._decide(cs[tr], st["sum4"][tr].copy())

    def _decide(self, trial, s4):
        """Accept / reject of rotation trials (BatchEngine._opt_trial)."""
        m, tr, frac = self.tm[trial], self.ttr[trial], self.tfrac[trial]
        j, Vj = self.tj[trial], self.tV[trial]
        pr_new = np.where(s4 > 1e-60, 1.0 / np.where(s4 > 1e-60, s4, 1.0), np.inf)
        pr = self.opr[trial]
        reject = pr_new >= pr * (1.0 - self.accept_rtol)
        if self.traces is not None:
            for r, c in enumerate(trial):
                self.traces[c].append(("rot", int(j[r]), bool(not reject[r]), float(pr[r]),
                                       float(pr_new[r]), float(frac[r])))
        acc = np.flatnonzero(~reject)
        self._count("reject", int(reject.sum()))
        self._count("accept", len(acc))
        self.act[trial] = _A_TRIAL
        if not len(acc):
            return
        a = trial[acc]
        ja = j[acc]
        self.cur[a] ^= 1
        self.nnz[a] = np.where(tr[acc], self.k[a], m[acc])
        self.pos[a] = self.tpos[a]
        self.sum2[a], self.sum4[a] = np.nan, s4[acc]
        self.U[a, ja] = self.U[a, ja] @ Vj[acc]
        self.opr[a] = pr_new[acc]
        f = frac[acc]
        self.gamma_run[a] *= np.sqrt(np.where(f > 1e-30, f, 0.0))
        self.oimproved[a] = True

    def _finish(self, cs):
        """Last gate done: copy the state to pinned host memory (returned once
        the copy is complete) and release the slab region."""
        if not len(cs):
            return
        import torch

        items = []
        for c in cs:
            c = int(c)
            n = int(self.nnz[c])
            buf = self.slabs[int(self.cur[c])]
            o = int(self.off[c])
            xp = torch.empty(max(n, 1), dtype=torch.int64, pin_memory=True)
            ap = torch.empty(max(n, 1), dtype=torch.complex128, pin_memory=True)
            if n:
                xp[:n].copy_(buf.x[o: o + n], non_blocking=True)
                ap[:n].copy_(buf.a[o: o + n], non_blocking=True)
            meta = {"nnz": n, "gamma": float(np.sqrt(max(self.gamma_run[c], 0))),
                    "U": [self.U[c, j].copy() for j in range(self.N[c])], "pos": self.pos[c, : self.N[c]].copy()}
            if self.traces is not None:
                meta["trace"] = self.traces[c]
            items.append((c, self.tags[c], meta, xp, ap))
            self.act[c] = _A_DONE
            self._release(o, int(self.k[c]))
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self._ctx.device))
        self._finishing.append((ev, items))

    def _collect(self, force=False):
        out = []
        keep = []
        for ev, items in self._finishing:
            if force:
                ev.synchronize()
            if force or ev.query():
                for c, tag, meta, xp, ap in items:
                    n = meta["nnz"]
                    w = xp[:n].numpy().view(np.uint64)
                    res = {"keys": stored_to_keys(w, meta.pop("pos")), "amps": ap[:n].numpy().copy(),
                           "nnz": n, "gamma": meta["gamma"], "U": meta["U"]}
                    if "trace" in meta:
                        res["trace"] = meta["trace"]
                    out.append((tag, res))
                    self.alive[c] = False
            else:
                keep.append((ev, items))
        self._finishing = keep
        return out

    # --------------------------------------------------------------- tick --
    def tick(self):
        """Launch one operation of every circuit that can go on, then apply
        the previous tick's results.  Returns [(tag, result)] of circuits
        whose final state has reached the host."""
        import torch

        slot = self._tickno & 1
        live = self.alive
        # pass ends (host only): another pass after an improvement, else finish
        tr = np.flatnonzero(live & (self.act == _A_TRIAL))
        if len(tr):
            end = tr[self.optr[tr] >= self.oclen[tr]]
            if len(end):
                again = self.oimproved[end] & (self.opass[end] < 2)
                a = end[again]
                self.opass[a] += 1
                self.ostage[a] = 0
                self.act[a] = _A_RDM
                f = end[~again]
                self.ostage[f] = 2
                need = np.isnan(self.sum4[f]) | np.isnan(self.sum2[f])
                self.act[f[need]] = _A_SUMS
                self.after[f[need]] = _C_FIN
                self.act[f[~need]] = _A_FIN
        g = np.flatnonzero(live & (self.act == _A_GATE))
        if len(g):
            self._issue_gates(g)
        s = np.flatnonzero(live & (self.act == _A_SUMS))
        if len(s):
            self.act[s] = _A_WAIT
            self._issue_sums(s)
        r = np.flatnonzero(live & (self.act == _A_RDM))
        if len(r):
            self.act[r] = _A_WAIT
            self._issue_rdm(r)
        t = np.flatnonzero(live & (self.act == _A_TRIAL))
        if len(t):
            self.act[t] = _A_WAIT
            self._issue_trials(t)
        kp = np.flatnonzero(live & (self.act == _A_KEEP))
        if len(kp):
            self.act[kp] = _A_WAIT
            self._issue_keep(kp)
        fn = np.flatnonzero(live & (self.act == _A_FIN))
        if len(fn):
            self.act[fn] = _A_WAIT
            self._issue_fin(fn)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self._ctx.device))
        self._events[slot] = ev
        prev = slot ^ 1
        if self._events[prev] is not None:
            self._events[prev].synchronize()
            self._events[prev] = None
            self._process(prev)
        self._tickno += 1
        # nothing launched and nothing in flight: the last results are in
        idle = not (len(g) or len(s) or len(r) or len(t) or len(kp) or len(fn))
        if idle:
            ev.synchronize()
            self._events[slot] = None
            self._process(slot)
        return self._collect(force=idle)
