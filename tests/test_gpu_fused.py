"""The one-call gate selects against the stored pair they replace (bass_lift,
then bass_truncate on the stored segmented output), through the C ABI: the
fused select (bass_gate_truncate, the gate output never stored) and the
pruned lift (bass_lift_truncate, only the outputs that can matter stored);
kept keys, amplitudes and stats bit for bit, with the speculative prediction
holding, missing and switched off, in plain and counted mode."""

import numpy as np
import pytest

from conftest import pkg_module, random_state, random_unitary

pytestmark = pytest.mark.gpu

FIELDS = ("nnz", "n_in", "total", "kept", "norm", "sum2", "sum4")


def _mods():
    return pkg_module("_ops"), pkg_module("_device"), pkg_module("core.sparse_state"), pkg_module("_lib")


def _stored(x, pos):
    out = np.zeros_like(x)
    for q, p in enumerate(pos):
        out |= ((x >> np.uint64(q)) & np.uint64(1)) << np.uint64(p)
    return out


def _upload(dev, ctx, w, a):
    import torch

    n = len(w)
    b = dev.Buffers(max(n, 1), ctx.device)
    b.x[:n] = torch.from_numpy(w.view(np.int64)).to(ctx.device)
    b.a[:n] = torch.from_numpy(a).to(ctx.device)
    return b


def _result(dev, dst, st):
    m = int(st.nnz)
    return dev.to_host_keys(dst.x, m).copy(), dev.to_host_amps(dst.a, m).copy(), st


def _pair(ops, dev, ctx, src, n, j, G, nq, mode, k, km, spec, count):
    """The stored path: lift into a segmented buffer, truncate it."""
    seg = dev.Buffers(1, ctx.device)
    dst = dev.Buffers(k, ctx.device)
    if count is not None:
        ops.counted_begin(ctx, count, n)
        tiles = ops.lift_counted(ctx, src, n, j, seg, G=G, nq=nq, mode=mode, spec=spec)
        ops.truncate_counted(ctx, seg, k, dst, tiles, key_map=km, spec=spec)
        log = ops.counted_read(ctx, 4)
        assert len(log) == 1 and log[0].err == 0
        return _result(dev, dst, log[0])
    _, tiles = ops.lift(ctx, src, n, j, seg, G=G, nq=nq, mode=mode, count=False, spec=spec)
    h = ops.truncate_launch(ctx, seg, None, k, dst, renorm=True, hist_tiles=tiles, segmented=True, key_map=km,
                            spec=spec)
    return _result(dev, dst, ops.truncate_finish(ctx, h))


def _fused(ops, dev, ctx, src, n, j, G, nq, mode, k, km, spec, count):
    dst = dev.Buffers(k, ctx.device)
    if count is not None:
        ops.counted_begin(ctx, count, n)
        ops.gate_truncate(ctx, src, n, j, dst, k, G, nq=nq, mode=mode, counted=True, spec=spec, key_map=km)
        log = ops.counted_read(ctx, 4)
        assert len(log) == 1 and log[0].err == 0
        return _result(dev, dst, log[0])
    h = ops.gate_truncate(ctx, src, n, j, dst, k, G, nq=nq, mode=mode, spec=spec, key_map=km)
    return _result(dev, dst, ops.truncate_finish(ctx, h))


def _same(got, ref, what):
    assert np.array_equal(got[0], ref[0]), what
    assert np.array_equal(got[1].view(np.float64), ref[1].view(np.float64)), what
    for f in FIELDS:
        assert getattr(got[2], f) == getattr(ref[2], f), (what, f)
    assert (got[2].flags & 3) == (ref[2].flags & 3), what


# (n, nq, mode, k, kind, count): kind "ties" = equal |a| everywhere and a gate
# that keeps them equal (massive ties split by the key); count = the counted
# mode's actual entry count (n bounds it), None = plain calls
CASES = [
    (70000, 1, 0, 50000, "rand", None),
    (300000, 1, 1, 123457, "rand", None),
    (60000, 2, 0, 90001, "rand", None),
    (40000, 2, 1, 70001, "rand", None),
    (5000, 2, 0, 30000, "rand", None),
    (200000, 1, 0, 150001, "ties", None),
    (120000, 2, 0, 200003, "ties", None),
    (70000, 1, 0, 50000, "rand", 65000),
    (60000, 2, 1, 90001, "rand", 51234),
    (5000000, 1, 0, 4000000, "rand", None),
    (2500000, 2, 1, 3000000, "rand", None),
]


@pytest.mark.parametrize("n,nq,mode,k,kind,count", CASES)
def test_gate_truncate_matches_stored_pair(gpu_ctx, n, nq, mode, k, kind, count):
    ops, dev, SS, L = _mods()
    print(f"[case] fused gate select n={n} nq={nq} mode={mode} k={k} {kind} count={count}", flush=True)
    nbits = 40
    rng = np.random.default_rng(n + 10 * nq + mode)
    x, a = random_state(n, nbits, seed=n + 7)
    if kind == "ties":
        a = np.exp(2j * np.pi * rng.random(n)) / np.sqrt(n)
    pos = rng.permutation(nbits)
    w = _stored(x, pos)
    o = np.argsort(w)
    w, aw = w[o], a[o]
    if kind == "ties":
        # outputs of equal |a|: a permutation (1 qubit) or a diagonal phase (2)
        G = np.array([[0, 1], [1, 0]], complex) if nq == 1 else np.diag(np.exp(1j * rng.random(4)))
    else:
        G = random_unitary(1 << nq, rng)
    if nq == 1:
        j = int(pos[int(rng.integers(nbits))])
        npos = SS.lifted_positions(pos, j, 0)
    else:
        q2 = int(np.flatnonzero(pos == 0)[0])
        q1 = int(rng.choice([q for q in range(nbits) if q != q2]))
        j = int(pos[q1])
        npos = SS.lifted_positions(pos, j, 1)
    km = ops.key_map_arg(npos)
    src = _upload(dev, gpu_ctx, w, aw)
    ref = _pair(ops, dev, gpu_ctx, src, n, j, G, nq, mode, k, km, False, count)
    keepall = (ref[2].flags & 1) == 0
    plain = _fused(ops, dev, gpu_ctx, src, n, j, G, nq, mode, k, km, False, count)
    _same(plain, ref, "no speculation")
    # prediction = this bucket: the speculation holds
    hit = _fused(ops, dev, gpu_ctx, src, n, j, G, nq, mode, k, km, True, count)
    assert (hit[2].flags & 4) == (0 if keepall else 4)
    _same(hit, ref, "hit")
    # prediction from a select 16 exponents lower: the digit-1 pass recomputed
    low = _upload(dev, gpu_ctx, w, aw * 2.0 ** -8)
    _fused(ops, dev, gpu_ctx, low, n, j, G, nq, mode, k, km, False, count)
    del low
    miss = _fused(ops, dev, gpu_ctx, src, n, j, G, nq, mode, k, km, True, count)
    assert (miss[2].flags & 4) == 0
    _same(miss, ref, "miss")
    print(f"  kept {ref[2].nnz} of {ref[2].n_in} (keep-all {keepall})", flush=True)


def test_gate_truncate_rejects_overlap(gpu_ctx):
    """The output may not share memory with the input (the compaction reads
    the input while it writes)."""
    ops, dev, _, L = _mods()
    x, a = random_state(1000, 20, seed=5)
    w = np.sort(x)
    src = _upload(dev, gpu_ctx, w, a[np.argsort(x)])
    with pytest.raises(L.BassError, match="overlaps"):
        ops.gate_truncate(gpu_ctx, src, 1000, 3, src, 500, random_unitary(2, np.random.default_rng(1)))


@pytest.mark.parametrize("family", ["kicked", "brickwork", "pairs"])
def test_driver_fused_changes_nothing(gpu_ctx, family):
    """The counted driver with its gates, rotation trials and 2-qubit pass
    fused (bass_gate_truncate) and stored (bass_lift + bass_truncate):
    identical keys, amplitudes, gamma, frames and decisions."""
    print(f"[case] driver fused on/off: {family}", flush=True)
    rc = pkg_module("utils.random_circuits")
    B = pkg_module("simulation.bass_simulator")
    kw = {}
    if family == "kicked":
        N, K = 18, 1 << 14
        circ, ends = rc.kicked_ising_circuit(N, 4)
    elif family == "brickwork":
        N, K = 16, 1 << 11
        circ, ends = rc.brickwork_haar_circuit(N, 8, np.random.default_rng(77))
    else:
        N, K = 12, 1 << 8
        circ, ends = rc.brickwork_haar_circuit(N, 4, np.random.default_rng(78))
        kw = dict(use_2qubit_rotations=True)
    # without a trace the driver runs its counted gate loop; with one, gate
    # by gate (recording every decision)
    for traced in (False, True):
        out = []
        for fused in (False, True):
            tr = [] if traced else None
            sim = B.BASS(N, K, device="cuda:0", optimize_at=ends, accept_rtol=1e-10, trace=tr, **kw)
            sim.fused = fused
            st = sim.simulate(circ)
            dec = [(e[0], e[1], e[2]) for e in (tr or []) if e[0] in ("rot", "rot2")]
            out.append((st.keys().copy(), st.alpha[: st.nnz].copy(), st.gamma, [u.copy() for u in sim.U], dec,
                        dict(sim.counters)))
        (k0, a0, g0, u0, d0, c0), (k1, a1, g1, u1, d1, c1) = out
        # (four kicked Ising steps at N = 18 stay below the optimiser's PR threshold)
        assert d0 == d1 and (len(d0) > 0 or not traced or family == "kicked")
        assert np.array_equal(k0, k1) and np.array_equal(a0.view(np.float64), a1.view(np.float64))
        assert g0 == g1 and all(np.array_equal(p, q) for p, q in zip(u0, u1))
        assert c0.get("accept", 0) == c1.get("accept", 0) and c0.get("reject", 0) == c1.get("reject", 0)
        print(f"  traced={traced}: {len(d0)} decisions, accepted {c1.get('accept', 0)}", flush=True)


# ------------------------------------------------------------ pruned lift --
def _pruned(ops, dev, ctx, src, n, j, G, nq, mode, k, km, spec, count, margin=1):
    seg = dev.Buffers(1, ctx.device)
    dst = dev.Buffers(k, ctx.device)
    if count is not None:
        ops.counted_begin(ctx, count, n)
        ops.lift_truncate(ctx, src, n, j, seg, dst, k, G, nq=nq, mode=mode, counted=True, spec=spec, key_map=km,
                          margin=margin)
        log = ops.counted_read(ctx, 4)
        assert len(log) == 1 and log[0].err == 0
        return _result(dev, dst, log[0])
    h = ops.lift_truncate(ctx, src, n, j, seg, dst, k, G, nq=nq, mode=mode, spec=spec, key_map=km, margin=margin)
    return _result(dev, dst, ops.truncate_finish(ctx, h))


@pytest.mark.parametrize("n,nq,mode,k,kind,count", CASES)
def test_lift_truncate_matches_stored_pair(gpu_ctx, n, nq, mode, k, kind, count):
    """bass_lift_truncate (a lift that stores only the outputs that can
    matter to the predicted threshold) against the stored pair, bit for bit:
    the prediction exact, one or two buckets low, one bucket high (the
    stored outputs a superset), far too high (the select needs dropped
    outputs: relift, stats flags bit 3), switched off, and keep-all after a
    truncating select (a relift whenever something was dropped); with the
    default margin (the binade below the prediction stored too) and margin 0
    (a prediction one bucket high then relifts)."""
    ops, dev, SS, L = _mods()
    print(f"[case] pruned lift n={n} nq={nq} mode={mode} k={k} {kind} count={count}", flush=True)
    nbits = 40
    rng = np.random.default_rng(n + 10 * nq + mode + 5)
    x, a = random_state(n, nbits, seed=n + 9)
    if kind == "ties":
        a = np.exp(2j * np.pi * rng.random(n)) / np.sqrt(n)
    pos = rng.permutation(nbits)
    w = _stored(x, pos)
    o = np.argsort(w)
    w, aw = w[o], a[o]
    if kind == "ties":
        G = np.array([[0, 1], [1, 0]], complex) if nq == 1 else np.diag(np.exp(1j * rng.random(4)))
    else:
        G = random_unitary(1 << nq, rng)
    if nq == 1:
        j = int(pos[int(rng.integers(nbits))])
        npos = SS.lifted_positions(pos, j, 0)
    else:
        q2 = int(np.flatnonzero(pos == 0)[0])
        q1 = int(rng.choice([q for q in range(nbits) if q != q2]))
        j = int(pos[q1])
        npos = SS.lifted_positions(pos, j, 1)
    km = ops.key_map_arg(npos)
    src = _upload(dev, gpu_ctx, w, aw)
    ref = _pair(ops, dev, gpu_ctx, src, n, j, G, nq, mode, k, km, False, count)
    keepall = (ref[2].flags & 1) == 0
    got = {"plain": _pruned(ops, dev, gpu_ctx, src, n, j, G, nq, mode, k, km, False, count)}
    # predictions set by a stored select on the state scaled by s: |a|^2
    # buckets shift by log2(s^2); the pruned call then predicts that bucket
    for margin in (1, 0):
        for name, scale in [("exact", 1.0), ("one low", 2.0 ** -0.5), ("two low", 0.5), ("one high", 2.0 ** 0.5),
                            ("far high", 16.0)]:
            pre = _upload(dev, gpu_ctx, w, aw * scale)
            _pair(ops, dev, gpu_ctx, pre, n, j, G, nq, mode, k, km, False, count)
            del pre
            got[f"{name} m{margin}"] = _pruned(ops, dev, gpu_ctx, src, n, j, G, nq, mode, k, km, True, count, margin)
    for name, g in got.items():
        _same(g, ref, name)
    if not keepall:
        for m in (1, 0):
            assert got[f"exact m{m}"][2].flags & 4 and not got[f"exact m{m}"][2].flags & 8
            assert not got[f"two low m{m}"][2].flags & 8
            assert got[f"far high m{m}"][2].flags & 8
        assert not got["one high m1"][2].flags & 8
        assert got["one high m0"][2].flags & 8
    print(f"  kept {ref[2].nnz} of {ref[2].n_in} (keep-all {keepall}); relifts "
          f"{[n for n, g in got.items() if g[2].flags & 8]}", flush=True)


@pytest.mark.parametrize("family", ["kicked", "brickwork"])
def test_driver_prune_changes_nothing(gpu_ctx, family):
    """The counted driver with its gate and trial lifts pruned and stored in
    full: identical keys, amplitudes, gamma, frames and counters."""
    print(f"[case] driver prune on/off: {family}", flush=True)
    rc = pkg_module("utils.random_circuits")
    B = pkg_module("simulation.bass_simulator")
    if family == "kicked":
        N, K = 18, 1 << 14
        circ, ends = rc.kicked_ising_circuit(N, 6)
    else:
        N, K = 16, 1 << 11
        circ, ends = rc.brickwork_haar_circuit(N, 8, np.random.default_rng(79))
    out = []
    for prune in (False, True):
        sim = B.BASS(N, K, device="cuda:0", optimize_at=ends, accept_rtol=1e-10)
        sim.prune = prune
        st = sim.simulate(circ)
        out.append((st.keys().copy(), st.alpha[: st.nnz].copy(), st.gamma, [u.copy() for u in sim.U],
                    dict(sim.counters)))
    (k0, a0, g0, u0, c0), (k1, a1, g1, u1, c1) = out
    assert np.array_equal(k0, k1) and np.array_equal(a0.view(np.float64), a1.view(np.float64))
    assert g0 == g1 and all(np.array_equal(p, q) for p, q in zip(u0, u1))
    assert {k: v for k, v in c0.items() if k != "relift"} == {k: v for k, v in c1.items() if k != "relift"}
    print(f"  truncations {c1.get('truncate', 0)}, relifts {c1.get('relift', 0)}", flush=True)
