"""The batched driver (many circuits per launch) against the single-circuit
driver: per circuit the same device operations run in the same order, so the
results must agree bit for bit."""

import numpy as np
import pytest

from conftest import pkg_module

pytestmark = pytest.mark.gpu


def _specs():
    rc = pkg_module("utils.random_circuits")
    out = []
    for t, (n, k, steps, h) in enumerate([(10, 64, 8, 0.9), (12, 256, 6, 1.7), (8, 32, 10, 0.4), (11, 128, 7, 1.2)]):
        circ = rc.generate_tfim_circuit(n, steps, J=1.0, h=h, dt=0.2, rng=np.random.default_rng(100 + t),
                                        disorder_strength_h=0.3)
        out.append((n, k, circ, None))
    for t, (n, k, layers) in enumerate([(10, 128, 5), (12, 64, 4)]):
        circ, ends = rc.brickwork_haar_circuit(n, layers, np.random.default_rng(200 + t))
        out.append((n, k, circ, ends))
    return out


# engine modes: pipelined with HIP graphs (the default), pipelined with direct
# launches, the synchronous engine
MODES = {"graphs": {}, "pipelined": {"graphs": False}, "sync": {"pipelined": False}}


@pytest.mark.parametrize("mode", list(MODES))
def test_batched_matches_single(gpu_ctx, mode):
    print(f"[case] batched vs single ({mode})", flush=True)
    Bm = pkg_module("simulation.batched")
    S = pkg_module("simulation.bass_simulator")
    specs = _specs()
    bat = Bm.BatchedBASS(specs, device="cuda:0", trace=True, **MODES[mode])
    res = bat.run()
    for c, (n, k, circ, ends) in enumerate(specs):
        tr = []
        kw = {"optimize_at": ends} if ends is not None else {}
        sim = S.BASS(n, k, device="cuda:0", trace=tr, **kw)
        st = sim.simulate(circ)
        keys = st.keys()
        amps = st.alpha[: st.nnz]
        r = res[c]
        print(f"  circuit {c}: nnz {r['nnz']} / {st.nnz}, gamma {r['gamma']:.15f} / {st.gamma:.15f}", flush=True)
        assert [e[:3] for e in bat.traces[c] if e[0] == "rot"] == [e[:3] for e in tr if e[0] == "rot"]
        assert r["nnz"] == st.nnz
        o1, o2 = np.argsort(r["keys"]), np.argsort(keys)
        assert np.array_equal(r["keys"][o1], keys[o2])
        assert np.array_equal(r["amps"][o1].view(np.float64), amps[o2].view(np.float64))
        assert r["gamma"] == st.gamma
        for j in range(n):
            assert np.array_equal(r["U"][j], sim.U[j])


def _single(spec):
    S = pkg_module("simulation.bass_simulator")
    n, k, circ, ends = spec
    kw = {"optimize_at": ends} if ends is not None else {}
    sim = S.BASS(n, k, device="cuda:0", **kw)
    st = sim.simulate(circ)
    return st.keys(), st.alpha[: st.nnz].copy(), st.nnz, st.gamma


@pytest.mark.parametrize("pipelined", [False, True])
def test_engine_staggered_admission(gpu_ctx, pipelined):
    """Circuits joining a running batch between ticks, into slots and slab
    regions released by finished ones (two slots, a slab that fits the two
    largest states only), give the single-circuit results bit for bit."""
    print(f"[case] engine staggered admission (pipelined={pipelined})", flush=True)
    Bm = pkg_module("simulation.batched")
    specs = _specs()
    want = [_single(s) for s in specs]
    ks = sorted(s[1] for s in specs)
    cls = Bm.PipelinedBatchEngine if pipelined else Bm.BatchEngine
    eng = cls(2, ks[-1] + ks[-2], device="cuda:0")
    queue = list(range(len(specs)))
    got = {}
    ticks = 0
    while queue or eng.active:
        # at most one admission per tick, so circuits start at different ticks
        if queue:
            c = queue[0]
            n, k, circ, ends = specs[c]
            if eng.admit(n, k, circ, ends, tag=c):
                queue.pop(0)
        for tag, res in eng.tick():
            got[tag] = res
        ticks += 1
        assert ticks < 10000
    assert not eng.alive.any() and eng.free == [(0, ks[-1] + ks[-2])]
    for c, (keys, amps, nnz, gamma) in enumerate(want):
        r = got[c]
        print(f"  circuit {c}: nnz {r['nnz']} / {nnz}", flush=True)
        assert r["nnz"] == nnz and r["gamma"] == gamma
        o1, o2 = np.argsort(r["keys"]), np.argsort(keys)
        assert np.array_equal(r["keys"][o1], keys[o2])
        assert np.array_equal(r["amps"][o1].view(np.float64), amps[o2].view(np.float64))


def test_batch_server_stream(gpu_ctx):
    """BatchServer over a short Poisson stream of small requests: every request
    completes, latency >= service time, and results match the plain driver."""
    print("[case] batch server stream", flush=True)
    Srv = pkg_module("serving.server")
    reqs = [Srv.CircuitRequest(rid=i, kind=("tfim" if i % 2 else "brickwork"), n=10 + i % 5, depth=3 + i % 4,
                               log2k=6 + i % 3, h=0.7, seed=50 + i, arrival=0.002 * i) for i in range(10)]
    with Srv.BatchServer(max_circuits=4, slab_entries=1 << 12, device="cuda:0") as srv:
        recs = srv.run(reqs)
    assert [r["rid"] for r in recs] == list(range(10))
    for q, r in zip(reqs, recs):
        assert r["latency"] >= r["service"] - 1e-9 and r["end"] >= r["start"]
        res = Srv.run_request(q, "cuda:0")
        assert res["nnz"] == r["nnz"] and res["gamma2"] == r["gamma2"]
    summ = Srv.summarize(recs, amp_updates=srv.engine.counters["amp_updates"])
    assert summ["circuits"] == 10 and summ["p99_latency_s"] >= summ["p50_latency_s"]


def test_batch_server_two_engines(gpu_ctx):
    """Two engine worker processes: every request is served once, by one of
    them, with the plain driver's result."""
    print("[case] batch server, two engine processes", flush=True)
    Srv = pkg_module("serving.server")
    reqs = [Srv.CircuitRequest(rid=i, kind=("brickwork" if i % 3 else "tfim"), n=9 + i % 4, depth=2 + i % 3,
                               log2k=5 + i % 3, h=1.1, seed=80 + i, arrival=0.001 * i) for i in range(8)]
    with Srv.BatchServer(workers=2, max_circuits=3, slab_entries=1 << 10, device="cuda:0") as srv:
        recs = srv.run(reqs)
    assert sorted(r["rid"] for r in recs) == list(range(8))
    assert {r["worker"] for r in recs} <= {0, 1}
    for q, r in zip(reqs, recs):
        res = Srv.run_request(q, "cuda:0")
        assert res["nnz"] == r["nnz"] and res["gamma2"] == r["gamma2"]
    assert srv.counters["amp_updates"] > 0


@pytest.mark.timeout(900)
def test_batched_config3_sample(gpu_ctx):
    """32 circuits of the configured batched stream (make_batched_requests:
    TFIM with n ~ U[16,48], steps ~ U[10,60], h ~ U[0.2,2], K = 2^U{10..18};
    every 16th of the 512) run as one batch, with the engine's launch
    sequences replayed from HIP graphs and launched directly: each circuit
    equals the single-circuit driver bit for bit (keys, amplitudes, gamma,
    frame) both ways."""
    print("[case] batched config-3 sample: 32 of 512 circuits", flush=True)
    Srv = pkg_module("serving.server")
    Bm = pkg_module("simulation.batched")
    S = pkg_module("simulation.bass_simulator")
    reqs = Srv.make_batched_requests(512)[::16]
    specs = []
    for r in reqs:
        circ, kw = Srv.build_circuit(r)
        specs.append((r.n, 1 << r.log2k, circ, kw.get("optimize_at")))
    print(f"  n {min(s[0] for s in specs)}..{max(s[0] for s in specs)}, K 2^{min(r.log2k for r in reqs)}.."
          f"2^{max(r.log2k for r in reqs)}, {sum(len(s[2]) for s in specs)} gates", flush=True)
    runs = {g: Bm.BatchedBASS(specs, device="cuda:0", graphs=g).run() for g in (True, False)}
    for c, (n, k, circ, oat) in enumerate(specs):
        sim = S.BASS(n, k, device="cuda:0", **({"optimize_at": oat} if oat else {}))
        st = sim.simulate(circ)
        keys, amps = st.keys(), st.alpha[: st.nnz]
        for g, res in runs.items():
            r = res[c]
            assert r["nnz"] == st.nnz and r["gamma"] == st.gamma, (c, g)
            o1, o2 = np.argsort(r["keys"]), np.argsort(keys)
            assert np.array_equal(r["keys"][o1], keys[o2]), (c, g)
            assert np.array_equal(r["amps"][o1].view(np.float64), amps[o2].view(np.float64)), (c, g)
            assert all(np.array_equal(r["U"][j], sim.U[j]) for j in range(n)), (c, g)


@pytest.mark.timeout(900)
def test_served_sample_alone_and_in_mix(gpu_ctx):
    """16 requests of the configured served stream (make_served_requests: 50/50
    TFIM and brickwork-Haar, n ~ U[16,48] / U[24,44], K = 2^U{12..16} here),
    arriving while others run (their arrival times compressed 100-fold), served
    by two engine processes: each result equals the same request run alone
    through the plain driver -- the same entry count, gamma^2 within 1e-12
    relative (equal bit for bit unless two reductions of different tile
    shapes round one sum apart)."""
    print("[case] served sample: 16 requests, alone vs in the mix", flush=True)
    Srv = pkg_module("serving.server")
    pool = [r for r in Srv.make_served_requests(512, 200.0) if r.log2k <= 16]
    reqs = pool[:: max(1, len(pool) // 16)][:16]
    for r in reqs:
        r.arrival /= 100.0
    assert {r.kind for r in reqs} == {"tfim", "brickwork"}
    with Srv.BatchServer(workers=2, max_circuits=8, slab_entries=1 << 20, device="cuda:0") as srv:
        recs = srv.run(reqs)
    assert sorted(r["rid"] for r in recs) == sorted(q.rid for q in reqs)
    by = {r["rid"]: r for r in recs}
    exact = 0
    for q in reqs:
        alone = Srv.run_request(q, "cuda:0")
        m = by[q.rid]
        assert alone["nnz"] == m["nnz"], q.rid
        assert abs(alone["gamma2"] - m["gamma2"]) <= 1e-12 * alone["gamma2"], (q.rid, alone["gamma2"], m["gamma2"])
        exact += alone["gamma2"] == m["gamma2"]
    summ = Srv.summarize(recs)
    print(f"  {exact} of {len(reqs)} bit for bit; p50 {summ['p50_latency_s']:.2f} s, "
          f"p99 {summ['p99_latency_s']:.2f} s", flush=True)
