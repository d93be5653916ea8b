#!/usr/bin/env python3
"""Plain-path benchmark of the BASS sparse-state core on one MI355X.

Workloads (--workload):
  kicked-ising  (default) periodic chain, n=56, each step RZZ(pi/4) on every
                bond then RX(0.3 + 0.01 i) on every qubit, 60 steps, K=2^26;
                the brickwork line below is reported beside it ("secondary");
  brickwork     random circuit, n=40 qubits, each layer a Haar-random U(2) on
                every qubit then CZ on alternating even/odd bonds, 30 layers,
                K=2^22.
Both: complex128 amplitudes, uint64 keys, top-k truncation after every
non-diagonal gate, the adaptive single-qubit basis optimisation at every
step boundary, the full configured circuit built.

--calibrate-copy MIB times a plain device-to-device copy of MIB MiB (torch
copy_) --steps times and prints its bytes; it is the reference kernel that
scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes.

A step is one circuit layer.  --warmup steps run untimed (the state saturates
at K during the first steps; each one's wall time is reported), then --steps
steps are timed between two device synchronisations.  value = amp-updates/s:
records produced by gate, rotation and diagonal kernels per second.

The JSON line also carries
  roofline     -- the dominant device operation's algorithmic bytes / its
                  average duration (HIP events recorded on its stream during
                  the timed region) against the 8 TB/s HBM peak, and its HBM
                  traffic per launch from the committed rocprofv3 summary;
  kernels      -- the same for every operation kind;
  cpu_baseline -- the oracle's single-thread C++ restatement of the
                  reference's 1-qubit step on a bounded sample of the same
                  state (one core, CPU model stated).
"""

import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = "bass-basis-adapative-sparse-simulations-_amd"
HBM_PEAK_GBS = 8000.0
SEED = 20260520 * 10 ** 6 + 1


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(keys, amps, n_qubits, k, gates, budget_s):
    """The reference's 1-qubit path (emit, 16-bit LSD radix sort, merge, top-k,
    renormalisation) as the oracle's single-thread C++ -O3 restatement, applied
    gate after gate to the saturated device state until the time budget is
    spent.  Counts the merged entries of each gate, as the device metric does."""
    import numpy as np

    from oracle import port

    n = len(keys)
    x = np.zeros(2 * n, np.uint64)
    a = np.zeros(2 * n, np.complex128)
    x[:n], a[:n] = keys, amps
    tx = np.zeros(2 * n, np.uint64)
    ta = np.zeros(2 * n, np.complex128)
    port.load()
    done = 0
    amp_updates = 0
    t0 = time.perf_counter()
    for g in gates:
        if 2 * n > len(x):
            raise RuntimeError("cpu baseline: state grew beyond its buffers")
        n, _, merged = port.step_1q(x, a, n, g.qubits[0], g.matrix, k, tx, ta)
        amp_updates += merged
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return amp_updates / dt, done, dt


def measured_traffic(workload, n, k, op):
    """HBM bytes per launch of `op` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (line "traffic <op> <bytes>"), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}{n}_K{k.bit_length() - 1}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic" and f[1] == op:
            return float(f[2])
    return None


def _spread(items, m):
    """m items evenly spaced over the list (the one-at-a-time sample)."""
    if not m or not items:
        return []
    step = max(1, len(items) // m)
    return items[::step][:m]


def bench_batched(args, emit=True):
    """Config 3: independent TFIM circuits (n ~ U[16,48], steps ~ U[10,60],
    h ~ U[0.2,2], J = 1, dt = 0.05, K = 2^U{10..18}) run as one batch; baseline: the same circuits one at a time through the plain driver
    (a --single-sample subset, timed, as circuits/s)."""
    import numpy as np
    import torch

    S = importlib.import_module(PKG + ".serving.server")
    Bm = importlib.import_module(PKG + ".simulation.batched")
    Bs = importlib.import_module(PKG + ".simulation.bass_simulator")
    importlib.import_module(PKG + "._lib").load()
    reqs = S.make_batched_requests(args.circuits)
    specs = []
    for r in reqs:
        circ, kw = S.build_circuit(r)
        specs.append((r.n, 1 << r.log2k, circ, kw.get("optimize_at")))
    log(f"batched: {len(specs)} circuits, {sum(len(s[2]) for s in specs)} gates, "
        f"sum K = {sum(s[1] for s in specs)}")
    # one at a time (plain driver), on an evenly spaced sample of the same stream
    sample = _spread(specs, args.single_sample)
    single_rate, t_single = None, 0.0
    if sample:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for n, k, circ, oat in sample:
            sim = Bs.BASS(n, k, device="cuda:0", **({"optimize_at": oat} if oat else {}))
            sim.simulate(circ)
        torch.cuda.synchronize()
        t_single = time.perf_counter() - t0
        single_rate = len(sample) / t_single
        log(f"one at a time: {len(sample)} circuits in {t_single:.2f} s = {single_rate:.2f} circuits/s")
    # the batch
    modes = {"graphs": {}, "pipelined": {"graphs": False}, "sync": {"pipelined": False}}
    engines = {}
    order = [args.batch_engine] + ([m for m in modes if m != args.batch_engine] if args.all_engines else [])
    for name in order:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b = Bm.BatchedBASS(specs, device="cuda:0", **modes[name])
        b.run()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        engines[name] = {"circuits_per_s": len(specs) / t, "seconds": t}
        log(f"batched ({name}): {len(specs)} circuits in {t:.2f} s = {len(specs) / t:.2f} circuits/s")
        if name == args.batch_engine:
            bat, t_batch = b, t
    rate = len(specs) / t_batch
    out = {
        "metric": "batched circuits/s", "value": rate, "unit": "circuits/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex128", "data": "synthetic",
        "config": {"workload": f"batched TFIM circuits={len(specs)} n~U[16,48] steps~U[10,60] h~U[0.2,2] "
                               f"K=2^U{{10..18}} (step = the whole batch)",
                   "circuits": len(specs), "gates": int(sum(len(s[2]) for s in specs)),
                   "amp_updates": int(bat.counters.get("amp_updates", 0)), "engine": args.batch_engine},
        "one_at_a_time": {"value": single_rate, "unit": "circuits/s", "sample": len(sample),
                          "seconds": t_single, "driver": "plain BASS, one circuit after another, "
                                                          "every (circuits/sample)-th circuit of the stream"},
        "speedup_vs_one_at_a_time": (rate / single_rate) if single_rate else None,
        "amp_updates_per_s": bat.counters.get("amp_updates", 0) / t_batch,
        # the same batch through each engine (graphs: pipelined with replayed
        # HIP graphs; pipelined: direct launches; sync: a host read per tick)
        "engines": engines,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_tfim20(args):
    """Config 0: uniform TFIM n=20 (open chain, J=h=1, dt=0.05, 40 steps =
    1,560 gates), K=2^14, optimize_every=39.  Times simulate() on the GPU
    (median of --steps runs after --warmup) and the numpy oracle's simulate()
    on one host core (one run); cache-resident sizes, so no roofline point."""
    import numpy as np
    import torch

    rc = importlib.import_module(PKG + ".utils.random_circuits")
    B = importlib.import_module(PKG + ".simulation.bass_simulator")
    importlib.import_module(PKG + "._lib").load()
    circ = rc.generate_tfim_circuit(20, 40, J=1.0, h=1.0, dt=0.05, rng=np.random.default_rng(0))
    K = 1 << 14

    def run():
        sim = B.BASS(20, K, device="cuda:0", optimize_every=39)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = sim.simulate(circ)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, sim, st

    for _ in range(args.warmup):
        run()
    times = []
    for _ in range(max(1, args.steps)):
        dt, sim, st = run()
        times.append(dt)
        log(f"gpu simulate {dt * 1e3:.1f} ms")
    t_gpu = float(np.median(times))
    amp = sim.counters.get("amp_updates", 0)
    # fidelity against the exact state, both computed on the device
    E = importlib.import_module(PKG + ".simulation.exact_simulator")
    Fm = importlib.import_module(PKG + ".benchmarking.fidelity")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sv = E.exact_statevector_device(20, circ, device="cuda:0")
    fid = Fm.fidelity_bass(sim, st, sv)
    t_exact = time.perf_counter() - t0
    log(f"exact state vector + fidelity on the device {t_exact * 1e3:.1f} ms, F = {fid:.12f}")
    cpu = None
    if not args.no_cpu:
        from oracle import drivers as OD

        ref = OD.BASS(20, K, selector=OD.key_tiebreak_selector, accept_rtol=sim.accept_rtol, optimize_every=39)
        t0 = time.perf_counter()
        rs = ref.simulate(circ)
        t_cpu = time.perf_counter() - t0
        same = bool(rs.nnz == st.nnz and abs(rs.gamma ** 2 - st.gamma ** 2) <= 1e-12)
        cpu = {"value": 1.0 / t_cpu, "unit": "circuits/s", "cores": 1, "kind": "port",
               "sample": f"the whole circuit, numpy oracle simulate() (key tie-break), {t_cpu:.2f} s; "
                         f"same nnz and gamma^2 as the GPU run: {same}; "
                         f"host affinity {len(os.sched_getaffinity(0))} cpus"}
        log(f"cpu oracle simulate {t_cpu:.2f} s")
    out = {
        "metric": "circuit simulations/s", "value": 1.0 / t_gpu, "unit": "circuits/s", "n_gpus": 1,
        "steps": len(times), "warmup": args.warmup, "ms_per_step": t_gpu * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex128", "data": "synthetic",
        "config": {"workload": "TFIM n=20 open J=h=1 dt=0.05 40 steps (1,560 gates) K=2^14 optimize_every=39 "
                               "(step = one simulate())", "amp_updates": amp,
                   "gamma2": st.gamma ** 2, "nnz": st.nnz},
        "fidelity": fid, "exact_plus_fidelity_ms": t_exact * 1e3,
        "amp_updates_per_s": amp / t_gpu,
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_sort(args):
    """bass_radix_sort on the records a 1-qubit gate emits from a saturated
    state: K random N-bit keys k, each emitted as (k & ~m, k | m) with m the
    middle qubit (2K records, the reference's dedup input,
    src/core/sparse_state.py:268-285), complex normal amplitudes.  Every step
    restores the unsorted input (a device copy outside the timed region) and
    sorts it; time = HIP events around the sort.  Algorithmic bytes: one key
    read for the histograms (8 B) + 48 B per digit pass (+ 48 B for the copy
    back of a single pass)."""
    import ctypes

    import numpy as np
    import torch

    L = importlib.import_module(PKG + "._lib")
    dev = importlib.import_module(PKG + "._device")
    lib = L.load()
    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    N = args.qubits or 40
    K = 1 << (args.log2k or 22)
    if N > 62:
        raise SystemExit("--workload sort: N <= 62")
    n = 2 * K
    ctx = dev.context("cuda:0")
    g = torch.Generator(device="cuda:0")
    g.manual_seed(SEED)
    k = torch.randint(0, 1 << N, (K,), device="cuda:0", dtype=torch.int64, generator=g)
    m = 1 << (N // 2)
    x0 = torch.stack((k & ~m, k | m), 1).reshape(-1).contiguous()
    del k
    a0 = torch.randn(n, dtype=torch.complex128, device="cuda:0", generator=g)
    x, a = x0.clone(), a0.clone()
    ctx.ensure_ws(n)
    if args.sort_digits:
        L.check(lib.bass_sort_digit_bits(args.sort_digits), "bass_sort_digit_bits")
    passes = ctypes.c_int(0)
    rb = lib.bass_sort_plan(N, ctypes.byref(passes))
    P = passes.value
    byt = 8 * n + 48 * n * (P + (1 if P == 1 else 0))  # one pass: copied back; odd counts >= 3 rotate

    def one():
        x.copy_(x0)
        a.copy_(a0)
        rc_ = lib.bass_radix_sort(x.data_ptr(), a.data_ptr(), n, N, ctx.wsp, ctx.ws_bytes, ctx.stream)
        L.check(rc_, "bass_radix_sort")

    for _ in range(args.warmup):
        one()
    torch.cuda.synchronize()
    L.timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cnt, ms = L.timing_read("sort")
    L.timing(False)
    ctx.read_stats("sort")
    avg = ms / cnt
    # the last sort against torch's stable sort
    ref_x, idx = torch.sort(x0, stable=True)
    ok = bool(torch.equal(ref_x, x) and torch.equal(a0[idx], a))
    log(f"sort n={n} bits={N} digits={rb}x{P}: {avg:.4f} ms per sort, matches torch stable sort: {ok}")
    if not ok:
        raise SystemExit("bass_radix_sort disagrees with torch.sort(stable=True)")
    gbs = byt / (avg * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": "sort", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS, "traffic": measured_traffic("sort", N, K, "sort"),
                "bytes_per_launch": byt, "avg_ms": avg}
    cpu = None
    if not args.no_cpu:
        from oracle import port

        cn = min(n, 1 << 22)
        hx = x0[:cn].cpu().numpy().view(np.uint64).copy()
        ha = a0[:cn].cpu().numpy().copy()
        port.load()
        reps, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and reps < 64:
            cx, ca = hx.copy(), ha.copy()
            c0 = time.perf_counter()
            port.radix_sort(cx, ca, cn)
            t_cpu += time.perf_counter() - c0
            reps += 1
        cpu = {"value": cn * reps / t_cpu, "unit": "records/s", "cores": 1, "kind": "port",
               "sample": f"{reps} sorts of the first {cn} records (reference radix_sort_uint64: 4 passes of "
                         f"16 bits), oracle C++ -O3 single-thread restatement, {t_cpu:.1f} s, "
                         f"host affinity {len(os.sched_getaffinity(0))} cpus"}
    out = {
        "metric": "radix sort records/s", "value": n / (avg * 1e-3), "unit": "records/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": avg, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64 keys + complex128 payload", "data": "synthetic",
        "config": {"workload": f"radix sort of 2K={n} emitted records, {N}-bit keys (step = one sort; "
                               f"input restore outside the timed region)",
                   "digit_bits": rb, "passes": P, "wall_s_including_restores": wall},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_served(args, emit=True):
    """Config 4: Poisson stream (--rate, or a sweep over --rates) of 50/50 TFIM
    and brickwork circuits, HBM-budget admission (200 GB), p50/p99 latency.
    --server engine: continuous batching (BatchServer, slabs sized from the
    budget); pool: worker processes running one circuit each (BassServer).
    Baseline: a --single-sample subset of the same requests one at a time
    through the plain driver, as circuits/s."""
    import torch

    S = importlib.import_module(PKG + ".serving.server")
    importlib.import_module(PKG + "._lib").load()
    rates = [float(r) for r in args.rates.split(",")] if args.rates else [args.rate]
    single_rate = None
    if args.single_sample:
        sample = _spread(S.make_served_requests(args.circuits, rates[0]), args.single_sample)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in sample:
            S.run_request(r, "cuda:0")
        t_single = time.perf_counter() - t0
        single_rate = len(sample) / t_single
        log(f"one at a time: {len(sample)} circuits in {t_single:.2f} s = {single_rate:.2f} circuits/s")
    sweep = []
    for rate in rates:
        reqs = S.make_served_requests(args.circuits, rate)
        prog = lambda d, t: log(f"served {d}/{t}") if d % 64 == 0 else None  # noqa: E731
        peak = None
        if args.server == "engine":
            ekw = {"graphs": {}, "pipelined": {"graphs": False}, "sync": {"pipelined": False}}[args.batch_engine]
            with S.BatchServer(workers=args.workers, max_circuits=args.max_circuits, hbm_budget=200e9,
                               **ekw) as srv:
                recs = srv.run(reqs, progress=prog)
            summ = S.summarize(recs, amp_updates=srv.counters.get("amp_updates", 0))
            peak = srv.peak_footprint_bytes
            cfg = {"server": "continuous batch", "engine": args.batch_engine, "engines": args.workers,
                   "max_circuits_per_engine": args.max_circuits, "slab_entries_per_engine": srv._cfg["slab_entries"]}
        else:
            with S.BassServer(workers=args.workers, hbm_budget=200e9) as srv:
                recs = srv.run(reqs, progress=prog)
            summ = S.summarize(recs)
            cfg = {"server": "worker pool", "workers": args.workers}
        row = {"rate_req_s": rate, "circuits_per_s": summ["circuits_per_s"], "p50_latency_s": summ["p50_latency_s"],
               "p99_latency_s": summ["p99_latency_s"], "mean_service_s": summ["mean_service_s"],
               "wall_s": summ["wall_s"], "amp_updates": summ["amp_updates"], "peak_footprint_bytes": peak}
        log(f"rate {rate:g} req/s: {summ['circuits_per_s']:.2f} circuits/s, p50 {summ['p50_latency_s']:.2f} s, "
            f"p99 {summ['p99_latency_s']:.2f} s, peak footprint {peak}")
        sweep.append(row)
    head = sweep[0]
    out = {
        "metric": "served circuits/s", "value": head["circuits_per_s"], "unit": "circuits/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": head["wall_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "complex128", "data": "synthetic",
        "config": {"workload": f"served Poisson {rates[0]:g} req/s, {args.circuits} requests, 50/50 TFIM "
                               f"(K=2^U{{12..20}}) / brickwork n~U[24,44] depth~U[10,30]",
                   "hbm_budget_bytes": 200e9, **cfg},
        "p50_latency_s": head["p50_latency_s"], "p99_latency_s": head["p99_latency_s"],
        "mean_service_s": head["mean_service_s"], "amp_updates": head["amp_updates"],
        "peak_footprint_bytes": head["peak_footprint_bytes"],
        "one_at_a_time_circuits_per_s": single_rate,
        "speedup_vs_one_at_a_time": (head["circuits_per_s"] / single_rate) if single_rate else None,
        "sweep": sweep,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_fixed(args):
    """FixedBasisSimulator (the reference's fixed-basis driver: group-table
    deduplicated 1- and 2-qubit gates in input order, top-k with the 2-qubit
    gates' gamma rule) on the brickwork circuit of the plain benchmark, n=40,
    K=2^22: --warmup layers untimed, --steps layers timed, per-operation
    HIP-event timing and algorithmic bytes (gate 24 (n_in + n_out), select
    40 n + 24 k)."""
    import numpy as np
    import torch

    lib = importlib.import_module(PKG + "._lib")
    rc = importlib.import_module(PKG + ".utils.random_circuits")
    Fs = importlib.import_module(PKG + ".simulation.simulator")
    lib.load()
    N, K = args.qubits or 40, 1 << (args.log2k or 22)
    total = 30
    if args.warmup + args.steps > total:
        raise SystemExit(f"--warmup + --steps exceeds the circuit's {total} layers")
    circ, ends = rc.brickwork_haar_circuit(N, total, np.random.default_rng(SEED))
    starts = [0] + [e + 1 for e in ends[:-1]]
    sim = Fs.FixedBasisSimulator(N, K, verbose=False, device="cuda:0")
    sim.begin(len(circ), seed=0)
    log(f"fixed-basis brickwork n={N} K=2^{K.bit_length() - 1}; timing layers {args.warmup}..")
    for layer in range(args.warmup):
        for i in range(starts[layer], ends[layer] + 1):
            sim.step(i, circ[i])
    torch.cuda.synchronize()
    c0 = dict(sim.counters)
    lib.timing(True)
    t0 = time.perf_counter()
    for layer in range(args.warmup, args.warmup + args.steps):
        for i in range(starts[layer], ends[layer] + 1):
            sim.step(i, circ[i])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    c1 = dict(sim.counters)
    delta = {k: c1.get(k, 0) - c0.get(k, 0) for k in c1}
    kernels = {}
    for op in ["gate1q", "gate2q", "select"]:
        cnt, ms = lib.timing_read(op)
        if cnt == 0:
            continue
        byt = delta.get("bytes_" + op, 0)
        gbs = byt / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        kernels[op] = {"launches": cnt, "avg_ms": ms / cnt, "total_ms": ms, "bytes_per_launch": byt / cnt,
                       "achieved_GBps": gbs, "frac_hbm": gbs / HBM_PEAK_GBS}
    lib.timing(False)
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    d = kernels[dom]
    out = {
        "metric": "sparse amp-updates/s", "value": delta.get("amp_updates", 0) / dt, "unit": "amp-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "complex128",
        "data": "synthetic",
        "config": {"workload": f"fixed-basis (hash-table gates) brickwork-haar n={N} layers={total} "
                               f"K=2^{K.bit_length() - 1} (step = 1 layer)",
                   "gate_counts": {k: v for k, v in delta.items() if not k.startswith("bytes_")}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": d["achieved_GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": d["frac_hbm"], "traffic": None, "bytes_per_launch": d["bytes_per_launch"],
                     "avg_ms": d["avg_ms"]},
        "kernels": kernels,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run_plain(workload, N, K, warmup, steps, fused=False, prune=True):
    """The plain driver on the full configured circuit (brickwork: 30 layers;
    kicked Ising: 60 steps): `warmup` steps untimed, then `steps` steps timed
    between two device synchronisations, with HIP-event timing of every
    device operation on its launch stream.  Returns the bench fields."""
    import numpy as np
    import torch

    lib = importlib.import_module(PKG + "._lib")
    rc = importlib.import_module(PKG + ".utils.random_circuits")
    B = importlib.import_module(PKG + ".simulation.bass_simulator")
    if workload == "brickwork":
        total = 30
        circ, ends = rc.brickwork_haar_circuit(N, total, np.random.default_rng(SEED))
        wname = f"brickwork-haar n={N} layers={total} K=2^{K.bit_length() - 1} (step = 1 layer)"
    else:
        total = 60
        circ, ends = rc.kicked_ising_circuit(N, total)
        wname = (f"kicked-ising n={N} periodic steps={total} K=2^{K.bit_length() - 1} "
                 f"(step = RZZ layer + RX layer)")
    if warmup + steps > total:
        raise SystemExit(f"--warmup + --steps exceeds the circuit's {total} steps")
    starts = [0] + [e + 1 for e in ends[:-1]]
    # the bench's fixed accept rule (deterministic on PR-neutral rotations)
    sim = B.BASS(N, K, optimize_at=ends, device="cuda:0", accept_rtol=1e-10)
    sim.fused = fused
    sim.prune = prune and not fused
    sim.begin(len(circ))
    log(f"workload {wname} gates={len(circ)}; timing steps {warmup}..{warmup + steps - 1}")
    wms = []
    for layer in range(warmup):
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        for i in range(starts[layer], ends[layer] + 1):
            sim.step(i, circ[i])
        sim.counters  # books the pending counted gates
        torch.cuda.synchronize()
        wms.append((time.perf_counter() - w0) * 1e3)
        log(f"warmup step {layer} done nnz={sim.state.nnz} {wms[-1]:.1f} ms")
    c0 = dict(sim.counters)
    lib.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for layer in range(warmup, warmup + steps):
        for i in range(starts[layer], ends[layer] + 1):
            sim.step(i, circ[i])
        log(f"step {layer} launched")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    c1 = dict(sim.counters)
    delta = {k: c1.get(k, 0) - c0.get(k, 0) for k in c1}
    amp = delta.get("amp_updates", 0)
    kernels = {}
    for op in ["gate1q", "gate2q", "rot1q", "move", "select", "diag2q", "rdm"]:
        cnt, ms = lib.timing_read(op)
        if cnt == 0:
            continue
        byt = delta.get("bytes_" + op, 0)
        ent = {"launches": cnt, "avg_ms": ms / cnt, "total_ms": ms}
        if byt:
            gbs = byt / (ms * 1e-3) / 1e9
            ent.update({"bytes_per_launch": byt / cnt, "achieved_GBps": gbs, "frac_hbm": gbs / HBM_PEAK_GBS})
        kernels[op] = ent
    lib.timing(False)
    dom = max((k for k in kernels if "achieved_GBps" in kernels[k]), key=lambda k: kernels[k]["total_ms"])
    d = kernels[dom]
    tr_name = f"hbm_traffic_{workload}{N}_K{K.bit_length() - 1}.md"
    roofline = {"bound": "hbm", "kernel": dom, "achieved": d["achieved_GBps"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": d["achieved_GBps"] / HBM_PEAK_GBS,
                "traffic": measured_traffic(workload, N, K, dom),
                "traffic_source": f"profiles/{tr_name}: rocprofv3 FETCH_SIZE / WRITE_SIZE of this workload's "
                                  f"bench run, scaled by a plain copy (not measured in this run)",
                "bytes_per_launch": d["bytes_per_launch"], "avg_ms": d["avg_ms"]}
    return {"value": amp / dt, "ms_per_step": dt * 1e3 / steps, "kernels": kernels, "roofline": roofline,
            "warmup_step_ms": wms,
            "config": {"workload": wname, "qubits": N, "k": K, "steps_timed": steps, "steps_untimed": warmup,
                       "optimize": "every step",
                       "gate_select": ("fused (output not stored)" if fused else
                                       "pruned lift + truncate" if prune else "stored lift + truncate"),
                       "gate_counts": {k: v for k, v in delta.items() if not k.startswith("bytes_")}},
            "_sim": sim, "_circ": circ}


def cpu_sample(sim, circ, workload, N, K, seconds):
    """cpu_baseline: the oracle's C++ restatement of the reference's 1-qubit
    step on the saturated device state (its largest 2^24 entries at larger K)."""
    import numpy as np

    dev = importlib.import_module(PKG + "._device")
    st = sim.state
    keys = st.keys()
    amps = dev.to_host_amps(st._buf.a, st.nnz)
    cpu_k = min(K, 1 << 24)
    if cpu_k < len(keys):
        top = np.argsort(-np.abs(amps))[:cpu_k]
        keys, amps = keys[top], amps[top]
        amps = amps / np.linalg.norm(amps)
    gates = [g for g in circ if g.n_qubits == 1][: 64]
    # the sample applies the frame-rotated gates the driver would see
    gl = []
    for g in gates:
        q = g.qubits[0]
        G = sim.U[q].conj().T @ g.matrix @ sim.U[q]
        gl.append(type("G", (), {"qubits": (q,), "matrix": G, "n_qubits": 1})())
    v, ng, cdt = cpu_baseline(keys, amps, N, min(K, cpu_k), gl, seconds)
    return {"value": v, "unit": "amp-updates/s", "cores": 1, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{ng} 1q gates of the circuit (emit+radix sort+merge+top-k+renorm) on the "
                      f"saturated {workload} n={N} state"
                      + (f" (largest {len(keys)} entries, K=2^{cpu_k.bit_length() - 1})" if cpu_k < K else
                         f" K=2^{K.bit_length() - 1}")
                      + f", oracle C++ -O3 single-thread restatement, {cdt:.1f} s, "
                        f"host affinity {len(os.sched_getaffinity(0))} cpus"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (circuit layers); default: the rest of the circuit after the warm-up "
                         "(kicked Ising 56 of 60, brickwork 25 of 30)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps before them (default 4; brickwork 5)")
    ap.add_argument("--workload", choices=["kicked-ising", "brickwork", "batched", "served", "tfim20", "sort",
                                           "fixed-brickwork"],
                    default="kicked-ising")
    ap.add_argument("--no-secondary", action="store_true",
                    help="kicked-ising: skip the lines reported beside it (brickwork n=40 K=2^22, the batched and "
                         "served configurations)")
    ap.add_argument("--circuits", type=int, default=512, help="batched / served: circuits in the stream")
    ap.add_argument("--single-sample", type=int, default=24,
                    help="batched / served: circuits timed one at a time for the baseline")
    ap.add_argument("--rate", type=float, default=200.0, help="served: Poisson arrival rate (req/s)")
    ap.add_argument("--rates", type=str, default=None, help="served: comma-separated arrival rates to sweep")
    ap.add_argument("--server", choices=["engine", "pool"], default="engine", help="served: server kind")
    ap.add_argument("--batch-engine", choices=["graphs", "pipelined", "sync"], default="graphs",
                    help="batched / served: pipelined engine with HIP graphs, pipelined with direct launches, "
                         "or the synchronous engine")
    ap.add_argument("--max-circuits", type=int, default=512, help="served (engine): circuits running at once")
    ap.add_argument("--all-engines", action="store_true",
                    help="batched: also time the batch through the other two engines")
    ap.add_argument("--workers", type=int, default=None,
                    help="served: worker processes (engine: one batch engine each, default 2; pool: default 8)")
    ap.add_argument("--qubits", type=int, default=None)
    ap.add_argument("--log2k", type=int, default=None)
    ap.add_argument("--calibrate-copy", type=float, default=None, metavar="MIB")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--sort-digits", type=int, default=0, choices=[0, 8, 10],
                    help="sort: force the digit width (0 = automatic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fused", action="store_true",
                    help="kicked-ising / brickwork: run gates and rotation trials through the fused gate select "
                         "(bass_gate_truncate, the gate output never stored) instead of the pruned lift + truncate")
    ap.add_argument("--stored", action="store_true",
                    help="kicked-ising / brickwork: store every gate output before the select (bass_lift + "
                         "bass_truncate) instead of only those the predicted threshold can keep "
                         "(bass_lift_truncate)")
    args = ap.parse_args()
    if args.calibrate_copy is not None or args.workload not in ("kicked-ising", "brickwork"):
        args.warmup = 4 if args.warmup is None else args.warmup
        args.steps = 10 if args.steps is None else args.steps
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs a sparse state on a single MI355X "
              "(independent circuits would shard by circuit across GPUs); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.calibrate_copy is not None:
        nbytes = int(args.calibrate_copy * 2 ** 20)
        src = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": args.steps}), flush=True)
        return

    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "tfim20":
        return bench_tfim20(args)
    if args.workload == "sort":
        return bench_sort(args)
    if args.workload == "fixed-brickwork":
        return bench_fixed(args)
    if args.workload == "served":
        if args.workers is None:
            args.workers = 2 if args.server == "engine" else 8
        return bench_served(args)

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    importlib.import_module(PKG + "._lib").load()
    if args.workload == "brickwork":
        N, K = args.qubits or 40, 1 << (args.log2k or 22)
        args.warmup = 5 if args.warmup is None else args.warmup
        args.steps = 25 if args.steps is None else args.steps
    else:
        N, K = args.qubits or 56, 1 << (args.log2k or 26)
        args.warmup = 4 if args.warmup is None else args.warmup
        args.steps = 56 if args.steps is None else args.steps
    res = run_plain(args.workload, N, K, args.warmup, args.steps, fused=args.fused, prune=not args.stored)
    sim, circ = res.pop("_sim"), res.pop("_circ")
    cpu = None if args.no_cpu else cpu_sample(sim, circ, args.workload, N, K, args.cpu_seconds)
    secondary = None
    if args.workload == "kicked-ising" and not args.no_secondary:
        # the brickwork configuration (n=40, K=2^22), reported beside the headline
        del sim
        torch.cuda.empty_cache()
        sec = run_plain("brickwork", 40, 1 << 22, 5, 25, fused=args.fused, prune=not args.stored)
        sec.pop("_sim"), sec.pop("_circ")
        secondary = {"metric": "sparse amp-updates/s", "value": sec["value"], "unit": "amp-updates/s",
                     "ms_per_step": sec["ms_per_step"], "steps": 25, "warmup": 5, "config": sec["config"],
                     "roofline": sec["roofline"], "kernels": sec["kernels"]}
    batched = served = None
    if args.workload == "kicked-ising" and not args.no_secondary:
        # the batched (512 TFIM circuits) and served (Poisson 200 req/s)
        # configurations, each against the same circuits one at a time
        torch.cuda.empty_cache()
        batched = bench_batched(args, emit=False)
        torch.cuda.empty_cache()
        if args.workers is None:
            args.workers = 2
        served = bench_served(args, emit=False)
    out = {
        "metric": "sparse amp-updates/s",
        "value": res["value"],
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic",
        "config": res["config"],
        "roofline": res["roofline"],
        "kernels": res["kernels"],
        "warmup_step_ms": res["warmup_step_ms"],
        "cpu_baseline": cpu,
        "secondary": secondary,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
