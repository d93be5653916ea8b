"""Circuit servers: many independent BASS circuits on one MI355X.

BatchServer (the default) runs every admitted circuit in a continuous batch
(simulation.batched.BatchEngine): each tick advances every running circuit by
one operation (a gate or an optimiser step) with one launch sequence per
operation kind, arrivals join between ticks, and a circuit leaves (its result
copied to the host) as soon as it is done.  Admission is FIFO while the
engine has a free slot and slab region and the summed HBM footprint of the
circuits in flight stays within the budget (footprint_bytes).  With several
workers, each worker process runs its own engine and the dispatcher balances
arrivals over them.

BassServer runs whole circuits one per worker process through the plain
device path (BASS driver), each worker with its own HIP stream, workspace and
buffers; it is kept as the one-circuit-per-stream comparison.

Requests carry the parameters of their circuit; build_circuit makes its gate
list when the request arrives (part of its latency) in both servers.

Timing uses time.monotonic() (CLOCK_MONOTONIC, shared by all processes):
latency = completion - arrival, so it includes queueing.
"""

import collections
import dataclasses
import multiprocessing as mp
import queue
import time

import numpy as np

PKG = __name__.rsplit(".", 2)[0]


@dataclasses.dataclass
class CircuitRequest:
    rid: int
    kind: str          # "tfim" or "brickwork"
    n: int             # qubits
    depth: int         # Trotter steps / brickwork layers
    log2k: int         # K = 2^log2k
    h: float = 1.0     # TFIM transverse field
    seed: int = 0
    arrival: float = 0.0  # seconds after the stream starts


FOOTPRINT_PER_ENTRY = 24 * 8 + 64 * 4  # bytes per state entry (K) of footprint_bytes
FOOTPRINT_FIXED = 8 << 20


def footprint_bytes(req):
    """HBM a request holds while it runs: state buffer (K), gate output (<= 4K
    slots, 2K for 1q), rotation buffer (2K), 24 B each, plus the workspace
    (64 B per entry of the largest operation, 4K, and 8 MB fixed)."""
    K = 1 << req.log2k
    return FOOTPRINT_PER_ENTRY * K + FOOTPRINT_FIXED


def slab_entries_for(budget):
    """Slab size (entries) of an engine with this HBM budget: large enough
    that the budget, not the slab, bounds admission (a request of K entries
    counts FOOTPRINT_PER_ENTRY * K bytes against the budget but takes K slab
    entries)."""
    return max(1 << 16, int(budget // FOOTPRINT_PER_ENTRY))


def make_batched_requests(count=512, seed=20260520):
    """Independent TFIM circuits: n ~ U[16,48], steps ~ U[10,60],
    h ~ U[0.2,2.0], J = 1, dt = 0.05, K = 2^U{10..18}; all arrive at 0."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        out.append(CircuitRequest(rid=i, kind="tfim", n=int(rng.integers(16, 49)), depth=int(rng.integers(10, 61)),
                                  log2k=int(rng.integers(10, 19)), h=float(rng.uniform(0.2, 2.0)),
                                  seed=seed * 1000 + i))
    return out


def make_served_requests(count, rate, seed=20260521):
    """Poisson stream (rate req/s): 50/50 TFIM as make_batched_requests with
    K = 2^U{12..20}, and brickwork-Haar circuits with n ~ U[24,44],
    depth ~ U[10,30], K = 2^U{12..20}."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        lk = int(rng.integers(12, 21))
        if rng.random() < 0.5:
            out.append(CircuitRequest(rid=i, kind="tfim", n=int(rng.integers(16, 49)), depth=int(rng.integers(10, 61)),
                                      log2k=lk, h=float(rng.uniform(0.2, 2.0)), seed=seed * 1000 + i, arrival=t))
        else:
            out.append(CircuitRequest(rid=i, kind="brickwork", n=int(rng.integers(24, 45)),
                                      depth=int(rng.integers(10, 31)), log2k=lk, seed=seed * 1000 + i, arrival=t))
    return out


def build_circuit(req):
    """(circuit, BASS keyword arguments) of a request."""
    import importlib

    rc = importlib.import_module(PKG + ".utils.random_circuits")
    if req.kind == "tfim":
        circ = rc.generate_tfim_circuit(req.n, req.depth, J=1.0, h=req.h, dt=0.05,
                                        rng=np.random.default_rng(req.seed))
        return circ, {}
    circ, ends = rc.brickwork_haar_circuit(req.n, req.depth, np.random.default_rng(req.seed))
    return circ, {"optimize_at": ends}


def run_request(req, device):
    """Run one request to completion on the calling process's GPU context."""
    import importlib

    import torch

    B = importlib.import_module(PKG + ".simulation.bass_simulator")
    circ, kw = build_circuit(req)
    sim = B.BASS(req.n, 1 << req.log2k, device=device, **kw)
    st = sim.simulate(circ)
    torch.cuda.current_stream(device).synchronize()
    return {"nnz": int(st.nnz), "gamma2": float(st.gamma ** 2), "gates": len(circ),
            "amp_updates": int(sim.counters.get("amp_updates", 0))}


def _worker(wid, device, inq, outq, blocking_sync):
    import ctypes
    import importlib

    if blocking_sync:
        # waits sleep instead of spinning, so idle workers leave the host cores
        # to the ones launching work (hipDeviceScheduleBlockingSync, before the
        # device is first used)
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipSetDeviceFlags(ctypes.c_uint(0x4))
    import torch

    importlib.import_module(PKG + "._lib").load()
    torch.cuda.set_device(device)
    outq.put(("ready", wid, None))
    while True:
        req = inq.get()
        if req is None:
            break
        t0 = time.monotonic()
        try:
            res = run_request(req, device)
            err = None
        except Exception as e:  # reported to the dispatcher, which raises
            res, err = None, f"{type(e).__name__}: {e}"
        t1 = time.monotonic()
        outq.put(("done", wid, {"rid": req.rid, "start": t0, "end": t1, "result": res, "error": err}))


def _request_cost(req):
    """Work estimate used to balance engine workers: K x qubits x depth."""
    return float(1 << req.log2k) * req.n * req.depth


class RejectedRequest(ValueError):
    """A request that no amount of waiting would admit (its state does not fit
    a slab, or its footprint exceeds the HBM budget)."""


def check_admissible(req, budget, slab_entries=None):
    """Raise RejectedRequest unless `req` fits an empty server: 1..63
    qubits, footprint within the budget and (engines) K within a slab."""
    if not 1 <= req.n <= 63:
        raise RejectedRequest(f"request {req.rid}: {req.n} qubits (1 to 63 supported)")
    fp = footprint_bytes(req)
    if fp > budget:
        raise RejectedRequest(f"request {req.rid}: footprint {fp:.3e} B exceeds the HBM budget {budget:.3e} B")
    if slab_entries is not None and (1 << req.log2k) > slab_entries:
        raise RejectedRequest(f"request {req.rid}: K = 2^{req.log2k} exceeds the slab ({slab_entries} entries)")


def _make_engine(cfg, device):
    """The engine of a server config: PipelinedBatchEngine (HIP graphs unless
    cfg["graphs"] is false) or, with cfg["pipelined"] false, BatchEngine."""
    import importlib

    Bm = importlib.import_module(PKG + ".simulation.batched")
    if cfg.get("pipelined", True):
        return Bm.PipelinedBatchEngine(cfg["max_circuits"], cfg["slab_entries"], device=device,
                                       graphs=cfg.get("graphs", True), **cfg["kw"])
    return Bm.BatchEngine(cfg["max_circuits"], cfg["slab_entries"], device=device, **cfg["kw"])


class _EngineLoop:
    """Admission + ticking of one BatchEngine (FIFO, HBM budget)."""

    def __init__(self, engine, budget):
        self.eng = engine
        self.budget = float(budget)
        self.pending = collections.deque()
        self.used = 0.0
        self.peak = 0.0
        self.inflight = {}  # rid -> (request, footprint, start, gates)

    def add(self, req):
        """Queue a request; RejectedRequest if it could never be admitted
        (FIFO admission would otherwise wait on it forever)."""
        check_admissible(req, self.budget, self.eng.slab_entries)
        circ, kw = build_circuit(req)
        self.pending.append((req, circ, kw))

    def admit(self):
        while self.pending:
            r, circ, kw = self.pending[0]
            fp = footprint_bytes(r)
            if self.eng.active and self.used + fp > self.budget:
                return
            if not self.eng.admit(r.n, 1 << r.log2k, circ, kw.get("optimize_at"), tag=r.rid):
                if not self.eng.active:
                    raise RuntimeError(f"request {r.rid} does not fit the empty engine")
                return
            self.pending.popleft()
            self.used += fp
            self.peak = max(self.peak, self.used)
            self.inflight[r.rid] = (r, fp, time.monotonic(), len(circ))

    def tick(self):
        """One engine tick; returns completion records (start, end, result)."""
        out = []
        for rid, res in self.eng.tick():
            r, fp, t0, gates = self.inflight.pop(rid)
            self.used -= fp
            out.append({"rid": rid, "start": t0, "end": time.monotonic(), "nnz": int(res["nnz"]),
                        "gamma2": res["gamma"] ** 2, "gates": gates})
        return out

    @property
    def busy(self):
        return bool(self.eng.active or self.pending)


def _engine_worker(wid, device, cfg, inq, outq):
    import importlib
    import traceback

    try:
        import torch

        importlib.import_module(PKG + "._lib").load()
        torch.cuda.set_device(device)
        loop = _EngineLoop(_make_engine(cfg, device), cfg["budget"])
        outq.put(("ready", wid, None))
        closing = False
        while True:
            block = not loop.busy and not closing
            try:
                while True:
                    msg = inq.get(timeout=0.5) if block else inq.get_nowait()
                    block = False
                    if msg is None:
                        closing = True
                        break
                    loop.add(msg)
            except queue.Empty:
                pass
            loop.admit()
            if loop.eng.active:
                for rec in loop.tick():
                    outq.put(("done", wid, rec))
            elif closing and not loop.pending:
                break
        outq.put(("counters", wid, {**loop.eng.counters, "peak_footprint_bytes": loop.peak}))
    except Exception:  # reported to the dispatcher, which raises
        outq.put(("error", wid, traceback.format_exc()))


class BatchServer:
    """Continuous-batching server on one GPU.

    workers = 1 runs one engine in this process; workers > 1 runs one engine
    per worker process (each with its own HIP stream and workspace, so one
    engine's host work overlaps the others' kernels) and a dispatcher that
    sends each arrival to the worker with the least outstanding work.
    max_circuits / slab_entries (24 B entries, two slabs) are per engine;
    hbm_budget is shared evenly; slab_entries=None sizes the slabs from it
    (slab_entries_for), so admission is bound by the budget and the slots.
    pipelined / graphs choose the engine (PipelinedBatchEngine with or without
    HIP graph replay, or BatchEngine); engine keywords (optimize_every, ...)
    pass through.  After a run,
    peak_footprint_bytes is the sum over engines of their peak admitted
    footprint.  Use as a context manager."""

    def __init__(self, workers=1, max_circuits=512, slab_entries=None, hbm_budget=200e9, device="cuda:0",
                 pipelined=True, graphs=True, **kw):
        self.workers = int(workers)
        self.hbm_budget = float(hbm_budget)
        self.device = device
        self.counters = {}
        self.peak_footprint_bytes = 0.0
        budget = self.hbm_budget / max(1, self.workers)
        if slab_entries is None:
            slab_entries = slab_entries_for(budget)
        self._cfg = {"max_circuits": int(max_circuits), "slab_entries": int(slab_entries),
                     "budget": budget, "kw": kw, "pipelined": bool(pipelined), "graphs": bool(graphs)}
        self._loop = None
        self._procs = []
        self._inqs = []

    def __enter__(self):
        if self.workers <= 1:
            self._loop = _EngineLoop(_make_engine(self._cfg, self.device), self._cfg["budget"])
            return self
        ctx = mp.get_context("spawn")
        self._outq = ctx.Queue()
        for w in range(self.workers):
            q = ctx.Queue()
            p = ctx.Process(target=_engine_worker, args=(w, self.device, self._cfg, q, self._outq), daemon=True)
            p.start()
            self._inqs.append(q)
            self._procs.append(p)
        ready = 0
        while ready < self.workers:
            kind, w, msg = self._outq.get(timeout=600)
            if kind == "error":
                raise RuntimeError(f"engine worker {w} failed to start:\n{msg}")
            ready += kind == "ready"
        return self

    def __exit__(self, *exc):
        if self._loop is not None:
            self.counters = dict(self._loop.eng.counters)
            self.peak_footprint_bytes = self._loop.peak
            return False
        for q in self._inqs:
            q.put(None)
        left = len(self._procs)
        tot = collections.Counter()
        while left and exc[0] is None:
            try:
                kind, w, msg = self._outq.get(timeout=120)
            except queue.Empty:
                break
            if kind == "counters":
                self.peak_footprint_bytes += msg.pop("peak_footprint_bytes", 0.0)
                tot.update(msg)
                left -= 1
        self.counters = dict(tot)
        for p in self._procs:
            p.join(timeout=60)
            if p.is_alive():
                p.terminate()
        return False

    @property
    def engine(self):
        """The in-process engine (workers = 1)."""
        return self._loop.eng if self._loop is not None else None

    def run(self, requests, progress=None):
        """Serve `requests` (arrival times relative to the start of the
        stream).  Returns one record per request (arrival, start, end, latency,
        service, result summary) in request order."""
        import torch

        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            check_admissible(r, self._cfg["budget"], self._cfg["slab_entries"])
        by_id = {r.rid: r for r in reqs}
        recs = {}
        torch.cuda.synchronize(self.device)
        t_base = time.monotonic()

        def record(w, m):
            r = by_id[m["rid"]]
            arrival = t_base + r.arrival
            recs[r.rid] = {"kind": r.kind, "n": r.n, "depth": r.depth, "log2k": r.log2k, "arrival": arrival,
                           "latency": m["end"] - arrival, "service": m["end"] - m["start"], "worker": w, **m}
            if progress:
                progress(len(recs), len(reqs))

        head = 0
        if self._loop is not None:
            loop = self._loop
            while len(recs) < len(reqs):
                now = time.monotonic() - t_base
                while head < len(reqs) and reqs[head].arrival <= now:
                    loop.add(reqs[head])
                    head += 1
                loop.admit()
                if loop.eng.active:
                    for m in loop.tick():
                        record(0, m)
                elif head < len(reqs):
                    time.sleep(max(0.0, min(0.01, reqs[head].arrival - (time.monotonic() - t_base))))
            return [recs[r.rid] for r in reqs]
        load = [0.0] * self.workers
        while len(recs) < len(reqs):
            now = time.monotonic() - t_base
            while head < len(reqs) and reqs[head].arrival <= now:
                r = reqs[head]
                w = int(np.argmin(load))
                load[w] += _request_cost(r)
                self._inqs[w].put(r)
                head += 1
            wait = 0.5 if head >= len(reqs) else max(1e-3, min(0.5, reqs[head].arrival - now))
            try:
                kind, w, msg = self._outq.get(timeout=wait)
            except queue.Empty:
                continue
            if kind == "error":
                raise RuntimeError(f"engine worker {w} failed:\n{msg}")
            if kind == "done":
                load[w] -= _request_cost(by_id[msg["rid"]])
                record(w, msg)
        return [recs[r.rid] for r in reqs]


class BassServer:
    """Worker pool + dispatcher.  Use as a context manager."""

    def __init__(self, workers=8, hbm_budget=200e9, device="cuda:0", blocking_sync=True):
        self.workers = int(workers)
        self.blocking_sync = bool(blocking_sync)
        self.hbm_budget = float(hbm_budget)
        self.device = device
        self._ctx = mp.get_context("spawn")
        self._outq = self._ctx.Queue()
        self._inqs = []
        self._procs = []

    def __enter__(self):
        for w in range(self.workers):
            q = self._ctx.Queue()
            p = self._ctx.Process(target=_worker, args=(w, self.device, q, self._outq, self.blocking_sync),
                                  daemon=True)
            p.start()
            self._inqs.append(q)
            self._procs.append(p)
        ready = 0
        while ready < self.workers:
            kind, _, _ = self._outq.get(timeout=600)
            ready += kind == "ready"
        return self

    def __exit__(self, *exc):
        for q in self._inqs:
            q.put(None)
        for p in self._procs:
            p.join(timeout=60)
            if p.is_alive():
                p.terminate()
        return False

    def run(self, requests, progress=None):
        """Serve `requests` (arrival times relative to now).  Returns one
        record per request: arrival, start, end (monotonic seconds), latency,
        worker and result."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            check_admissible(r, self.hbm_budget)
        t_base = time.monotonic()
        pending = list(reqs)
        idle = list(range(self.workers))
        inflight = {}  # rid -> footprint
        used = 0.0
        recs = {}
        by_id = {r.rid: r for r in reqs}
        while len(recs) < len(reqs):
            now = time.monotonic() - t_base
            # admit in arrival order while a worker is idle and the budget allows
            while pending and idle and pending[0].arrival <= now:
                fp = footprint_bytes(pending[0])
                if inflight and used + fp > self.hbm_budget:
                    break
                r = pending.pop(0)
                w = idle.pop(0)
                inflight[r.rid] = fp
                used += fp
                self._inqs[w].put(r)
            wait = 0.5
            if pending and idle:
                wait = max(0.0, min(wait, pending[0].arrival - now))
            try:
                kind, w, msg = self._outq.get(timeout=wait if wait > 0 else 1e-3)
            except queue.Empty:
                continue
            if kind != "done":
                continue
            if msg["error"]:
                raise RuntimeError(f"request {msg['rid']} failed in worker {w}: {msg['error']}")
            idle.append(w)
            used -= inflight.pop(msg["rid"])
            r = by_id[msg["rid"]]
            arrival = t_base + r.arrival
            recs[msg["rid"]] = {"rid": r.rid, "kind": r.kind, "n": r.n, "depth": r.depth, "log2k": r.log2k,
                                "arrival": arrival, "start": msg["start"], "end": msg["end"],
                                "latency": msg["end"] - arrival, "service": msg["end"] - msg["start"],
                                "worker": w, **msg["result"]}
            if progress:
                progress(len(recs), len(reqs))
        return [recs[r.rid] for r in reqs]


def summarize(records, amp_updates=None):
    """circuits/s over the stream, p50 / p99 latency (s); amp_updates: the
    total when records do not carry per-request counts."""
    t0 = min(r["arrival"] for r in records)
    t1 = max(r["end"] for r in records)
    lat = np.array([r["latency"] for r in records])
    return {"circuits": len(records), "wall_s": t1 - t0, "circuits_per_s": len(records) / (t1 - t0),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_service_s": float(np.mean([r["service"] for r in records])),
            "amp_updates": int(amp_updates if amp_updates is not None
                               else sum(r.get("amp_updates", 0) for r in records))}
