"""Circuit server: many independent BASS circuits on one MI355X.

A pool of worker processes shares the GPU.  Each worker owns a HIP stream,
a library workspace and its state buffers, and runs whole circuits through
the plain device path (BASS driver, key-ordered kernels); independent
circuits overlap on the GPU because every worker submits to its own queues,
and the per-gate host work of different circuits runs in parallel instead of
behind one interpreter.

The dispatcher admits a request when a worker is idle and the summed HBM
footprint of the requests in flight stays within the budget (footprint_bytes:
state, gate-output and rotation buffers, workspace).  Requests carry the
parameters of their circuit, not the gates; the worker builds the circuit.

Timing uses time.monotonic() (CLOCK_MONOTONIC, shared by all processes):
latency = completion - arrival, so it includes queueing.
"""

import dataclasses
import math
import multiprocessing as mp
import os
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


def footprint_bytes(req):
    """HBM a request holds while it runs: state buffer (K), gate output (<= 4K
    slots, 2K for 1q), rotation buffer (2K), 24 B each, plus the workspace
    (64 B per entry of the largest operation, 4K, and 8 MB fixed)."""
    K = 1 << req.log2k
    return 24 * (K + 4 * K + 2 * K + K) + 64 * 4 * K + (8 << 20)


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


def summarize(records):
    """circuits/s over the stream, p50 / p99 latency (s)."""
    t0 = min(r["arrival"] for r in records)
    t1 = max(r["end"] for r in records)
    lat = np.array([r["latency"] for r in records])
    return {"circuits": len(records), "wall_s": t1 - t0, "circuits_per_s": len(records) / (t1 - t0),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_service_s": float(np.mean([r["service"] for r in records])),
            "amp_updates": int(sum(r["amp_updates"] for r in records))}
