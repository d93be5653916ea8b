"""Autotuner for the per-size-class knobs (see tuning/__init__.py).

    python -m <package>.tuning.autotune [--out FILE] [--reps R]

For each size class it times, on the current GPU with HIP events around the
calls, a 1-qubit and a 2-qubit deduplicating gate (bass_gate, dense output)
on a random state of a representative size, for every table factor and every
emission tile size, and the radix sort of the 2n records a 1-qubit gate emits
for both digit widths; the fastest value of each knob is written to the table
(with the measured times).  Results do not depend on the knobs, so a table
from another device is still correct, only possibly slower.
"""

import argparse
import ctypes
import importlib
import json
import os
import sys

import numpy as np

PKG = __package__.rsplit(".", 1)[0]
REPRESENTATIVE_N = (1 << 12, 1 << 16, 1 << 20, 1 << 23)
NBITS = 40


def _state(n, seed):
    import torch

    rng = np.random.default_rng(seed)
    keys = np.unique(rng.integers(0, 1 << NBITS, size=n + n // 16, dtype=np.uint64))[:n]
    rng.shuffle(keys)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a /= np.linalg.norm(a)
    return (torch.from_numpy(keys.view(np.int64)).to("cuda:0"), torch.from_numpy(a).to("cuda:0"))


def _time(fn, reps):
    import torch

    fn()  # warm-up (compiles nothing; first-touch of buffers)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e3 / reps  # microseconds


def main(argv=None):
    import torch

    from . import KNOBS, TABLE_PATH, VALUES, validate

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)
    L = importlib.import_module(PKG + "._lib")
    dev = importlib.import_module(PKG + "._device")
    lib = L.load()
    ctx = dev.context("cuda:0")
    table = {k: [0] * len(REPRESENTATIVE_N) for k in KNOBS}
    times = {}
    rng = np.random.default_rng(7)
    for c, n in enumerate(REPRESENTATIVE_N):
        x, a = _state(n, 100 + c)
        out = dev.Buffers(4 * n, ctx.device)
        ctx.ensure_ws(4 * n)
        ht = ctypes.c_int64(0)

        def gate(nq):
            G = np.linalg.qr(rng.standard_normal((1 << nq, 1 << nq)) + 1j * rng.standard_normal((1 << nq, 1 << nq)))[0]
            g = np.ascontiguousarray(G.astype(np.complex128)).view(np.float64)
            qs = (11, 27) if nq == 2 else (19, 0)

            def run():
                rc = lib.bass_gate(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(a.data_ptr()), n, nq, qs[0], qs[1],
                                   g.ctypes.data_as(ctypes.c_void_p), L.MODE_GATE, 1, 0, out.xp, out.ap, out.cap,
                                   ctypes.byref(ht), ctx.wsp, ctx.ws_bytes, ctx.stream)
                L.check(rc, "bass_gate")
            return run

        def best(knob, cands, fns):
            res = {}
            for v in cands:
                lib.bass_tune(KNOBS[knob], c, v)
                res[v] = sum(_time(f, args.reps) for f in fns)
            v = min(res, key=res.get)
            lib.bass_tune(KNOBS[knob], c, v)
            table[knob][c] = v
            times[f"{knob}[{c}]"] = {str(k): round(t, 2) for k, t in res.items()}
            print(f"class {c} (n={n}): {knob} -> {v}  {times[f'{knob}[{c}]']} us", flush=True)

        g1, g2 = gate(1), gate(2)
        best("table_factor", VALUES["table_factor"], [g1, g2])
        best("emit_items_1q", VALUES["emit_items_1q"], [g1])
        best("emit_items_2q", VALUES["emit_items_2q"], [g2])
        # the sort of the records a 1-qubit gate emits (2n keys, pairs)
        m = 1 << (NBITS // 2)
        k0 = torch.stack((x & ~m, x | m), 1).reshape(-1).contiguous()
        v0 = torch.randn(2 * n, dtype=torch.complex128, device="cuda:0")
        ks, vs = k0.clone(), v0.clone()
        ctx.ensure_ws(4 * n)

        def sort():
            ks.copy_(k0)
            vs.copy_(v0)
            L.check(lib.bass_radix_sort(ctypes.c_void_p(ks.data_ptr()), ctypes.c_void_p(vs.data_ptr()), 2 * n, NBITS,
                                        ctx.wsp, ctx.ws_bytes, ctx.stream), "bass_radix_sort")
        best("sort_digits", VALUES["sort_digits"], [sort])
        del out, x, a, k0, v0, ks, vs
        torch.cuda.empty_cache()
    ctx.read_stats("autotune")
    doc = {"device": torch.cuda.get_device_name(0), "representative_n": list(REPRESENTATIVE_N), **table,
           "times_us": times, "reps": args.reps}
    validate(doc)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
