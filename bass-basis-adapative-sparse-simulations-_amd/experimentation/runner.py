"""Experiment harness: circuit factories, single trials, sweeps, statistics and
the on-disk result cache of the reference's benchmark notebooks, running every
simulation on the device.

Follows src/experimentation/runner.py of the reference:
  factories             make_brickwork (:114-120), make_brickwork_2d (:123-159),
                        make_qft (:168-185), make_tfim (:188-207),
                        make_rfim (:210-234), make_qaoa (:466-609)
  exact reference       exact_statevector (:615-617), fidelity_bass (:620-626)
  trials                run_trial_fixed (:697-708), run_trial (:711-753),
                        run_trial_timed (:756-860), run_trial_bass (:863-877)
  sweeps                run_comparison (:880-945), run_sweep_k (:948-1026),
                        run_n_scaling_full (:1029-1144), run_sweep_full (:1147-1268)
  statistics            compute_metrics (:632-694), bootstrap_ci (:1271-1297),
                        trial_stats (:1300-1391), ratio_stats (:1394-1440)
and the notebooks' cache pattern (tests/runtimeOverhead.ipynb cell 3,
run_or_load): a result is stored as a CSV of rows and/or an .npz of arrays
next to a _meta.json holding the configuration; it is reloaded only when the
stored configuration equals the requested one.

Seeds, the order of random draws and the gate order are the reference's, so a
sweep reproduces the reference's stored per-trial results (RFIM and QAOA use
numpy Generators only).  Brickwork circuits draw their Haar 2-qubit gates with
scipy's unitary_group from a per-gate seed, as the reference does; the matrices
then depend on the scipy version (the reference pins 1.17.1).

Timings bracket simulate() with device synchronisations, so they measure the
whole device run.  Plots and the notebooks' table printers are not part of the
harness.
"""

import csv
import json
import os
import time

import numpy as np
import torch
from scipy import stats as _sstats
from scipy.stats import unitary_group

from ..benchmarking.fidelity import compute_fidelity, compute_participation_ratio
from ..benchmarking.fidelity import fidelity_bass as _fidelity_bass_device
from ..core.gates import HGate, RXGate, RZZGate, TwoQubitGate
from ..simulation.bass_simulator import BASS
from ..simulation.exact_simulator import exact_statevector as _exact_device
from ..simulation.simulator import FixedBasisSimulator
from ..utils.random_circuits import generate_rfim_circuit, generate_tfim_circuit

# ------------------------------------------------------------------ factories


def _haar_2q(qubit1, qubit2, seed):
    """Haar-random 2-qubit gate drawn like the reference's RandomTwoQubitGate
    (gates.py:142-149): scipy's unitary_group from a RandomState seeded with
    `seed` (the reference seeds numpy's global state, which is the same stream;
    here the global state is left alone)."""
    m = unitary_group.rvs(4, random_state=np.random.RandomState(seed))
    return TwoQubitGate(qubit1, qubit2, np.asarray(m, dtype=np.complex128))


def make_brickwork(N, depth, rng):
    """runner.py:114-120: per layer, Haar 2-qubit gates on (i, i+1) for
    i = layer % 2, layer % 2 + 2, ...; one seed drawn from rng per gate."""
    gates = []
    for layer in range(depth):
        for i in range(layer % 2, N - 1, 2):
            gates.append(_haar_2q(i, i + 1, int(rng.integers(0, 2 ** 31))))
    return gates


def make_brickwork_2d(rows, cols, depth, rng):
    """runner.py:123-159: rows x cols grid, row-major qubits; layer type
    layer % 4 = horizontal even / horizontal odd / vertical even / vertical
    odd pairs."""
    gates = []
    for layer in range(depth):
        kind = layer % 4
        if kind < 2:
            pairs = [(r * cols + c, r * cols + c + 1) for r in range(rows) for c in range(kind, cols - 1, 2)]
        else:
            pairs = [(r * cols + c, (r + 1) * cols + c) for r in range(kind - 2, rows - 1, 2) for c in range(cols)]
        for q1, q2 in pairs:
            gates.append(_haar_2q(q1, q2, int(rng.integers(0, 2 ** 31))))
    return gates


def make_haar(N, depth, rng):
    """runner.py:162-165: random-pairing Haar circuit; one seed drawn from rng
    reseeds numpy's legacy global generator, which generate_random_circuit
    draws from."""
    from ..utils.random_circuits import generate_random_circuit

    np.random.seed(int(rng.integers(0, 2 ** 31)))
    return generate_random_circuit(N, depth)


def make_qft(N, rng=None):
    """runner.py:168-185: H on qubit i, then controlled phases pi / 2^(j-i) to
    the next four qubits."""
    gates = []
    for i in range(N):
        gates.append(HGate(i))
        for j in range(i + 1, min(i + 5, N)):
            cp = np.eye(4, dtype=np.complex128)
            cp[3, 3] = np.exp(1j * (np.pi / (2 ** (j - i))))
            gates.append(TwoQubitGate(i, j, cp))
    return gates


def make_tfim(N, rng, layers=5, J=1.0, h=3.0, dt=0.3, disorder_strength_J=0.0, disorder_strength_h=0.0):
    """runner.py:188-207 (generate_tfim_circuit)."""
    return generate_tfim_circuit(N, layers, J=J, h=h, dt=dt, rng=rng, disorder_strength_J=disorder_strength_J,
                                 disorder_strength_h=disorder_strength_h)


def make_rfim(N, rng, layers=5, J=1.0, h0=1.0, W=2.0, dt=0.2):
    """runner.py:210-234 (generate_rfim_circuit): one disorder realisation per
    call."""
    return generate_rfim_circuit(N, layers, J=J, h0=h0, W=W, dt=dt, rng=rng)


def _random_3regular_graph(N, rng, max_tries=500):
    """runner.py:466-497: pairing model.  Three stubs per vertex, shuffled by
    rng, paired in order; retried on a self-loop or a repeated edge.  The edge
    list is the list of the set the pairs were inserted into."""
    if N % 2:
        raise ValueError(f"3-regular graph requires even N; got {N}")
    if N < 4:
        raise ValueError(f"3-regular graph requires N >= 4; got {N}")
    for _ in range(max_tries):
        stubs = list(range(N)) * 3
        rng.shuffle(stubs)
        edges = set()
        ok = True
        for s in range(0, len(stubs), 2):
            u, v = stubs[s], stubs[s + 1]
            e = (min(u, v), max(u, v))
            if u == v or e in edges:
                ok = False
                break
            edges.add(e)
        if ok:
            return list(edges)
    raise RuntimeError(f"no 3-regular graph on N={N} vertices after {max_tries} attempts")


def make_qaoa(N, rounds, rng):
    """runner.py:500-609: H on every qubit, then per round RZZ(gamma_k / 2) on
    every edge of one random 3-regular graph and RX(-beta_k) on every qubit;
    gamma_k = (pi/4) sin((2k-1) pi / 4p) + U(-0.05, 0.05), then
    beta_k = (pi/4) cos((2k-1) pi / 4p) + U(-0.05, 0.05) (all gammas drawn
    before the betas)."""
    p = rounds
    edges = _random_3regular_graph(N, rng)
    amp = np.pi / 4.0
    gammas = [amp * np.sin((2 * k - 1) * np.pi / (4 * p)) + rng.uniform(-0.05, 0.05) for k in range(1, p + 1)]
    betas = [amp * np.cos((2 * k - 1) * np.pi / (4 * p)) + rng.uniform(-0.05, 0.05) for k in range(1, p + 1)]
    circuit = [HGate(q) for q in range(N)]
    for k in range(p):
        circuit += [RZZGate(i, j, gammas[k] / 2.0) for i, j in edges]
        circuit += [RXGate(q, -betas[k]) for q in range(N)]
    return circuit


# ------------------------------------------------------------ exact reference


def exact_statevector(N, circuit):
    """runner.py:615-617: dense state vector (host array) from the device's
    exact simulator; N <= 24."""
    return _exact_device(N, circuit)


def fidelity_bass(sim, state, exact_sv):
    """runner.py:620-626: true fidelity of a BASS state (0 when N > 24)."""
    if sim.N > 24:
        return 0.0
    return _fidelity_bass_device(sim, state, exact_sv)


# ----------------------------------------------------------------- statistics


def compute_metrics(data_array, n_boot=4000, seed=42):
    """runner.py:632-694: percentiles, arithmetic mean and SEM, geometric mean
    of the positive values with a bootstrap 95 % interval of the mean log."""
    data = np.asarray(data_array, dtype=float)
    data = data[~np.isnan(data)]
    keys = ["median", "iqr_25", "iqr_75", "p10", "p90", "arithmetic_mean", "sem", "geometric_mean", "gm_ci95_lo",
            "gm_ci95_hi"]
    if len(data) == 0:
        return {k: np.nan for k in keys}
    p10, p25, med, p75, p90 = np.percentile(data, [10, 25, 50, 75, 90])
    sem = _sstats.sem(data) if len(data) > 1 else 0.0
    pos = data[data > 0]
    if len(pos):
        logs = np.log(pos)
        gm = np.exp(np.mean(logs))
        if len(pos) > 1:
            boots = np.random.default_rng(seed=seed).choice(logs, size=(n_boot, len(logs)), replace=True)
            bm = np.mean(boots, axis=1)
            lo, hi = np.exp(np.percentile(bm, 2.5)), np.exp(np.percentile(bm, 97.5))
        else:
            lo = hi = gm
    else:
        gm = lo = hi = np.nan
    vals = [med, p25, p75, p10, p90, np.mean(data), sem, gm, lo, hi]
    return dict(zip(keys, vals))


def bootstrap_ci(data, stat_fn=None, n_boot=4000, ci_level=95, seed=0):
    """runner.py:1271-1297: percentile bootstrap interval of stat_fn (default:
    the median) over resamples of the finite values."""
    if stat_fn is None:
        stat_fn = lambda x: np.median(x, axis=1)  # noqa: E731
    data = np.asarray(data, dtype=float)
    data = data[np.isfinite(data)]
    if len(data) == 0:
        return np.nan, np.nan
    if len(data) == 1:
        v = float(np.asarray(stat_fn(data[np.newaxis, :])).ravel()[0])
        return v, v
    boots = np.random.default_rng(seed).choice(data, size=(n_boot, len(data)), replace=True)
    bs = stat_fn(boots)
    a = (100 - ci_level) / 2.0
    return float(np.percentile(bs, a)), float(np.percentile(bs, 100 - a))


_TRIAL_FIELDS = ["n", "median", "q10", "q25", "q75", "q90", "mean", "std", "sem", "geomean", "gm_ci_lo", "gm_ci_hi",
                 "med_ci_lo", "med_ci_hi"]


def trial_stats(data, n_boot=4000, ci_level=95, seed=0):
    """runner.py:1300-1391: distribution summary of one array of per-trial
    values (finite values only)."""
    data = np.asarray(data, dtype=float)
    data = data[np.isfinite(data)]
    n = len(data)
    if n == 0:
        d = {k: np.nan for k in _TRIAL_FIELDS}
        d["n"] = 0
        return d
    p10, p25, med, p75, p90 = np.percentile(data, [10, 25, 50, 75, 90])
    std = float(np.std(data, ddof=1)) if n > 1 else 0.0
    med_lo, med_hi = bootstrap_ci(data, n_boot=n_boot, ci_level=ci_level, seed=seed)
    pos = data[data > 0]
    if len(pos) >= 2:
        gm = float(np.exp(np.mean(np.log(pos))))
        gm_lo, gm_hi = bootstrap_ci(pos, stat_fn=lambda x: np.exp(np.mean(np.log(np.maximum(x, 1e-300)), axis=1)),
                                    n_boot=n_boot, ci_level=ci_level, seed=seed)
    elif len(pos) == 1:
        gm = gm_lo = gm_hi = float(pos[0])
    else:
        gm = gm_lo = gm_hi = np.nan
    return dict(n=n, median=float(med), q10=float(p10), q25=float(p25), q75=float(p75), q90=float(p90),
                mean=float(np.mean(data)), std=std, sem=std / np.sqrt(n), geomean=gm, gm_ci_lo=float(gm_lo),
                gm_ci_hi=float(gm_hi), med_ci_lo=float(med_lo), med_ci_hi=float(med_hi))


def ratio_stats(adaptive_arr, fixed_arr, n_boot=4000, ci_level=95, seed=0):
    """runner.py:1394-1440: paired adaptive / fixed ratios over the trials where
    both are finite and fixed > the smallest normal double."""
    a = np.asarray(adaptive_arr, dtype=float)
    f = np.asarray(fixed_arr, dtype=float)
    keep = np.isfinite(a) & np.isfinite(f) & (f > np.finfo(float).tiny)
    a, f = a[keep], f[keep]
    if len(a) == 0:
        return {"win_rate": np.nan, "n_trials": 0, "ratio_trials": np.array([])}
    ratio = a / f
    win = a > f
    rs = trial_stats(ratio, n_boot=n_boot, ci_level=ci_level, seed=seed)
    out = dict(ratio_trials=ratio, delta_trials=a - f, win_rate=float(np.mean(win)), win_count=int(np.sum(win)),
               n_trials=len(ratio))
    out.update({f"ratio_{k}": v for k, v in rs.items()})
    return out


# --------------------------------------------------------------------- trials


def _timed(fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0


def run_trial_fixed(N, k, circuit, exact_sv=None):
    """runner.py:697-708: one fixed-basis simulation: runtime, fidelity, PR."""
    sim = FixedBasisSimulator(N, k, verbose=False)
    state, dt = _timed(lambda: sim.simulate(circuit))
    m = {"runtime_s": dt}
    if exact_sv is not None:
        m["fidelity"] = compute_fidelity(state, exact_sv)
    m["pr"] = compute_participation_ratio(state)
    return m


def run_trial_bass(N, k, circuit, exact_sv=None, **bass_kwargs):
    """runner.py:863-877: one BASS simulation: runtime, fidelity, PR."""
    sim = BASS(N, k, verbose=False, **bass_kwargs)
    state, dt = _timed(lambda: sim.simulate(circuit))
    m = {"runtime_s": dt}
    if exact_sv is not None:
        m["fidelity"] = fidelity_bass(sim, state, exact_sv)
    m["pr"] = compute_participation_ratio(state)
    return m


def run_trial(circuit_gen, N, k, seed=0, bass_kwargs=None, _exact_sv=None):
    """runner.py:711-753: one circuit from circuit_gen(N, default_rng(seed)),
    simulated by the fixed-basis simulator and by BASS; returns
    (f_fixed, f_bass, pr_fixed, pr_bass)."""
    r = run_trial_timed(circuit_gen, N, k, seed=seed, bass_kwargs=bass_kwargs, _exact_sv=_exact_sv)
    return r["f_fixed"], r["f_bass"], r["pr_fixed"], r["pr_bass"]


def run_trial_timed(circuit_gen, N, k, seed=0, bass_kwargs=None, _exact_sv=None):
    """runner.py:756-860: as run_trial, with the simulate() times of both
    simulators; every metric comes from the timed runs themselves."""
    bass_kwargs = bass_kwargs or {}
    circuit = circuit_gen(N, np.random.default_rng(seed))
    if _exact_sv is None:
        _exact_sv = exact_statevector(N, circuit)
    fsim = FixedBasisSimulator(N, k, verbose=False)
    fstate, t_fixed = _timed(lambda: fsim.simulate(circuit))
    f_fixed = compute_fidelity(fstate, _exact_sv)
    pr_fixed = compute_participation_ratio(fstate)
    bsim = BASS(N, k, **bass_kwargs)
    bstate, t_bass = _timed(lambda: bsim.simulate(circuit))
    f_bass = fidelity_bass(bsim, bstate, _exact_sv)
    pr_bass = compute_participation_ratio(bstate)
    return {"f_fixed": float(f_fixed), "f_bass": float(f_bass), "pr_fixed": float(pr_fixed),
            "pr_bass": float(pr_bass), "t_fixed": float(t_fixed), "t_bass": float(t_bass), "circuit": circuit}


# --------------------------------------------------------------------- sweeps


def run_comparison(N, k, depth, n_trials, circuit_type="brickwork", bass_kwargs=None):
    """runner.py:880-945: n_trials brickwork circuits from one
    default_rng(42), fixed vs BASS; win rate and compute_metrics of every
    quantity (flattened as <quantity>_<metric>)."""
    if circuit_type != "brickwork":
        raise ValueError("run_comparison: only circuit_type='brickwork' is provided")
    bass_kwargs = bass_kwargs or {}
    rng = np.random.default_rng(42)
    cols = {name: np.zeros(n_trials) for name in ("f_fixed", "f_bass", "pr_fixed", "pr_bass", "ratio")}
    for i in range(n_trials):
        circ = make_brickwork(N, depth, rng)
        sv = exact_statevector(N, circ)
        rf = run_trial_fixed(N, k, circ, sv)
        rb = run_trial_bass(N, k, circ, sv, **bass_kwargs)
        cols["f_fixed"][i], cols["f_bass"][i] = rf["fidelity"], rb["fidelity"]
        cols["pr_fixed"][i], cols["pr_bass"][i] = rf["pr"], rb["pr"]
        cols["ratio"][i] = rb["fidelity"] / rf["fidelity"] if rf["fidelity"] > 1e-12 else np.nan
    out = {"N": N, "k": k, "depth": depth, "trials": n_trials,
           "win_rate": float(np.sum(cols["f_bass"] > cols["f_fixed"]) / n_trials)}
    for name, arr in cols.items():
        for key, val in compute_metrics(arr).items():
            out[f"{name}_{key}"] = val
    return out


def run_sweep_k(circuit_gen, N, k_values, n_trials=1, seed_base=0, bass_kwargs=None, verbose=True,
                cache_exact=True):
    """runner.py:948-1026: trial t uses seed seed_base * 10^6 + t at every k;
    returns the per-k means of F and PR."""
    bass_kwargs = bass_kwargs or {}
    exact = {}
    if cache_exact:
        for t in range(n_trials):
            exact[t] = exact_statevector(N, circuit_gen(N, np.random.default_rng(seed_base * 10 ** 6 + t)))
    res = {"k_values": k_values, "f_fixed": [], "f_bass": [], "pr_fixed": [], "pr_bass": []}
    for k in k_values:
        acc = [[], [], [], []]
        for t in range(n_trials):
            vals = run_trial(circuit_gen, N, k, seed=seed_base * 10 ** 6 + t, bass_kwargs=bass_kwargs,
                             _exact_sv=exact.get(t))
            for lst, v in zip(acc, vals):
                lst.append(v)
            if verbose:
                ff, fb = vals[0], vals[1]
                ratio = fb / ff if ff > 1e-30 else float("nan")
                print(f"  k={k:>8,}  trial={t}  FixedBasis={ff:.4f}  BASS={fb:.4f}  ratio={ratio:.2f}x  "
                      f"pr_fixed={vals[2]:.0f}  pr_bass={vals[3]:.0f}", flush=True)
        for key, lst in zip(("f_fixed", "f_bass", "pr_fixed", "pr_bass"), acc):
            res[key].append(np.mean(lst))
    return res


def _column_stats(f_fixed, f_bass, pr_fixed, pr_bass, col, n_boot):
    ratio = f_bass[:, col] / np.maximum(f_fixed[:, col], 1e-300)
    return {"f_fixed": trial_stats(f_fixed[:, col], n_boot=n_boot),
            "f_bass": trial_stats(f_bass[:, col], n_boot=n_boot),
            "pr_fixed": trial_stats(pr_fixed[:, col], n_boot=n_boot),
            "pr_bass": trial_stats(pr_bass[:, col], n_boot=n_boot),
            "ratio": trial_stats(ratio, n_boot=n_boot)}


def _progress(label, f_fixed, f_bass, col):
    r = f_bass[:, col] / np.maximum(f_fixed[:, col], 1e-300)
    pos = r[r > 0]
    gm = float(np.exp(np.mean(np.log(pos)))) if len(pos) else np.nan
    win = float(np.mean(f_bass[:, col] > f_fixed[:, col]))
    print(f"  {label}: ratio_gm={gm:.3f}x  win={100 * win:.0f}%", flush=True)


def run_n_scaling_full(circuit_gen, N_values, k_fn, n_trials, seed_base, bass_kwargs=None, verbose=True,
                       n_boot=2000):
    """runner.py:1029-1144: for every N (index ni) and trial t the circuit of
    seed seed_base * 10^6 + ni * 1000 + t at k = k_fn(N); per-trial arrays of
    shape (n_trials, len(N_values)) and trial_stats per N."""
    bass_kwargs = bass_kwargs or {}
    k_values = [int(k_fn(N)) for N in N_values]
    shape = (n_trials, len(N_values))
    f_fixed, f_bass, pr_fixed, pr_bass = (np.full(shape, np.nan) for _ in range(4))
    for ni, N in enumerate(N_values):
        seeds = [seed_base * 10 ** 6 + ni * 1000 + t for t in range(n_trials)]
        svs = [exact_statevector(N, circuit_gen(N, np.random.default_rng(s))) for s in seeds]
        for t, s in enumerate(seeds):
            f_fixed[t, ni], f_bass[t, ni], pr_fixed[t, ni], pr_bass[t, ni] = run_trial(
                circuit_gen, N, k_values[ni], seed=s, bass_kwargs=bass_kwargs, _exact_sv=svs[t])
        if verbose:
            _progress(f"N={N} k={k_values[ni]}", f_fixed, f_bass, ni)
    st = {N: _column_stats(f_fixed, f_bass, pr_fixed, pr_bass, ni, n_boot) for ni, N in enumerate(N_values)}
    return dict(N_values=N_values, k_values=k_values, f_fixed=f_fixed, f_bass=f_bass, pr_fixed=pr_fixed,
                pr_bass=pr_bass, stats=st)


def run_sweep_full(circuit_gen, N, k_values, n_trials, seed_base, bass_kwargs=None, verbose=True, n_boot=2000):
    """runner.py:1147-1268: trial t is the circuit of seed seed_base * 10^6 + t
    at every k; per-trial arrays of shape (n_trials, len(k_values)) and
    trial_stats per k."""
    bass_kwargs = bass_kwargs or {}
    k_values = np.asarray(k_values, dtype=int)
    shape = (n_trials, len(k_values))
    f_fixed, f_bass, pr_fixed, pr_bass = (np.full(shape, np.nan) for _ in range(4))
    seeds = [seed_base * 10 ** 6 + t for t in range(n_trials)]
    svs = []
    for t, s in enumerate(seeds):
        svs.append(exact_statevector(N, circuit_gen(N, np.random.default_rng(s))))
        if verbose:
            print(f"  [exact] trial {t}/{n_trials}", flush=True)
    for ki, k in enumerate(k_values):
        for t, s in enumerate(seeds):
            f_fixed[t, ki], f_bass[t, ki], pr_fixed[t, ki], pr_bass[t, ki] = run_trial(
                circuit_gen, N, int(k), seed=s, bass_kwargs=bass_kwargs, _exact_sv=svs[t])
        if verbose:
            _progress(f"k={int(k):>8,}", f_fixed, f_bass, ki)
    st = {int(k): _column_stats(f_fixed, f_bass, pr_fixed, pr_bass, ki, n_boot) for ki, k in enumerate(k_values)}
    return dict(k_values=k_values, f_fixed=f_fixed, f_bass=f_bass, pr_fixed=pr_fixed, pr_bass=pr_bass, stats=st)


# ---------------------------------------------------------------- result cache


def cache_paths(name, cache_dir):
    """(csv, npz, meta) paths of a named result in cache_dir."""
    base = os.path.join(cache_dir, name)
    return base + ".csv", base + ".npz", base + "_meta.json"


def _jsonable(v):
    if isinstance(v, dict):
        return {str(k): _jsonable(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_jsonable(x) for x in v]
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, np.generic):
        return v.item()
    return v


def _parse_cell(v, int_cols, str_cols, k):
    if k in str_cols:
        return v
    if k in int_cols:
        return int(v)
    return float(v)


def run_or_load(name, meta, compute, cache_dir, force=False, int_cols=(), str_cols=()):
    """Notebook cache pattern (tests/runtimeOverhead.ipynb cell 3, run_or_load).

    compute() -> (rows, arrays): rows a list of dicts (written as <name>.csv) or
    None, arrays a dict of numpy arrays (written as <name>.npz) or None.  The
    result is reloaded when every file the previous run wrote exists and
    <name>_meta.json equals `meta` (compared after a JSON round trip); else it
    is recomputed and the files are rewritten.  CSV cells load as float except
    the columns named in int_cols / str_cols.  Returns (rows, arrays)."""
    csv_p, npz_p, meta_p = cache_paths(name, cache_dir)
    want = json.loads(json.dumps(_jsonable(meta)))
    if not force and os.path.exists(meta_p):
        stored = json.load(open(meta_p, encoding="utf-8"))
        files = stored.get("_files", [])
        if {k: v for k, v in stored.items() if k != "_files"} == want and all(
                os.path.exists(p) for p in ([csv_p] if "csv" in files else []) + ([npz_p] if "npz" in files else [])):
            rows = None
            if "csv" in files:
                with open(csv_p, newline="", encoding="utf-8") as fh:
                    rows = [{k: _parse_cell(v, int_cols, str_cols, k) for k, v in r.items()}
                            for r in csv.DictReader(fh)]
            arrays = dict(np.load(npz_p, allow_pickle=False)) if "npz" in files else None
            return rows, arrays
    rows, arrays = compute()
    os.makedirs(cache_dir, exist_ok=True)
    files = []
    if rows:
        with open(csv_p, "w", newline="", encoding="utf-8") as fh:
            w = csv.DictWriter(fh, fieldnames=list(rows[0].keys()))
            w.writeheader()
            w.writerows(rows)
        files.append("csv")
    if arrays:
        np.savez_compressed(npz_p, **{k: np.asarray(v) for k, v in arrays.items()})
        files.append("npz")
    with open(meta_p, "w", encoding="utf-8") as fh:
        json.dump(dict(want, _files=files), fh, indent=2)
    return rows, arrays


def sweep_arrays(result, prefix=""):
    """The per-trial arrays of a run_sweep_full / run_n_scaling_full result
    under the fixture naming of the reference's notebooks
    (<prefix>f_fixed, <prefix>f_bass, <prefix>pr_fixed, <prefix>pr_bass)."""
    return {prefix + key: np.asarray(result[key]) for key in ("f_fixed", "f_bass", "pr_fixed", "pr_bass")}
