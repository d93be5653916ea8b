"""The experiment harness on the device drivers.

Against per-trial results the reference stored (tests/golden/reference_results):
  nscaling_fixed_k.npz     run_n_scaling_full, RFIM W=2, k = 8192 >= 2^N (no
                           truncation): every trial's F and PR of both simulators;
  nscaling_adaptive_k.npz  run_n_scaling_full, RFIM N=10, k = 2^(N-1) (truncating):
                           100-trial means.  From |0...0> the RFIM state's |a|^2
                           values tie exactly at the truncation thresholds; the
                           reference breaks those ties with np.argpartition, this
                           build by the smaller key, so single trials differ and
                           the means must agree;
  phase_diagram_extended   N=16 RFIM and QAOA exact-state PR and entropy per
                           trial; the RFIM rows of run_trial_timed (K=4096,
                           default BASS arguments) as 20-trial means.
Against the oracle on the same circuits:
  run_sweep_full on brickwork circuits (Haar 2-qubit gates: no |a|^2 ties, so
  the trajectories do not hinge on last-bit differences), truncating at two
  budgets: per trial F and PR of both simulators.  (Exact ties, as from
  |0...0> in the RFIM runs, are decided by the last bit of each |a|^2, which
  differs between numpy on the host and the device, so those are compared in
  the mean above, or gate by gate in the driver tests.)
"""

import os

import numpy as np
import pytest

from conftest import REF_RESULTS, pkg_module

pytestmark = pytest.mark.gpu

SEED_BASE = 20260520
KW = {"optimize_every": 5, "buffer_factor": 1}


def _rfim(N, rng):
    return pkg_module("experimentation.runner").make_rfim(N, rng, layers=5, W=2.0, dt=0.2)


def test_n_scaling_full_no_truncation(gpu_ctx):
    print("[case] harness n-scaling N=10,12", flush=True)
    R = pkg_module("experimentation.runner")
    d = np.load(os.path.join(REF_RESULTS, "nscaling_fixed_k.npz"), allow_pickle=False)
    T = 4
    res = R.run_n_scaling_full(_rfim, [10, 12], lambda N: 8192, T, SEED_BASE + 1, bass_kwargs=KW, verbose=False,
                               n_boot=200)
    assert res["k_values"] == [8192, 8192]
    for key, fx in [("f_fixed", "RFIM_(Weq2)_f_fixed"), ("f_bass", "RFIM_(Weq2)_f_bass"),
                    ("pr_fixed", "RFIM_(Weq2)_pr_fixed"), ("pr_bass", "RFIM_(Weq2)_pr_bass")]:
        np.testing.assert_allclose(res[key], d[fx][:T, :2], rtol=1e-10, atol=1e-12, err_msg=key)
    assert res["stats"][10]["f_bass"]["n"] == T


def test_n_scaling_adaptive_means(gpu_ctx):
    print("[case] harness n-scaling adaptive N=10 k=512, 100 trials", flush=True)
    R = pkg_module("experimentation.runner")
    d = np.load(os.path.join(REF_RESULTS, "nscaling_adaptive_k.npz"), allow_pickle=False)
    T = 100
    res = R.run_n_scaling_full(_rfim, [10], lambda N: 2 ** (N - 1), T, SEED_BASE, bass_kwargs=KW, verbose=False,
                               n_boot=200)
    for key, fx in [("f_fixed", "RFIM_(Weq2)_f_fixed"), ("f_bass", "RFIM_(Weq2)_f_bass")]:
        assert abs(np.mean(res[key][:, 0]) - np.mean(d[fx][:T, 0])) < 2e-3, key
        # per-trial spread is the tie-resolution noise, not a systematic offset
        assert np.median(np.abs(res[key][:, 0] - d[fx][:T, 0])) < 5e-3, key


def test_sweep_full_brickwork_vs_oracle(gpu_ctx, tmp_path):
    print("[case] harness k-sweep brickwork N=12 vs oracle", flush=True)
    from oracle import drivers as OD
    from oracle import exact as OE

    R = pkg_module("experimentation.runner")
    N, depth, T, ks = 12, 10, 3, [48, 160]

    def gen(n, rng):
        return R.make_brickwork(n, depth, rng)

    def compute():
        res = R.run_sweep_full(gen, N, ks, T, 7, bass_kwargs={**KW, "accept_rtol": 1e-10}, verbose=False,
                                n_boot=200)
        return None, R.sweep_arrays(res)

    meta = {"N": N, "depth": depth, "k_values": ks, "n_trials": T, "seed_base": 7, "bass_kw": KW}
    _, arr = R.run_or_load("brickwork_sweep", meta, compute, str(tmp_path))
    rotated = 0
    for t in range(T):
        circ = gen(N, np.random.default_rng(7 * 10 ** 6 + t))
        sv = OE.exact_statevector(N, circ)
        for ki, k in enumerate(ks):
            fs = OD.FixedBasisSimulator(N, k, selector=OD.key_tiebreak_selector).simulate(circ, seed=0)
            trace = []
            b = OD.BASS(N, k, selector=OD.key_tiebreak_selector, accept_rtol=1e-10, trace=trace, **KW)
            bs = b.simulate(circ)
            rotated += sum(1 for e in trace if e[0] == "rot" and e[2])
            want = {"f_fixed": OE.compute_fidelity(fs.x[: fs.nnz], fs.alpha[: fs.nnz], sv),
                    "pr_fixed": OE.compute_participation_ratio(fs.alpha[: fs.nnz]),
                    "f_bass": OE.fidelity_bass(N, bs.x[: bs.nnz], bs.alpha[: bs.nnz], b.U, sv),
                    "pr_bass": OE.compute_participation_ratio(bs.alpha[: bs.nnz])}
            for key, v in want.items():
                assert abs(arr[key][t, ki] / v - 1) < 1e-10, (t, k, key, arr[key][t, ki], v)
    assert rotated > 0, "the optimiser accepted rotations in these runs"
    _, again = R.run_or_load("brickwork_sweep", meta, lambda: (_ for _ in ()).throw(AssertionError("recomputed")),
                             str(tmp_path))
    assert np.array_equal(again["f_bass"], arr["f_bass"])


def test_trial_timed_and_comparison(gpu_ctx):
    print("[case] harness timed trial + comparison", flush=True)
    R = pkg_module("experimentation.runner")
    r = R.run_trial_timed(_rfim, 10, 8192, seed=(SEED_BASE + 1) * 10 ** 6 + 0, bass_kwargs=KW)
    assert r["t_fixed"] > 0 and r["t_bass"] > 0 and abs(r["f_fixed"] - 1) < 1e-10 and abs(r["f_bass"] - 1) < 1e-10
    c = R.run_comparison(8, 16, 4, 3, bass_kwargs=KW)
    assert c["trials"] == 3 and 0 <= c["win_rate"] <= 1 and 0 < c["f_bass_median"] <= 1 + 1e-12
    # no truncation (k >= 2^N): both simulators are exact
    c2 = R.run_comparison(6, 64, 4, 2, bass_kwargs=KW)
    assert abs(c2["f_fixed_median"] - 1) < 1e-10 and abs(c2["f_bass_median"] - 1) < 1e-10


def test_phase_diagram_rows(gpu_ctx):
    """PhaseDiagram.ipynb c2-c4 (N=16, K=4096, seed fi*1000 + t): the device's
    exact states reproduce the stored exact PR and half-chain entropy of the
    RFIM (fi=0) and QAOA (fi=4) rows per trial; run_trial_timed with default
    BASS arguments reproduces the RFIM rows' 20-trial means (the trials
    truncate through exact |a|^2 ties, see the module docstring)."""
    print("[case] harness phase-diagram rows N=16 K=4096", flush=True)
    R = pkg_module("experimentation.runner")
    d = np.load(os.path.join(REF_RESULTS, "phase_diagram_extended.npz"), allow_pickle=False)
    N, K, T = 16, 4096, 20
    fams = {0: lambda n, rng: R.make_rfim(n, rng, layers=5, W=2.0, dt=0.2),
            4: lambda n, rng: R.make_qaoa(n, 3, rng)}
    got = []
    for fi, gen in fams.items():
        for t in range(T):
            seed = fi * 1000 + t
            sv = R.exact_statevector(N, gen(N, np.random.default_rng(seed)))
            p = np.abs(sv) ** 2
            psi = sv.reshape(2 ** (N // 2), -1)
            ev = np.linalg.eigvalsh(psi @ psi.conj().T)
            ev = ev[ev > 1e-15]
            assert abs(1.0 / np.sum(p * p) / 2 ** N - d["pr_exact"][fi, t]) < 1e-12, (fi, t)
            assert abs(float(-np.sum(ev * np.log2(ev))) / (N // 2) - d["entropy_exact"][fi, t]) < 1e-12, (fi, t)
            if fi == 0:
                r = R.run_trial_timed(gen, N, K, seed=seed, bass_kwargs={}, _exact_sv=sv)
                got.append((r["f_fixed"], r["f_bass"], r["pr_fixed"] / 2 ** N, r["pr_bass"] / 2 ** N))
    got = np.array(got)
    ref = np.stack([d["fidelity_fixed"][0], d["fidelity_bass"][0], d["pr_fixed"][0], d["pr_bass"][0]], 1)
    assert np.all(np.abs(got[:, :2].mean(0) - ref[:, :2].mean(0)) < 2e-3), (got.mean(0), ref.mean(0))
    assert np.all(np.abs(got[:, 2:].mean(0) / ref[:, 2:].mean(0) - 1) < 1.5e-2), (got.mean(0), ref.mean(0))
    assert np.abs(got[:, :2] - ref[:, :2]).max() < 1e-2


def _near_tie(e, rel=1e-12):
    return e[0] in ("rot", "rot2") and abs(e[4] - e[3]) <= rel * abs(e[3])


def _pin_until_near_tie(trace, circ, N, K, **kw):
    """A run the stored result cannot pin (a PR-neutral decision, taken by
    summation order, may go either way) is pinned against the oracle instead,
    decision by decision up to and including that first near-tie decision:
    the same truncations (kept / total within 1e-10), the same rotations tried
    and accepted (PR before / after and kept fraction within 1e-10).  Returns
    the number of records compared."""
    from oracle import drivers as OD

    to = []
    OD.BASS(N, K, selector=OD.key_tiebreak_selector, trace=to, accept_rtol=0.0, **kw).simulate(circ)
    first = next(i for i, e in enumerate(trace) if _near_tie(e))
    assert len(to) > first, (len(to), first)
    for i, (d, o) in enumerate(zip(trace[: first + 1], to[: first + 1])):
        assert d[0] == o[0] and np.array_equal(np.asarray(d[1]), np.asarray(o[1])), (i, d, o)
        if d[0] != "trunc" and i < first:
            assert d[2] == o[2], (i, d, o)
        vals = (d[2], d[3]) if d[0] == "trunc" else (d[3], d[4], d[5])
        ref = (o[2], o[3]) if o[0] == "trunc" else (o[3], o[4], o[5])
        for a, b in zip(vals, ref):
            assert abs(a - b) <= 1e-10 * abs(b), (i, d, o)
    return first + 1


@pytest.mark.parametrize("fi", [1, 2, 3])
def test_phase_diagram_haar_rows_device(gpu_ctx, fi):
    """PhaseDiagram.ipynb rows built from Haar 2-qubit gates (1 brickwork 1D,
    2 brickwork 2D, 3 random pairing; N=16, K=4096, default arguments with the
    reference's accept rule): the device drivers reproduce the stored F and PR
    per trial (1e-10); BASS rows are compared on every trial whose rotation
    decisions stand clear of PR-neutral relabelings (decided by summation
    order in either code), at least 14 of 20; the others are pinned against the
    oracle decision by decision up to their first such decision."""
    print(f"[case] harness phase-diagram Haar family {fi} N=16 K=4096", flush=True)
    R = pkg_module("experimentation.runner")
    B = pkg_module("simulation.bass_simulator")
    S = pkg_module("simulation.simulator")
    F = pkg_module("benchmarking.fidelity")
    d = np.load(os.path.join(REF_RESULTS, "phase_diagram_extended.npz"), allow_pickle=False)
    N, K, T = 16, 4096, 20
    gen = {1: lambda rng: R.make_brickwork(N, 6, rng),
           2: lambda rng: R.make_brickwork_2d(4, 4, 4, rng),
           3: lambda rng: R.make_haar(N, 3, rng)}[fi]
    near = 0
    for t in range(T):
        circ = gen(np.random.default_rng(fi * 1000 + t))
        sv = R.exact_statevector(N, circ)
        fst = S.FixedBasisSimulator(N, K).simulate(circ)
        assert abs(F.compute_fidelity(fst, sv) - d["fidelity_fixed"][fi, t]) < 1e-10, t
        assert abs(F.compute_participation_ratio(fst) / 2 ** N / d["pr_fixed"][fi, t] - 1) < 1e-10, t
        trace = []
        sim = B.BASS(N, K, device="cuda:0", trace=trace, accept_rtol=0.0)
        st = sim.simulate(circ)
        if any(_near_tie(e) for e in trace):
            near += 1
            _pin_until_near_tie(trace, circ, N, K)
            continue
        assert abs(F.fidelity_bass(sim, st, sv) - d["fidelity_bass"][fi, t]) < 1e-10, t
        assert abs(F.compute_participation_ratio(st) / 2 ** N / d["pr_bass"][fi, t] - 1) < 1e-10, t
    assert T - near >= 14, near


@pytest.mark.parametrize("fam", ["brickwork", "brickwork_2d"])
def test_fidelity_pr_brickwork_rows_device(gpu_ctx, fam):
    """fidelity_pr_raw.npz Haar rows (N=20, seed 20260520e6 + t,
    optimize_every=5, buffer_factor=1, the reference's accept rule) on the
    device drivers: per trial and budget the fixed-basis F and PR (1e-10
    relative), BASS's on every run whose rotation decisions stand clear of a
    PR-neutral relabeling (a run with one may take the other branch; most still
    match; one that does not is pinned against the oracle decision by decision
    up to its first such decision)."""
    print(f"[case] harness fidelity_pr {fam} N=20", flush=True)
    R = pkg_module("experimentation.runner")
    B = pkg_module("simulation.bass_simulator")
    S = pkg_module("simulation.simulator")
    F = pkg_module("benchmarking.fidelity")
    d = np.load(os.path.join(REF_RESULTS, "fidelity_pr_raw.npz"), allow_pickle=False)
    N = 20
    gen = (lambda rng: R.make_brickwork(N, 6, rng)) if fam == "brickwork" else \
        (lambda rng: R.make_brickwork_2d(4, 5, 4, rng))
    ks = [100, 200, 500, 1000, 2000, 5000, 10000, 20000, 50000]
    compared = 0
    for t in range(3):
        circ = gen(np.random.default_rng(SEED_BASE * 10 ** 6 + t))
        sv = R.exact_statevector(N, circ)
        for ki in (0, 3, 6):
            k = ks[ki]
            fst = S.FixedBasisSimulator(N, k).simulate(circ)
            assert abs(F.compute_fidelity(fst, sv) / d[fam + "_f_fixed"][t, ki] - 1) < 1e-10, (t, k)
            assert abs(F.compute_participation_ratio(fst) / d[fam + "_pr_fixed"][t, ki] - 1) < 1e-10, (t, k)
            trace = []
            sim = B.BASS(N, k, device="cuda:0", trace=trace, accept_rtol=0.0, **KW)
            st = sim.simulate(circ)
            near = any(_near_tie(e) for e in trace)
            ok = (abs(F.fidelity_bass(sim, st, sv) / d[fam + "_f_bass"][t, ki] - 1) < 1e-10
                  and abs(F.compute_participation_ratio(st) / d[fam + "_pr_bass"][t, ki] - 1) < 1e-10)
            # a run with a PR-neutral rotation decision may follow the other branch
            assert ok or near, (t, k)
            if not ok:
                _pin_until_near_tie(trace, circ, N, k, **KW)
            compared += ok
    assert compared >= 4, compared


@pytest.mark.parametrize("fi", [0, 1])
def test_gamma_diagnostics_2q_pass_rows_device(gpu_ctx, fi):
    """gamma_diagnostics.npz tie-free families (fi = 0 random-pairing Haar
    depth 3, 1 brickwork depth 5; use_2qubit_rotations=True, the reference's
    accept rule) on the device driver, N = 14 and 16: per run the stored
    gamma^2 and true F (1e-10), the 2-qubit brick-wall pass included; a run with
    a PR-neutral decision may take the other branch, and is then pinned against
    the oracle decision by decision up to that decision."""
    print(f"[case] harness gamma diagnostics family {fi}", flush=True)
    R = pkg_module("experimentation.runner")
    B = pkg_module("simulation.bass_simulator")
    F = pkg_module("benchmarking.fidelity")
    d = np.load(os.path.join(REF_RESULTS, "gamma_diagnostics.npz"), allow_pickle=False)
    matched = 0
    for ni, N in ((0, 14), (1, 16)):
        gen = (lambda rng: R.make_haar(N, 3, rng)) if fi == 0 else (lambda rng: R.make_brickwork(N, 5, rng))
        for t in range(3):
            circ = gen(np.random.default_rng(42 * 10 ** 6 + ni * 1000 + fi * 100 + t))
            sv = R.exact_statevector(N, circ)
            for ki, K in ((0, 512), (1, 1024)):
                trace = []
                sim = B.BASS(N, K, device="cuda:0", trace=trace, accept_rtol=0.0, use_2qubit_rotations=True)
                st = sim.simulate(circ)
                assert any(e[0] == "rot2" for e in trace)
                near = any(_near_tie(e) for e in trace)
                ok = (abs(st.gamma ** 2 - d["gamma_sq"][fi, ni, ki, t]) < 1e-10
                      and abs(F.fidelity_bass(sim, st, sv) - d["true_f"][fi, ni, ki, t]) < 1e-10)
                assert ok or near, (N, t, K)
                if not ok:
                    _pin_until_near_tie(trace, circ, N, K, use_2qubit_rotations=True)
                matched += ok
    assert matched >= 9, matched
