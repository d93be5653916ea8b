"""The CPU oracle against results the reference itself stored
(tests/golden/reference_results, copied from the reference's tests/results).

What each fixture pins:
  rotation_diag.csv      TFIM generator + RX/RZZ conventions + dense simulator
                         (Ordered TFIM, N=12: PRZ and mean single-qubit angle);
  nscaling_fixed_k.npz   N=10, 12 with k=8192 >= 2^N: no truncation, so every
                         trial must reproduce the stored participation ratio of
                         the full state (RFIM generator, sparse 1q/2q/diagonal
                         kernels, both drivers) and F = 1;
  fidelity_pr_raw.npz    N=20 RFIM, BASS k=100: trials whose trajectory has no
                         |a|^2 tie at a truncation threshold are reproduced to
                         1e-10 (drivers, truncation, RDMs, eigh gauge, the
                         coordinate-descent optimiser with accepted rotations);
  nscaling_adaptive_k    N=10, k=512 (truncating, tie-prone): the reference
                         resolves exact ties of product-state amplitudes by
                         floating-point noise, so per-trial values differ; the
                         100-trial means must agree;
  phase_diagram_extended N=16 RFIM and QAOA: exact-state PR and half-chain
                         entropy per trial (generators + dense simulator); the
                         RFIM rows of both drivers at K=4096 as 20-trial means.
"""

import csv
import json
import os

import numpy as np
import pytest

from conftest import REF_RESULTS, pkg_module
from oracle import circuits as OC
from oracle import drivers as OD
from oracle import exact as OE

SEED_BASE = 20260520


def _load(name):
    return np.load(os.path.join(REF_RESULTS, name), allow_pickle=False)


def _rfim(N, seed):
    return OC.rfim(N, 5, W=2.0, dt=0.2, rng=np.random.default_rng(seed))


def test_rotation_diag_ordered_tfim():
    row = next(r for r in csv.DictReader(open(os.path.join(REF_RESULTS, "rotation_diag.csv")))
               if r["family"] == "Ordered TFIM")
    meta = json.load(open(os.path.join(REF_RESULTS, "rotation_diag_meta.json")))
    N = meta["N"]
    # RotationDiagnostics.ipynb c2: make_tfim(N, rng, layers=5, J=1.0, h=0.5, dt=0.1)
    circ = OC.tfim(N, 5, J=1.0, h=0.5, dt=0.1, rng=np.random.default_rng(SEED_BASE + 10_000))
    sv = OE.exact_statevector(N, circ)
    p = np.abs(sv) ** 2
    assert abs(1.0 / np.sum(p ** 2) - float(row["prz_median"])) < 1e-10
    # single-qubit diagnostics (c3): angle of the dominant RDM eigenvector to Z
    psi = sv.reshape([2] * N)
    phi = []
    for q in range(N):
        m = np.moveaxis(psi, q, 0).reshape(2, -1)
        ev, evec = np.linalg.eigh(m @ m.conj().T)
        amp0 = min(float(abs(evec[:, np.argmax(ev)][0])), 1.0)
        th0 = 2.0 * np.arccos(amp0)
        phi.append(min(th0, np.pi - th0))
    assert abs(np.degrees(np.mean(phi)) - float(row["mean_phi_deg"])) < 1e-9


# accept_rtol=1e-10 is the device driver's deterministic rotation-accept rule:
# the reference's recorded results do not change under it.
@pytest.mark.parametrize("rtol", [0.0, 1e-10])
@pytest.mark.parametrize("ni,N", [(0, 10), (1, 12)])
def test_nscaling_fixed_k_no_truncation(ni, N, rtol):
    d = _load("nscaling_fixed_k.npz")
    for t in range(12):
        circ = _rfim(N, (SEED_BASE + 1) * 10 ** 6 + ni * 1000 + t)
        sv = OE.exact_statevector(N, circ)
        fs = OD.FixedBasisSimulator(N, 8192).simulate(circ, seed=0)
        b = OD.BASS(N, 8192, optimize_every=5, buffer_factor=1, accept_rtol=rtol)
        bs = b.simulate(circ)
        assert abs(OE.compute_fidelity(fs.x[: fs.nnz], fs.alpha[: fs.nnz], sv) - d["RFIM_(Weq2)_f_fixed"][t, ni]) < 1e-10
        assert abs(OE.fidelity_bass(N, bs.x[: bs.nnz], bs.alpha[: bs.nnz], b.U, sv) - d["RFIM_(Weq2)_f_bass"][t, ni]) < 1e-10
        pr_f = OE.compute_participation_ratio(fs.alpha[: fs.nnz])
        pr_b = OE.compute_participation_ratio(bs.alpha[: bs.nnz])
        assert abs(pr_f / d["RFIM_(Weq2)_pr_fixed"][t, ni] - 1) < 1e-10
        assert abs(pr_b / d["RFIM_(Weq2)_pr_bass"][t, ni] - 1) < 1e-10


@pytest.mark.parametrize("rtol", [0.0, 1e-10])
@pytest.mark.parametrize("t", [0, 2, 6])
def test_fidelity_pr_rfim_bass_k100(t, rtol):
    d = _load("fidelity_pr_raw.npz")
    N, k = 20, 100
    circ = _rfim(N, SEED_BASE * 10 ** 6 + t)
    sv = OE.exact_statevector(N, circ)
    trace = []
    b = OD.BASS(N, k, optimize_every=5, buffer_factor=1, trace=trace, accept_rtol=rtol)
    bs = b.simulate(circ)
    f = OE.fidelity_bass(N, bs.x[: bs.nnz], bs.alpha[: bs.nnz], b.U, sv)
    pr = OE.compute_participation_ratio(bs.alpha[: bs.nnz])
    assert abs(f / d["rfim_f_bass"][t, 0] - 1) < 1e-10
    assert abs(pr / d["rfim_pr_bass"][t, 0] - 1) < 1e-10
    assert any(e[0] == "rot" and e[2] for e in trace), "optimiser accepted rotations in this trial"


def test_nscaling_adaptive_means():
    d = _load("nscaling_adaptive_k.npz")
    N, k, ni, T = 10, 512, 0, 100
    ff, fb = [], []
    for t in range(T):
        circ = _rfim(N, SEED_BASE * 10 ** 6 + ni * 1000 + t)
        sv = OE.exact_statevector(N, circ)
        fs = OD.FixedBasisSimulator(N, k).simulate(circ, seed=0)
        ff.append(OE.compute_fidelity(fs.x[: fs.nnz], fs.alpha[: fs.nnz], sv))
        b = OD.BASS(N, k, optimize_every=5, buffer_factor=1)
        bs = b.simulate(circ)
        fb.append(OE.fidelity_bass(N, bs.x[: bs.nnz], bs.alpha[: bs.nnz], b.U, sv))
    assert abs(np.mean(ff) - np.mean(d["RFIM_(Weq2)_f_fixed"][:T, ni])) < 2e-3
    assert abs(np.mean(fb) - np.mean(d["RFIM_(Weq2)_f_bass"][:T, ni])) < 2e-3
    # per-trial spread is the tie-resolution noise, not a systematic offset
    assert np.median(np.abs(np.array(fb) - d["RFIM_(Weq2)_f_bass"][:T, ni])) < 5e-3


# PhaseDiagram.ipynb c2-c4: N=16, K=4096, 20 trials, seed fi*1000 + t, family
# index fi (0 = RFIM W=2, 4 = QAOA p=3 on a random 3-regular graph; the other
# families need scipy's unitary_group or are not restated here).  pr_exact is
# PR/2^N of the exact state and entropy_exact the half-chain entanglement
# entropy over N/2.
PD_N, PD_K, PD_T = 16, 4096, 20
PD_FAMILIES = {0: lambda rng: OC.rfim(16, 5, W=2.0, dt=0.2, rng=rng),
               4: lambda rng: OC.qaoa(16, 3, rng)}


def _half_chain_entropy(sv, N):
    psi = sv.reshape(2 ** (N // 2), 2 ** (N - N // 2))
    ev = np.linalg.eigvalsh(psi @ psi.conj().T)
    ev = ev[ev > 1e-15]
    return float(-np.sum(ev * np.log2(ev))) / (N // 2)


@pytest.mark.parametrize("fi", sorted(PD_FAMILIES))
def test_phase_diagram_exact_states(fi):
    """Per trial, tie-free: the generator at the notebook's arguments and seeds
    plus the dense simulator reproduce the stored exact PR and entropy."""
    d = _load("phase_diagram_extended.npz")
    for t in range(PD_T):
        sv = OE.exact_statevector(PD_N, PD_FAMILIES[fi](np.random.default_rng(fi * 1000 + t)))
        p = np.abs(sv) ** 2
        assert abs(1.0 / np.sum(p * p) / 2 ** PD_N - d["pr_exact"][fi, t]) < 1e-12
        assert abs(_half_chain_entropy(sv, PD_N) - d["entropy_exact"][fi, t]) < 1e-12


def test_phase_diagram_rfim_means():
    """Both drivers with default BASS arguments (run_trial_timed, bass_kwargs={}).
    RFIM truncates from |0...0> through exact |a|^2 ties, so single trials
    follow the host's tie order; the 20-trial means must agree (F within 2e-3,
    PR/2^N within 1.5 %) and no trial may be off by more than 1e-2 in F.
    The QAOA rows start from the uniform superposition over all 2^16 keys: which
    4096 entries survive is decided by tie order alone (on identical inputs,
    numpy's argpartition on this host gives per-trial F 2e-3 apart between its
    AVX-512, AVX2 and generic dispatch), so they pin only the exact states."""
    d = _load("phase_diagram_extended.npz")
    fi = 0
    got = []
    for t in range(PD_T):
        circ = PD_FAMILIES[fi](np.random.default_rng(fi * 1000 + t))
        sv = OE.exact_statevector(PD_N, circ)
        fs = OD.FixedBasisSimulator(PD_N, PD_K).simulate(circ, seed=0)
        b = OD.BASS(PD_N, PD_K)
        bs = b.simulate(circ)
        got.append((OE.compute_fidelity(fs.x[: fs.nnz], fs.alpha[: fs.nnz], sv),
                    OE.fidelity_bass(PD_N, bs.x[: bs.nnz], bs.alpha[: bs.nnz], b.U, sv),
                    OE.compute_participation_ratio(fs.alpha[: fs.nnz]) / 2 ** PD_N,
                    OE.compute_participation_ratio(bs.alpha[: bs.nnz]) / 2 ** PD_N))
    got = np.array(got)
    ref = np.stack([d["fidelity_fixed"][fi], d["fidelity_bass"][fi], d["pr_fixed"][fi], d["pr_bass"][fi]], 1)
    assert np.all(np.abs(got[:, :2].mean(0) - ref[:, :2].mean(0)) < 2e-3)
    assert np.all(np.abs(got[:, 2:].mean(0) / ref[:, 2:].mean(0) - 1) < 1.5e-2)
    assert np.abs(got[:, :2] - ref[:, :2]).max() < 1e-2


def _pd_haar_families():
    """The phase-diagram families built from Haar 2-qubit gates (scipy
    unitary_group, whose draws match the reference's stored exact states):
    fi = 1 brickwork 1D (depth 6), 2 brickwork 2D (4 x 4, depth 4), 3 random
    pairing (depth 3).  No |a|^2 ties, so every trial is a per-trial pin."""
    R = pkg_module("experimentation.runner")
    return {1: lambda rng: R.make_brickwork(PD_N, 6, rng),
            2: lambda rng: R.make_brickwork_2d(4, 4, 4, rng),
            3: lambda rng: R.make_haar(PD_N, 3, rng)}


def near_tie_rotation(trace, rel=1e-12):
    """A rotation decision (1- or 2-qubit) whose PR change is within rounding:
    the rotation only relabels the state, and accept / reject follows
    summation order."""
    return any(e[0] in ("rot", "rot2") and abs(e[4] - e[3]) <= rel * abs(e[3]) for e in trace)


@pytest.mark.parametrize("fi", [1, 2, 3])
def test_phase_diagram_haar_rows(fi):
    """Per trial: the generators plus the dense simulator give the stored exact
    PR and entropy (1e-12); the fixed-basis driver gives the stored F (1e-10)
    and PR (1e-10 relative) at K=4096; BASS with default arguments does too --
    optimiser passes and truncations included -- on every trial whose
    rotation decisions all stand clear of PR-neutral relabelings (those are
    decided by summation order in either code; at least 14 of the 20 trials
    are compared)."""
    d = _load("phase_diagram_extended.npz")
    gen = _pd_haar_families()[fi]
    near = 0
    for t in range(PD_T):
        circ = gen(np.random.default_rng(fi * 1000 + t))
        sv = OE.exact_statevector(PD_N, circ)
        p = np.abs(sv) ** 2
        assert abs(1.0 / np.sum(p * p) / 2 ** PD_N - d["pr_exact"][fi, t]) < 1e-12, t
        assert abs(_half_chain_entropy(sv, PD_N) - d["entropy_exact"][fi, t]) < 1e-12, t
        fs = OD.FixedBasisSimulator(PD_N, PD_K).simulate(circ, seed=0)
        assert abs(OE.compute_fidelity(fs.x[: fs.nnz], fs.alpha[: fs.nnz], sv) - d["fidelity_fixed"][fi, t]) < 1e-10, t
        pr_f = OE.compute_participation_ratio(fs.alpha[: fs.nnz]) / 2 ** PD_N
        assert abs(pr_f / d["pr_fixed"][fi, t] - 1) < 1e-10, t
        trace = []
        b = OD.BASS(PD_N, PD_K, trace=trace)
        bs = b.simulate(circ)
        if near_tie_rotation(trace):
            near += 1
            continue
        assert abs(OE.fidelity_bass(PD_N, bs.x[: bs.nnz], bs.alpha[: bs.nnz], b.U, sv) - d["fidelity_bass"][fi, t]) < 1e-10, t
        pr_b = OE.compute_participation_ratio(bs.alpha[: bs.nnz]) / 2 ** PD_N
        assert abs(pr_b / d["pr_bass"][fi, t] - 1) < 1e-10, t
    assert PD_T - near >= 14, near


@pytest.mark.parametrize("fam", ["brickwork", "brickwork_2d"])
def test_fidelity_pr_brickwork_rows(fam):
    """fidelity_pr_raw.npz Haar rows (Fidelity_Pr.ipynb: N=20, brickwork depth
    6 or 4 x 5 depth 4, seed 20260520e6 + t, optimize_every=5,
    buffer_factor=1): per trial and budget, the fixed-basis F and PR (1e-10
    relative) and BASS's on every run whose rotation decisions stand clear of
    a PR-neutral relabeling (a run with one may take the other branch; most
    still match)."""
    d = _load("fidelity_pr_raw.npz")
    R = pkg_module("experimentation.runner")
    N = 20
    gen = (lambda rng: R.make_brickwork(N, 6, rng)) if fam == "brickwork" else \
        (lambda rng: R.make_brickwork_2d(4, 5, 4, rng))
    ks = [100, 200, 500, 1000, 2000, 5000, 10000, 20000, 50000]
    compared = 0
    for t in range(2):
        circ = gen(np.random.default_rng(SEED_BASE * 10 ** 6 + t))
        sv = OE.exact_statevector(N, circ)
        for ki in (0, 3):
            k = ks[ki]
            fs = OD.FixedBasisSimulator(N, k).simulate(circ, seed=0)
            f = OE.compute_fidelity(fs.x[: fs.nnz], fs.alpha[: fs.nnz], sv)
            assert abs(f / d[fam + "_f_fixed"][t, ki] - 1) < 1e-10, (t, k)
            assert abs(OE.compute_participation_ratio(fs.alpha[: fs.nnz]) / d[fam + "_pr_fixed"][t, ki] - 1) < 1e-10
            trace = []
            b = OD.BASS(N, k, optimize_every=5, buffer_factor=1, trace=trace)
            bs = b.simulate(circ)
            fb = OE.fidelity_bass(N, bs.x[: bs.nnz], bs.alpha[: bs.nnz], b.U, sv)
            pb = OE.compute_participation_ratio(bs.alpha[: bs.nnz])
            ok = abs(fb / d[fam + "_f_bass"][t, ki] - 1) < 1e-10 and abs(pb / d[fam + "_pr_bass"][t, ki] - 1) < 1e-10
            # a run with a PR-neutral rotation decision may follow the other branch
            assert ok or near_tie_rotation(trace), (t, k)
            compared += ok
    assert compared >= 2, compared


@pytest.mark.parametrize("fi", [0, 1])
def test_gamma_diagnostics_2q_pass_rows(fi):
    """gamma_diagnostics.npz (GammaOverstimationDiagnostics.ipynb: BASS with
    use_2qubit_rotations=True, seed 42e6 + ni*1000 + fi*100 + t) on its
    tie-free families, fi = 0 random-pairing Haar (depth 3) and 1 brickwork
    (depth 5), N = 14: per run the stored gamma^2 and true F (1e-10), the
    2-qubit brick-wall pass included; a run with a PR-neutral decision may
    take the other branch."""
    d = _load("gamma_diagnostics.npz")
    R = pkg_module("experimentation.runner")
    N, ni = 14, 0
    gen = (lambda rng: R.make_haar(N, 3, rng)) if fi == 0 else (lambda rng: R.make_brickwork(N, 5, rng))
    matched = 0
    for t in range(4):
        circ = gen(np.random.default_rng(42 * 10 ** 6 + ni * 1000 + fi * 100 + t))
        sv = OE.exact_statevector(N, circ)
        for ki, K in ((0, 512), (1, 1024)):
            trace = []
            b = OD.BASS(N, K, use_2qubit_rotations=True, trace=trace)
            bs = b.simulate(circ)
            assert any(e[0] == "rot2" for e in trace)
            f = OE.fidelity_bass(N, bs.x[: bs.nnz], bs.alpha[: bs.nnz], b.U, sv)
            ok = abs(bs.gamma ** 2 - d["gamma_sq"][fi, ni, ki, t]) < 1e-10 and abs(f - d["true_f"][fi, ni, ki, t]) < 1e-10
            assert ok or near_tie_rotation(trace), (t, K)
            matched += ok
    assert matched >= 6, matched


SCHMIDT_FAMILIES = ("haar_L3", "haar_L5", "brickwork_L3", "brickwork_L5")


def schmidt_circuit(ci, t):
    """SchmidtWeightedTruncation.ipynb c1 / c4: N=16, family ci (Haar random
    pairing or brickwork of Haar 2-qubit gates, depth 3 or 5), trial t, seed
    20260520 * 10^6 + ci * 10^4 + t."""
    R = pkg_module("experimentation.runner")
    rng = np.random.default_rng(SEED_BASE * 10 ** 6 + ci * 10_000 + t)
    depth = 3 if ci % 2 == 0 else 5
    return R.make_haar(16, depth, rng) if ci < 2 else R.make_brickwork(16, depth, rng)


@pytest.mark.parametrize("ci", [0, 1, 2, 3])
def test_schmidt_benchmark_topk_rows(ci):
    """schmidt_benchmark.npz, method "top-k" (SchmidtWeightedTruncation.ipynb
    c2 TruncationSimulator: every gate through the reference's emit / radix
    sort / merge kernels, truncation to the K=8192 largest |a|^2 with
    renormalisation and gamma *= sqrt(kept) -- the SparseState top-k path,
    src/core/sparse_state.py:268-400): per trial the stored fidelity against
    the exact state and gamma^2, to 1e-10.  Haar gates: no |a|^2 ties, so
    every trial is a pin.  The first 25 trials of each family here; the device
    test replays all 100."""
    d = _load("schmidt_benchmark.npz")
    meta = json.load(open(os.path.join(REF_RESULTS, "schmidt_benchmark_meta.json")))
    assert meta["circuits"][ci] == SCHMIDT_FAMILIES[ci] and meta["methods"][0] == "top-k"
    N, K = meta["N"], meta["K"]
    for t in range(25):
        circ = schmidt_circuit(ci, t)
        sv = OE.exact_statevector(N, circ)
        s = OD.SparseState(N, K)
        for g in circ:
            s.apply_gate(OD._G(g.qubits, g.matrix), "top-k")
        f = OE.compute_fidelity(s.x[: s.nnz], s.alpha[: s.nnz], sv)
        assert abs(f - d["fidelity"][ci, 0, t]) < 1e-10, t
        assert abs(s.gamma ** 2 - d["gamma_sq"][ci, 0, t]) < 1e-10, t
