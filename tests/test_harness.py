"""Host side of the experiment harness (src/experimentation/runner.py of the
reference): circuit factories against the oracle's restatements and closed
forms, the statistics helpers against closed forms and invariants, and the
result cache's reload / recompute rules.  No device needed."""

import json
import os

import numpy as np
import pytest
from scipy.stats import unitary_group

from conftest import pkg_module
from oracle import circuits as OC

R = pkg_module("experimentation.runner")


def _same_circuit(a, b):
    assert len(a) == len(b)
    for ga, gb in zip(a, b):
        assert tuple(ga.qubits) == tuple(gb.qubits)
        np.testing.assert_array_equal(np.asarray(ga.matrix), np.asarray(gb.matrix))


@pytest.mark.parametrize("seed", [0, 20260520 * 10 ** 6 + 6])
def test_rfim_tfim_qaoa_factories_match_oracle(seed):
    _same_circuit(R.make_rfim(12, np.random.default_rng(seed), layers=5, W=2.0, dt=0.2),
                  OC.rfim(12, 5, W=2.0, dt=0.2, rng=np.random.default_rng(seed)))
    _same_circuit(R.make_tfim(10, np.random.default_rng(seed), layers=3, J=1.0, h=0.5, dt=0.1),
                  OC.tfim(10, 3, J=1.0, h=0.5, dt=0.1, rng=np.random.default_rng(seed)))
    _same_circuit(R.make_qaoa(12, 3, np.random.default_rng(seed)), OC.qaoa(12, 3, np.random.default_rng(seed)))


def test_random_3regular_graph():
    for seed in range(5):
        edges = R._random_3regular_graph(14, np.random.default_rng(seed))
        assert len(edges) == 21 and len(set(edges)) == 21
        deg = np.bincount(np.array(edges).ravel(), minlength=14)
        assert np.all(deg == 3) and all(i < j for i, j in edges)
        assert edges == OC.random_3regular_graph(14, np.random.default_rng(seed))
    with pytest.raises(ValueError):
        R._random_3regular_graph(7, np.random.default_rng(0))


def test_brickwork_factories():
    rng = np.random.default_rng(3)
    circ = R.make_brickwork(7, 4, rng)
    # layers alternate (0,1),(2,3),(4,5) and (1,2),(3,4),(5,6)
    assert [g.qubits for g in circ[:6]] == [(0, 1), (2, 3), (4, 5), (1, 2), (3, 4), (5, 6)]
    assert len(circ) == 12
    # per-gate seed drawn from rng, matrix from scipy's unitary_group seeded
    # the way the reference seeds numpy's global state
    seed0 = int(np.random.default_rng(3).integers(0, 2 ** 31))
    np.random.seed(seed0)
    np.testing.assert_array_equal(circ[0].matrix, unitary_group.rvs(4))
    for g in circ:
        np.testing.assert_allclose(g.matrix.conj().T @ g.matrix, np.eye(4), atol=1e-12)
    c2 = R.make_brickwork_2d(3, 4, 4, np.random.default_rng(0))
    pairs = [g.qubits for g in c2]
    # type 0: (r,0)-(r,1), (r,2)-(r,3); type 1: (r,1)-(r,2); type 2: rows 0-1; type 3: rows 1-2
    assert len(pairs) == 3 * 2 + 3 * 1 + 4 + 4
    assert pairs[:2] == [(0, 1), (2, 3)] and pairs[6] == (1, 2) and pairs[9] == (0, 4) and pairs[13] == (4, 8)


def test_qft_factory():
    c = R.make_qft(6)
    assert len(c) == 6 + 4 + 4 + 3 + 2 + 1
    assert c[1].qubits == (0, 1) and abs(c[1].matrix[3, 3] - 1j) < 1e-15
    assert abs(c[2].matrix[3, 3] - np.exp(1j * np.pi / 4)) < 1e-15


def test_trial_stats_closed_forms():
    d = np.array([1.0, 2.0, 4.0, 8.0, np.nan, np.inf])
    s = R.trial_stats(d, n_boot=500)
    f = np.array([1.0, 2.0, 4.0, 8.0])
    assert s["n"] == 4 and s["median"] == 3.0 and s["mean"] == 3.75
    assert abs(s["geomean"] - 2 ** 1.5) < 1e-12
    assert abs(s["std"] - np.std(f, ddof=1)) < 1e-12 and abs(s["sem"] - s["std"] / 2) < 1e-12
    assert s["q25"] == np.percentile(f, 25) and s["q90"] == np.percentile(f, 90)
    assert s["med_ci_lo"] <= s["median"] <= s["med_ci_hi"] and s["gm_ci_lo"] <= s["geomean"] <= s["gm_ci_hi"]
    assert R.trial_stats([])["n"] == 0 and np.isnan(R.trial_stats([])["median"])
    one = R.trial_stats([0.5])
    assert one["geomean"] == one["gm_ci_lo"] == 0.5 and one["std"] == 0.0
    # a constant sample: every bootstrap statistic is the constant
    assert R.bootstrap_ci(np.full(30, 7.0)) == (7.0, 7.0)
    # bootstrap intervals narrow with more data
    rng = np.random.default_rng(1)
    w_small = np.diff(R.bootstrap_ci(rng.normal(size=20), n_boot=2000))[0]
    w_large = np.diff(R.bootstrap_ci(rng.normal(size=2000), n_boot=2000))[0]
    assert w_large < w_small / 3


def test_ratio_and_metrics():
    a = np.array([2.0, 1.0, 3.0, np.nan, 5.0])
    f = np.array([1.0, 2.0, 3.0, 1.0, 0.0])
    r = R.ratio_stats(a, f, n_boot=200)
    assert r["n_trials"] == 3 and r["win_count"] == 1 and abs(r["win_rate"] - 1 / 3) < 1e-15
    np.testing.assert_array_equal(r["ratio_trials"], [2.0, 0.5, 1.0])
    assert abs(r["ratio_geomean"] - 1.0) < 1e-12
    m = R.compute_metrics([1.0, 100.0, 10.0, np.nan])
    assert m["median"] == 10.0 and abs(m["geometric_mean"] - 10.0) < 1e-12
    assert m["gm_ci95_lo"] <= 10.0 <= m["gm_ci95_hi"] and abs(m["arithmetic_mean"] - 37.0) < 1e-12
    assert all(np.isnan(v) for v in R.compute_metrics([np.nan]).values())


def test_run_or_load_cache(tmp_path):
    calls = []

    def compute():
        calls.append(1)
        rows = [{"N": 10, "regime": "a", "f": 0.5}, {"N": 12, "regime": "b", "f": 0.25}]
        return rows, {"f_bass": np.arange(6.0).reshape(3, 2)}

    meta = {"N_values": [10, 12], "k": 8192, "kw": {"optimize_every": 5}}
    d = str(tmp_path)
    rows, arr = R.run_or_load("sweep", meta, compute, d, int_cols=("N",), str_cols=("regime",))
    rows2, arr2 = R.run_or_load("sweep", dict(meta), compute, d, int_cols=("N",), str_cols=("regime",))
    assert len(calls) == 1, "a matching configuration reloads"
    assert rows2 == rows and np.array_equal(arr2["f_bass"], arr["f_bass"])
    R.run_or_load("sweep", dict(meta, k=4096), compute, d)
    assert len(calls) == 2, "a different configuration recomputes"
    R.run_or_load("sweep", dict(meta, k=4096), compute, d, force=True)
    assert len(calls) == 3
    csv_p, npz_p, meta_p = R.cache_paths("sweep", d)
    assert json.load(open(meta_p))["k"] == 4096
    os.remove(npz_p)
    R.run_or_load("sweep", dict(meta, k=4096), compute, d)
    assert len(calls) == 4, "a missing file recomputes"
    # arrays only (no rows)
    _, a3 = R.run_or_load("arr", {"x": 1}, lambda: (None, {"a": np.ones(3)}), d)
    _, a4 = R.run_or_load("arr", {"x": 1}, lambda: (_ for _ in ()).throw(AssertionError("recomputed")), d)
    assert np.array_equal(a3["a"], a4["a"]) and not os.path.exists(R.cache_paths("arr", d)[0])


def test_server_rejects_requests_that_never_fit():
    """A request whose K exceeds the engine's slab, whose footprint exceeds the
    HBM budget, or whose qubit count is out of range is rejected when it is
    queued (FIFO admission would otherwise wait on it forever); requests that
    fit are queued."""
    S = pkg_module("serving.server")

    class _Engine:  # what _EngineLoop reads of a BatchEngine, without a device
        slab_entries = 1 << 16
        active = 0

        def admit(self, *a, **k):
            return True

    loop = S._EngineLoop(_Engine(), budget=200e9)
    ok = S.CircuitRequest(rid=0, kind="tfim", n=12, depth=2, log2k=10)
    loop.add(ok)
    assert len(loop.pending) == 1
    with pytest.raises(S.RejectedRequest):
        loop.add(S.CircuitRequest(rid=1, kind="tfim", n=12, depth=2, log2k=17))  # K > slab
    with pytest.raises(S.RejectedRequest):
        loop.add(S.CircuitRequest(rid=2, kind="tfim", n=70, depth=2, log2k=10))  # qubits
    small = S._EngineLoop(_Engine(), budget=S.footprint_bytes(ok) - 1)
    with pytest.raises(S.RejectedRequest):
        small.add(ok)  # footprint > budget
    assert len(loop.pending) == 1
    loop.admit()
    assert not loop.pending and loop.used == S.footprint_bytes(ok) and loop.peak == loop.used
