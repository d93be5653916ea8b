"""The SparseState top-k path on the device against the reference-held
schmidt_benchmark.npz (SchmidtWeightedTruncation.ipynb: N=16, K=8192, 100
trials each of Haar random-pairing and brickwork circuits of depth 3 and 5):
per trial, the stored fidelity and gamma^2 to 1e-10, with the reference's
sort-and-merge gates (dedup="sort": emit, radix sort, merge) and with the
group table (dedup="table").  Pins emit -> radix sort -> merge -> top-k
end to end (src/core/sparse_state.py:70-172, :268-400)."""

import json
import os

import numpy as np
import pytest

from conftest import REF_RESULTS, pkg_module
from oracle import exact as OE
from test_oracle_fixtures import SCHMIDT_FAMILIES, schmidt_circuit

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dedup", ["sort", "table"])
@pytest.mark.parametrize("ci", [0, 1, 2, 3])
def test_schmidt_benchmark_topk_rows_device(gpu_ctx, ci, dedup):
    print(f"[case] schmidt benchmark top-k rows {SCHMIDT_FAMILIES[ci]} dedup={dedup}", flush=True)
    SS = pkg_module("core.sparse_state")
    d = np.load(os.path.join(REF_RESULTS, "schmidt_benchmark.npz"), allow_pickle=False)
    meta = json.load(open(os.path.join(REF_RESULTS, "schmidt_benchmark_meta.json")))
    N, K, T = meta["N"], meta["K"], meta["n_trials"]
    worst = 0.0
    for t in range(T):
        circ = schmidt_circuit(ci, t)
        sv = OE.exact_statevector(N, circ)
        s = SS.SparseState(N, K, device="cuda:0", dedup=dedup)
        for g in circ:
            s.apply_gate(g, "top-k")
        f = OE.compute_fidelity(s.keys(), s.alpha[: s.nnz], sv)
        e = max(abs(f - d["fidelity"][ci, 0, t]), abs(s.gamma ** 2 - d["gamma_sq"][ci, 0, t]))
        assert e < 1e-10, (t, f, d["fidelity"][ci, 0, t], s.gamma ** 2, d["gamma_sq"][ci, 0, t])
        worst = max(worst, e)
    print(f"  {T} trials, worst |dF|, |dgamma^2| = {worst:.2e}", flush=True)
