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


def test_batched_matches_single(gpu_ctx):
    print("[case] batched vs single", flush=True)
    Bm = pkg_module("simulation.batched")
    S = pkg_module("simulation.bass_simulator")
    specs = _specs()
    bat = Bm.BatchedBASS(specs, device="cuda:0", trace=True)
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
