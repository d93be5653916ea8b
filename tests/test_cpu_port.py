"""The oracle's single-thread C++ 1-qubit step (the CPU baseline of the
benchmark) against the numpy restatement of the same reference functions."""

import numpy as np
import pytest

from conftest import random_state, random_unitary
from oracle import drivers as OD
from oracle import kernels as OK
from oracle import port


def _numpy_step(x, a, q, G, k):
    n = len(x)
    ox, oa = np.empty(2 * n, np.uint64), np.empty(2 * n, np.complex128)
    m = OK.apply_1qubit_gate_kernel_opt(x, a, n, ox, oa, G, q)
    OK.radix_sort_uint64(ox, oa, m)
    m = OK.merge_sorted_kernel(ox, oa, m)
    ox, oa = ox[:m], oa[:m]
    merged = m
    gf = 1.0
    if 0 < k < m:
        p = oa.real * oa.real + oa.imag * oa.imag
        idx = OD.key_tiebreak_selector(ox, oa, p, k)
        gf = p[idx].sum() / p.sum()
        ox, oa = ox[idx], oa[idx]
    oa = oa / np.sqrt(np.sum(oa.real * oa.real + oa.imag * oa.imag))
    return ox, oa, gf, merged


@pytest.fixture(scope="module", autouse=True)
def built():
    try:
        port.load()
    except ImportError:
        from oracle.port import build as B

        B.build()


@pytest.mark.parametrize("n,nbits,q,k", [(1, 4, 0, 0), (50, 8, 3, 0), (3000, 20, 19, 0), (3000, 20, 5, 2000),
                                         (40000, 30, 11, 30000)])
def test_port_matches_numpy_restatement(n, nbits, q, k):
    rng = np.random.default_rng(n + q)
    x, a = random_state(n, nbits, seed=n + 1)
    G = random_unitary(2, rng)
    G[0, 1] = 0.0  # exercise the exact-zero emission rule
    cx, ca = np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.complex128)
    cx[:n], ca[:n] = x, a
    tx, ta = np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.complex128)
    m, gf, merged = port.step_1q(cx, ca, n, q, G, k, tx, ta)
    ox, oa, og, omerged = _numpy_step(x, a, q, G, k)
    assert merged == omerged
    assert m == len(ox)
    if k == 0 or k >= merged:
        # no truncation: same keys in the same (ascending) order
        assert np.array_equal(cx[:m], ox)
        # the renormalising sums run in different orders (numpy: pairwise)
        np.testing.assert_allclose(ca[:m], oa, rtol=1e-14, atol=1e-300)
    else:
        d = dict(zip(cx[:m].tolist(), ca[:m].tolist()))
        assert set(d) == set(ox.tolist())
        np.testing.assert_allclose(np.array([d[v] for v in ox.tolist()]), oa, rtol=1e-13, atol=1e-300)
    assert abs(gf - og) <= 1e-13


def test_port_exact_ties_keep_smaller_keys():
    """All |a|^2 equal after an identity gate: the kept set is the k smallest
    keys (the oracle's tie rule)."""
    n, k = 1000, 300
    rng = np.random.default_rng(7)
    x = np.sort(rng.choice(1 << 20, n, replace=False)).astype(np.uint64)
    a = np.full(n, 1 / np.sqrt(n), np.complex128)
    cx, ca = np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.complex128)
    cx[:n], ca[:n] = x, a
    tx, ta = np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.complex128)
    m, gf, merged = port.step_1q(cx, ca, n, 4, np.eye(2), k, tx, ta)
    ox, oa, og, _ = _numpy_step(x, a, 4, np.eye(2), k)
    assert (m, merged) == (k, n)
    assert set(cx[:m].tolist()) == set(ox.tolist()) == set(x[:k].tolist())
    assert abs(gf - og) <= 1e-15


def test_port_radix_sort_is_stable():
    rng = np.random.default_rng(3)
    n = 50000
    x = rng.integers(0, 1 << 40, size=n, dtype=np.uint64)
    x[: n // 4] = x[n // 4: n // 2]
    a = (np.arange(n) + 1j * rng.standard_normal(n)).astype(np.complex128)
    gx, ga = x.copy(), a.copy()
    port.radix_sort(gx, ga, n)
    order = np.argsort(x, kind="stable")
    assert np.array_equal(gx, x[order]) and np.array_equal(ga, a[order])
