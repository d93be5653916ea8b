"""Comparison helpers for device-vs-oracle parity tests."""

import numpy as np

AMP_RTOL = 1e-12      # relative agreement of amplitudes (complex128)
AMP_ATOL_REL = 1e-15  # absolute floor, relative to the largest amplitude (cancellation)
BAND = 1e-12          # |a|^2 band around the k-th threshold where kept sets may differ


def numpy_uses_fma_cmul():
    """True when numpy's complex128 multiply rounds like the device's
    fma-based product (true on AVX2/AVX-512 hosts)."""
    rng = np.random.default_rng(0)
    a = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    b = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    c = a * b
    # exact emulation of fma(ar, br, -(ai*bi)) with Python integers (Fractions)
    from fractions import Fraction as F

    ok = True
    for x, y, z in zip(a, b, c):
        re = float(F(x.real) * F(y.real) + F(-(x.imag * y.imag)))
        im = float(F(x.real) * F(y.imag) + F(x.imag * y.real))
        ok &= (re == z.real) and (im == z.imag)
    return bool(ok)


def as_dict(keys, amps):
    keys = np.asarray(keys, dtype=np.uint64)
    amps = np.asarray(amps, dtype=np.complex128)
    assert len(np.unique(keys)) == len(keys), "duplicate keys"
    return dict(zip(keys.tolist(), amps.tolist()))


def assert_same_entries(gk, ga, ok_, oa, rtol=AMP_RTOL, exact=False):
    """Same key set, amplitudes equal (bitwise if exact) or within rtol."""
    g = as_dict(gk, ga)
    o = as_dict(ok_, oa)
    assert set(g) == set(o), f"key sets differ: {len(set(g) ^ set(o))} keys"
    if not o:
        return
    ref = np.array([o[k] for k in g])
    got = np.array([g[k] for k in g])
    if exact:
        bad = np.flatnonzero((ref.real != got.real) | (ref.imag != got.imag))
        assert len(bad) == 0, f"{len(bad)} amplitudes differ bitwise, e.g. {got[bad[0]]} vs {ref[bad[0]]}"
    floor = AMP_ATOL_REL * np.max(np.abs(ref))
    err = np.abs(got - ref)
    lim = rtol * np.abs(ref) + floor
    assert np.all(err <= lim), f"max amplitude error {np.max(err / np.maximum(np.abs(ref), floor)):.3e}"


def assert_topk_sets(gk, ok_, keys_all, amps_all, k, band=BAND):
    """Kept sets identical except entries whose |a|^2 lies within `band`
    (relative) of the k-th largest |a|^2."""
    gs, os_ = set(np.asarray(gk).tolist()), set(np.asarray(ok_).tolist())
    assert len(gs) == k and len(os_) == k
    diff = gs ^ os_
    if not diff:
        return 0
    p = dict(zip(np.asarray(keys_all).tolist(),
                 (np.asarray(amps_all).real ** 2 + np.asarray(amps_all).imag ** 2).tolist()))
    thr = np.sort(np.array(list(p.values())))[::-1][k - 1]
    for key in diff:
        assert abs(p[key] - thr) <= band * thr, f"key {key} outside the tie band"
    return len(diff)
