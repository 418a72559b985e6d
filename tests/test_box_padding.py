"""The traversal's box test stays conservative with a hardware reciprocal.

The kernels compute each slab distance as fl(fl(lo - o) * r) in float32, where
r is v_rcp_f32 of the direction component (within 1 ulp of 1/d, not correctly
rounded), and accept a box when entry <= exit * kBoxPad with
kBoxPad = 1 + 5 * 2^-23 (helios_amd/csrc/traverse.hpp). A box that the exact
ray enters must never be culled, or a hit could be lost. Here rays are aimed at
box corners and edges (the grazing cases, where exact entry and exit nearly
coincide); the exact decision uses rationals, and the float32 test is run with
every reciprocal within 1 ulp of 1/d. Without the padding the same rays cull
boxes they enter, so the check can fail.
"""
from fractions import Fraction

import numpy as np

PAD = np.float32(1.0 + 5.0 * 2.0 ** -23)
F32 = np.float32


def _rcp_candidates(d):
    """float32 values within 1 ulp (of themselves) of the exact 1/d."""
    exact = 1 / Fraction(float(d))
    a = F32(1.0) / F32(d)   # correctly rounded
    out = []
    for r in (np.nextafter(a, F32(-np.inf)), a, np.nextafter(a, F32(np.inf))):
        ulp = abs(float(np.nextafter(r, F32(np.inf))) - float(r))
        if abs(Fraction(float(r)) - exact) <= Fraction(ulp):
            out.append(F32(r))
    return out


def _exact_hit(lo, hi, o, d, tmax):
    near, far = Fraction(0), Fraction(tmax)
    for k in range(3):
        t0 = (Fraction(float(lo[k])) - Fraction(float(o[k]))) / Fraction(float(d[k]))
        t1 = (Fraction(float(hi[k])) - Fraction(float(o[k]))) / Fraction(float(d[k]))
        near = max(near, min(t0, t1))
        far = min(far, max(t0, t1))
    return near <= far


def _f32_hit(lo, hi, o, inv, tmax, pad):
    t0 = (lo - o).astype(F32) * inv
    t1 = (hi - o).astype(F32) * inv
    a = max(F32(0.0), *np.minimum(t0, t1))
    b = min(F32(tmax), *np.maximum(t0, t1))
    return bool(a <= F32(b * pad))


def _grazing_cases(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        c = rng.uniform(-50, 50, 3)
        lo = (c - rng.uniform(0.01, 8, 3)).astype(F32)
        hi = (c + rng.uniform(0.01, 8, 3)).astype(F32)
        # a corner, or a point on an edge, of the box
        target = np.where(rng.random(3) < 0.5, lo, hi).astype(np.float64)
        if rng.random() < 0.5:
            k = rng.integers(3)
            target[k] = rng.uniform(lo[k], hi[k])
        o = (target + rng.normal(size=3) * rng.uniform(1, 200)).astype(F32)
        d = (target - o.astype(np.float64)).astype(F32)
        d /= F32(np.abs(d).max())
        if np.any(d == 0):
            continue
        yield lo, hi, o, d


def test_padding_keeps_grazing_boxes():
    tmax = 1e30
    checked = culled_unpadded = 0
    for lo, hi, o, d in _grazing_cases(6000, 7):
        if not _exact_hit(lo, hi, o, d, tmax):
            continue
        cands = [_rcp_candidates(d[k]) for k in range(3)]
        for rx in cands[0]:
            for ry in cands[1]:
                for rz in cands[2]:
                    inv = np.array([rx, ry, rz], F32)
                    assert _f32_hit(lo, hi, o, inv, tmax, PAD), (lo, hi, o, d, inv)
                    checked += 1
                    culled_unpadded += not _f32_hit(lo, hi, o, inv, tmax, F32(1.0))
    assert checked > 10000
    # the grazing cases do exercise the padding
    assert culled_unpadded > 0
