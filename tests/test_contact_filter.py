"""The step kernels' FP32 contact filter (law.hpp: maybe_touch, host entry dem_contact_filter).

The fused and split step kernels decide contact with the FP64 test of the law
((ra + rb) - |xb - xa| > 0, law_touch); the fused step first drops candidate pairs whose FP32
position copies put them certainly apart.  The filter must never drop a pair that the FP64 test
keeps, whatever the distance from the origin, and must drop pairs clearly apart (or it saves
nothing).  The FP64 decision is computed here with numpy in the same expression order (IEEE
subtraction, products, sums and a correctly rounded square root, as in the C code compiled
without contraction).
"""
import ctypes

import numpy as np
import pytest


def fp64_touch(q):
    dx = q[:, 4] - q[:, 0]
    dy = q[:, 5] - q[:, 1]
    dz = q[:, 6] - q[:, 2]
    dist = np.sqrt(dx * dx + dy * dy + dz * dz)
    return (q[:, 3] + q[:, 7]) - dist > 0.0


def run_filter(dem, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    keep = np.zeros(len(q), dtype=np.int32)
    rc = dem.lib().dem_contact_filter(q.ctypes.data_as(ctypes.c_void_p), len(q),
                                      keep.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    return keep.astype(bool)


def pairs_at(rng, n, extent, gap):
    """n pairs around a point within `extent` of the origin, radii 0.25..0.65, centre distance
    (ra + rb) + gap[i] along a random direction."""
    xa = rng.uniform(-extent, extent, size=(n, 3))
    ra = rng.uniform(0.25, 0.65, n)
    rb = rng.uniform(0.25, 0.65, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    dist = (ra + rb) + gap
    xb = xa + dist[:, None] * u
    return np.column_stack([xa, ra, xb, rb])


@pytest.mark.parametrize("extent", [1.0, 50.0, 250.0, 1000.0])
def test_filter_keeps_every_fp64_contact(dem, extent):
    rng = np.random.default_rng(int(extent) + 7)
    n = 200_000
    # gaps straddling contact at every scale, down to exact ties and one-ulp distances
    scale = rng.choice([0.0, 1e-12, 1e-9, 1e-7, 1e-5, 1e-4, 1e-3], size=n)
    gap = rng.uniform(-1.0, 1.0, n) * scale
    q = pairs_at(rng, n, extent, gap)
    touch = fp64_touch(q)
    keep = run_filter(dem, q)
    assert touch.sum() > n // 4, "the sample must contain contacts"
    missed = np.flatnonzero(touch & ~keep)
    assert missed.size == 0, f"{missed.size} FP64 contacts dropped, e.g. {q[missed[:3]]}"


@pytest.mark.parametrize("extent", [1.0, 250.0, 1000.0])
def test_filter_drops_pairs_clearly_apart(dem, extent):
    rng = np.random.default_rng(int(extent) + 11)
    n = 100_000
    # the margin is below 1e-3 d within 1000 d of the origin
    gap = rng.uniform(1e-2, 1.0, n)
    q = pairs_at(rng, n, extent, gap)
    assert not fp64_touch(q).any()
    assert not run_filter(dem, q).any()


def test_filter_keeps_nan(dem):
    q = np.array([[0.0, 0.0, 0.0, 0.5, np.nan, 0.0, 0.0, 0.5],
                  [0.0, 0.0, 0.0, 0.5, 3.0, 0.0, 0.0, 0.5]])
    assert list(run_filter(dem, q)) == [True, False]
