"""The tiled backprojection's tile accumulator against the oracle.

The tiled scheme sums a tile's contributions w * r in 64-bit fixed point
(a power-of-two scale from the update's largest r and the work item's run
count) when every r of the update is finite and non-negative and the largest
is at most 2^(t-26) times the geometric mean of the positive ones
(t = min(51, 64 - log2(runs per item))), so that a typical contribution spans
at least 2^24 units; otherwise in binary64.  Both must reproduce the oracle's MLEM update
(projector.rs:67-74,135-149, mlem.rs:45-52): fixed point to the same
tolerance as binary64, binary64 with the reference's inf/NaN propagation for
zero projections.
"""
import numpy as np
import pytest

import oracle_ffi as O
import petalo_amd as P
from helpers import rel_l2

pytestmark = pytest.mark.gpu

FULL, N = [60.0, 60.0, 60.0], [20, 20, 20]
TOF = (42.47, 3.0)


def _say(msg):
    print(msg, flush=True)


def _lors(seed, n=60000):
    return P.synth_phantom_lors(seed, n, 200.0)


def _max_rel(g, o):
    """Largest relative difference over voxels holding at least 1e-3 of the
    image maximum."""
    big = o >= 1e-3 * o.max()
    return float(np.max(np.abs(g[big].astype(np.float64) - o[big]) / o[big]))


def _steps_vs_oracle(lors, iters, tof=TOF, image=None):
    """(units, relL2 per update, last GPU image, last oracle image)."""
    fov = P.FOV(FULL, N)
    ctx = P.Mlem(fov, lors, tof=P.Tof(*tof) if tof else None, scheme=P.SCHEME_TILE)
    x = np.ones(fov.nvox, np.float32)
    if image is not None:
        ctx.set_image(image)
        x = image.astype(np.float32).copy()
    units, errs = [], []
    for _ in range(iters):
        img, _ = ctx.step()
        units.append(ctx.accumulator_unit())
        O.mlem_update(FULL, N, lors, x, tof=tof, threads=16)
        errs.append(rel_l2(img.data, x) if np.all(np.isfinite(x)) else float("nan"))
    g = ctx.image().data.copy()
    ctx.close()
    return units, errs, g, x


@pytest.mark.parametrize("tof", [None, TOF])
def test_fixed_point_tiles_match_oracle(gpu, tof):
    """Ordinary phantom data: every update takes the fixed-point tiles and
    stays within 1e-6 relative L2 of the binary32 oracle."""
    lors = _lors(31)
    units, errs, g, o = _steps_vs_oracle(lors, 5, tof=tof)
    m = _max_rel(g, o)
    _say(f"tof={tof}: units {['%.3e' % u for u in units]}, relL2 {['%.1e' % e for e in errs]}, voxel max {m:.1e}")
    assert all(u > 0 for u in units)
    assert max(errs) <= 1e-6
    assert m <= 1e-5


@pytest.mark.parametrize("ac,fixed", [(1e-4, True), (1e-10, False)])
def test_outlier_r(gpu, ac, fixed):
    """A few LORs with a tiny additive correction: in the first update their
    r = 1/(p*ac) is 1e4 times the others', within the guard for these small
    work items (largest r at most 2^(t-26) = 2^25 geometric-mean r): fixed
    point; 1e10 times takes binary64 tiles.  (The update raises the voxels
    those LORs cross, so their p grows and later updates see ordinary r.)
    Every image within 1e-6 of the oracle, also voxel by voxel (1e-5)."""
    lors = _lors(32)
    lors["ac"][::20000] = np.float32(ac)
    units, errs, g, o = _steps_vs_oracle(lors, 3)
    m = _max_rel(g, o)
    _say(f"outliers ac={ac}: units {units}, relL2 {['%.1e' % e for e in errs]}, voxel max {m:.1e}")
    assert (units[0] > 0.0) == fixed
    assert max(errs) <= 1e-6
    assert m <= 1e-5


def test_moderate_r_spread_stays_fixed_point(gpu):
    """Corrections spread over [1, 1000]: fixed point, within 1e-6 of the
    oracle."""
    lors = _lors(33)
    rng = np.random.default_rng(5)
    lors["ac"] = (10.0 ** rng.uniform(0.0, 3.0, len(lors))).astype(np.float32)
    units, errs, g, o = _steps_vs_oracle(lors, 3)
    m = _max_rel(g, o)
    _say(f"spread ac: units {['%.3e' % u for u in units]}, relL2 {['%.1e' % e for e in errs]}, voxel max {m:.1e}")
    assert all(u > 0 for u in units)
    assert max(errs) <= 1e-6
    assert m <= 1e-5


def test_zero_projection_takes_binary64(gpu):
    """An image that is zero where some LORs run gives them p = 0 and r = inf
    (projector.rs:145): binary64 tiles, with the oracle's non-finite voxels."""
    lors = _lors(34)
    x0 = np.ones(int(np.prod(N)), np.float32)
    x0[: 20 * 20 * 4] = 0.0  # the first four slices
    units, _, g, o = _steps_vs_oracle(lors, 1, image=x0)
    _say(f"zero slab: unit {units[0]}, non-finite voxels {int((~np.isfinite(o)).sum())}")
    assert units[0] == 0.0
    assert np.array_equal(np.isnan(g), np.isnan(o))
    fin = np.isfinite(o)
    assert np.array_equal(np.isfinite(g), fin)
    assert rel_l2(g[fin], o[fin]) <= 1e-5
