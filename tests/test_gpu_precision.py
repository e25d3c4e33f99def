"""The binary64 precision modes of a context against the oracle's binary64
twins, through the C ABI.

  PRECISION_F64       vs orc_mlem_update_f64 (the oracle's binary64 walk)
  PRECISION_F32_WALK  vs orc_mlem_update_mixed (binary32 walk decisions,
                      binary64 arithmetic)

Both sides compute the same binary64 operations; only the order of the
binary64 backprojection sums (device atomics, oracle thread buffers) and the
last bit of exp (device vs host libm, with TOF) differ, so images agree to
~1e-14.  A walk that took a different voxel anywhere would differ by ~1e-3.
On phantom LORs the binary32 path then sits within rounding (1e-5 on a
sparse set) of the mixed mode and a shifted-walk geometry (1e-1) away from
the binary64 mode.
"""
import numpy as np
import pytest

import oracle_ffi as O
import petalo_amd as P
from helpers import random_lors, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _gpu(fov, lors, iters, tof, precision, subsets=1, sens=None):
    ctx = P.Mlem(fov, lors, tof=P.Tof(*tof) if tof else None, n_subsets=subsets, precision=precision,
                 sensitivity=sens)
    out = []
    for _ in range(iters * subsets):
        ctx.step(copy_image=False)
        out.append(ctx.image_f64())
    ctx.close()
    return out


def _oracle(full, n, lors, iters, tof, mixed, subsets=1, sens=None):
    x = np.ones(int(np.prod(n)), np.float64)
    size = len(lors) // subsets
    out = []
    s = None if sens is None else np.asarray(sens, np.float64)
    for _ in range(iters):
        for k in range(subsets):
            O.mlem_update(full, n, lors[k * size:(k + 1) * size], x, sens=s, tof=tof, threads=16, f64=not mixed,
                          mixed=mixed)
            out.append(x.copy())
    return out


@pytest.mark.parametrize("precision", [P.PRECISION_F64, P.PRECISION_F32_WALK])
@pytest.mark.parametrize("tof", [None, (42.47, 3.0)])
def test_binary64_modes_match_oracle(gpu, precision, tof):
    full, n = [48.0, 48.0, 40.0], [12, 12, 10]
    rng = np.random.default_rng(8)
    lors = random_lors(rng, 20_003, dt_sigma_ps=84.93 if tof else 0.0)
    lors["ac"] = rng.uniform(1.0, 2.0, len(lors)).astype(np.float32)
    sens = rng.uniform(0.5, 1.5, int(np.prod(n))).astype(np.float32)
    sens[::97] = 0.0
    fov = P.FOV(full, n)
    mixed = precision == P.PRECISION_F32_WALK
    g = _gpu(fov, lors, 3, tof, precision, subsets=2, sens=sens)
    o = _oracle(full, n, lors, 3, tof, mixed, subsets=2, sens=sens)
    worst = max(rel_l2(a, b) for a, b in zip(g, o))
    _say(f"precision {precision} tof {tof}: worst relL2 vs the oracle's twin over 6 updates {worst:.2e}")
    assert worst <= 1e-11
    # the binary32 narrowing of the same image
    ctx = P.Mlem(fov, lors[:5000], tof=P.Tof(*tof) if tof else None, precision=precision)
    img, _ = ctx.step()
    assert np.array_equal(img.data, ctx.image_f64().astype(np.float32))
    ctx.close()


def test_binary64_modes_on_phantom_decompose_the_gap(gpu):
    """120^3 TOF phantom LORs (16.5 % of walks shifted): binary32 tiled GPU
    path vs the GPU mixed mode is rounding; the GPU binary64 mode is a walk
    geometry away; each GPU mode equals its oracle twin."""
    full, n = [180.0] * 3, [120] * 3
    tof = (42.47, 3.0)
    lors = P.synth_phantom_lors(2, 200_000, 200.0)
    fov = P.FOV(full, n)
    it32 = P.Mlem(fov, lors, tof=P.Tof(*tof))
    x32 = []
    for _ in range(2):
        img, _ = it32.step()
        x32.append(img.data.astype(np.float64))
    it32.close()
    gm = _gpu(fov, lors, 2, tof, P.PRECISION_F32_WALK)
    gd = _gpu(fov, lors, 2, tof, P.PRECISION_F64)
    om = _oracle(full, n, lors, 2, tof, True)
    od = _oracle(full, n, lors, 2, tof, False)
    for k in range(2):
        r_round, r_geom = rel_l2(x32[k], gm[k]), rel_l2(gm[k], gd[k])
        r_m, r_d = rel_l2(gm[k], om[k]), rel_l2(gd[k], od[k])
        _say(f"phantom 120^3 iteration {k + 1}: f32-mixed {r_round:.2e}, mixed-f64 {r_geom:.2e}; "
             f"GPU vs oracle: mixed {r_m:.2e}, f64 {r_d:.2e}")
        assert r_m <= 1e-10 and r_d <= 1e-10
        # 0.12 LORs per voxel: rounding of the binary32 weights shows at 1e-5
        assert r_round <= 1e-4 and r_geom >= 1000 * r_round


def test_precision_argument_errors(gpu):
    fov = P.FOV([48.0] * 3, [8] * 3)
    lors = random_lors(np.random.default_rng(1), 100)
    with pytest.raises(P.PetaloError):
        P.Mlem(fov, lors, precision=P.PRECISION_F64, autotune=True)
    with pytest.raises(P.PetaloError):
        P.Mlem(fov, lors, precision=7)
    ctx = P.Mlem(fov, lors)
    with pytest.raises(P.PetaloError):
        ctx.image_f64()  # a binary32 context has no binary64 image
    ctx.close()
