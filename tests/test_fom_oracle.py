"""The oracle's figures of merit (src/fom.rs, src/bin/foms.rs) pinned by the
reference's own known answers (region voxel counts, which-voxel patterns,
mean and sigma) and by closed forms: an ideal phantom image, whose hot
spheres, background and lung insert hold exact activities, has CRC 100 %,
zero background variability, infinite SNR and AOC equal to the insert's
activity (the reference divides by the 37 mm sphere's CRC)."""
import json
import math
import os

import numpy as np
import pytest

import oracle_ffi as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
KATS = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))["fom_roi_counts"]
MAGIC = float(np.float32(10.0) / np.float32(9.0) * np.float32(200.0))
KIND = {"sphere": 0, "cylinder_x": 1, "cylinder_y": 2, "cylinder_z": 3, "disc_z": 4}


def _roi(kind, c, r):
    c = list(c) + [0.0] * (3 - len(c))
    return np.array([(kind, c, r)], O.ROI_DTYPE)


@pytest.mark.parametrize("case", KATS["sphere"])
def test_voxels_in_sphere(case):
    n, c, r, expected = case
    cnt, _, _ = O.roi_stats([MAGIC] * 3, [n] * 3, np.ones(n ** 3, np.float32), _roi(0, c, r))
    assert int(cnt[0]) == expected


@pytest.mark.parametrize("case", KATS["z_cylinder"])
def test_voxels_in_z_cylinder(case):
    n, c, r, expected = case
    cnt, _, _ = O.roi_stats([MAGIC] * 3, [n] * 3, np.ones(n ** 3, np.float32), _roi(3, c, r))
    assert int(cnt[0]) == expected


@pytest.mark.parametrize("case", KATS["which_voxels"]["cases"])
def test_which_voxels(case):
    kind, c, expected = case
    val = {"I": 1.0, "O": -1.0}
    img = np.array([1, 2, 4, 8, 16, 32, 64, 128], np.float32)
    cnt, mean, _ = O.roi_stats([4.0] * 3, [2] * 3, img, _roi(KIND[kind], [val[v] for v in c], 1.0))
    assert round(float(cnt[0]) * mean[0]) == expected


@pytest.mark.parametrize("f64", [False, True])
def test_mean_and_sigma(f64):
    for values, mu in KATS["mean"]:
        n = len(values)
        cnt, m, _ = O.roi_stats([float(n), 1.0, 1.0], [n, 1, 1], np.array(values, np.float32),
                                _roi(3, [0.0, 0.0], 1e9), f64=f64)
        assert cnt[0] == n and m[0] == mu
    (v1, mu1, s1), (v2, mu2, _) = KATS["mu_and_sigma"]
    _, m, s = O.roi_stats([3.0, 1.0, 1.0], [3, 1, 1], np.array(v1, np.float32), _roi(3, [0, 0], 1e9), f64=f64)
    assert m[0] == mu1 and s[0] == s1
    _, m, s = O.roi_stats([5.0, 1.0, 1.0], [5, 1, 1], np.array(v2, np.float32), _roi(3, [0, 0], 1e9), f64=f64)
    assert m[0] == mu2
    assert s[0] == (float(np.sqrt(np.float32(2.0))) if not f64 else math.sqrt(2.0))


def test_disc_needs_an_exact_slice_centre():
    """DiscZ compares z exactly: a disc between slice centres is empty."""
    img = np.ones(8 ** 3, np.float32)
    full, n = [16.0] * 3, [8] * 3  # centres at -7, -5, ..., 7
    cnt, _, _ = O.roi_stats(full, n, img, np.concatenate([_roi(4, [0, 0, 1.0], 3.0), _roi(4, [0, 0, 2.0], 3.0)]))
    assert cnt[0] == 4 and cnt[1] == 0  # centres (+-1, +-1): x^2 + y^2 = 2 < 9; no slice centre at z = 2


def _grid(full, n):
    hw = np.float32(full) / np.float32(2)
    vs = hw * np.float32(2) / np.asarray(n, np.float32)
    axes = [(np.arange(n[a], dtype=np.float32) + np.float32(0.5)) * vs[a] - hw[a] for a in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return x.astype(np.float64).ravel(), y.astype(np.float64).ravel(), z.astype(np.float64).ravel()


def ideal_phantom(phantom, full, n, lung=0.25):
    """Background 1, hot spheres 4 (1 mm wider than the analysis discs, at the
    analysis height), NEMA7 lung insert (r 16 mm, all z) at `lung`."""
    x, y, z = _grid(full, n)
    img = np.ones(x.size, np.float32)
    if phantom == 0:
        ring, zc, diam, pos = 57.2, 0.0, [10.0, 13.0, 17.0, 22.0, 28.0, 37.0], [1, 2, 3, 4, 5, 0]
        img[x ** 2 + y ** 2 < 16.0 ** 2] = lung
    else:
        ring, zc, diam, pos = 54.0, 34.0, [9.5, 12.7, 15.9, 19.1, 25.4, 31.8], [0, 1, 2, 3, 4, 5]
    for d, p in zip(diam, pos):
        a = 2 * math.pi * p / 6
        sx, sy = ring * math.cos(a), ring * math.sin(a)
        img[(x - sx) ** 2 + (y - sy) ** 2 + (z - zc) ** 2 < (d / 2 + 1.0) ** 2] = 4.0
    return img


NEMA_GRID = ([260.0, 260.0, 182.0], [65, 65, 91])     # 4 x 4 x 2 mm; slice centres at 0 and every 2 mm
JASZ_GRID = ([200.0, 200.0, 182.0], [50, 50, 91])


@pytest.mark.parametrize("f64", [False, True])
def test_ideal_nema7(f64):
    full, n = NEMA_GRID
    foms, aocs = O.phantom_foms(full, n, ideal_phantom(0, full, n), 0, f64=f64)
    np.testing.assert_array_equal(foms[:, 0], [5.0, 6.5, 8.5, 11.0, 14.0, 18.5])
    np.testing.assert_array_equal(foms[:, 1], 100.0)   # CRC
    np.testing.assert_array_equal(foms[:, 2], 0.0)     # background variability
    assert np.isinf(foms[:, 3]).all() and (foms[:, 3] > 0).all()  # sigma 0 inside each sphere
    # slices whose centres lie in [-80, 40] mm: 61 at 2 mm; mean 0.25 / CRC 100 %, x 100
    assert len(aocs) == 61
    np.testing.assert_array_equal(aocs, 0.25)


@pytest.mark.parametrize("f64", [False, True])
def test_ideal_jaszczak(f64):
    full, n = JASZ_GRID
    foms, aocs = O.phantom_foms(full, n, ideal_phantom(1, full, n), 1, f64=f64)
    np.testing.assert_allclose(foms[:, 0], [4.75, 6.35, 7.95, 9.55, 12.7, 15.9], rtol=1e-6)
    np.testing.assert_array_equal(foms[:, 1], 100.0)
    np.testing.assert_array_equal(foms[:, 2], 0.0)
    assert len(aocs) == 0


def test_noisy_phantom_f32_close_to_f64():
    """Reference-order binary32 and binary64 statistics agree on a noisy
    image; a cold sphere scores crc = 100 (1 - measured / background)."""
    full, n = NEMA_GRID
    rng = np.random.default_rng(5)
    img = ideal_phantom(0, full, n) * rng.gamma(20.0, 1 / 20.0, size=int(np.prod(n))).astype(np.float32)
    a, aa = O.phantom_foms(full, n, img, 0)
    b, bb = O.phantom_foms(full, n, img, 0, f64=True)
    np.testing.assert_allclose(a, b, rtol=2e-4)
    np.testing.assert_allclose(aa, bb, rtol=2e-4)
    assert 80 < b[:, 1].min() and b[:, 1].max() < 120 and (b[:, 2] > 0).all()


def test_image_foms_closed_form():
    full, n = [20.0] * 3, [10] * 3
    x, y, z = _grid(full, n)
    img = np.ones(x.size, np.float32)
    hot = (x - 5) ** 2 + y ** 2 + z ** 2 < 9
    cold = (x + 5) ** 2 + y ** 2 + z ** 2 < 9
    img[hot] = 3.0
    img[cold] = 0.5
    rois = np.concatenate([_roi(0, [5, 0, 0], 3.0), _roi(0, [-5, 0, 0], 3.0)])
    bg = np.concatenate([_roi(0, [0, 6, 0], 2.5), _roi(0, [0, -6, 5], 2.5)])
    for f64 in (False, True):
        crcs, snrs, bgm = O.image_foms(full, n, img, rois, [3.0, 0.0], bg, 1.0, f64=f64)
        assert bgm == 1.0
        np.testing.assert_array_equal(crcs, [100.0, 50.0])  # hot: (3/1-1)/(3/1-1); cold: 1 - 0.5/1
        assert np.isinf(snrs).all()
