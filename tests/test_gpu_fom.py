"""Figures of merit on the device (petalo_roi_stats, petalo_image_foms,
petalo_phantom_foms) against the oracle's restatement of src/fom.rs and
src/bin/foms.rs.  Region membership is integer work: voxel counts must be
identical.  Statistics accumulate in binary64 on both sides (different
summation order): means within 1e-12, sigmas within 1e-9 relative; against
the reference-order binary32 oracle within 2e-4."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
import petalo_amd as P
from test_fom_oracle import JASZ_GRID, NEMA_GRID, ideal_phantom

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _random_rois(rng, fov, k):
    full = np.asarray(fov.full_size)
    hw = full / 2
    vs = full / np.asarray(fov.n)
    rois = np.zeros(k, P.ROI_DTYPE)
    rois["kind"] = rng.integers(0, 5, k)
    rois["c"] = rng.uniform(-1.2, 1.2, (k, 3)) * hw
    rois["r"] = rng.uniform(-2.0, 0.6 * hw.max(), k)
    # half of the discs on an exact slice centre
    disc = np.flatnonzero(rois["kind"] == P.ROI_DISC_Z)[::2]
    iz = rng.integers(0, fov.n[2], len(disc))
    rois["c"][disc, 2] = ((iz.astype(np.float32) + np.float32(0.5)) * np.float32(vs[2]) - np.float32(hw[2]))
    return rois


def test_roi_stats_match_oracle(gpu):
    rng = np.random.default_rng(8)
    fov = P.FOV([80.0, 72.0, 60.0], [40, 36, 30])
    img = P.Image(fov, rng.gamma(3.0, 1.0, fov.nvox).astype(np.float32))
    rois = _random_rois(rng, fov, 300)
    g = P.roi_stats(img, rois)
    cnt, mean, sig = O.roi_stats(fov.full_size, fov.n, img.data, rois, f64=True)
    assert np.array_equal(g["count"], cnt)
    full = cnt > 0
    assert full.sum() > 150 and (~full).sum() > 10
    assert np.isnan(g["mean"][~full]).all() and np.isnan(g["sigma"][~full]).all()
    np.testing.assert_allclose(g["mean"][full], mean[full], rtol=1e-12)
    np.testing.assert_allclose(g["sigma"][full], sig[full], rtol=1e-9, atol=1e-12)
    _, m32, s32 = O.roi_stats(fov.full_size, fov.n, img.data, rois)
    np.testing.assert_allclose(g["mean"][full], m32[full], rtol=2e-4)


@pytest.mark.parametrize("phantom", [P.PHANTOM_NEMA7, P.PHANTOM_JASZCZAK])
def test_ideal_phantom_closed_form(gpu, phantom):
    full, n = NEMA_GRID if phantom == P.PHANTOM_NEMA7 else JASZ_GRID
    img = P.Image(P.FOV(full, n), ideal_phantom(phantom, full, n))
    foms, aocs = P.phantom_foms(img, phantom)
    np.testing.assert_array_equal(foms[:, 1], 100.0)
    np.testing.assert_array_equal(foms[:, 2], 0.0)
    assert np.isinf(foms[:, 3]).all()
    if phantom == P.PHANTOM_NEMA7:
        assert len(aocs) == 61
        np.testing.assert_array_equal(aocs, 0.25)
    else:
        assert len(aocs) == 0


@pytest.mark.parametrize("phantom", [P.PHANTOM_NEMA7, P.PHANTOM_JASZCZAK])
def test_noisy_phantom_matches_oracle(gpu, phantom):
    full, n = NEMA_GRID if phantom == P.PHANTOM_NEMA7 else JASZ_GRID
    rng = np.random.default_rng(40 + phantom)
    data = ideal_phantom(phantom, full, n) * rng.gamma(10.0, 0.1, size=int(np.prod(n))).astype(np.float32)
    img = P.Image(P.FOV(full, n), data)
    g, ga = P.phantom_foms(img, phantom)
    o, oa = O.phantom_foms(full, n, data, phantom, f64=True)
    o32, oa32 = O.phantom_foms(full, n, data, phantom)
    np.testing.assert_array_equal(g[:, 0], o[:, 0])
    np.testing.assert_allclose(g, o, rtol=1e-5)
    np.testing.assert_allclose(ga, oa, rtol=1e-5)
    np.testing.assert_allclose(g, o32, rtol=2e-3)
    np.testing.assert_allclose(ga, oa32, rtol=2e-3)


def test_image_foms_match_oracle(gpu):
    rng = np.random.default_rng(3)
    fov = P.FOV([60.0] * 3, [30] * 3)
    data = rng.gamma(5.0, 0.2, fov.nvox).astype(np.float32)
    img = P.Image(fov, data)
    rois = np.array([P.roi(P.ROI_SPHERE, [10, 0, 0], 6.0), P.roi(P.ROI_CYLINDER_X, [5, -3], 4.0),
                     P.roi(P.ROI_CYLINDER_Y, [-7, 2], 5.0), P.roi(P.ROI_DISC_Z, [0, 0, 1.0], 9.0)], P.ROI_DTYPE)
    bg = np.array([P.roi(P.ROI_SPHERE, [-15, 10, 5], 5.0), P.roi(P.ROI_CYLINDER_Z, [15, 15], 4.0)], P.ROI_DTYPE)
    act = [4.0, 2.0, 0.5, 0.0]
    crcs, snrs, bgm = P.image_foms(img, rois, act, bg, 1.0)
    oc, osn, obg = O.image_foms(fov.full_size, fov.n, data, rois, act, bg, 1.0, f64=True)
    np.testing.assert_allclose(crcs, oc, rtol=1e-5)
    np.testing.assert_allclose(snrs, osn, rtol=1e-5)
    assert abs(bgm - obg) <= 1e-6 * abs(obg)


def test_empty_region_is_an_error(gpu):
    fov = P.FOV([100.0] * 3, [20] * 3)  # smaller than the NEMA7 background ring
    img = P.Image(fov, np.ones(fov.nvox, np.float32))
    with pytest.raises(P.PetaloError) as e:
        P.phantom_foms(img, P.PHANTOM_NEMA7)
    assert "no voxel" in str(e.value)


def test_foms_cli(gpu, tmp_path):
    full, n = NEMA_GRID
    img = P.Image(P.FOV(full, n), ideal_phantom(P.PHANTOM_NEMA7, full, n))
    path = str(tmp_path / "nema.raw")
    P.write_image3d(path, img)
    r = subprocess.run([os.path.join(ROOT, "petalo_amd", "bin", "foms"), "nema7", path], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].split() == ["Sphere", "diameter", "/", "mm", "CRC", "%", "bg", "var", "%", "SNR", "%"]
    assert lines[1].split()[:3] == ["10.0", "100.0", "0.0"] and lines[6].split()[:3] == ["37.0", "100.0", "0.0"]
    assert lines[-1].startswith("AOCs: [0.2, 0.2,") or lines[-1].startswith("AOCs: [0.3, 0.3,")
