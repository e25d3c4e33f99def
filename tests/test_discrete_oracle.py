"""The oracle's discretized-detector LORs (discrete.rs:150-176 element
centres, make_sensitivity_image/discrete.rs:40-62 pairs and adjustment),
pinned by closed forms: element count and order, pair enumeration, centres
that the Centre adjustment reproduces, and draws that stay inside their
element under the random adjustments."""
import math

import numpy as np
import pytest

import oracle_ffi as O

TAU = 2.0 * math.pi


def _geometry(r_min, dr, dz, da, length):
    r_mid = r_min + dr / 2.0
    n_az = int(round(TAU * r_mid / da))
    n_half = int(round(length / 2.0 / dz))
    return r_mid, n_az, n_half


def _centres(r_min, dr, dz, da, length):
    r_mid, n_az, n_half = _geometry(r_min, dr, dz, da, length)
    d_phi = TAU / n_az
    out = []
    for a in range(n_az):
        for nz in range(-n_half, n_half + 1):
            out.append((r_mid * math.cos(a * d_phi), r_mid * math.sin(a * d_phi), nz * dz))
    return np.array(out), n_az, d_phi


@pytest.mark.parametrize("geom", [(30.0, 2.0, 3.0, 4.0, 20.0), (100.0, 10.0, 5.0, 7.5, 31.0), (5.0, 1.0, 1.0, 1.0, 0.4)])
def test_pairs_of_element_centres(geom):
    """Adjustment none: every pair (i, k >= i) of element centres, k-major
    within row i, LOR from element k to element i, dt 0, correction 1."""
    c, _, _ = _centres(*geom)
    lors, ne = O.discrete_lors(*geom, adjust=0)
    n = len(c)
    assert ne == n and len(lors) == n * (n + 1) // 2
    i, k = np.triu_indices(n)  # row-major: i outer, k from i
    np.testing.assert_allclose(lors["p1"], c[k], atol=2e-4 * max(1.0, np.abs(c).max()))
    np.testing.assert_allclose(lors["p2"], c[i], atol=2e-4 * max(1.0, np.abs(c).max()))
    assert (lors["dt"] == 0).all() and (lors["ac"] == 1).all()


def test_centre_adjustment_is_idempotent_on_centres():
    geom = (30.0, 2.0, 3.0, 4.0, 20.0)
    a, _ = O.discrete_lors(*geom, adjust=0)
    b, _ = O.discrete_lors(*geom, adjust=1, seed=7)
    np.testing.assert_allclose(b["p1"], a["p1"], atol=1e-4)
    np.testing.assert_allclose(b["p2"], a["p2"], atol=1e-4)


@pytest.mark.parametrize("adjust", [2, 3])
def test_random_adjustments_stay_in_their_element(adjust):
    geom = (30.0, 2.0, 3.0, 4.0, 20.0)
    r_min, dr, dz, _, _ = geom
    c, n_az, d_phi = _centres(*geom)
    r_mid = r_min + dr / 2.0
    lors, n = O.discrete_lors(*geom, adjust=adjust, seed=11)
    i, k = np.triu_indices(n)
    for pts, idx in ((lors["p1"], k), (lors["p2"], i)):
        ctr = c[idx]
        assert np.abs(pts[:, 2] - ctr[:, 2]).max() <= dz / 2 + 1e-4
        dphi = np.angle(np.exp(1j * (np.arctan2(pts[:, 1], pts[:, 0]) - np.arctan2(ctr[:, 1], ctr[:, 0]))))
        assert np.abs(dphi).max() <= d_phi / 2 + 1e-5
        r = np.hypot(pts[:, 0], pts[:, 1])
        if adjust == 2:
            assert r.min() >= r_mid - dr / 2 - 1e-4 and r.max() <= r_mid + dr / 2 + 1e-4
            assert r.std() > dr / 10  # the radial draw is used
        else:
            np.testing.assert_allclose(r, r_mid, rtol=1e-5)  # random z-phi keeps the radius
        # uniform draws: spread over most of the element
        assert np.abs(pts[:, 2] - ctr[:, 2]).max() > 0.4 * dz
    # draws differ between seeds and between the two ends of one pair
    other, _ = O.discrete_lors(*geom, adjust=adjust, seed=12)
    assert not np.array_equal(other["p1"], lors["p1"])
    diag = i == k
    assert not np.allclose(lors["p1"][diag], lors["p2"][diag])


def test_capacity_truncates():
    full, n = O.discrete_lors(30.0, 2.0, 3.0, 4.0, 20.0, adjust=3, seed=1)
    part, _ = O.discrete_lors(30.0, 2.0, 3.0, 4.0, 20.0, adjust=3, seed=1, capacity=100)
    assert len(part) == 100 and np.array_equal(part, full[:100])
