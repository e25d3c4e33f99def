"""Oracle pins of the LOR ingestion (src/io/hdf5.rs:97-228, src/discrete.rs,
src/lorogram.rs): the reference's own properties of element adjustment and
smearing (discrete.rs:226-268), the Philox known-answer vectors, cut
semantics, event slicing and the scattergram's value rule."""
import math

import numpy as np
import pytest

import oracle_ffi as O
import petalo_amd as P


def _records(n, rng, r=350.0, zhalf=150.0, e=511.0):
    rec = np.zeros(n, P.H5LOR_DTYPE)
    a1, a2 = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    rec["x1"], rec["y1"] = r * np.cos(a1), r * np.sin(a1)
    rec["x2"], rec["y2"] = r * np.cos(a2), r * np.sin(a2)
    rec["z1"], rec["z2"] = rng.uniform(-zhalf, zhalf, n), rng.uniform(-zhalf, zhalf, n)
    rec["dt"] = np.arange(n, dtype=np.float32) * np.float32(1e-3)  # unique per record
    rec["q1"], rec["q2"] = rng.uniform(0, 100, n), rng.uniform(0, 100, n)
    rec["E1"], rec["E2"] = np.full(n, e, np.float32), np.full(n, e, np.float32)
    return rec


def test_oracle_philox_known_answers():
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0])) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_element_centres_invariant_under_centre_adjustment():
    """discrete.rs:234-251 restated: element centres are fixed points of the
    adjustment (Centre here; the reference proptest uses the same centres)."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        r_min, dr, dz, da = rng.uniform(300, 400), rng.uniform(10, 30), rng.uniform(1, 5), rng.uniform(1, 5)
        r = np.float32(np.float32(r_min) + np.float32(dr) / np.float32(2))
        n_az = round(float(np.float32(np.float32(2 * math.pi) * r) / np.float32(da)))
        dphi = np.float32(np.float32(6.2831855) / np.float32(n_az))
        rad = np.float32(np.float32(r / np.float32(dr)) * np.float32(dr))
        k = rng.integers(0, n_az, 50)
        nz = rng.integers(-10, 10, 50)
        phi = (k * dphi).astype(np.float32)
        rec = np.zeros(50, P.H5LOR_DTYPE)
        rec["x1"], rec["y1"], rec["z1"] = rad * np.cos(phi), rad * np.sin(phi), (nz * np.float32(dz))
        rec["x2"], rec["y2"], rec["z2"] = -rec["x1"], -rec["y1"], rec["z1"]
        rec["E1"] = rec["E2"] = 520.0  # above the pre-cut of 511 keV at fwhm 0
        p = P.ingest_params(discretize=(r_min, dr, dz, da), adjust=P.ADJUST_CENTRE)
        out, n = O.ingest(rec, p)
        assert len(out) == 50
        assert np.allclose(out["p1"][:, 0], rec["x1"], atol=1e-3)
        assert np.allclose(out["p1"][:, 1], rec["y1"], atol=1e-3)
        assert np.allclose(out["p1"][:, 2], rec["z1"], atol=1e-3)


def test_random_z_phi_stays_in_element_and_keeps_radius():
    """RandomZPhi (discrete.rs:107-116): z and phi uniform over the element,
    r kept; smear statistics as discrete.rs:254-268 (spread 1, mean at the
    centre, in element units)."""
    n = 20000
    r_min, dr, dz, da = 350.0, 20.0, 4.0, 4.0
    rec = np.zeros(n, P.H5LOR_DTYPE)
    x, y, z = 359.0, 12.0, 21.0
    rec["x1"], rec["y1"], rec["z1"] = x, y, z
    rec["x2"], rec["y2"], rec["z2"] = -x, -y, -z
    rec["E1"] = rec["E2"] = 520.0
    out, _ = O.ingest(rec, P.ingest_params(discretize=(r_min, dr, dz, da), seed=9))
    p = out["p1"].astype(np.float64)
    rr = np.hypot(p[:, 0], p[:, 1])
    assert np.allclose(rr, math.hypot(x, y), rtol=1e-6)
    nz = round(z / dz)
    zs = p[:, 2] / dz
    assert abs((zs.max() - zs.min()) - 1.0) < 0.01 and abs(zs.mean() - nz) < 0.01
    assert np.all((zs >= nz - 0.5) & (zs < nz + 0.5 + 1e-6))
    n_az = round(2 * math.pi * (r_min + dr / 2) / da)
    dphi = 2 * math.pi / n_az
    ph = np.arctan2(p[:, 1], p[:, 0]) / dphi
    n_phi = round(math.atan2(y, x) / dphi)
    assert abs((ph.max() - ph.min()) - 1.0) < 0.01 and abs(ph.mean() - n_phi) < 0.01


def test_energy_smearing_statistics():
    """hdf5.rs:230-238: Normal(E, fwhm * E / 2.35)."""
    rng = np.random.default_rng(2)
    rec = _records(100000, rng, e=511.0)
    fwhm = 0.1
    p = P.ingest_params(smear_energy_fwhm=fwhm, energy=(None, None), adjust=P.ADJUST_NONE, seed=3)
    # the energies are not in the LOR; recover them through the energy cut:
    # the fraction above E = 511 + k sigma must follow the normal tail
    sigma = fwhm * 511.0 / 2.35
    for k in (-1.0, 0.0, 1.0):
        cut = 511.0 + k * sigma
        q = P.ingest_params(smear_energy_fwhm=fwhm, energy=(cut, None), adjust=P.ADJUST_NONE, seed=3)
        out, n = O.ingest(rec, q)
        tail = 0.5 * math.erfc(k / math.sqrt(2))
        assert abs(len(out) / n - tail * tail) < 0.01  # both ends above the cut, independent draws
    out, n = O.ingest(rec, p)
    assert len(out) == n


def test_cuts_and_event_range():
    rng = np.random.default_rng(4)
    rec = _records(1000, rng, e=511.0)
    rec["E1"][::5] = 400.0  # below the pre-cut 511 * (1 - 1.6 * 0.05) = 470.12
    p = P.ingest_params(events=(100, 900), charge=(10.0, 90.0), smear_energy_fwhm=0.05, z_max=120.0,
                        adjust=P.ADJUST_NONE, energy=(None, None))
    out, n = O.ingest(rec, p)
    assert n == 800
    sel = rec[100:900]
    m = ((np.abs(sel["z1"]) < 120) & (np.abs(sel["z2"]) < 120) & (sel["q1"] >= 10) & (sel["q1"] < 90) &
         (sel["q2"] >= 10) & (sel["q2"] < 90) & (sel["E1"] > 470.12) & (sel["E2"] > 470.12))
    assert len(out) == int(m.sum())
    assert np.array_equal(out["dt"], (sel["dt"][m] * np.float32(1000)).astype(np.float32))
    assert np.all(out["ac"] == 1.0)


def test_scattergram_value_rule():
    """lorogram.rs:65-74: (scatters + trues) / trues per bin; FLT_MAX for a
    bin without trues.  One r bin over [0, 30) mm plus flow bins."""
    rng = np.random.default_rng(6)
    rec = _records(400, rng, e=511.0)
    # make LOR i's distance from the axis ~ 10 mm (bin 1) or ~ 50 mm (overflow)
    t = rng.uniform(0, 2 * np.pi, 400)
    d = np.where(np.arange(400) % 2 == 0, 10.0, 50.0)
    ux, uy = np.cos(t), np.sin(t)
    rec["x1"], rec["y1"] = d * -uy + 300 * ux, d * ux + 300 * uy
    rec["x2"], rec["y2"] = d * -uy - 300 * ux, d * ux - 300 * uy
    # Classification must not depend on the draws: with fwhm 0.3 the pre-cut
    # is 511 * (1 - 0.48) = 265.7 keV, and a Box-Muller normal from 24-bit
    # uniforms is bounded by sqrt(-2 ln 2^-24) = 5.77, so E = 280 smears to
    # at most 280 + 5.77 * 0.3 * 280 / 2.35 = 486 < 510 (scatter) and
    # E = 2000 to at least 528 > 510 (true).
    rec["E1"] = rec["E2"] = 2000.0
    idx = np.arange(400)
    rec["E2"][idx % 4 == 0] = 280.0   # half the near LORs are scatters
    rec["E2"][idx % 2 == 1] = 280.0   # every far LOR is a scatter: no trues there
    spec = P.scattergram_spec(r=(1, 30.0))
    p = P.ingest_params(adjust=P.ADJUST_NONE, energy=(None, None), smear_energy_fwhm=0.3)
    out, _ = O.ingest(rec, p, spec)
    near = out["ac"][0::2]
    far = out["ac"][1::2]
    assert np.allclose(near, (100 + 100) / 100)  # 100 scatters, 100 trues among the 200 near LORs
    assert np.all(far == np.float32(3.4028235e38))
