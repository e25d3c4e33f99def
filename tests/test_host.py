"""Host-side logic of the C ABI (no GPU): config TOML, Image3D, FOV strings,
synthetic workloads, the counter-based generator, and the exported symbols."""
import json
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
import petalo_amd as P
from helpers import ulps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
KATS = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))

FULL = """
[input]
file = "f.h5"
dataset = "reco_info/lors"
[iterations]
number = 2
subsets = 3
[fov]
nvoxels = [10, 20, 30]
size = ["123 mm", "456 mm", "78 cm"]
"""


# ---- config (src/config/mlem.rs:265-630) ----------------------------------
def test_config_file_fixture():
    exp = KATS["config_file"]
    c = P.read_config(os.path.join(GOLDEN, "mlem-config.toml"))
    assert c.input_file == exp["input_file"]
    assert c.input_dataset == exp["input_dataset"]
    assert c.iterations_number == exp["iterations_number"]
    assert c.iterations_subsets == exp["iterations_subsets"]
    assert c.tof.sigma_ps == exp["tof_sigma_ps"] and c.tof.cutoff == exp["tof_cutoff"]
    assert list(c.fov_nvoxels) == exp["fov_nvoxels"]
    assert np.allclose(c.fov_size, exp["fov_size_mm"])
    assert c.sensitivity_image == exp["sensitivity_image"]
    assert c.energy == (434.0, 600.0) and c.charge == (123.0, 456.0) and c.events == (100000, 200000)
    assert c.has_scatter_correction
    # the fixture's [scatter_correction] table
    assert c.scatter_correction == {"phi_bins": 30, "r_bins": 31, "r_max_mm": 32.0, "dz_bins": 33, "dz_max_mm": 34.0,
                                    "dt_bins": 35, "dt_max_ps": 36.0, "z_bins": 37, "z_length_mm": 380.0}
    assert c.smear_energy_fwhm is None and c.detector_full_axial_length is None


def test_config_scattergram_kat():
    """config_scattergram (src/config/mlem.rs:474-510)."""
    c = P.parse_config(FULL + """
[scatter_correction]
phi.bins = 12
r  .bins = 34
dz .bins = 99
dt .bins = 98
z  .bins = 97

r  .max  = "56 mm"
dz .max  = "78 mm"
dt .max  = "90 ps"
z.length = "38 cm"
""")
    assert c.scatter_correction == {"phi_bins": 12, "r_bins": 34, "r_max_mm": 56.0, "dz_bins": 99, "dz_max_mm": 78.0,
                                    "dt_bins": 98, "dt_max_ps": 90.0, "z_bins": 97, "z_length_mm": 380.0}
    part = P.parse_config(FULL + '[scatter_correction]\nphi.bins = 7\n')
    assert part.scatter_correction["phi_bins"] == 7 and part.scatter_correction["r_bins"] == 0
    with pytest.raises(P.PetaloError):  # BinsMax needs both fields
        P.parse_config(FULL + '[scatter_correction]\nr.bins = 7\n')


def test_config_to_ingestion_parameters():
    """read_lors's use of the config (hdf5.rs:97-130, 240-253): bounds pass
    through, z_max = axial length / 2, RandomZPhi with the dataset's
    discretization; the smearing table is mandatory."""
    text = open(os.path.join(GOLDEN, "mlem-config.toml")).read()
    with pytest.raises(P.PetaloError):
        P.parse_config(text).ingest_params((350.0, 30.0, 3.0, 3.0))
    c = P.parse_config(text + '\n[smear_energy]\nfwhm = "15 %"\n[detector_full_axial_length]\ndz = "1 m"\n')
    assert abs(c.smear_energy_fwhm - 0.15) < 1e-7 and c.detector_full_axial_length == 1000.0
    p = c.ingest_params((350.0, 30.0, 3.0, 3.0), seed=9)
    assert (p.has_events_min, p.has_events_max, p.events_min, p.events_max) == (1, 1, 100000, 200000)
    assert (p.energy_min, p.energy_max, p.charge_min, p.charge_max) == (434.0, 600.0, 123.0, 456.0)
    assert p.has_z_max == 1 and p.z_max_mm == 500.0 and abs(p.smear_energy_fwhm - 0.15) < 1e-7
    assert (p.r_min_mm, p.dr_mm, p.dz_mm, p.da_mm) == (350.0, 30.0, 3.0, 3.0)
    assert p.adjust == P.ADJUST_RANDOM_Z_PHI and p.seed == 9


def test_config_fov_units():
    c = P.parse_config(FULL)
    assert c.fov_nvoxels == (10, 20, 30)
    assert c.fov_size == (123.0, 456.0, 780.0)
    assert c.tof is None and c.sensitivity_image is None


def test_config_tof_default_cutoff_and_units():
    c = P.parse_config(FULL + '[tof]\nsigma = "123.4 ps"\n')
    assert c.tof.cutoff == 3.0 and abs(c.tof.sigma_ps - 123.4) < 1e-4
    c = P.parse_config(FULL + '[tof]\nsigma = "2 ns"\ncutoff = 4\n')
    assert c.tof.sigma_ps == 2000.0 and c.tof.cutoff == 4.0


def test_config_inline_table_and_dotted_keys():
    text = "iterations = { number = 4, subsets = 20 }\n" + FULL.replace("[iterations]\nnumber = 2\nsubsets = 3\n", "")
    c = P.parse_config(text)
    assert (c.iterations_number, c.iterations_subsets) == (4, 20)
    c = P.parse_config(FULL.replace('dataset = "reco_info/lors"',
                                    'dataset = "x"\nenergy.min = 123\ncharge .min = 444\nevents.max = 100_000'))
    assert c.energy == (123.0, None) and c.charge == (444.0, None) and c.events == (None, 100000)


@pytest.mark.parametrize("extra", ["unknown_field = 666\n", "[tof]\nsigma = \"1 ps\"\nwidth = 3\n",
                                   "[attenuation_correction]\nimage = \"x\"\n"])
def test_config_rejects_unknown_fields(extra):
    with pytest.raises(P.PetaloError) as e:
        P.parse_config(extra + FULL if extra.startswith("unknown") else FULL + extra)
    assert e.value.code == 5


@pytest.mark.parametrize("drop", ["[iterations]\nnumber = 2\nsubsets = 3\n", 'file = "f.h5"\n',
                                  'size = ["123 mm", "456 mm", "78 cm"]\n'])
def test_config_mandatory_fields(drop):
    with pytest.raises(P.PetaloError):
        P.parse_config(FULL.replace(drop, ""))


def test_quantities():
    assert P.parse_quantity("2 ps", "T") == 2.0
    assert P.parse_quantity("2 ns", "T") == 2000.0
    assert P.parse_quantity("3 mm", "L") == 3.0
    assert P.parse_quantity("38 cm", "L") == 380.0
    for bad in ("3", "3mm", "3 parsecs"):
        with pytest.raises(P.PetaloError):
            P.parse_quantity(bad, "L")


def test_config_attenuation_and_scatter_tables():
    c = P.parse_config(FULL + '[attenuation_correction]\nsensitivity_image = "s.raw"\n'
                       '[scatter_correction]\nphi.bins = 12\nr.bins = 34\nr.max = "56 mm"\n')
    assert c.sensitivity_image == "s.raw" and c.has_scatter_correction
    with pytest.raises(P.PetaloError):
        P.parse_config(FULL + '[scatter_correction]\nr.bins = 34\n')  # r.max missing


# ---- FOV -------------------------------------------------------------------
def test_fov_from_str():
    f = P.FOV.from_str("1 mm  2 mm  0.2 cm  100 100 100")
    assert f.full_size == (1.0, 2.0, 2.0) and f.n == (100, 100, 100)


def test_fov_derive_matches_oracle():
    rng = np.random.default_rng(4)
    for _ in range(200):
        full = rng.uniform(1, 500, 3).astype(np.float32)
        n = rng.integers(1, 300, 3)
        f = P.FOV(full, n)
        hw, vs = O.fov_derive(full, n)
        assert np.array_equal(f.half_width, hw) and np.array_equal(f.voxel_size, vs)


# ---- Image3D (src/io/raw.rs:78-191) ----------------------------------------
def test_image3d_guinea_pig(tmp_path):
    g = KATS["image3d_guinea_pig"]
    fov = P.FOV(g["mm"], g["pixels"])
    img = P.Image(fov, np.array(g["data"], np.float32))
    path = tmp_path / "gp.raw"
    P.write_image3d(str(path), img)
    expected = struct.pack(">3H", *g["pixels"]) + struct.pack(">3f", *g["mm"]) + struct.pack(">6f", *g["data"])
    assert path.read_bytes() == expected
    back = P.read_image3d(str(path))
    assert back.fov.n == fov.n and back.fov.full_size == fov.full_size
    assert np.array_equal(back.data, img.data)


def test_image3d_errors(tmp_path):
    path = tmp_path / "short.raw"
    path.write_bytes(struct.pack(">3H", 2, 2, 2) + struct.pack(">3f", 1, 1, 1) + b"\0" * 8)
    with pytest.raises(P.PetaloError):
        P.read_image3d(str(path))
    with pytest.raises(P.PetaloError):
        P.Image(P.FOV([1, 1, 1], [2, 2, 2]), np.zeros(7, np.float32))


# ---- synthetic workloads ---------------------------------------------------
def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32_10."""
    assert list(P.philox4x32_10([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(P.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2)) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(P.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                [0xa4093822, 0x299f31d0])) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_phantom_lors():
    a = P.synth_phantom_lors(5, 20000, 200.0)
    b = P.synth_phantom_lors(5, 1000, 200.0)
    assert np.array_equal(a[:1000], b)  # LOR i depends only on (seed, i)
    for end in ("p1", "p2"):
        r = np.hypot(a[end][:, 0].astype(np.float64), a[end][:, 1])
        assert np.all(np.abs(r - 350.0) < 1e-3)
        assert np.all(np.abs(a[end][:, 2]) <= 150.0 + 1e-3)
    # exact dt (no noise): the emission point C*dt/2 from the midpoint lies in the phantom
    c = P.synth_phantom_lors(6, 5000, 0.0)
    p1, p2 = c["p1"].astype(np.float64), c["p2"].astype(np.float64)
    d = p2 - p1
    L = np.linalg.norm(d, axis=1)
    t = 0.5 - 0.299792458 * c["dt"].astype(np.float64) / 2 / L
    e = p1 + t[:, None] * d
    assert np.all(np.hypot(e[:, 0], e[:, 1]) <= 80.0 + 0.05) and np.all(np.abs(e[:, 2]) <= 75.0 + 0.05)
    # hot spheres are 4x denser: counts inside the largest sphere vs water
    inside = np.hypot(e[:, 0] - 50 * np.cos(5 * np.pi / 3), e[:, 1] - 50 * np.sin(5 * np.pi / 3)) ** 2 + e[:, 2] ** 2 <= 15 ** 2
    vol_sphere = 4 / 3 * np.pi * 15 ** 3
    density_ratio = (inside.sum() / vol_sphere) / ((~inside).sum() / (np.pi * 80 ** 2 * 150 - vol_sphere))
    assert 2.5 < density_ratio < 6.0


def test_density_and_attenuation():
    fov = P.FOV([180.0] * 3, [36] * 3)
    rho = P.synth_density(fov)
    vals = set(np.unique(rho.data).tolist())
    assert vals == {0.0, 1000.0, np.float32(1791.67).item()}
    mu = P.density_to_attenuation(P.Image(fov, rho.data.copy()), 0.096)
    f = O.rho_to_mu_factor(0.096)
    assert np.array_equal(mu.data, (rho.data * np.float32(f)).astype(np.float32))
    assert abs(np.max(mu.data) - 0.0172) < 1e-5 and abs(1000 * f - 0.0096) < 1e-8


# ---- C ABI surface -----------------------------------------------------------
def _declared_functions():
    text = open(os.path.join(ROOT, "include", "petalo_hip.h")).read()
    return sorted(set(re.findall(r"\b(petalo_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = _declared_functions()
    assert len(names) >= 25
    out = subprocess.check_output(["nm", "-D", "--defined-only", P.LIB_PATH]).decode()
    exported = set(re.findall(r"\bT (petalo_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = P.lib()
    for n in names:
        assert hasattr(L, n)


def test_no_gpu_reports_error_not_crash():
    if P.device_count() > 0:
        pytest.skip("a GPU is visible")
    lors = P.make_lors([[-100, 0, 0]], [[100, 0, 0]])
    with pytest.raises(P.PetaloError) as e:
        P.Mlem(P.FOV([10.0] * 3, [2] * 3), lors)
    assert e.value.code == 2


def test_mlem_option_checks_come_first():
    """Argument errors are reported before any device work (no GPU needed)."""
    lors = P.make_lors([[-100, 0, 0]], [[100, 0, 0]])
    fov = P.FOV([10.0] * 3, [2] * 3)
    with pytest.raises(P.PetaloError) as e:
        P.Mlem(fov, lors, tile=(8, -1, 4))
    assert e.value.code == 1 and "tile" in str(e.value)
    with pytest.raises(P.PetaloError) as e:
        P.Mlem(fov, lors, scheme=7)
    assert e.value.code == 1


def test_ingest_argument_checks_come_first():
    rec = np.zeros((3, 11), np.float32)
    with pytest.raises(P.PetaloError) as e:
        P.ingest_lors(rec, P.ingest_params(adjust=7))
    assert e.value.code == 1
    with pytest.raises(P.PetaloError) as e:
        P.ingest_lors(rec, P.ingest_params(adjust=P.ADJUST_RANDOM_Z_PHI, discretize=(350.0, 0.0, 3.0, 3.0)))
    assert e.value.code == 1
    out, n = P.ingest_lors(rec, P.ingest_params(events=(5, 9), adjust=P.ADJUST_NONE))
    assert n == 0 and len(out) == 0  # empty event range: nothing to do, no device needed


def test_discrete_sensitivity_argument_checks_come_first():
    fov = P.FOV([60.0] * 3, [6] * 3)
    mu = P.Image(fov, np.zeros(fov.nvox, np.float32))
    bad = [dict(r_min=100.0, dr=10.0, dz=10.0, da=10.0, adjust=9),
           dict(r_min=100.0, dr=0.0, dz=10.0, da=10.0),
           dict(r_min=100.0, dr=10.0, dz=-1.0, da=10.0),
           dict(r_min=0.0, dr=1.0, dz=10.0, da=1000.0)]  # one element: no pairs
    for kw in bad:
        with pytest.raises(P.PetaloError) as e:
            P.make_sensitivity_image_discrete(mu, 10.0, **kw)
        assert e.value.code == 1, kw


def test_fom_argument_checks_come_first():
    fov = P.FOV([60.0] * 3, [6] * 3)
    img = P.Image(fov, np.ones(fov.nvox, np.float32))
    with pytest.raises(P.PetaloError) as e:
        P.phantom_foms(img, 7)
    assert e.value.code == 1
    for bad in (P.roi(9, [0, 0, 0], 1.0), P.roi(P.ROI_SPHERE, [0, np.nan, 0], 1.0),
                P.roi(P.ROI_SPHERE, [0, 0, 0], np.inf)):
        with pytest.raises(P.PetaloError) as e:
            P.roi_stats(img, [bad])
        assert e.value.code == 1
    with pytest.raises(P.PetaloError):
        P.image_foms(img, [P.roi(P.ROI_SPHERE, [0, 0, 0], 5.0)], [2.0], [], 1.0)  # no background
    assert len(P.roi_stats(img, np.zeros(0, P.ROI_DTYPE))) == 0  # nothing to do, no device needed


def test_tools_help():
    for tool in ("mlem", "make_sensitivity_image", "foms"):
        exe = os.path.join(ROOT, "petalo_amd", "bin", tool)
        out = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=30)
        assert out.returncode == 0 and "usage" in out.stdout
