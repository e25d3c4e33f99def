"""The command-line drop-ins (petalo_amd/bin/mlem, make_sensitivity_image)
run end to end on the device and write what the library's entry points
compute for the same arguments (src/bin/mlem.rs:47-105,
src/bin/make_sensitivity_image/main.rs:6-74)."""
import os
import subprocess

import numpy as np
import pytest

import petalo_amd as P
from helpers import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "petalo_amd", "bin")

pytestmark = pytest.mark.gpu


def _mu(rho, rho_to_mu):
    """Attenuation from a copy (the conversion works in place)."""
    return P.density_to_attenuation(P.Image(rho.fov, rho.data.copy()), rho_to_mu)


def _run(args, timeout=120):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


@pytest.fixture(scope="module")
def density(tmp_path_factory):
    d = tmp_path_factory.mktemp("sens")
    fov = P.FOV([120.0] * 3, [24] * 3)
    rho = P.synth_density(fov)
    path = str(d / "density.raw")
    P.write_image3d(path, rho)
    return d, rho


def test_make_sensitivity_image_continuous_cli(gpu, density):
    d, rho = density
    out = str(d / "cont.raw")
    _run([os.path.join(BIN, "make_sensitivity_image"), "-i", str(d / "density.raw"), "-o", out,
          "-l", "300 mm", "-r", "35 cm", "--rho-to-mu", "0.096", "--seed", "4", "continuous", "200000"])
    img = P.read_image3d(out)
    ref = P.make_sensitivity_image(_mu(rho, 0.096), 200_000, 300.0, 700.0, seed=4)
    assert list(img.fov.n) == list(ref.fov.n) and rel_l2(img.data, ref.data) <= 1e-6


@pytest.mark.parametrize("adjust,code", [("centre", P.ADJUST_CENTRE), ("random-z-phi", P.ADJUST_RANDOM_Z_PHI)])
def test_make_sensitivity_image_discrete_cli(gpu, density, adjust, code):
    d, rho = density
    out = str(d / f"disc-{adjust}.raw")
    txt = _run([os.path.join(BIN, "make_sensitivity_image"), "-i", str(d / "density.raw"), "-o", out,
                "-l", "10 cm", "-r", "100 mm", "--rho-to-mu", "0.096", "--seed", "2",
                "discrete", "--dr", "1 cm", "--dz", "10 mm", "--da", "1 cm", "--adjust", adjust])
    assert "263901 LORs between detector elements" in txt
    img = P.read_image3d(out)
    ref = P.make_sensitivity_image_discrete(_mu(rho, 0.096), 100.0, 100.0, 10.0, 10.0, 10.0,
                                            adjust=code, seed=2)
    assert rel_l2(img.data, ref.data) <= 1e-6


def test_mlem_cli_from_records(gpu, tmp_path):
    """--records: the config's cuts, smearing, axial limit and scatter
    correction applied to raw Hdf5Lor records, then the iterations."""
    rng = np.random.default_rng(31)
    lors = P.synth_phantom_lors(5, 300_000, 200.0)
    n = len(lors)
    rec = np.zeros(n, P.H5LOR_DTYPE)
    rec["x1"], rec["y1"], rec["z1"] = lors["p1"].T
    rec["x2"], rec["y2"], rec["z2"] = lors["p2"].T
    rec["dt"] = lors["dt"] / np.float32(1000.0)
    rec["q1"], rec["q2"] = rng.uniform(0, 3000, n), rng.uniform(0, 3000, n)
    rec["E1"], rec["E2"] = rng.normal(511, 15, n), rng.normal(511, 15, n)
    rec_path = str(tmp_path / "records.bin")
    rec.tofile(rec_path)
    cfg = f"""
[input]
file = "{tmp_path}/unused.h5"
dataset = "reco_info/lors"
energy.min = 440
energy.max = 600
charge.min = 100
events.min = 1000
[iterations]
number = 2
subsets = 3
[fov]
nvoxels = [60, 60, 60]
size = ["180 mm", "180 mm", "180 mm"]
[tof]
sigma = "84.93 ps"
cutoff = 3
[smear_energy]
fwhm = "10 %"
[detector_full_axial_length]
dz = "30 cm"
[scatter_correction]
phi.bins = 8
r.bins = 10
r.max = "120 mm"
"""
    cfg_path = str(tmp_path / "cfg.toml")
    open(cfg_path, "w").write(cfg)
    outdir = str(tmp_path / "out") + "/"
    txt = _run([os.path.join(BIN, "mlem"), cfg_path, outdir, "--records", rec_path,
                "--discretize", "350 20 3 3", "--seed", "5"])
    c = P.read_config(cfg_path)
    ing, n_read = P.ingest_lors(rec, c.ingest_params((350.0, 20.0, 3.0, 3.0), seed=5), c.scattergram_spec())
    assert f"{n_read} read, {len(ing)} selected" in txt and n_read == n - 1000 and len(ing) > n // 4
    it = P.mlem(c.fov, ing, tof=c.tof, n_subsets=3)
    for k in range(6):
        img, osem = next(it)
        got = P.read_image3d(f"{outdir}{osem.iteration:02d}-{osem.subset:02d}.raw")
        assert rel_l2(got.data, img.data) <= 1e-6, (k, rel_l2(got.data, img.data))
