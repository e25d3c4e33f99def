"""The tile-jump binning of the tiled scheme (bin_lor_jump, which fills the
runs of walks without TOF; tile_runs, which counts them) against the
step-by-step binning walk (bin_lor), compiled for the host from the
library's own headers: identical runs (every field, bit for bit), step
counts and tile sequences on 8 FOV/tile cases x 20 000 LORs, the same run
length class in both jumps (the sensitivity segments are sized by one and
filled by the other) and, for LORs in general position, a class that
matches the run's length to within the estimate's 3 steps; and the jumps'
repeated binary32 addition (repeat_add) against one addition at a time on
400 000 cases (ties, binade crossings, long runs)."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_axes_binning_matches_step_walk():
    with tempfile.TemporaryDirectory(prefix="petalo_axes_") as tmp:
        exe = os.path.join(tmp, "check")
        subprocess.check_call([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                               "-w", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "petalo_amd", "csrc"),
                               os.path.join(ROOT, "tests", "cpp", "binning_axes_check.cpp"), "-o", exe])
        out = subprocess.run([exe, "20000"], capture_output=True, text=True, timeout=600)
        print(out.stdout)
        assert out.returncode == 0, out.stdout
        assert "repeat_add mismatches 0" in out.stdout
        assert ", mismatches 0, tile mismatches 0" in out.stdout
        assert "class far from the run length 0" in out.stdout
