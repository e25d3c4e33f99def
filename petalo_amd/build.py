"""Build recipe for the HIP library (gfx950) and its command-line tools.

    python -m petalo_amd.build            # library + tools
    python -m petalo_amd.build --lib-only

Every source is compiled by hipcc for gfx950 only.  Device code is built
with -ffp-contract=off (no a*b+c contraction, so the binary32 voxel walk is
bit-identical to the reference's) and with packed-FP32 instructions disabled.
Products land next to the sources (in-tree), so they travel with the tree.
"""
import argparse
import os
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libpetalo_hip.so")
TOOLS = {"mlem": "mlem.cpp", "make_sensitivity_image": "make_sensitivity_image.cpp", "foms": "foms.cpp"}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",          # exact binary32 walk: no FMA contraction
    "-munsafe-fp-atomics",        # hardware float/double atomic add
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]

SOURCES = ["capi.hip", "ingest.hip", "fom.hip", "host.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib():
    with tempfile.TemporaryDirectory(prefix="petalo_build_") as tmp:
        objs = []
        for src in SOURCES:
            obj = os.path.join(tmp, os.path.splitext(src)[0] + ".o")
            _run([HIPCC, "-x", "hip", f"--offload-arch={ARCH}", *COMMON, "-c",
                  os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        tmp_lib = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp_lib, *objs, "-lpthread"])
        os.replace(tmp_lib, LIB)
    return LIB


def build_tools():
    out = []
    for name, src in TOOLS.items():
        exe = os.path.join(PKG, "bin", name)
        os.makedirs(os.path.dirname(exe), exist_ok=True)
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
              os.path.join(PKG, "tools", src), "-o", exe + ".tmp",
              "-L", PKG, "-lpetalo_hip", "-Wl,-rpath,$ORIGIN/..",
              "-Wl,-rpath,/opt/rocm/lib", "-lpthread"])
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    a = ap.parse_args(argv)
    build_lib()
    if not a.lib_only:
        build_tools()
    return 0


if __name__ == "__main__":
    sys.exit(main())
