#!/usr/bin/env python3
"""Entry-point study on the oracle (test infrastructure; CPU only).

How much the unpinned binary32 arithmetic of the FOV entry (the ncollide3d
0.33 slab cast behind FOV::entry, src/fov.rs:55-71, restated from its
published algorithm) can move a reconstruction, and what the binary32-vs-
binary64 gap is made of.  For the MLEM workloads of the 60^3 x 3 mm (1e6
LORs, no TOF) and 120^3 x 1.5 mm (1e7 LORs, TOF 42.47 ps, cutoff 3) phantom
configurations, from the same seeded LORs as bench.py, with a sensitivity
image of ones, it runs ITER iterations of

  f32     the binary32 oracle (the reference's walk and arithmetic);
  exact   the same with the entry variant whose hit-face coordinate is exact
          (no walk starts shifted);
  mixed   binary32 walk decisions with binary64 arithmetic;
  f64     the binary64 oracle;

and prints one JSON object: the shifted-row fraction and, per iteration,
relative L2 and normalised max differences f32-exact (the bound on what the
entry arithmetic can change), f32-mixed (rounding alone), mixed-f64 (walk
geometry alone) and f32-f64.

    python oracle/entry_study.py [--iters 10] [--threads 8] [--configs 1,2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_ffi as O  # noqa: E402
import petalo_amd as P  # noqa: E402

CONFIGS = {
    1: dict(full=180.0, n=60, lors=1_000_000, fwhm=0.0, tof=None),
    2: dict(full=180.0, n=120, lors=10_000_000, fwhm=200.0, tof=(42.47, 3.0)),
}


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b)), float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    ap.add_argument("--configs", default="1,2")
    a = ap.parse_args()
    out = {}
    for k in (int(c) for c in a.configs.split(",")):
        c = CONFIGS[k]
        full, n = [c["full"]] * 3, [c["n"]] * 3
        lors = P.synth_phantom_lors(1, c["lors"], c["fwhm"])
        st = O.walk_stats(full, n, lors, tof=c["tof"], threads=a.threads)
        nv = int(np.prod(n))
        x = {"f32": np.ones(nv, np.float32), "exact": np.ones(nv, np.float32),
             "mixed": np.ones(nv, np.float64), "f64": np.ones(nv, np.float64)}
        rows = []
        t0 = time.time()
        for it in range(1, a.iters + 1):
            O.mlem_update(full, n, lors, x["f32"], tof=c["tof"], threads=a.threads)
            with O.entry_variant(1):
                O.mlem_update(full, n, lors, x["exact"], tof=c["tof"], threads=a.threads)
            O.mlem_update(full, n, lors, x["mixed"], tof=c["tof"], threads=a.threads, mixed=True)
            O.mlem_update(full, n, lors, x["f64"], tof=c["tof"], threads=a.threads, f64=True)
            r = {"iteration": it}
            for name, (u, v) in {"f32_vs_exact_entry": ("f32", "exact"), "f32_vs_mixed": ("f32", "mixed"),
                                 "mixed_vs_f64": ("mixed", "f64"), "f32_vs_f64": ("f32", "f64")}.items():
                l2, mx = rel(x[u], x[v])
                r[name] = {"rel_l2": l2, "max": mx}
            rows.append(r)
            print(f"config {k} iteration {it} ({time.time() - t0:.0f} s): "
                  + ", ".join(f"{m} {r[m]['rel_l2']:.2e}" for m in r if m != "iteration"), file=sys.stderr,
                  flush=True)
        out[f"config{k}"] = {"voxels": n, "lors": c["lors"], "tof": c["tof"],
                             "shifted_rows": st["n_shifted"], "shifted_fraction": st["n_shifted"] / st["n_lors"],
                             "iterations": rows}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
