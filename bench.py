#!/usr/bin/env python3
"""Benchmark of the reconstruction hot path on one MI355X.

Workload (default): TOF-MLEM on a 120^3 x 1.5 mm image (180 mm FOV) with
1e7 synthetic phantom LORs (200 ps FWHM timing, reconstruction sigma 42.47 ps,
cutoff 3), sensitivity image from 1e8 device-generated random LORs (mu = 0).
One step = one MLEM iteration (1 subset) over all LORs.

Prints ONE JSON line:
  value      MLEM iterations/s (inputs resident in HBM before timing; the
             updates replayed from their HIP graphs, back to back, after 20
             warm-up iterations)
  roofline   the dominant kernel (the longer of the tiled forward kernel
             k_fwd_tile and the tiled backprojection k_scatter) -- its
             algorithmic bytes per launch (forward: 32 B per LOR + 4 B per
             pushed element; backprojection: 4 B per pushed element) / its
             average duration from HIP events on the context's stream over
             as many updates again (launched directly, an event around each
             kernel), against 8 TB/s; traffic from the committed rocprofv3
             FETCH_SIZE/WRITE_SIZE profile of the same workload
  cpu_baseline  the CPU oracle (binary32 restatement of the reference, all
             host threads, 10 000-LOR jobs with private buffers) on a
             bounded sample of the same workload, scaled to iterations/s
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (fov mm, voxels, n_lors, tof_fwhm_ps, tof (sigma_ps, cutoff) or None, sens_lors, subsets)
    "tof_mlem_120": (180.0, 120, 10_000_000, 200.0, (42.47, 3.0), 100_000_000, 1),
    "plain_mlem_60": (180.0, 60, 1_000_000, 0.0, None, 0, 1),
    "precision_200": (200.0, 200, 50_000_000, 150.0, (31.85, 3.0), 0, 1),
}
METRICS = {
    "tof_mlem_120": "TOF-MLEM iter/s + Siddon voxel-steps/s (1e7 LORs, 120^3) at <=1e-4 rel L2",
    "plain_mlem_60": "MLEM iter/s + Siddon voxel-steps/s (1e6 LORs, 60^3, no TOF)",
    "precision_200": "TOF-MLEM iter/s + Siddon voxel-steps/s (5e7 LORs, 200^3, 150 ps)",
}


def host_cpus():
    """(logical CPUs of the host, physical cores, CPUs this process may use)."""
    phys = set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ":" in ln:
                    k, v = (t.strip() for t in ln.split(":", 1))
                    if k in ("physical id", "core id"):
                        cur[k] = v
                elif cur:
                    phys.add((cur.get("physical id"), cur.get("core id")))
                    cur = {}
        if cur:
            phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    return os.cpu_count(), (len(phys) or None), len(os.sched_getaffinity(0))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def batch_workload(P, n_rec=32, seed=4):
    """The batched workload: 32 reconstructions, LOR counts log-uniform in
    [1e5, 1e7], images 60^3 .. 120^3 over a 180 mm FOV, TOF 200 ps."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([60, 72, 84, 96, 108, 120], n_rec)
    counts = np.round(10 ** (5 + 2 * rng.uniform(size=n_rec))).astype(np.int64)
    fovs = [P.FOV([180.0] * 3, [int(v)] * 3) for v in sizes]
    lors = [P.synth_phantom_lors(100 + k, int(counts[k]), 200.0) for k in range(n_rec)]
    return fovs, lors, sizes, counts


def measure_batched(P, steps, warmup, singles=True):
    """32 independent TOF-OSEM reconstructions (20 subsets) advanced together;
    a step is one full iteration (20 subset updates of all 32).  With singles,
    the same reconstructions run one at a time as single contexts, on the
    default (tiled) path and on the plain one-lane atomic path: one warm-up
    full iteration each (captures its graphs), then one timed full iteration
    each; their sums are the per-iteration times the batch is compared with."""
    n_rec, subsets = 32, 20
    t0 = time.time()
    fovs, lors, sizes, counts = batch_workload(P, n_rec)
    log(f"[bench] batched: {counts.sum()} LORs over {n_rec} reconstructions in {time.time() - t0:.1f} s")
    tof = P.Tof(42.47, 3.0)
    t0 = time.perf_counter()
    b = P.Batch(fovs, lors, tofs=[tof] * n_rec, n_subsets=subsets)
    create_s = time.perf_counter() - t0
    log(f"[bench] batch ready in {create_s:.1f} s")
    for _ in range(warmup * subsets):
        b.step()
    P.lib().petalo_device_synchronize(0)
    t0 = time.perf_counter()
    for _ in range(steps * subsets):
        b.step()
    P.lib().petalo_device_synchronize(0)
    el = time.perf_counter() - t0
    b.set_options(timing=True, graphs=False)
    ms_direct = []
    for _ in range(subsets):
        b.step()
        ms_direct.append(b.last_ms())
    # the persistent schedule (one launch per subset update taking every
    # phase's tasks from a queue), timed as the default one
    b.set_options(timing=False, graphs=True)
    b.set_schedule(True)
    for _ in range(subsets):
        b.step()
    P.lib().petalo_device_synchronize(0)
    t0 = time.perf_counter()
    for _ in range(steps * subsets):
        b.step()
    P.lib().petalo_device_synchronize(0)
    persistent_ms = (time.perf_counter() - t0) / steps * 1e3
    b.image(0)  # fails if a wait of the persistent launch ran into its time limit
    b.close()
    full_ms = el / steps * 1e3
    log(f"[bench] batched: {full_ms:.2f} ms per full iteration of all {n_rec} "
        f"({persistent_ms:.2f} ms with the persistent schedule)")
    out = {"value": steps / el, "unit": "iter/s", "steps": steps, "warmup": warmup,
           "ms_per_full_iteration": full_ms, "subset_update_ms_graphs": full_ms / subsets,
           "subset_update_ms_direct_launches": float(np.mean(ms_direct)), "create_s": create_s,
           "persistent_schedule_ms_per_full_iteration": persistent_ms,
           "config": {"workload": "batched_tof_osem_32", "reconstructions": n_rec, "subsets": subsets,
                      "voxels": [int(v) for v in sizes], "n_lors": [int(c) for c in counts],
                      "total_lors": int(counts.sum()), "tof_sigma_ps": 42.47, "tof_cutoff": 3.0}}
    if singles:
        for name, scheme in (("tiled", P.SCHEME_TILE), ("atomic", P.SCHEME_ATOMIC)):
            tot = 0.0
            for k in range(n_rec):
                c = P.Mlem(fovs[k], lors[k], tof=tof, n_subsets=subsets, scheme=scheme)
                for _ in range(subsets):
                    c.step(copy_image=False)
                P.lib().petalo_device_synchronize(0)
                t1 = time.perf_counter()
                for _ in range(subsets):
                    c.step(copy_image=False)
                P.lib().petalo_device_synchronize(0)
                tot += time.perf_counter() - t1
                c.close()
            out[f"single_{name}_full_iteration_ms"] = tot * 1e3
            out[f"speedup_vs_single_{name}"] = tot * 1e3 / full_ms
            log(f"[bench] batched: one at a time on the {name} path {tot * 1e3:.1f} ms per full iteration "
                f"({tot * 1e3 / full_ms:.1f}x the batch)")
    return out


def timed_kernels(P, ctx, n):
    """Per-kernel device times of n more back-to-back updates: HIP events on
    the context's stream around every kernel (the updates are then launched
    directly instead of replayed from their graphs), read in batches of the
    context's ring of 256 event sets."""
    ctx.set_timing(True)
    kern = []
    for k in range(n):
        ctx.step(copy_image=False)
        if (k + 1) % 256 == 0:
            kern += ctx.drain_kernel_timings(256)
    P.lib().petalo_device_synchronize(0)
    kern += ctx.drain_kernel_timings(256)
    ctx.set_timing(False)
    assert len(kern) == n
    return kern


def measure_workload(P, name, steps, warmup):
    """One of the other workloads, timed as the headline is (steady state:
    `warmup` untimed iterations, then `steps` back to back with HIP-event
    kernel times), without a sensitivity image: its iterations/s, voxel
    steps/s and the dominant kernel's roofline fraction."""
    full, nvx, n_lors, fwhm, tof, _, subsets = WORKLOADS[name]
    fov = P.FOV([full] * 3, [nvx] * 3)
    lors = P.synth_phantom_lors(1, n_lors, fwhm)
    t0 = time.perf_counter()
    ctx = P.Mlem(fov, lors, tof=P.Tof(*tof) if tof else None, n_subsets=subsets)
    create_ms = (time.perf_counter() - t0) * 1e3
    ws = ctx.walk_stats()
    for _ in range(warmup * subsets):
        ctx.step(copy_image=False)
    P.lib().petalo_device_synchronize(0)
    t0 = time.perf_counter()
    for _ in range(steps * subsets):
        ctx.step(copy_image=False)
    P.lib().petalo_device_synchronize(0)
    el = time.perf_counter() - t0
    kern = timed_kernels(P, ctx, steps * subsets)
    ctx.close()
    kt = {k: float(np.mean([d[k] for d in kern])) for k in kern[0]}
    fwd_bytes = 32 * ws["n_lors"] + 4 * ws["n_nz"]
    sct_bytes = 4 * ws["n_nz"]
    cands = [("k_fwd_tile", fwd_bytes, kt["forward_tiles"]), ("k_scatter", sct_bytes, kt["scatter"])]
    dom, dom_bytes, dom_ms = max(cands, key=lambda c: c[2])
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    r = {"metric": METRICS[name], "value": steps / el, "unit": "iter/s", "steps": steps, "warmup": warmup,
         "ms_per_iteration": el / steps * 1e3, "voxel_steps_per_s": ws["n_steps"] * steps / el,
         "context_create_ms": create_ms, "lors": n_lors, "image": [nvx] * 3, "tof": tof,
         "kernel_ms": kt,
         "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": dom_bytes, "kernel_ms": dom_ms}}
    log(f"[bench] {name}: {r['value']:.1f} iter/s ({r['ms_per_iteration']:.3f} ms), {dom} at "
        f"{achieved / HBM_PEAK_GBS:.3f} of the HBM roofline")
    return r


def run_batched(P, a):
    """The batched workload on its own line."""
    r = measure_batched(P, a.steps, a.warmup, singles=not a.no_cpu_baseline)
    out = {
        "metric": "batched TOF-OSEM full iterations/s (32 reconstructions, 20 subsets)",
        "value": r["value"], "unit": "iter/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": r["ms_per_full_iteration"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded phantom LORs)", "config": r["config"],
        "batched": r,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_precision_study(P, a):
    """The precision workload (200^3 x 1 mm, 5e7 TOF LORs, 150 ps FWHM data,
    reconstruction sigma 31.85 ps) for `iters` iterations on three contexts:
    the binary32 path (tiled, the product), the full binary64 GPU path and the
    binary32-walk/binary64-arithmetic path; per iteration the relative L2 and
    normalised max difference of the binary32 image against each binary64
    image, and between the two binary64 images (walk geometry alone)."""
    full, nvx, n_lors, fwhm, tof, _, _ = WORKLOADS["precision_200"]
    fov = P.FOV([full] * 3, [nvx] * 3)
    t0 = time.time()
    lors = P.synth_phantom_lors(1, n_lors, fwhm)
    log(f"[precision] {n_lors} LORs in {time.time() - t0:.1f} s")
    tof_obj = P.Tof(*tof)
    ctx = {name: P.Mlem(fov, lors, tof=tof_obj, precision=prec) for name, prec in
           (("f32", P.PRECISION_F32), ("f64", P.PRECISION_F64), ("f32_walk_f64", P.PRECISION_F32_WALK))}
    ws = ctx["f32"].walk_stats()

    def rel(u, v):
        d = u - v
        return float(np.linalg.norm(d) / np.linalg.norm(v)), float(np.max(np.abs(d)) / np.max(np.abs(v)))

    rows = []
    times = {k: 0.0 for k in ctx}
    for it in range(1, a.precision_iters + 1):
        for k, c in ctx.items():
            P.lib().petalo_device_synchronize(0)
            t1 = time.perf_counter()
            c.step(copy_image=False)
            P.lib().petalo_device_synchronize(0)
            times[k] += time.perf_counter() - t1
        x32 = ctx["f32"].image().data.astype(np.float64)
        x64 = ctx["f64"].image_f64()
        xm = ctx["f32_walk_f64"].image_f64()
        r = {"iteration": it}
        for name, (u, v) in {"f32_vs_f64": (x32, x64), "f32_vs_f32_walk_f64": (x32, xm),
                             "f32_walk_f64_vs_f64": (xm, x64)}.items():
            l2, mx = rel(u, v)
            r[name] = {"rel_l2": l2, "max": mx}
        rows.append(r)
        log(f"[precision] iteration {it}: " + ", ".join(f"{k} {r[k]['rel_l2']:.2e} (max {r[k]['max']:.2e})"
                                                         for k in r if k != "iteration"))
    for c in ctx.values():
        c.close()
    out = {"metric": "precision study: relative L2 / max error per iteration of the binary32 path",
           "config": {"workload": "precision_200", "fov_mm": full, "voxels": [nvx] * 3, "n_lors": n_lors,
                      "tof_fwhm_ps": fwhm, "tof_sigma_ps": tof[0], "tof_cutoff": tof[1],
                      "iterations": a.precision_iters},
           "walk": ws, "shifted_fraction": ws["n_shifted"] / ws["n_lors"],
           "ms_per_iteration": {k: v / a.precision_iters * 1e3 for k, v in times.items()},
           "iterations": rows}
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # back-to-back iterations run through a clock transient of about 25
    # iterations (30 ms) after an idle GPU: 20 warm-up iterations put the
    # timed ones in the steady state; end_to_end_ms keeps the cold 10
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="tof_mlem_120", choices=sorted(WORKLOADS))
    ap.add_argument("--scheme", type=int, default=0, help="0 auto, 1 atomic, 2 tiled")
    ap.add_argument("--tile", default=None, help="tiled scheme tile shape x,y,z (default: automatic)")
    ap.add_argument("--autotune", action="store_true",
                    help="let the library time the atomic scheme and five tile shapes and keep the fastest")
    ap.add_argument("--batched", action="store_true",
                    help="measure 32 batched TOF-OSEM reconstructions (20 subsets) instead")
    ap.add_argument("--no-batched", action="store_true",
                    help="leave the batched workload out of the headline line")
    ap.add_argument("--precision-study", action="store_true",
                    help="binary32 vs the binary64 GPU paths on the 200^3 precision workload instead")
    ap.add_argument("--precision-iters", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-lors", type=int, default=1_000_000)
    ap.add_argument("--no-sensitivity", action="store_true")
    ap.add_argument("--no-other-workloads", action="store_true",
                    help="skip the plain 60^3 and precision 200^3 workloads timed beside the headline")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run a plain device copy of this size first (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported in this build "
              "(the path shards over LORs with one all-reduce of the backprojection; not built here)",
              file=sys.stderr)
        return 2

    import petalo_amd as P

    if a.batched:
        return run_batched(P, a)
    if a.precision_study:
        return run_precision_study(P, a)

    if a.copy_probe_mb:
        ms = P.copy_probe(a.copy_probe_mb << 20, 3)
        log(f"[bench] copy probe {a.copy_probe_mb} MiB: {ms:.3f} ms per copy "
            f"({2 * (a.copy_probe_mb << 20) / ms / 1e6:.0f} GB/s read+write)")

    full, nvx, n_lors, fwhm, tof, n_sens, subsets = WORKLOADS[a.workload]
    fov = P.FOV([full] * 3, [nvx] * 3)
    tof_obj = P.Tof(*tof) if tof else None

    log(f"[bench] generating {n_lors} phantom LORs")
    t0 = time.time()
    lors = P.synth_phantom_lors(1, n_lors, fwhm)
    log(f"[bench] LORs ready in {time.time() - t0:.1f} s")

    sens_img, sens_ms, sens_stats, discrete, sens_att = None, None, None, None, None
    sens_calls, first_ms = None, None
    if n_sens and not a.no_sensitivity:
        # each image is timed as a caller makes it (no work counters: those
        # cost a walk over every step), median of three calls; the counters
        # behind the algorithmic bytes come from one more call with them
        def timed(fn, reps=3):
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                out = fn()
                ts.append((time.perf_counter() - t0) * 1e3)
            return out, float(np.median(ts)), ts
        mu = P.Image(fov, np.zeros(fov.nvox, np.float32))
        # the first call of a process also allocates the buffers the library
        # keeps for later calls: timed on its own, after a small call that
        # loads the kernels
        P.make_sensitivity_image(mu, 1_000_000, 300.0, 700.0, seed=7)
        P.release_workspaces()
        t0 = time.perf_counter()
        P.make_sensitivity_image(mu, n_sens, 300.0, 700.0, seed=7)
        first_ms = (time.perf_counter() - t0) * 1e3
        sens_img, sens_ms, ts = timed(lambda: P.make_sensitivity_image(mu, n_sens, 300.0, 700.0, seed=7))
        sens_calls = ts
        t0 = time.perf_counter()
        img_st, sens_stats = P.make_sensitivity_image(mu, n_sens, 300.0, 700.0, seed=7, return_stats=True)
        stats_ms = (time.perf_counter() - t0) * 1e3
        ref = sens_img.data.astype(np.float64)
        same = float(np.linalg.norm(img_st.data - ref) / max(np.linalg.norm(ref), 1e-300))
        assert same < 1e-6, f"sensitivity images with and without work counters differ: relL2 {same:.2e}"
        log(f"[bench] sensitivity image from {n_sens} LORs: {sens_ms:.1f} ms "
            f"({', '.join(f'{t:.1f}' for t in ts)}; first call {first_ms:.1f} ms; "
            f"{stats_ms:.1f} ms with work counters, relL2 {same:.1e})")
        # the same LORs through the synthetic body's attenuation (forward half included)
        att = P.density_to_attenuation(P.synth_density(fov), 0.096)
        _, att_ms, ts = timed(lambda: P.make_sensitivity_image(att, n_sens, 300.0, 700.0, seed=7))
        sens_att = {"lors": n_sens, "ms": att_ms, "ms_calls": ts}
        log(f"[bench] attenuated sensitivity image from {n_sens} LORs: {att_ms:.1f} ms")
        # discrete detector: all pairs of 10 mm elements of a 700 mm ring, 600 mm long
        dgeom = dict(r_min=350.0, dr=20.0, dz=10.0, da=10.0)
        _, dms, ts = timed(lambda: P.make_sensitivity_image_discrete(att, 600.0, **dgeom))
        _, dst = P.make_sensitivity_image_discrete(att, 600.0, return_stats=True, **dgeom)
        discrete = {"detector": dict(dgeom, length=600.0), "pairs": dst["n_lors"], "ms": dms, "ms_calls": ts,
                    "n_steps": dst["n_steps"], "n_nz": dst["n_nz"], "n_missed": dst["n_missed"]}
        log(f"[bench] discrete sensitivity image, {dst['n_lors']} element pairs: {dms:.1f} ms")

    t0 = time.perf_counter()
    tile = tuple(int(v) for v in a.tile.split(",")) if a.tile else None
    ctx = P.Mlem(fov, lors, tof=tof_obj, sensitivity=sens_img, n_subsets=subsets, scheme=a.scheme, tile=tile,
                 autotune=a.autotune)
    log(f"[bench] context ready in {time.perf_counter() - t0:.2f} s")
    ws = ctx.walk_stats()
    log(f"[bench] walk stats {ws}")

    for _ in range(a.warmup):
        ctx.step(copy_image=False)
    P.lib().petalo_device_synchronize(0)

    steps_per_item, n_items = ctx.work_items()
    info = dict(ctx.info(), tile=list(ctx.tile_shape()), autotune=bool(a.autotune),
                item_steps=steps_per_item, work_items=n_items)
    unit, r_max, r_gm = ctx.accumulator()
    info["accumulator"] = {"kind": "fixed64" if unit > 0 else "binary64", "unit": unit, "r_max": r_max,
                           "r_geomean": r_gm}
    log(f"[bench] scheme {info}")
    # the timed region: the updates as a caller runs them, each replayed
    # from its captured HIP graph, back to back
    P.lib().petalo_device_synchronize(0)
    t0 = time.perf_counter()
    for _ in range(a.steps * subsets):
        ctx.step(copy_image=False)
    P.lib().petalo_device_synchronize(0)
    elapsed = time.perf_counter() - t0
    # then as many updates again with HIP events around every kernel, for
    # the kernel times and the roofline
    kern = timed_kernels(P, ctx, a.steps * subsets)
    fwd_ms = [d["forward_tiles"] + d["forward_reduce"] for d in kern]
    sct_ms = [d["scatter"] for d in kern]
    proj_ms = [f + s_ for f, s_ in zip(fwd_ms, sct_ms)]
    upd_ms = [d["update"] for d in kern]

    it_per_s = a.steps / elapsed
    steps_per_s = ws["n_steps"] * a.steps / elapsed

    # End to end, as the reference's caller runs it: create the context from
    # host LORs (upload, LOR ordering, tile binning), 10 iterations, read the
    # image back.
    ctx.close()
    P.lib().petalo_device_synchronize(0)
    t0 = time.perf_counter()
    e2e = P.Mlem(fov, lors, tof=tof_obj, sensitivity=sens_img, n_subsets=subsets, scheme=a.scheme, tile=tile)
    t_create = time.perf_counter() - t0
    for _ in range(10 * subsets - 1):
        e2e.step(copy_image=False)
    e2e.step(copy_image=True)
    end_to_end_ms = (time.perf_counter() - t0) * 1e3
    e2e.close()
    log(f"[bench] end to end (create {t_create * 1e3:.0f} ms, 10 iterations, image read back): {end_to_end_ms:.1f} ms")
    proj_avg = float(np.mean(proj_ms))
    # Algorithmic bytes: per LOR its 32-byte record, per pushed element one
    # 4-byte image gather (forward) and one 4-byte backprojection add (scatter).
    fwd_bytes = 32 * ws["n_lors"] + 4 * ws["n_nz"]
    sct_bytes = 4 * ws["n_nz"]
    tiled = info["scheme"] == P.SCHEME_TILE
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", f"traffic_{a.workload}.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        if tj.get("scheme") == info["scheme"]:
            traffic = {k: v["read_bytes"] + v["write_bytes"] for k, v in tj["kernels"].items()}
    kt = {k: float(np.mean([d[k] for d in kern])) for k in kern[0]}
    if tiled:
        cands = [("k_fwd_tile", fwd_bytes, kt["forward_tiles"]), ("k_scatter", sct_bytes, kt["scatter"])]
        dom, dom_bytes, dom_ms = max(cands, key=lambda c: c[2])
    else:
        dom, dom_bytes, dom_ms = "k_mlem_project", fwd_bytes + sct_bytes, proj_avg
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    log(f"[bench] {it_per_s:.2f} iter/s; projection {proj_avg:.3f} ms (forward {np.mean(fwd_ms):.3f}, "
        f"scatter {np.mean(sct_ms):.3f}), update {np.mean(upd_ms):.3f} ms")

    # LOR ingestion of 1e7 reco_info/lors records (cuts, energy and position
    # smearing, 5-axis scattergram), host arrays in and out
    ingest = None
    if not a.no_sensitivity:
        rng = np.random.default_rng(9)
        m = 10_000_000
        rec = np.zeros(m, P.H5LOR_DTYPE)
        ang = rng.uniform(0, 2 * np.pi, (2, m)).astype(np.float32)
        rec["x1"], rec["y1"] = 360 * np.cos(ang[0]), 360 * np.sin(ang[0])
        rec["x2"], rec["y2"] = 360 * np.cos(ang[1]), 360 * np.sin(ang[1])
        rec["z1"], rec["z2"] = rng.uniform(-150, 150, (2, m)).astype(np.float32)
        rec["E1"], rec["E2"] = rng.normal(511, 20, (2, m)).astype(np.float32)
        rec["q1"] = rec["q2"] = 1000.0
        p = P.ingest_params(energy=(434.0, 580.0), smear_energy_fwhm=0.1, z_max=150.0,
                            discretize=(350.0, 20.0, 3.0, 3.0), seed=1)
        spec = P.scattergram_spec(phi=30, r=(30, 120.0), z=(30, 300.0), dz=(30, 300.0))
        P.ingest_lors(rec[:100_000], p, spec)
        t0 = time.perf_counter()
        got, nread = P.ingest_lors(rec, p, spec)
        ims = (time.perf_counter() - t0) * 1e3
        ingest = {"records": m, "kept": int(len(got)), "ms_with_transfers": ims,
                  "records_per_s": m / (ims * 1e-3)}
        log(f"[bench] ingestion of {m} records: {ims:.1f} ms with host transfers ({len(got)} LORs kept)")

    # NEMA7 figures of merit (457 regions) of a noisy 2 mm image, 130 x 130 x 91
    foms = None
    if not a.no_sensitivity:
        nema = P.FOV([260.0, 260.0, 182.0], [130, 130, 91])
        hw = np.float32(130.0), np.float32(91.0)
        ax = (np.arange(130, dtype=np.float32) + np.float32(0.5)) * np.float32(2.0) - hw[0]
        az = (np.arange(91, dtype=np.float32) + np.float32(0.5)) * np.float32(2.0) - hw[1]
        zz, yy, xx = np.meshgrid(az, ax, ax, indexing="ij")
        act = np.ones(zz.shape, np.float32)
        act[xx ** 2 + yy ** 2 < 256.0] = 0.25
        for d, k in zip([10.0, 13.0, 17.0, 22.0, 28.0, 37.0], [1, 2, 3, 4, 5, 0]):
            sx, sy = 57.2 * np.cos(np.pi * k / 3), 57.2 * np.sin(np.pi * k / 3)
            act[(xx - sx) ** 2 + (yy - sy) ** 2 + zz ** 2 < (d / 2) ** 2] = 4.0
        act *= np.random.default_rng(2).gamma(10.0, 0.1, act.shape).astype(np.float32)
        img = P.Image(nema, act.ravel())
        P.phantom_foms(img, P.PHANTOM_NEMA7)
        t0 = time.perf_counter()
        fv, aocs = P.phantom_foms(img, P.PHANTOM_NEMA7)
        fms = (time.perf_counter() - t0) * 1e3
        foms = {"phantom": "nema7", "voxels": nema.nvox, "regions": 6 * 66 + len(aocs), "ms_with_transfers": fms,
                "crc": [round(float(c), 2) for c in fv[:, 1]]}
        log(f"[bench] NEMA7 figures of merit ({foms['regions']} regions, {nema.nvox} voxels): {fms:.2f} ms")

    batched = None
    if not a.no_batched and a.workload == "tof_mlem_120":
        batched = measure_batched(P, 2, 1)

    # the parity-shape workloads, briefly, in the same run
    others = None
    if not a.no_other_workloads and a.workload == "tof_mlem_120":
        others = {name: measure_workload(P, name, a.steps, a.warmup) for name in ("plain_mlem_60", "precision_200")}

    cpu = None
    parity = None
    rows = None
    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_ffi as O
        m = min(a.cpu_sample_lors, n_lors)
        sample = lors[:m]
        n_cpus, n_phys, threads = host_cpus()
        x = np.ones(fov.nvox, np.float32)
        s = sens_img.data if sens_img is not None else None
        log(f"[bench] CPU oracle: one update over {m} LORs on {threads} threads")
        t0 = time.perf_counter()
        O.mlem_update(fov.full_size, fov.n, sample, x, sens=s, tof=tof, threads=threads)
        t_cpu = time.perf_counter() - t0
        cpu_it_s = 1.0 / (t_cpu * n_lors / m)
        # one thread, on a tenth of the sample
        m1 = max(1, m // 10)
        x1 = np.ones(fov.nvox, np.float32)
        t0 = time.perf_counter()
        O.mlem_update(fov.full_size, fov.n, lors[:m1], x1, sens=s, tof=tof, threads=1)
        t_one = time.perf_counter() - t0
        steps_per_lor = ws["n_steps"] / n_lors
        model = None
        try:
            with open("/proc/cpuinfo") as fh:
                model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), None)
        except OSError:
            pass
        cpu = {"value": cpu_it_s, "unit": "iter/s", "cores": threads, "host_logical_cpus": n_cpus,
               "host_physical_cores": n_phys, "kind": "port",
               "sample": f"one MLEM update over the first {m} of the {n_lors} LORs "
                         f"({t_cpu:.2f} s), scaled by {n_lors / m:g} to one iteration",
               "voxel_steps_per_s": steps_per_lor * m / t_cpu,
               "one_thread": {"value": 1.0 / (t_one * n_lors / m1), "unit": "iter/s",
                              "voxel_steps_per_s": steps_per_lor * m1 / t_one,
                              "sample": f"the first {m1} LORs ({t_one:.2f} s)"},
               "cpu_model": model}
        log(f"[bench] CPU oracle: {cpu_it_s:.3f} iter/s on {threads} threads, "
            f"{cpu['one_thread']['value']:.4f} iter/s on one ({model})")
        # the same oracle beside the paths either side of the reconstruction
        if ingest is not None:
            k = 1_000_000
            t0 = time.perf_counter()
            O.ingest(rec[:k], p, spec)
            t_i = time.perf_counter() - t0
            cpu["ingest"] = {"records_per_s": k / t_i, "threads": 1, "sample": f"the first {k} records ({t_i:.2f} s)"}
            log(f"[bench] CPU oracle ingestion: {k / t_i:.3g} records/s")
        if discrete is not None:
            k = 1_000_000
            pl, _ = O.discrete_lors(dgeom["r_min"], dgeom["dr"], dgeom["dz"], dgeom["da"], 600.0, 3, seed=0,
                                    capacity=k)
            t0 = time.perf_counter()
            O.sensitivity(fov.full_size, fov.n, att.data, pl, threads=threads)
            t_d = time.perf_counter() - t0
            cpu["discrete_sensitivity"] = {"pairs_per_s": len(pl) / t_d, "threads": threads,
                                           "sample": f"the first {len(pl)} element pairs ({t_d:.2f} s; "
                                                     f"their generation not timed)"}
            log(f"[bench] CPU oracle discrete sensitivity: {len(pl) / t_d:.3g} pairs/s")
        if foms is not None:
            t0 = time.perf_counter()
            O.phantom_foms(img.fov.full_size, img.fov.n, img.data, 0)
            t_f = time.perf_counter() - t0
            cpu["foms"] = {"ms": t_f * 1e3, "threads": 1, "sample": "the same NEMA7 analysis, binary32 oracle"}
            log(f"[bench] CPU oracle NEMA7 figures of merit: {t_f * 1e3:.0f} ms")
        # Row dump on a sample: the GPU's voxel sequences against the oracle's
        # walk (identical sequences and step counts, non-TOF weights bit for
        # bit), with the boundary ties of the sample (steps of zero length).
        k = 2000
        idx, w, cnt, stp = P.system_matrix_rows(fov, lors[:k])
        same = ties = 0
        for i in range(k):
            oi, ow, os_ = O.row(fov.full_size, fov.n, lors[i])
            mm = int(cnt[i])
            ok = mm == len(oi) and np.array_equal(idx[i, :mm], oi) and np.array_equal(w[i, :mm], ow) and stp[i] == os_
            same += int(ok)
            ties += int(stp[i]) - mm
        rows = {"sample": k, "tof": False, "sequence_match_rate": same / k, "ties": ties,
                "shifted": int(ws["n_shifted"])}
        log(f"[bench] row dump: {same} of {k} rows identical to the oracle's, {ties} boundary ties")
        # Parity on the same sample: one update from ones on the GPU.
        g = P.Mlem(fov, sample, tof=tof_obj, sensitivity=sens_img, n_subsets=1)
        img, _ = g.step()
        g.close()
        num = np.linalg.norm(img.data.astype(np.float64) - x)
        parity = float(num / np.linalg.norm(x.astype(np.float64)))
        log(f"[bench] parity on sample: relL2 {parity:.2e}")

    out = {
        "metric": METRICS[a.workload],
        "value": it_per_s,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded phantom LORs, device-generated sensitivity LORs)",
        "config": {
            "workload": a.workload,
            "fov_mm": full, "voxels": [nvx] * 3, "n_lors": n_lors, "subsets": subsets,
            "tof_fwhm_ps": fwhm, "tof_sigma_ps": tof[0] if tof else None,
            "tof_cutoff": tof[1] if tof else None, "sensitivity_lors": n_sens if sens_img is not None else 0,
            "scheme": a.scheme, "tile": list(tile) if tile else None,
        },
        "voxel_steps_per_s": steps_per_s,
        "walk": {"n_steps": ws["n_steps"], "n_nz": ws["n_nz"], "n_skipped": ws["n_skipped"],
                 "n_shifted": ws["n_shifted"], "shifted_fraction": ws["n_shifted"] / max(ws["n_lors"], 1)},
        "end_to_end_ms": end_to_end_ms,
        "end_to_end": {"what": "context creation from host LORs + 10 iterations + image read back",
                       "create_ms": t_create * 1e3, "iter_per_s": 10.0 / (end_to_end_ms * 1e-3)},
        "rows": rows,
        "batched": batched,
        "other_workloads": others,
        "projection_ms": proj_avg,
        "kernel_ms": kt,
        "forward_ms": float(np.mean(fwd_ms)),
        "scatter_ms": float(np.mean(sct_ms)),
        "scheme": info,
        "update_ms": float(np.mean(upd_ms)),
        "sensitivity_ms": sens_ms,
        "sensitivity": ({"lors": n_sens, "ms": sens_ms, "ms_calls": sens_calls, "first_call_ms": first_ms,
                         "n_steps": sens_stats["n_steps"], "n_nz": sens_stats["n_nz"],
                         "algorithmic_bytes": 8 * sens_stats["n_nz"] + 4 * fov.nvox,
                         "achieved_gbs": (8 * sens_stats["n_nz"] + 4 * fov.nvox) / (sens_ms * 1e-3) / 1e9,
                         "frac": (8 * sens_stats["n_nz"] + 4 * fov.nvox) / (sens_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
                        if sens_stats else None),
        "sensitivity_attenuated": sens_att,
        "discrete_sensitivity": discrete,
        "parity_rel_l2_sample": parity,
        "ingest": ingest,
        "foms": foms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic.get(dom), "kernel": dom,
                     "algorithmic_bytes": dom_bytes, "kernel_ms": dom_ms},
        "roofline_kernels": ({
            "k_fwd_tile": {"achieved": fwd_bytes / (kt["forward_tiles"] * 1e-3) / 1e9, "unit": "GB/s",
                           "algorithmic_bytes": fwd_bytes, "kernel_ms": kt["forward_tiles"],
                           "traffic": traffic.get("k_fwd_tile")},
            "k_scatter": {"achieved": sct_bytes / (kt["scatter"] * 1e-3) / 1e9, "unit": "GB/s",
                          "algorithmic_bytes": sct_bytes, "kernel_ms": kt["scatter"],
                          "traffic": traffic.get("k_scatter")}} if tiled else None),
        "roofline_projection": {"achieved": (fwd_bytes + sct_bytes) / (proj_avg * 1e-3) / 1e9, "unit": "GB/s",
                                "frac": (fwd_bytes + sct_bytes) / (proj_avg * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
