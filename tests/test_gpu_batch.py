"""Batched OSEM in the shape of the batched workload, against the oracle.

32 independent TOF reconstructions (images 60^3 .. 120^3 over a 180 mm FOV,
LOR counts log-uniform in [1e4, 1e6] -- the workload's [1e5, 1e7] scaled down
tenfold so that the oracle finishes in seconds), 20 subsets x 5 iterations,
advanced together by one batch; every reconstruction is compared with the
binary32 oracle (src/mlem.rs:8-31, 90-95: contiguous subsets, remainder
dropped) after full iterations 1 and 5.
"""
import time

import numpy as np
import pytest

import oracle_ffi as O
import petalo_amd as P
from helpers import max_rel, rel_l2_finite

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def batch_workload(scale=0.1, n_rec=32, seed=4):
    """(fovs, lors, sizes, counts): the batched workload's shapes, LOR counts
    log-uniform in [1e5, 1e7] * scale."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([60, 72, 84, 96, 108, 120], n_rec)
    counts = np.round(scale * 10 ** (5 + 2 * rng.uniform(size=n_rec))).astype(np.int64)
    fovs = [P.FOV([180.0] * 3, [int(v)] * 3) for v in sizes]
    lors = [P.synth_phantom_lors(100 + k, int(counts[k]), 200.0) for k in range(n_rec)]
    return fovs, lors, sizes, counts


def test_batched_workload_matches_oracle(gpu):
    """Every reconstruction after subset updates 1, 2 and 3 and after full
    iterations 1 and 5.  Without a sensitivity image, a voxel that no LOR of
    a subset crosses is zeroed (mlem.rs:45-52) and a later LOR crossing only
    zeroed voxels projects to 0, leaving 0 * inf = NaN: with these sparse
    subsets the images fill with zeros and NaN within the first iteration, as
    the reference's would.  The comparison is NaN-aware (the same NaN and inf
    voxels) and the early updates, where most voxels are still finite and
    non-zero, carry the weight of the check."""
    tof = (42.47, 3.0)
    subsets, iters = 20, 5
    fovs, lors, sizes, counts = batch_workload()
    _say(f"batch: {len(fovs)} reconstructions, {counts.sum()} LORs ({counts.min()}..{counts.max()}), "
         f"images {sorted(set(int(v) for v in sizes))}")
    b = P.Batch(fovs, lors, tofs=[P.Tof(*tof)] * len(fovs), n_subsets=subsets)
    x = [np.ones(f.nvox, np.float32) for f in fovs]
    t_or = 0.0
    for it in range(1, iters + 1):
        for sub in range(subsets):
            osem = b.step()
            assert (osem.iteration, osem.subset) == (it, sub + 1)
            t0 = time.time()
            for k, f in enumerate(fovs):
                m = len(lors[k]) // subsets
                O.mlem_update(f.full_size, f.n, lors[k][sub * m:(sub + 1) * m], x[k], tof=tof, threads=16)
            t_or += time.time() - t0
            if not ((it == 1 and sub < 3) or sub == subsets - 1 and it in (1, iters)):
                continue
            worst, live = (0.0, 0), 0
            for k in range(len(fovs)):
                g = b.image(k).data
                e = rel_l2_finite(g, x[k])
                live += int(np.count_nonzero(np.isfinite(x[k]) & (x[k] != 0)))
                if e >= worst[0]:
                    worst = (e, k)
                assert e <= 1e-4, f"reconstruction {k} ({sizes[k]}^3, {counts[k]} LORs) update {it}.{sub + 1}: {e:.2e}"
            k = worst[1]
            gk = b.image(k).data
            fin = np.isfinite(x[k])
            _say(f"update {it}.{sub + 1}: worst relL2 {worst[0]:.2e} (reconstruction {k}, {sizes[k]}^3, "
                 f"{counts[k]} LORs, max {max_rel(gk[fin], x[k][fin]) if np.any(x[k][fin]) else 0.0:.2e}); "
                 f"{live} finite non-zero voxels compared over the batch; oracle {t_or:.1f} s so far")
            if it == 1 and sub < 3:
                assert live >= 1_000_000
    b.close()


def test_autotuned_batch_matches_untuned(gpu):
    """A batch created with autotune (work granularity timed on the batch,
    then images and subset position reset) reconstructs as the batch with the
    automatic rule, and both match the oracle after the first subset updates."""
    fovs, lors, sizes, counts = batch_workload(scale=0.02, n_rec=8, seed=13)
    tofs = [P.Tof(42.47, 3.0)] * len(fovs)
    tuned = P.Batch(fovs, lors, tofs=tofs, n_subsets=4, autotune=True)
    plain = P.Batch(fovs, lors, tofs=tofs, n_subsets=4)
    steps, n_items = tuned.work_items()
    _say(f"batch autotune: {steps} steps per merged item ({n_items} items over 4 subsets; "
         f"automatic rule {plain.work_items()[1]} items)")
    assert steps == 0 or (1 << 14) <= steps <= (1 << 20)
    x = [np.ones(f.nvox, np.float32) for f in fovs]
    for sub in range(3):
        a, b = tuned.step(), plain.step()
        assert (a.iteration, a.subset) == (b.iteration, b.subset) == (1, sub + 1)
        for k, f in enumerate(fovs):
            m = len(lors[k]) // 4
            O.mlem_update(f.full_size, f.n, lors[k][sub * m:(sub + 1) * m], x[k], tof=(42.47, 3.0), threads=16)
            gt, gp = tuned.image(k).data, plain.image(k).data
            assert rel_l2_finite(gt, gp) <= 1e-6, (sub, k)
            assert rel_l2_finite(gt, x[k]) <= 1e-4, (sub, k)
    tuned.close()
    plain.close()


def test_persistent_and_phase_schedules_agree(gpu):
    """The persistent launch (one task queue per subset update) and the
    launch-per-phase schedule give the same images (tile sums flushed in
    another order, rounded to binary32), for a batch mixing TOF and non-TOF
    reconstructions, graphs on and off; no wait of the persistent launch ran
    into its time limit (image reads would fail)."""
    fovs, lors, sizes, counts = batch_workload(scale=0.02, n_rec=12, seed=9)
    tofs = [P.Tof(42.47, 3.0) if k % 3 else None for k in range(len(fovs))]
    imgs = {}
    for persistent in (True, False):
        for graphs in (True, False):
            b = P.Batch(fovs, lors, tofs=tofs, n_subsets=4)
            b.set_schedule(persistent)
            b.set_options(timing=False, graphs=graphs)
            for _ in range(6):
                b.step()
            imgs[(persistent, graphs)] = [b.image(k).data.copy() for k in range(len(fovs))]
            b.close()
    ref = imgs[(False, False)]
    for key, got in imgs.items():
        for k in range(len(fovs)):
            e = rel_l2_finite(got[k], ref[k])
            assert e <= 1e-6, (key, k, e)
    _say(f"schedules agree over {len(fovs)} reconstructions ({counts.sum()} LORs), 6 subset updates")
