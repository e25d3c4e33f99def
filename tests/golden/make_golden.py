"""Writes tests/golden/oracle_rows.npz and tests/golden/oracle_mlem8.npz from
the CPU oracle (run from the repository root after `make -C oracle`):

  python tests/golden/make_golden.py

oracle_rows.npz: 1000 LORs on three small grids (8^3, 12x10x16, 16^3) --
random oblique chords through the grid, axis-parallel and face-diagonal lines through voxel
corners (boundary ties), endpoints on voxel boundaries, misses -- and for
each its row (voxel indices, binary32 weights), walk steps and tie count
(steps whose length is 0, where two boundaries coincide); 200 of them also
carry a TOF dt and a TOF row.

oracle_mlem8.npz: 1000 LORs through an 8^3 grid of 3 mm voxels, three MLEM
iterations from ones in binary32 and binary64, without and with TOF.

The fixtures pin the oracle against later changes (tests/test_oracle.py
regenerates them) and give the GPU tests committed inputs and outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_ffi as O  # noqa: E402

GRIDS = [([24.0, 24.0, 24.0], [8, 8, 8]), ([36.0, 20.0, 32.0], [12, 10, 16]), ([32.0, 32.0, 32.0], [16, 16, 16])]
TOF = (42.47, 3.0)


def row_lors(rng):
    """(grid index, LOR) pairs, 1000 in all."""
    out = []
    for g, (full, n) in enumerate(GRIDS):
        h = np.asarray(full) / 2
        vs = np.asarray(full) / np.asarray(n)
        # random oblique chords through a random point of the grid
        k = 242
        u = rng.normal(size=(k, 3))
        c = rng.uniform(-h, h, size=(k, 3))
        rad = 1.5 * np.linalg.norm(h)
        d = rad * u / np.linalg.norm(u, axis=1, keepdims=True)
        p1, p2 = c + d, c - d
        out += [(g, a, b) for a, b in zip(p1, p2)]
        # axis-parallel lines on voxel boundaries and centres
        for axis in range(3):
            for _ in range(20):
                q = np.array([rng.integers(-n[i] // 2, n[i] // 2 + 1) * vs[i] for i in range(3)], float)
                if rng.random() < 0.5:
                    q += 0.5 * vs
                a, b = q.copy(), q.copy()
                a[axis], b[axis] = -2 * h[axis], 2 * h[axis]
                if rng.random() < 0.5:
                    a, b = b, a
                out.append((g, a, b))
        # face and body diagonals through voxel corners (ties)
        for _ in range(25):
            c = np.array([rng.integers(-n[i] // 2, n[i] // 2 + 1) * vs[i] for i in range(3)], float)
            d = np.array([rng.choice([-1, 0, 1]) * vs[i] for i in range(3)], float)
            if not d.any():
                d[0] = vs[0]
            out.append((g, c - 40 * d, c + 40 * d))
        # misses
        for _ in range(8):
            off = np.array([3 * h[0], rng.uniform(-h[1], h[1]), rng.uniform(-h[2], h[2])])
            out.append((g, off + [0, -100, 0], off + [0, 100, 0]))
    return out[:1000]


def make_rows(rng):
    sel = row_lors(rng)
    assert len(sel) == 1000
    grid = np.array([g for g, _, _ in sel], np.int32)
    lors = O.make_lors(np.array([a for _, a, _ in sel]), np.array([b for _, _, b in sel]))
    tof_mask = np.zeros(len(lors), bool)
    tof_mask[rng.choice(len(lors), 200, replace=False)] = True
    lors["dt"][tof_mask] = rng.normal(0, 84.93, tof_mask.sum()).astype(np.float32)
    idx, w, off, steps = [], [], [0], []
    tidx, tw, toff = [], [], [0]
    for k, l in enumerate(lors):
        full, n = GRIDS[grid[k]]
        i, ww, st = O.row(full, n, l)
        idx.append(i.astype(np.int32))
        w.append(ww)
        off.append(off[-1] + len(i))
        steps.append(st)
        if tof_mask[k]:
            i2, w2, _ = O.row(full, n, l, tof=TOF)
            tidx.append(i2.astype(np.int32))
            tw.append(w2)
            toff.append(toff[-1] + len(i2))
    steps = np.array(steps, np.int64)
    lens = np.diff(off)
    return dict(grid=grid, lors=lors.view(np.float32).reshape(-1, 8), idx=np.concatenate(idx), w=np.concatenate(w),
                offsets=np.array(off, np.int64), steps=steps, ties=np.where(steps >= 0, steps - lens, 0),
                tof_lors=np.flatnonzero(tof_mask).astype(np.int32), tof_idx=np.concatenate(tidx),
                tof_w=np.concatenate(tw), tof_offsets=np.array(toff, np.int64),
                grids_full=np.array([g[0] for g in GRIDS], np.float32),
                grids_n=np.array([g[1] for g in GRIDS], np.uint32),
                tof=np.array(TOF, np.float32))


def make_mlem(rng):
    full, n = [24.0] * 3, [8] * 3
    t1, t2 = rng.uniform(0, 2 * np.pi, (2, 1000))
    z1, z2 = rng.uniform(-14, 14, (2, 1000))
    p1 = np.stack([30 * np.cos(t1), 30 * np.sin(t1), z1], 1)
    p2 = np.stack([30 * np.cos(t2), 30 * np.sin(t2), z2], 1)
    lors = O.make_lors(p1, p2)
    lors["dt"] = rng.normal(0, 84.93, 1000).astype(np.float32)
    out = dict(lors=lors.view(np.float32).reshape(-1, 8), full=np.array(full, np.float32),
               n=np.array(n, np.uint32), tof=np.array(TOF, np.float32))
    for name, tof in (("notof", None), ("tof", TOF)):
        for f64 in (False, True):
            x = np.ones(512, np.float64 if f64 else np.float32)
            imgs = []
            for _ in range(3):
                O.mlem_update(full, n, lors, x, tof=tof, threads=1, f64=f64)
                imgs.append(x.copy())
            out[f"{name}_{'f64' if f64 else 'f32'}"] = np.stack(imgs)
    return out


def main():
    np.savez_compressed(os.path.join(HERE, "oracle_rows.npz"), **make_rows(np.random.default_rng(2024)))
    np.savez_compressed(os.path.join(HERE, "oracle_mlem8.npz"), **make_mlem(np.random.default_rng(88)))


if __name__ == "__main__":
    main()
