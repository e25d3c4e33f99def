"""Shared test helpers: seeded LOR sets and error metrics."""
import numpy as np

import oracle_ffi as O

C_MM_PER_PS = np.float32(0.299792458)


def rel_l2(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def max_rel(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def proptest_lors(rng, n):
    """LORs and FOVs drawn like siddon.rs:440-490 (sum_of_weights proptest):
    r in [200,300), angles, z in [-200,200), FOV sizes and voxel counts."""
    out = []
    for _ in range(n):
        r = np.float32(rng.uniform(200, 300))
        a1 = np.float32(rng.uniform(0, 1))
        d2 = np.float32(rng.uniform(0.1, 0.9))
        z1 = np.float32(rng.uniform(-200, 200))
        z2 = np.float32(rng.uniform(-200, 200))
        full = np.array([rng.uniform(100, 150), rng.uniform(100, 150), rng.uniform(100, 190)], np.float32)
        nv = np.array([rng.integers(5, 50), rng.integers(5, 50), rng.integers(5, 90)], np.uint32)
        t1 = np.float32(a1 * np.float32(2 * np.pi))
        t2 = np.float32(t1 + d2 * np.float32(2 * np.pi))
        p1 = np.array([r * np.cos(t1), r * np.sin(t1), z1], np.float32)
        p2 = np.array([r * np.cos(t2), r * np.sin(t2), z2], np.float32)
        out.append((full, nv, p1, p2))
    return out


def random_lors(rng, n, radius=60.0, zhalf=40.0, dt_sigma_ps=0.0):
    """Endpoints uniform on a cylinder around a small FOV, plus TOF dt."""
    t1 = rng.uniform(0, 2 * np.pi, n)
    t2 = rng.uniform(0, 2 * np.pi, n)
    p1 = np.stack([radius * np.cos(t1), radius * np.sin(t1), rng.uniform(-zhalf, zhalf, n)], 1)
    p2 = np.stack([radius * np.cos(t2), radius * np.sin(t2), rng.uniform(-zhalf, zhalf, n)], 1)
    lors = O.make_lors(p1, p2)
    if dt_sigma_ps:
        lors["dt"] = rng.normal(0, dt_sigma_ps, n).astype(np.float32)
    return lors


def ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


def rel_l2_finite(x, ref):
    """Relative L2 over the voxels finite in both, after checking that the
    non-finite voxels (NaN from 0 * inf when a LOR projects to 0, inf where a
    projection of 0 meets a non-zero voxel, as in the reference) coincide."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    assert np.array_equal(np.isnan(x), np.isnan(ref)), "NaN voxels differ"
    assert np.array_equal(np.isinf(x), np.isinf(ref)), "infinite voxels differ"
    fin = np.isfinite(ref)
    if not np.any(ref[fin]):
        return float(np.abs(x[fin]).max(initial=0.0))
    return rel_l2(x[fin], ref[fin])
