"""ctypes loader for the CPU oracle (oracle/liboracle.so).

Test infrastructure: tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use this module to run the checker.
The product (``petalo_amd``) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

LOR_DTYPE = np.dtype([("p1", "<f4", 3), ("p2", "<f4", 3), ("dt", "<f4"), ("ac", "<f4")])


class Stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in
                ("n_lors", "n_steps", "n_nz", "n_skipped", "n_missed", "n_runaway", "n_shifted")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Tof(C.Structure):
    _fields_ = [("sigma_ps", C.c_float), ("cutoff", C.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_row_f32.restype = C.c_int64
        L.orc_row_f32.argtypes = [P, P, P, P, P, P, C.c_int64, P]
        L.orc_row_f64.restype = C.c_int64
        L.orc_row_f64.argtypes = [P, P, P, P, P, P, C.c_int64, P]
        L.orc_classic_siddon_f64.restype = C.c_int64
        L.orc_classic_siddon_f64.argtypes = [P, P, P, P, P, P, C.c_int64]
        L.orc_bruteforce_lengths_f64.argtypes = [P, P, P, P, P]
        L.orc_chord_f64.restype = C.c_double
        L.orc_chord_f64.argtypes = [P, P, P]
        L.orc_fov_entry_f32.restype = C.c_int
        L.orc_fov_entry_f32.argtypes = [P, P, P, P, P]
        L.orc_fov_derive_f32.argtypes = [P, P, P, P]
        L.orc_gauss_f32.restype = C.c_float
        L.orc_gauss_f32.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_gauss_f64.restype = C.c_double
        L.orc_gauss_f64.argtypes = [C.c_double, C.c_double, C.c_double]
        for name in ("orc_mlem_update_f32", "orc_mlem_update_f64"):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [P, P, P, P, C.c_size_t, P, P, C.c_int, P]
        L.orc_backproject_mlem_f32.restype = C.c_int
        L.orc_backproject_mlem_f32.argtypes = [P, P, P, P, C.c_size_t, P, P, C.c_int, P]
        for name in ("orc_sensitivity_f32", "orc_sensitivity_f64"):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [P, P, P, P, P, P, C.c_size_t, C.c_int, P, P]
        L.orc_rho_to_mu_factor_f32.restype = C.c_float
        L.orc_rho_to_mu_factor_f32.argtypes = [C.c_float]
        L.orc_walk_stats_f32.restype = C.c_int
        L.orc_walk_stats_f32.argtypes = [P, P, P, P, C.c_size_t, C.c_int, P]
        L.orc_row_mixed.restype = C.c_int64
        L.orc_row_mixed.argtypes = [P, P, P, P, P, P, C.c_int64, P]
        L.orc_mlem_update_mixed.restype = C.c_int
        L.orc_mlem_update_mixed.argtypes = [P, P, P, P, C.c_size_t, P, P, C.c_int, P]
        L.orc_set_entry_variant.restype = C.c_int
        L.orc_set_entry_variant.argtypes = [C.c_int]
        L.orc_entry_shifted_f32.restype = C.c_int
        L.orc_entry_shifted_f32.argtypes = [P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _fov(full, n):
    return np.asarray(full, np.float32), np.asarray(n, np.uint32)


def _tof(tof):
    if tof is None:
        return None
    return C.byref(Tof(float(tof[0]), float(tof[1])))


def make_lors(p1, p2, dt=0.0, ac=1.0):
    p1 = np.atleast_2d(np.asarray(p1, np.float32))
    p2 = np.atleast_2d(np.asarray(p2, np.float32))
    out = np.zeros(len(p1), LOR_DTYPE)
    out["p1"] = p1
    out["p2"] = p2
    out["dt"] = dt
    out["ac"] = ac
    return out


def row(full, n, lor, tof=None, f64=False, mixed=False):
    """One system-matrix row: (indices, weights, steps); runaway -> steps < 0.
    mixed: binary32 walk decisions with binary64 weights."""
    full, n = _fov(full, n)
    cap = int(2 * n.sum() + 64)
    idx = np.zeros(cap, np.int64)
    w = np.zeros(cap, np.float64 if (f64 or mixed) else np.float32)
    steps = C.c_uint64(0)
    lor = np.ascontiguousarray(np.asarray(lor, LOR_DTYPE).reshape(1))
    fn = lib().orc_row_mixed if mixed else (lib().orc_row_f64 if f64 else lib().orc_row_f32)
    m = fn(_p(full), _p(n), _tof(tof), _p(lor), _p(idx), _p(w), cap, C.byref(steps))
    runaway = m < 0
    if runaway:
        m = -1 - m
    return idx[:m].copy(), w[:m].copy(), (-1 if runaway else int(steps.value))


def classic_siddon(full, n, p1, p2):
    full = np.asarray(full, np.float64)
    n = np.asarray(n, np.uint32)
    p1 = np.asarray(p1, np.float64)
    p2 = np.asarray(p2, np.float64)
    cap = int(4 * n.sum() + 64)
    idx = np.zeros(cap, np.int64)
    ln = np.zeros(cap, np.float64)
    m = lib().orc_classic_siddon_f64(_p(full), _p(n), _p(p1), _p(p2), _p(idx), _p(ln), cap)
    return idx[:m].copy(), ln[:m].copy()


def bruteforce(full, n, p1, p2):
    full = np.asarray(full, np.float64)
    n = np.asarray(n, np.uint32)
    out = np.zeros(int(np.prod(n.astype(np.int64))), np.float64)
    lib().orc_bruteforce_lengths_f64(_p(full), _p(n), _p(np.asarray(p1, np.float64)),
                                     _p(np.asarray(p2, np.float64)), _p(out))
    return out


def chord(full, p1, p2):
    return lib().orc_chord_f64(_p(np.asarray(full, np.float64)), _p(np.asarray(p1, np.float64)),
                               _p(np.asarray(p2, np.float64)))


def fov_entry(full, n, p1, p2):
    full, n = _fov(full, n)
    out = np.zeros(3, np.float32)
    hit = lib().orc_fov_entry_f32(_p(full), _p(n), _p(np.asarray(p1, np.float32)),
                                  _p(np.asarray(p2, np.float32)), _p(out))
    return out if hit else None


def fov_derive(full, n):
    full, n = _fov(full, n)
    hw = np.zeros(3, np.float32)
    vs = np.zeros(3, np.float32)
    lib().orc_fov_derive_f32(_p(full), _p(n), _p(hw), _p(vs))
    return hw, vs


def gauss(sigma_ps, cutoff, dx, f64=False):
    return (lib().orc_gauss_f64 if f64 else lib().orc_gauss_f32)(sigma_ps, cutoff, dx)


def mlem_update(full, n, lors, image, sens=None, tof=None, threads=1, f64=False, mixed=False):
    """One subset update in place; returns the work counters.  mixed: binary32
    walk decisions with binary64 arithmetic (binary64 image)."""
    full, n = _fov(full, n)
    dt = np.float64 if (f64 or mixed) else np.float32
    assert image.dtype == dt and image.flags.c_contiguous
    lors = np.ascontiguousarray(lors, LOR_DTYPE)
    st = Stats()
    s = None if sens is None else _p(np.ascontiguousarray(sens, dt))
    fn = lib().orc_mlem_update_mixed if mixed else (lib().orc_mlem_update_f64 if f64 else lib().orc_mlem_update_f32)
    fn(_p(full), _p(n), _tof(tof), _p(lors), len(lors), s, _p(image), threads, C.byref(st))
    return st.as_dict()


def backproject(full, n, lors, image, tof=None, threads=1):
    full, n = _fov(full, n)
    lors = np.ascontiguousarray(lors, LOR_DTYPE)
    bp = np.zeros(int(np.prod(n.astype(np.int64))), np.float32)
    st = Stats()
    lib().orc_backproject_mlem_f32(_p(full), _p(n), _tof(tof), _p(lors), len(lors),
                                   _p(np.ascontiguousarray(image, np.float32)), _p(bp), threads,
                                   C.byref(st))
    return bp, st.as_dict()


def sensitivity(att_full, att_n, mu, lors, out_full=None, out_n=None, threads=1, f64=False):
    af, an = _fov(att_full, att_n)
    mu = np.ascontiguousarray(mu, np.float32)
    lors = np.ascontiguousarray(lors, LOR_DTYPE)
    if out_full is not None:
        of, on = _fov(out_full, out_n)
        nv = int(np.prod(on.astype(np.int64)))
        ofp, onp = _p(of), _p(on)
    else:
        nv = int(np.prod(an.astype(np.int64)))
        ofp = onp = None
    out = np.zeros(nv, np.float64 if f64 else np.float32)
    st = Stats()
    fn = lib().orc_sensitivity_f64 if f64 else lib().orc_sensitivity_f32
    fn(_p(af), _p(an), _p(mu), ofp, onp, _p(lors), len(lors), threads, _p(out), C.byref(st))
    return out, st.as_dict()


def ingest(records, params, spec=None):
    """Oracle LOR ingestion; params / spec are the library's ctypes structs
    (same layout as orc_ingest_t / orc_sgram_spec_t)."""
    rec = np.ascontiguousarray(records)
    rec = rec.view(np.float32).reshape(-1, 11) if rec.dtype != np.float32 else rec.reshape(-1, 11)
    rec = np.ascontiguousarray(rec, np.float32)
    out = np.zeros(len(rec), LOR_DTYPE)
    n_read = C.c_uint64(0)
    L = lib()
    L.orc_ingest_f32.restype = C.c_int64
    L.orc_ingest_f32.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    m = L.orc_ingest_f32(_p(rec), len(rec), C.byref(params), C.byref(spec) if spec is not None else None,
                         _p(out), C.byref(n_read))
    return out[:m].copy(), int(n_read.value)


def discrete_lors(r_min, dr, dz, da, length, adjust, seed=0, capacity=None):
    """(LORs, number of elements) of a discretized detector, all pairs."""
    L = lib()
    L.orc_discrete_lors.restype = C.c_int64
    L.orc_discrete_lors.argtypes = [C.c_float] * 5 + [C.c_int, C.c_uint64, C.c_void_p, C.c_size_t, C.c_void_p]
    ne = C.c_uint64(0)
    m = L.orc_discrete_lors(r_min, dr, dz, da, length, adjust, seed, None, 0, C.byref(ne))
    cap = m if capacity is None else min(m, capacity)
    out = np.zeros(cap, LOR_DTYPE)
    L.orc_discrete_lors(r_min, dr, dz, da, length, adjust, seed, _p(out) if cap else None, cap, C.byref(ne))
    return out, int(ne.value)


ROI_DTYPE = np.dtype([("kind", np.int32), ("c", np.float32, 3), ("r", np.float32)])


def roi_stats(full, n, image, rois, f64=False):
    """(counts, means, sigmas) of regions (ROI_DTYPE records) of an image."""
    L = lib()
    L.orc_roi_stats.restype = C.c_int
    f, nn = _fov(full, n)
    rois = np.ascontiguousarray(rois, ROI_DTYPE)
    img = np.ascontiguousarray(image, np.float32)
    k = len(rois)
    cnt, mean, sig = np.zeros(k, np.uint64), np.zeros(k), np.zeros(k)
    L.orc_roi_stats(_p(f), _p(nn), _p(img), _p(rois), C.c_size_t(k), C.c_int(int(f64)), _p(cnt), _p(mean), _p(sig))
    return cnt, mean, sig


def image_foms(full, n, image, rois, activities, bg_rois, bg_activity, f64=False):
    L = lib()
    f, nn = _fov(full, n)
    rois = np.ascontiguousarray(rois, ROI_DTYPE)
    bg = np.ascontiguousarray(bg_rois, ROI_DTYPE)
    act = np.ascontiguousarray(activities, np.float32)
    img = np.ascontiguousarray(image, np.float32)
    crcs, snrs, bgm = np.zeros(len(rois), np.float32), np.zeros(len(rois), np.float32), C.c_float(0)
    rc = L.orc_image_foms(_p(f), _p(nn), _p(img), _p(rois), _p(act), C.c_size_t(len(rois)), _p(bg),
                          C.c_size_t(len(bg)), C.c_float(bg_activity), C.c_int(int(f64)), _p(crcs), _p(snrs),
                          C.byref(bgm))
    if rc != 0:
        raise ValueError("a region holds no voxel")
    return crcs, snrs, bgm.value


def phantom_foms(full, n, image, phantom, f64=False, aoc_capacity=256):
    """(foms [6 x (r, crc, bg variability, snr)], aocs) of the foms tool."""
    L = lib()
    f, nn = _fov(full, n)
    img = np.ascontiguousarray(image, np.float32)
    foms = np.zeros((6, 4), np.float32)
    aocs = np.zeros(aoc_capacity, np.float32)
    na = C.c_size_t(0)
    rc = L.orc_phantom_foms(_p(f), _p(nn), _p(img), C.c_int(phantom), C.c_int(int(f64)), _p(foms), _p(aocs),
                            C.c_size_t(aoc_capacity), C.byref(na))
    if rc != 0:
        raise ValueError("a region holds no voxel")
    return foms, aocs[:min(na.value, aoc_capacity)].copy()


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    L = lib()
    L.orc_philox4x32_10.restype = None
    L.orc_philox4x32_10.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def rho_to_mu_factor(rho_to_mu):
    return lib().orc_rho_to_mu_factor_f32(rho_to_mu)


def walk_stats(full, n, lors, tof=None, threads=8):
    full, n = _fov(full, n)
    lors = np.ascontiguousarray(lors, LOR_DTYPE)
    st = Stats()
    lib().orc_walk_stats_f32(_p(full), _p(n), _tof(tof), _p(lors), len(lors), threads, C.byref(st))
    return st.as_dict()


def set_entry_variant(v):
    """0: the restated ncollide3d cast (default); 1: the hit face's coordinate exact."""
    assert lib().orc_set_entry_variant(int(v)) == 0


class entry_variant:
    """Context manager selecting an entry variant for the calls inside it."""

    def __init__(self, v):
        self.v = v

    def __enter__(self):
        set_entry_variant(self.v)

    def __exit__(self, *a):
        set_entry_variant(0)


def entry_shifted(full, n, lors):
    """Per LOR: 1 if the binary32 walk starts shifted, 0 if not, -1 on a miss."""
    full, n = _fov(full, n)
    lors = np.ascontiguousarray(lors, LOR_DTYPE)
    return np.array([lib().orc_entry_shifted_f32(_p(full), _p(n), _p(lors[i:i + 1])) for i in range(len(lors))],
                    np.int32)


def sensitivity_lors(full, n, n_lors, seed, length, diameter):
    """(LORs, attempts per LOR) of the library's sensitivity stream, restated."""
    f, nn = _fov(full, n)
    L = lib()
    L.orc_sensitivity_lors.restype = C.c_int64
    L.orc_sensitivity_lors.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint64, C.c_float, C.c_float,
                                       C.c_void_p, C.c_void_p]
    out = np.zeros(int(n_lors), LOR_DTYPE)
    att = np.zeros(int(n_lors), np.uint32)
    L.orc_sensitivity_lors(_p(f), _p(nn), int(n_lors), int(seed), float(length), float(diameter), _p(out), _p(att))
    return out, att
