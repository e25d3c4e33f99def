"""petalo_amd -- MI355X (gfx950) implementation of petalo's reconstruction hot path.

Python mirror of the reference interface over the C ABI of
``libpetalo_hip.so`` (``include/petalo_hip.h``):

  FOV, Image, Osem, Tof ........... src/fov.rs, src/image.rs, src/mlem.rs:70-96,
                                    src/config/mlem.rs:177-188
  mlem(...) ....................... petalo::mlem::mlem (src/mlem.rs:8-31): an
                                    infinite iterator of (Image, Osem), one item
                                    per OSEM subset update
  make_sensitivity_image(...) ..... continuous::sensitivity_image
                                    (src/bin/make_sensitivity_image/continuous.rs:3-17)
  system_matrix_rows(...) ......... Siddon::new_system_matrix_row (siddon.rs:69-73)
  Image3D read/write, Config ...... src/io/raw.rs:78-126, src/config/mlem.rs

All numerical work runs in the HIP library; there is no CPU fallback.  Loading
fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "FOV", "Image", "Osem", "Tof", "Stats", "LOR_DTYPE", "PetaloError", "lib", "mlem", "Mlem",
    "make_sensitivity_image", "sensitivity_lors", "system_matrix_rows", "density_to_attenuation",
    "synth_phantom_lors", "synth_density", "philox4x32_10", "read_config", "parse_config", "parse_quantity",
    "write_image3d", "read_image3d", "device_count", "make_lors",
    "SCHEME_AUTO", "SCHEME_ATOMIC", "SCHEME_TILE", "PRECISION_F32", "PRECISION_F64", "PRECISION_F32_WALK", "Batch", "ingest_lors", "ingest_params",
    "scattergram_spec", "H5LOR_DTYPE", "make_sensitivity_image_discrete",
    "roi", "roi_stats", "image_foms", "phantom_foms", "ROI_DTYPE", "ROI_STATS_DTYPE",
]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libpetalo_hip.so")

LOR_DTYPE = np.dtype([("p1", "<f4", 3), ("p2", "<f4", 3), ("dt", "<f4"), ("ac", "<f4")])
assert LOR_DTYPE.itemsize == 32

SCHEME_AUTO, SCHEME_ATOMIC, SCHEME_TILE = 0, 1, 2


class PetaloError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code, message):
        super().__init__(f"petalo error {code}: {message}")
        self.code = code


class _Fov(C.Structure):
    _fields_ = [("full_size_mm", C.c_float * 3), ("n", C.c_uint32 * 3)]


class _Tof(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("sigma_ps", C.c_float), ("cutoff", C.c_float)]


class _Osem(C.Structure):
    _fields_ = [("n_subsets", C.c_uint64), ("iteration", C.c_uint64), ("subset", C.c_uint64)]


class _Stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in
                ("n_lors", "n_steps", "n_nz", "n_skipped", "n_missed", "n_runaway", "n_shifted")]


class _Options(C.Structure):
    _fields_ = [("scheme", C.c_int32), ("sort_lors", C.c_int32), ("device", C.c_int32),
                ("tile", C.c_int32 * 3), ("autotune", C.c_int32), ("precision", C.c_int32),
                ("item_steps", C.c_int32)]


PRECISION_F32, PRECISION_F64, PRECISION_F32_WALK = 0, 1, 2


class _Ingest(C.Structure):
    _fields_ = [("events_min", C.c_uint64), ("events_max", C.c_uint64),
                ("has_events_min", C.c_int32), ("has_events_max", C.c_int32),
                ("has_charge_min", C.c_int32), ("has_charge_max", C.c_int32),
                ("charge_min", C.c_float), ("charge_max", C.c_float),
                ("has_energy_min", C.c_int32), ("has_energy_max", C.c_int32),
                ("energy_min", C.c_float), ("energy_max", C.c_float),
                ("smear_energy_fwhm", C.c_float), ("has_z_max", C.c_int32), ("z_max_mm", C.c_float),
                ("r_min_mm", C.c_float), ("dr_mm", C.c_float), ("dz_mm", C.c_float), ("da_mm", C.c_float),
                ("adjust", C.c_int32), ("seed", C.c_uint64)]


class _SgramSpec(C.Structure):
    _fields_ = [("phi_bins", C.c_uint32), ("z_bins", C.c_uint32), ("z_length_mm", C.c_float),
                ("dz_bins", C.c_uint32), ("dz_max_mm", C.c_float), ("r_bins", C.c_uint32),
                ("r_max_mm", C.c_float), ("dt_bins", C.c_uint32), ("dt_max_ps", C.c_float)]


ADJUST_NONE, ADJUST_CENTRE, ADJUST_RANDOM, ADJUST_RANDOM_Z_PHI = 0, 1, 2, 3
H5LOR_DTYPE = np.dtype([(k, "<f4") for k in ("dt", "x1", "y1", "z1", "x2", "y2", "z2", "q1", "q2", "E1", "E2")])


def ingest_params(events=(None, None), charge=(None, None), energy=(None, None), smear_energy_fwhm=0.0,
                  z_max=None, discretize=(0.0, 0.0, 0.0, 0.0), adjust=ADJUST_RANDOM_Z_PHI, seed=0):
    """Parameters of ingest_lors; bounds are (min, max) pairs with None for open."""
    p = _Ingest()
    p.has_events_min, p.events_min = (0, 0) if events[0] is None else (1, int(events[0]))
    p.has_events_max, p.events_max = (0, 0) if events[1] is None else (1, int(events[1]))
    p.has_charge_min, p.charge_min = (0, 0.0) if charge[0] is None else (1, float(charge[0]))
    p.has_charge_max, p.charge_max = (0, 0.0) if charge[1] is None else (1, float(charge[1]))
    p.has_energy_min, p.energy_min = (0, 0.0) if energy[0] is None else (1, float(energy[0]))
    p.has_energy_max, p.energy_max = (0, 0.0) if energy[1] is None else (1, float(energy[1]))
    p.smear_energy_fwhm = float(smear_energy_fwhm)
    p.has_z_max, p.z_max_mm = (0, 0.0) if z_max is None else (1, float(z_max))
    p.r_min_mm, p.dr_mm, p.dz_mm, p.da_mm = (float(v) for v in discretize)
    p.adjust = int(adjust)
    p.seed = int(seed)
    return p


def scattergram_spec(phi=0, z=(0, 0.0), dz=(0, 0.0), r=(0, 0.0), dt=(0, 0.0)):
    """[scatter_correction] axes: phi bins; (bins, length or max) for z, dz, r, dt (0 = unset)."""
    return _SgramSpec(int(phi), int(z[0]), float(z[1]), int(dz[0]), float(dz[1]), int(r[0]), float(r[1]),
                      int(dt[0]), float(dt[1]))


def ingest_lors(records, params, scattergram=None):
    """reco_info/lors records (H5LOR_DTYPE or n x 11 float32) -> (LORs, records read)
    (src/io/hdf5.rs:97-228).  The LORs are a view of a buffer sized for every
    record (the kept ones first)."""
    rec = np.ascontiguousarray(records)
    rec = rec.view(np.float32).reshape(-1, 11) if rec.dtype != np.float32 else rec.reshape(-1, 11)
    rec = np.ascontiguousarray(rec, np.float32)
    out = np.zeros(len(rec), LOR_DTYPE)
    n_out = C.c_size_t(0)
    n_read = C.c_uint64(0)
    _check(lib().petalo_ingest_lors(_ptr(rec), len(rec), C.byref(params),
                                    C.byref(scattergram) if scattergram is not None else None,
                                    _ptr(out), len(out), C.byref(n_out), C.byref(n_read)))
    return out[:n_out.value], int(n_read.value)


class _Config(C.Structure):
    _fields_ = [
        ("input_file", C.c_char * 512), ("input_dataset", C.c_char * 256),
        ("iterations_number", C.c_uint64), ("iterations_subsets", C.c_uint64),
        ("fov", _Fov), ("has_tof", C.c_int32), ("tof", _Tof),
        ("has_attenuation_correction", C.c_int32), ("sensitivity_image", C.c_char * 512),
        ("has_energy_min", C.c_int32), ("has_energy_max", C.c_int32),
        ("has_charge_min", C.c_int32), ("has_charge_max", C.c_int32),
        ("has_events_min", C.c_int32), ("has_events_max", C.c_int32),
        ("energy_min", C.c_float), ("energy_max", C.c_float),
        ("charge_min", C.c_float), ("charge_max", C.c_float),
        ("events_min", C.c_uint64), ("events_max", C.c_uint64),
        ("has_scatter_correction", C.c_int32), ("scatter", _SgramSpec),
        ("has_smear_energy", C.c_int32), ("smear_energy_fwhm", C.c_float),
        ("has_detector_full_axial_length", C.c_int32), ("detector_full_axial_length_mm", C.c_float),
    ]


# name -> (restype, argtypes)
_P = C.c_void_p
_SIGS = {
    "petalo_last_error": (C.c_char_p, []),
    "petalo_version": (C.c_int, []),
    "petalo_device_count": (C.c_int, [_P]),
    "petalo_device_synchronize": (C.c_int, [C.c_int]),
    "petalo_copy_probe": (C.c_int, [C.c_size_t, C.c_int, _P]),
    "petalo_fov_derive": (C.c_int, [_P, _P, _P]),
    "petalo_fov_parse": (C.c_int, [C.c_char_p, _P]),
    "petalo_fov_voxel_centre": (C.c_int, [_P, _P, _P]),
    "petalo_mlem_create": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, C.c_size_t, _P]),
    "petalo_mlem_next": (C.c_int, [_P, _P, _P]),
    "petalo_mlem_get_image": (C.c_int, [_P, _P]),
    "petalo_mlem_get_image_f64": (C.c_int, [_P, _P]),
    "petalo_mlem_set_image": (C.c_int, [_P, _P]),
    "petalo_mlem_walk_stats": (C.c_int, [_P, C.c_size_t, _P]),
    "petalo_mlem_last_skipped": (C.c_int, [_P, _P]),
    "petalo_mlem_accumulator": (C.c_int, [_P, _P, _P, _P]),
    "petalo_mlem_set_timing": (C.c_int, [_P, C.c_int]),
    "petalo_mlem_last_timing": (C.c_int, [_P, _P, _P]),
    "petalo_mlem_last_phase_timing": (C.c_int, [_P, _P, _P, _P]),
    "petalo_mlem_last_kernel_timing": (C.c_int, [_P, _P, C.c_size_t]),
    "petalo_mlem_drain_kernel_timings": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "petalo_mlem_info": (C.c_int, [_P, _P, _P]),
    "petalo_mlem_destroy": (C.c_int, [_P]),
    "petalo_mlem_set_graphs": (C.c_int, [_P, C.c_int]),
    "petalo_batch_create": (C.c_int, [_P, C.c_size_t, _P, _P, _P, _P, _P, C.c_size_t, _P]),
    "petalo_batch_next": (C.c_int, [_P, _P]),
    "petalo_batch_work_items": (C.c_int, [_P, _P, _P]),
    "petalo_mlem_work_items": (C.c_int, [_P, _P, _P]),
    "petalo_batch_get_image": (C.c_int, [_P, C.c_size_t, _P]),
    "petalo_batch_set_options": (C.c_int, [_P, C.c_int, C.c_int]),
    "petalo_batch_set_schedule": (C.c_int, [_P, C.c_int]),
    "petalo_batch_last_ms": (C.c_int, [_P, _P]),
    "petalo_batch_destroy": (C.c_int, [_P]),
    "petalo_make_sensitivity_image": (C.c_int, [_P, _P, _P, _P, C.c_size_t, C.c_uint64, C.c_float,
                                                C.c_float, _P, _P]),
    "petalo_sensitivity_lors": (C.c_int, [_P, C.c_size_t, C.c_uint64, C.c_float, C.c_float, _P]),
    "petalo_sensitivity_image_noalloc": (C.c_int, [_P, _P, _P, C.c_size_t, _P, _P]),
    "petalo_release_workspaces": (C.c_int, []),
    "petalo_density_to_attenuation": (C.c_int, [_P, C.c_size_t, C.c_float]),
    "petalo_system_matrix_rows": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_size_t, _P, _P, _P, _P]),
    "petalo_image3d_write": (C.c_int, [C.c_char_p, _P, _P]),
    "petalo_image3d_read_header": (C.c_int, [C.c_char_p, _P]),
    "petalo_image3d_read": (C.c_int, [C.c_char_p, _P, _P, C.c_size_t]),
    "petalo_config_parse": (C.c_int, [C.c_char_p, _P]),
    "petalo_config_read": (C.c_int, [C.c_char_p, _P]),
    "petalo_parse_quantity": (C.c_int, [C.c_char_p, C.c_char, _P]),
    "petalo_synth_phantom_lors": (C.c_int, [C.c_uint64, C.c_size_t, C.c_float, _P]),
    "petalo_synth_density": (C.c_int, [_P, _P]),
    "petalo_philox4x32_10": (C.c_int, [_P, _P, _P]),
    "petalo_ingest_lors": (C.c_int, [_P, C.c_size_t, _P, _P, _P, C.c_size_t, _P, _P]),
    "petalo_roi_stats": (C.c_int, [_P, _P, _P, C.c_size_t, _P]),
    "petalo_image_foms": (C.c_int, [_P, _P, _P, _P, C.c_size_t, _P, C.c_size_t, C.c_float, _P, _P, _P]),
    "petalo_phantom_foms": (C.c_int, [_P, _P, C.c_int32, _P, _P, C.c_size_t, _P]),
    "petalo_mlem_tile_shape": (C.c_int, [_P, _P]),
    "petalo_config_ingest": (C.c_int, [_P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint64, _P]),
    "petalo_make_sensitivity_image_discrete": (C.c_int, [_P, _P, _P, C.c_float, C.c_float, C.c_float, C.c_float,
                                                         C.c_float, C.c_int32, C.c_uint64, _P, _P]),
}

_lib = None


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m petalo_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().petalo_last_error()
        raise PetaloError(rc, msg.decode() if msg else "")


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def copy_probe(nbytes, iters=3):
    """Plain device copy of nbytes (calibration for traffic counters); ms per copy."""
    ms = C.c_double(0)
    _check(lib().petalo_copy_probe(int(nbytes), int(iters), C.byref(ms)))
    return ms.value


def device_count():
    n = C.c_int(0)
    _check(lib().petalo_device_count(C.byref(n)))
    return n.value


# ---------------------------------------------------------------------------
class FOV:
    """Field of view: full size (mm) and voxel counts (src/fov.rs:13-39)."""

    def __init__(self, full_size_mm, n):
        self.full_size = tuple(float(np.float32(x)) for x in full_size_mm)
        self.n = tuple(int(x) for x in n)
        hw = np.zeros(3, np.float32)
        vs = np.zeros(3, np.float32)
        _check(lib().petalo_fov_derive(C.byref(self._c()), _ptr(hw), _ptr(vs)))
        self.half_width = hw
        self.voxel_size = vs

    def _c(self):
        return _Fov((C.c_float * 3)(*self.full_size), (C.c_uint32 * 3)(*self.n))

    @property
    def nvox(self):
        return self.n[0] * self.n[1] * self.n[2]

    @classmethod
    def from_str(cls, text):
        """'1 mm  2 mm  0.2 cm  100 100 100' (FOV::from_str, src/fov.rs:75-91)."""
        f = _Fov()
        _check(lib().petalo_fov_parse(text.encode(), C.byref(f)))
        return cls(list(f.full_size_mm), list(f.n))

    def voxel_centre(self, i3):
        c = np.zeros(3, np.float32)
        i = np.asarray(i3, np.uint32)
        _check(lib().petalo_fov_voxel_centre(C.byref(self._c()), _ptr(i), _ptr(c)))
        return c

    def __repr__(self):
        return f"FOV(full_size_mm={self.full_size}, n={self.n})"


class Image:
    """Voxel image on a FOV, x fastest (src/image.rs:13-57)."""

    def __init__(self, fov, data):
        data = np.ascontiguousarray(data, np.float32).reshape(-1)
        if data.size != fov.nvox:
            raise PetaloError(6, f"Image data length {data.size} does not match dimensions "
                                 f"{fov.n[0]}x{fov.n[1]}x{fov.n[2]} = {fov.nvox}")
        self.fov = fov
        self.data = data

    @classmethod
    def ones(cls, fov):
        return cls(fov, np.ones(fov.nvox, np.float32))

    def as_zyx(self):
        return self.data.reshape(self.fov.n[2], self.fov.n[1], self.fov.n[0])


class Osem:
    """(n_subsets, iteration, subset), 1-based (src/mlem.rs:70-96)."""

    def __init__(self, n_subsets, iteration, subset):
        self.n_subsets, self.iteration, self.subset = int(n_subsets), int(iteration), int(subset)

    def __repr__(self):
        return f"Osem(n_subsets={self.n_subsets}, iteration={self.iteration}, subset={self.subset})"


class Tof:
    """[tof] config table: sigma in ps, cutoff in sigmas (default 3)."""

    def __init__(self, sigma_ps, cutoff=3.0):
        self.sigma_ps = float(sigma_ps)
        self.cutoff = float(cutoff)

    def _c(self):
        return _Tof(1, self.sigma_ps, self.cutoff)


def _tof_ptr(tof):
    return C.byref(tof._c()) if tof is not None else None


class Stats(dict):
    pass


def _stats(s):
    return Stats({k: int(getattr(s, k)) for k, _ in _Stats._fields_})


def make_lors(p1, p2, dt=0.0, ac=1.0):
    """LOR records from endpoint arrays (src/lor.rs:22-25)."""
    p1 = np.atleast_2d(np.asarray(p1, np.float32))
    p2 = np.atleast_2d(np.asarray(p2, np.float32))
    out = np.zeros(len(p1), LOR_DTYPE)
    out["p1"], out["p2"], out["dt"], out["ac"] = p1, p2, dt, ac
    return out


# ---------------------------------------------------------------------------
class Mlem:
    """Device-resident MLEM/OSEM state; iterating yields (Image, Osem) per
    subset update, like petalo::mlem::mlem (src/mlem.rs:8-31)."""

    def __init__(self, fov, lors, tof=None, sensitivity=None, n_subsets=1, scheme=SCHEME_AUTO,
                 sort_lors=True, device=0, tile=None, autotune=False, precision=PRECISION_F32, item_steps=0):
        self.fov = fov
        lors = np.ascontiguousarray(lors, LOR_DTYPE)
        sens = None
        if sensitivity is not None:
            sdata = sensitivity.data if isinstance(sensitivity, Image) else sensitivity
            sens = np.ascontiguousarray(sdata, np.float32).reshape(-1)
            if sens.size != fov.nvox:
                raise PetaloError(6, "sensitivity image size does not match the FOV")
        self._ctx = C.c_void_p()
        opt = _Options(int(scheme), 1 if sort_lors else 0, int(device), (C.c_int32 * 3)(*(tile or (0, 0, 0))),
                       1 if autotune else 0, int(precision), int(item_steps))
        _check(lib().petalo_mlem_create(C.byref(self._ctx), C.byref(fov._c()), _tof_ptr(tof),
                                        _ptr(lors), len(lors), _ptr(sens), int(n_subsets),
                                        C.byref(opt)))

    def step(self, copy_image=True):
        """One subset update; returns (Image or None, Osem)."""
        osem = _Osem()
        out = np.empty(self.fov.nvox, np.float32) if copy_image else None
        _check(lib().petalo_mlem_next(self._ctx, _ptr(out), C.byref(osem)))
        img = Image(self.fov, out) if copy_image else None
        return img, Osem(osem.n_subsets, osem.iteration, osem.subset)

    def __iter__(self):
        return self

    def __next__(self):
        return self.step(True)

    def image(self):
        out = np.empty(self.fov.nvox, np.float32)
        _check(lib().petalo_mlem_get_image(self._ctx, _ptr(out)))
        return Image(self.fov, out)

    def image_f64(self):
        """The binary64 image of a binary64-precision context."""
        out = np.empty(self.fov.nvox, np.float64)
        _check(lib().petalo_mlem_get_image_f64(self._ctx, _ptr(out)))
        return out

    def set_image(self, data):
        data = np.ascontiguousarray(data, np.float32).reshape(-1)
        if data.size != self.fov.nvox:
            raise PetaloError(6, "image size does not match the FOV")
        _check(lib().petalo_mlem_set_image(self._ctx, _ptr(data)))

    def tile_shape(self):
        """The tile shape in voxels, (0, 0, 0) for the atomic scheme."""
        t = (C.c_int32 * 3)()
        _check(lib().petalo_mlem_tile_shape(self._ctx, t))
        return tuple(int(v) for v in t)

    def work_items(self):
        """(replayed steps per work item, 0 for the automatic rule; work items over all subsets)."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        _check(lib().petalo_mlem_work_items(self._ctx, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def walk_stats(self, subset=0):
        s = _Stats()
        _check(lib().petalo_mlem_walk_stats(self._ctx, int(subset), C.byref(s)))
        return _stats(s)

    def last_skipped(self):
        v = C.c_uint64(0)
        _check(lib().petalo_mlem_last_skipped(self._ctx, C.byref(v)))
        return int(v.value)

    def accumulator(self):
        """The last tiled update's backprojection accumulator: (fixed-point
        unit of its first work item, 0.0 for binary64 tile sums; largest r;
        geometric mean of the positive r)."""
        u, m, g = C.c_double(0.0), C.c_double(0.0), C.c_double(0.0)
        _check(lib().petalo_mlem_accumulator(self._ctx, C.byref(u), C.byref(m), C.byref(g)))
        return float(u.value), float(m.value), float(g.value)

    def accumulator_unit(self):
        """The last tiled update's fixed-point unit (0.0: binary64 tile sums)."""
        return self.accumulator()[0]

    def set_timing(self, on=True):
        _check(lib().petalo_mlem_set_timing(self._ctx, 1 if on else 0))

    def set_graphs(self, on=True):
        """Replay each subset update as a captured HIP graph (default) or launch directly."""
        _check(lib().petalo_mlem_set_graphs(self._ctx, 1 if on else 0))

    def last_timing(self):
        a, b = C.c_double(0), C.c_double(0)
        _check(lib().petalo_mlem_last_timing(self._ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def last_phase_timing(self):
        a, b, d = C.c_double(0), C.c_double(0), C.c_double(0)
        _check(lib().petalo_mlem_last_phase_timing(self._ctx, C.byref(a), C.byref(b), C.byref(d)))
        return a.value, b.value, d.value

    def last_kernel_timing(self):
        """{'forward_tiles', 'forward_reduce', 'scatter', 'update'} in ms."""
        buf = (C.c_double * 4)()
        _check(lib().petalo_mlem_last_kernel_timing(self._ctx, buf, 4))
        return dict(zip(("forward_tiles", "forward_reduce", "scatter", "update"), list(buf)))

    def drain_kernel_timings(self, max_updates=256):
        """Kernel timings of the timed updates since timing was switched on or
        last drained, oldest first: a list of dicts as last_kernel_timing."""
        buf = (C.c_double * (4 * max_updates))()
        n = C.c_size_t(0)
        _check(lib().petalo_mlem_drain_kernel_timings(self._ctx, buf, max_updates, C.byref(n)))
        names = ("forward_tiles", "forward_reduce", "scatter", "update")
        return [dict(zip(names, list(buf[4 * u:4 * u + 4]))) for u in range(n.value)]

    def info(self):
        sch, nr = C.c_int32(0), C.c_uint64(0)
        _check(lib().petalo_mlem_info(self._ctx, C.byref(sch), C.byref(nr)))
        return {"scheme": int(sch.value), "n_runs": int(nr.value)}

    def close(self):
        if self._ctx:
            lib().petalo_mlem_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent reconstructions advanced together, one subset update of
    each per step(), with one forward and one backprojection launch for all."""

    def __init__(self, fovs, lors_list, tofs=None, sensitivities=None, n_subsets=1, sort_lors=True, device=0,
                 tile=None, autotune=False, item_steps=0):
        n = len(fovs)
        if len(lors_list) != n:
            raise PetaloError(1, "one LOR array per FOV")
        self.fovs = list(fovs)
        self._lors = [np.ascontiguousarray(l, LOR_DTYPE) for l in lors_list]
        cfov = (_Fov * n)(*[f._c() for f in fovs])
        ctof = None
        if tofs is not None:
            ctof = (_Tof * n)()
            for k, t in enumerate(tofs):
                if t is not None:
                    ctof[k] = t._c()
        ptrs = (C.c_void_p * n)(*[l.ctypes.data if len(l) else None for l in self._lors])
        counts = (C.c_size_t * n)(*[len(l) for l in self._lors])
        csens = None
        self._sens = []
        if sensitivities is not None:
            csens = (C.c_void_p * n)()
            for k, sv in enumerate(sensitivities):
                if sv is None:
                    continue
                a = np.ascontiguousarray(sv.data if isinstance(sv, Image) else sv, np.float32).reshape(-1)
                if a.size != fovs[k].nvox:
                    raise PetaloError(6, "sensitivity image size does not match the FOV")
                self._sens.append(a)
                csens[k] = a.ctypes.data
        opt = _Options(SCHEME_TILE, 1 if sort_lors else 0, int(device), (C.c_int32 * 3)(*(tile or (0, 0, 0))),
                       1 if autotune else 0, PRECISION_F32, int(item_steps))
        self._b = C.c_void_p()
        _check(lib().petalo_batch_create(C.byref(self._b), n, cfov, ctof, ptrs, counts, csens, int(n_subsets),
                                         C.byref(opt)))

    def __len__(self):
        return len(self.fovs)

    def step(self):
        osem = _Osem()
        _check(lib().petalo_batch_next(self._b, C.byref(osem)))
        return Osem(osem.n_subsets, osem.iteration, osem.subset)

    def work_items(self):
        """(replayed steps per merged work item, 0 for the automatic rule; work items over all subsets)."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        _check(lib().petalo_batch_work_items(self._b, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def image(self, r):
        out = np.empty(self.fovs[r].nvox, np.float32)
        _check(lib().petalo_batch_get_image(self._b, int(r), _ptr(out)))
        return Image(self.fovs[r], out)

    def set_options(self, timing=False, graphs=True):
        _check(lib().petalo_batch_set_options(self._b, 1 if timing else 0, 1 if graphs else 0))

    def set_schedule(self, persistent=True):
        """persistent: one launch per subset update taking every phase's
        tasks from a queue; False: one launch per phase."""
        _check(lib().petalo_batch_set_schedule(self._b, 1 if persistent else 0))

    def last_ms(self):
        v = C.c_double(0)
        _check(lib().petalo_batch_last_ms(self._b, C.byref(v)))
        return v.value

    def close(self):
        if self._b:
            lib().petalo_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mlem(fov, lors, tof=None, sensitivity=None, n_subsets=1, **options):
    """Infinite iterator of (Image, Osem), one per subset update (src/mlem.rs:8-31)."""
    return Mlem(fov, lors, tof=tof, sensitivity=sensitivity, n_subsets=n_subsets, **options)


# ---------------------------------------------------------------------------
def make_sensitivity_image(attenuation, n_lors, detector_length=None, detector_diameter=None,
                           output_fov=None, lors=None, seed=0, return_stats=False):
    """Sensitivity image (continuous::sensitivity_image, continuous.rs:3-17).

    ``attenuation`` is an Image of mu [1/mm].  With ``lors`` None the LORs are
    generated on the device from (seed, detector_length, detector_diameter).
    ``return_stats`` also counts every step and pushed step, which costs a
    walk over every step of every LOR.
    """
    mu = np.ascontiguousarray(attenuation.data, np.float32)
    out_fov = output_fov or attenuation.fov
    out = np.empty(out_fov.nvox, np.float32)
    st = _Stats()
    lp = None
    if lors is not None:
        lors = np.ascontiguousarray(lors, LOR_DTYPE)
        n_lors = len(lors)
        lp = _ptr(lors)
    L = float(detector_length or 0.0)
    D = float(detector_diameter or 0.0)
    _check(lib().petalo_make_sensitivity_image(
        C.byref(attenuation.fov._c()), _ptr(mu),
        C.byref(output_fov._c()) if output_fov is not None else None,
        lp, int(n_lors), int(seed), L, D, _ptr(out), C.byref(st) if return_stats else None))
    img = Image(out_fov, out)
    return (img, _stats(st)) if return_stats else img


def release_workspaces():
    """Free the device buffers the sensitivity calls keep between calls (one
    set per device, about 10 GB for their 2^25-LOR batches)."""
    _check(lib().petalo_release_workspaces())


def make_sensitivity_image_discrete(attenuation, detector_length, r_min, dr, dz, da,
                                    adjust=ADJUST_RANDOM_Z_PHI, output_fov=None, seed=0, return_stats=False):
    """Sensitivity image of a discretized detector (discrete::sensitivity_image,
    src/bin/make_sensitivity_image/discrete.rs:3-62): all element pairs."""
    att_fov = attenuation.fov
    ofov = output_fov or att_fov
    out = np.zeros(ofov.nvox, np.float32)
    st = _Stats()
    mu = np.ascontiguousarray(attenuation.data, np.float32)
    _check(lib().petalo_make_sensitivity_image_discrete(
        C.byref(att_fov._c()), _ptr(mu), C.byref(output_fov._c()) if output_fov is not None else None,
        float(detector_length), float(r_min), float(dr), float(dz), float(da), int(adjust), int(seed),
        _ptr(out), C.byref(st) if return_stats else None))
    img = Image(ofov, out)
    return (img, _stats(st)) if return_stats else img


# ---------------------------------------------------------------------------
# Figures of merit (src/fom.rs, src/bin/foms.rs)
ROI_SPHERE, ROI_CYLINDER_X, ROI_CYLINDER_Y, ROI_CYLINDER_Z, ROI_DISC_Z = 0, 1, 2, 3, 4
PHANTOM_NEMA7, PHANTOM_JASZCZAK = 0, 1
ROI_DTYPE = np.dtype([("kind", np.int32), ("c", np.float32, 3), ("r", np.float32)])
ROI_STATS_DTYPE = np.dtype([("count", np.uint64), ("mean", np.float64), ("sigma", np.float64)])


def roi(kind, centre, r):
    """One region: ROI::Sphere((x, y, z), r), CylinderX((y, z), r),
    CylinderY((x, z), r), CylinderZ((x, y), r) or DiscZ((x, y, z), r)."""
    c = list(centre) + [0.0] * (3 - len(centre))
    return np.array([(kind, c, r)], ROI_DTYPE)[0]


def roi_stats(image, rois):
    """Per region: voxel count, mean, population sigma (binary64)."""
    rois = np.ascontiguousarray(np.atleast_1d(np.asarray(rois, ROI_DTYPE)))
    out = np.zeros(len(rois), ROI_STATS_DTYPE)
    data = np.ascontiguousarray(image.data, np.float32)
    _check(lib().petalo_roi_stats(C.byref(image.fov._c()), _ptr(data), _ptr(rois), len(rois), _ptr(out)))
    return out


def image_foms(image, rois, activities, background_rois, background_activity):
    """Image::foms (fom.rs:426-450): (crcs %, snrs, background measured)."""
    rois = np.ascontiguousarray(np.atleast_1d(np.asarray(rois, ROI_DTYPE)))
    bg = np.ascontiguousarray(np.atleast_1d(np.asarray(background_rois, ROI_DTYPE)))
    act = np.ascontiguousarray(activities, np.float32)
    if len(act) != len(rois):
        raise PetaloError(1, "one activity per region")
    crcs, snrs, bgm = np.zeros(len(rois), np.float32), np.zeros(len(rois), np.float32), C.c_float(0)
    data = np.ascontiguousarray(image.data, np.float32)
    _check(lib().petalo_image_foms(C.byref(image.fov._c()), _ptr(data), _ptr(rois), _ptr(act), len(rois),
                                   _ptr(bg), len(bg), float(background_activity), _ptr(crcs), _ptr(snrs),
                                   C.byref(bgm)))
    return crcs, snrs, bgm.value


def phantom_foms(image, phantom, aoc_capacity=256):
    """The foms tool's NEMA7 / Jaszczak analysis (src/bin/foms.rs):
    (foms [6 x (r mm, crc %, bg variability %, snr %)], AOCs % (NEMA7))."""
    foms = np.zeros((6, 4), np.float32)
    aocs = np.zeros(aoc_capacity, np.float32)
    na = C.c_size_t(0)
    data = np.ascontiguousarray(image.data, np.float32)
    _check(lib().petalo_phantom_foms(C.byref(image.fov._c()), _ptr(data), int(phantom), _ptr(foms), _ptr(aocs),
                                     aoc_capacity, C.byref(na)))
    return foms, aocs[:min(na.value, aoc_capacity)].copy()


def sensitivity_lors(fov, n_lors, seed, detector_length, detector_diameter):
    """The LORs make_sensitivity_image generates on the device."""
    out = np.zeros(int(n_lors), LOR_DTYPE)
    _check(lib().petalo_sensitivity_lors(C.byref(fov._c()), int(n_lors), int(seed),
                                         float(detector_length), float(detector_diameter), _ptr(out)))
    return out


def density_to_attenuation(image, rho_to_mu=0.095):
    """kg/m^3 -> 1/mm in place (main.rs:80-94); returns the image."""
    _check(lib().petalo_density_to_attenuation(_ptr(image.data), image.data.size, float(rho_to_mu)))
    return image


def system_matrix_rows(fov, lors, tof=None, cap=None):
    """Rows of the system matrix: (idx[n,cap], w[n,cap], count[n], steps[n])."""
    lors = np.ascontiguousarray(lors, LOR_DTYPE)
    n = len(lors)
    if cap is None:
        cap = 2 * sum(fov.n) + 16
    idx = np.zeros((n, cap), np.int32)
    w = np.zeros((n, cap), np.float32)
    count = np.zeros(n, np.uint32)
    steps = np.zeros(n, np.uint32)
    _check(lib().petalo_system_matrix_rows(C.byref(fov._c()), _tof_ptr(tof), _ptr(lors), n, cap,
                                           _ptr(idx), _ptr(w), _ptr(count), _ptr(steps)))
    return idx, w, count, steps


# ---------------------------------------------------------------------------
def synth_phantom_lors(seed, n, tof_fwhm_ps=0.0):
    out = np.zeros(int(n), LOR_DTYPE)
    _check(lib().petalo_synth_phantom_lors(int(seed), int(n), float(tof_fwhm_ps), _ptr(out)))
    return out


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    _check(lib().petalo_philox4x32_10(_ptr(c), _ptr(k), _ptr(out)))
    return out


def synth_density(fov):
    out = np.zeros(fov.nvox, np.float32)
    _check(lib().petalo_synth_density(C.byref(fov._c()), _ptr(out)))
    return Image(fov, out)


def write_image3d(path, image):
    _check(lib().petalo_image3d_write(os.fsencode(path), C.byref(image.fov._c()), _ptr(image.data)))


def read_image3d(path):
    f = _Fov()
    _check(lib().petalo_image3d_read_header(os.fsencode(path), C.byref(f)))
    fov = FOV(list(f.full_size_mm), list(f.n))
    data = np.empty(fov.nvox, np.float32)
    _check(lib().petalo_image3d_read(os.fsencode(path), C.byref(f), _ptr(data), data.size))
    return Image(fov, data)


class Config:
    """Parsed MLEM config (src/config/mlem.rs:73-102)."""

    def __init__(self, c):
        self.input_file = c.input_file.decode()
        self.input_dataset = c.input_dataset.decode()
        self.iterations_number = int(c.iterations_number)
        self.iterations_subsets = int(c.iterations_subsets)
        self.fov = FOV(list(c.fov.full_size_mm), list(c.fov.n))
        self.fov_size = tuple(float(x) for x in c.fov.full_size_mm)
        self.fov_nvoxels = tuple(int(x) for x in c.fov.n)
        self.tof = Tof(c.tof.sigma_ps, c.tof.cutoff) if c.has_tof else None
        self.sensitivity_image = c.sensitivity_image.decode() if c.has_attenuation_correction else None

        def b(has, v):
            return v if has else None
        self.energy = (b(c.has_energy_min, float(c.energy_min)), b(c.has_energy_max, float(c.energy_max)))
        self.charge = (b(c.has_charge_min, float(c.charge_min)), b(c.has_charge_max, float(c.charge_max)))
        self.events = (b(c.has_events_min, int(c.events_min)), b(c.has_events_max, int(c.events_max)))
        self.has_scatter_correction = bool(c.has_scatter_correction)
        g = c.scatter
        self.scatter_correction = ({
            "phi_bins": g.phi_bins, "z_bins": g.z_bins, "z_length_mm": g.z_length_mm, "dz_bins": g.dz_bins,
            "dz_max_mm": g.dz_max_mm, "r_bins": g.r_bins, "r_max_mm": g.r_max_mm, "dt_bins": g.dt_bins,
            "dt_max_ps": g.dt_max_ps} if c.has_scatter_correction else None)
        self.smear_energy_fwhm = float(c.smear_energy_fwhm) if c.has_smear_energy else None
        self.detector_full_axial_length = (float(c.detector_full_axial_length_mm)
                                           if c.has_detector_full_axial_length else None)
        self._c = c

    def ingest_params(self, discretize, seed=0):
        """Ingestion parameters of read_lors for this config (hdf5.rs:97-130,
        240-262); ``discretize`` = the dataset's (r_min, dr, dz, da) in mm."""
        p = _Ingest()
        _check(lib().petalo_config_ingest(C.byref(self._c), *(float(v) for v in discretize), int(seed),
                                          C.byref(p)))
        return p

    def scattergram_spec(self):
        """The [scatter_correction] axes, or None."""
        return self._c.scatter if self.has_scatter_correction else None


def parse_config(text):
    c = _Config()
    _check(lib().petalo_config_parse(text.encode(), C.byref(c)))
    return Config(c)


def read_config(path):
    c = _Config()
    _check(lib().petalo_config_read(os.fsencode(path), C.byref(c)))
    return Config(c)


def parse_quantity(text, kind):
    v = C.c_float(0)
    _check(lib().petalo_parse_quantity(text.encode(), kind.encode(), C.byref(v)))
    return v.value
