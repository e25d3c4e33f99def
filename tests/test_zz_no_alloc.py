"""petalo_sensitivity_image_noalloc against petalo_make_sensitivity_image.

The allocation-free entry point takes device pointers (made here with the HIP
runtime's own hipMalloc, outside the call) and the caller's stream; its
binary64 accumulator lives in the library's static scratch array.  Two calls
in a row on one device -- the second with fewer LORs and a smaller image, so
that the scratch's leftovers from the first would show -- each match the
allocating entry point on the same LORs (binary64 sums in another order,
rounded to binary32: within 1e-6 relative L2, 1e-5 max).  Images beyond the
scratch's 2^21 voxels are refused.
"""
import ctypes as C

import numpy as np
import pytest

import petalo_amd as P
from helpers import max_rel, rel_l2

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        try:
            _hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        except OSError:
            _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return _hip


H2D, D2H = 1, 2


class Dev:
    """A device copy of a numpy array (hipMalloc outside the tested call)."""

    def __init__(self, a):
        self.nbytes = a.nbytes
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(self.nbytes, 1)) == 0
        if self.nbytes:
            assert hip().hipMemcpy(self.p, a.ctypes.data, self.nbytes, H2D) == 0

    def get(self, dtype, n):
        out = np.empty(n, dtype)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


def noalloc(fov, mu, lors, stream):
    f = P.lib().petalo_sensitivity_image_noalloc
    dmu = Dev(np.ascontiguousarray(mu, np.float32))
    dl = Dev(np.ascontiguousarray(lors, P.LOR_DTYPE))
    dout = Dev(np.full(fov.nvox, np.nan, np.float32))
    rc = f(C.byref(fov._c()), dmu.p, dl.p, len(lors), dout.p, stream)
    assert rc == 0, P.lib().petalo_last_error()
    assert hip().hipStreamSynchronize(stream) == 0
    out = dout.get(np.float32, fov.nvox)
    for d in (dmu, dl, dout):
        d.free()
    return out


def test_noalloc_matches_allocating_entry_point(gpu):
    stream = C.c_void_p()
    assert hip().hipStreamCreate(C.byref(stream)) == 0
    for k, (nv, n, seed) in enumerate([(40, 400_000, 3), (24, 50_000, 4)]):
        fov = P.FOV([180.0] * 3, [nv] * 3)
        mu = P.density_to_attenuation(P.synth_density(fov), 0.096)
        lors = P.sensitivity_lors(fov, n, seed, 300.0, 700.0)
        got = noalloc(fov, mu.data, lors, stream)
        ref = P.make_sensitivity_image(mu, n, lors=lors).data
        e, m = rel_l2(got, ref), max_rel(got, ref)
        print(f"call {k + 1}: {nv}^3, {n} LORs: relL2 {e:.2e}, max {m:.2e}", flush=True)
        assert e <= 1e-6 and m <= 1e-5
    hip().hipStreamDestroy(stream)


def test_noalloc_default_stream_and_limits(gpu):
    fov = P.FOV([180.0] * 3, [16] * 3)
    mu = np.zeros(fov.nvox, np.float32)
    lors = P.sensitivity_lors(fov, 20_000, 9, 300.0, 700.0)
    got = noalloc(fov, mu, lors, None)
    ref = P.make_sensitivity_image(P.Image(fov, mu), len(lors), lors=lors).data
    assert rel_l2(got, ref) <= 1e-6
    big = P.FOV([180.0] * 3, [129, 128, 128])  # one voxel column over 2^21 voxels
    f = P.lib().petalo_sensitivity_image_noalloc
    rc = f(C.byref(big._c()), None, None, 0, None, None)
    assert rc != 0  # NULL pointers refused first
    dmu = Dev(np.zeros(big.nvox, np.float32))
    dout = Dev(np.zeros(big.nvox, np.float32))
    rc = f(C.byref(big._c()), dmu.p, None, 0, dout.p, None)
    assert rc == 6 and b"exceeds" in P.lib().petalo_last_error()
    dmu.free()
    dout.free()
