// hip_util.hpp -- error handling, launch sizes and device buffers shared by
// the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/petalo_hip.h"
#include "capi_internal.hpp"

namespace petalo {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& m) : std::runtime_error(m) {}
};

#define HIP_CHECK(expr)                                                                              \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

template <class F> inline int guarded(F&& f) {
  try {
    return f();
  } catch (const HipError& e) {
    return fail(PETALO_ERR_HIP, e.what());
  } catch (const std::bad_alloc&) {
    return fail(PETALO_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(PETALO_ERR_ARG, e.what());
  } catch (...) {
    return fail(PETALO_ERR_ARG, "unknown error");
  }
}

constexpr int kBlock = 256;
inline unsigned grid_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <class T> struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
};

// A FOV the kernels accept: positive finite sizes, voxels on every axis,
// fewer than 2^31 voxels (32-bit voxel indices).
inline bool fov_ok(const petalo_fov_t* f, std::string& why) {
  if (!f) { why = "fov is NULL"; return false; }
  for (int i = 0; i < 3; ++i) {
    if (f->n[i] == 0) { why = "fov has zero voxels along an axis"; return false; }
    if (!(f->full_size_mm[i] > 0.0f) || !std::isfinite(f->full_size_mm[i])) {
      why = "fov size must be positive and finite";
      return false;
    }
  }
  const uint64_t nv = (uint64_t)f->n[0] * f->n[1] * f->n[2];
  if (nv >= (1ull << 31)) { why = "fov has 2^31 or more voxels"; return false; }
  return true;
}

inline void require_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw HipError("no HIP device available");
}

}  // namespace petalo
