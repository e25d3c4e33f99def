// hip_util.hpp -- error handling, launch sizes and device buffers shared by
// the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

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

inline void require_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw HipError("no HIP device available");
}

}  // namespace petalo
