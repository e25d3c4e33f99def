// philox.hpp -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable on the
// host and on the device.  Streams are keyed by a 64-bit seed and counted by
// (item index, attempt, lane, tag), so any item can be regenerated anywhere
// without sequential state.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace petalo {

struct U32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    const U32x4 o = {(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
                     (uint32_t)p0};
    c = o;
  }
  return c;
}

__host__ __device__ __forceinline__ U32x4 philox(uint64_t seed, uint32_t item, uint32_t attempt, uint32_t lane,
                                                 uint32_t tag) {
  return philox4x32_10(U32x4{item, attempt, lane, tag}, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Uniform binary32 in [0, 1) from the top 24 bits.
__host__ __device__ __forceinline__ float unit_f32(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

// Uniform binary64 in [0, 1) from 32 bits.
__host__ __device__ __forceinline__ double unit_f64(uint32_t u) { return (double)u * (1.0 / 4294967296.0); }

}  // namespace petalo
