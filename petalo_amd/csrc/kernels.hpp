// kernels.hpp -- HIP kernels of the reconstruction hot path (gfx950).
//
// Layout in HBM
//   LORs          32-byte records (Lor), one lane per LOR, read once per
//                 launch as two 16-byte loads: fully coalesced.
//   image         binary32, x fastest (src/index.rs:15-20).
//   backprojection binary32 (MLEM) or binary64 (sensitivity) accumulators of
//                 the same layout, reset by the update kernel.
//
// The MLEM projection is one fused kernel per subset update:
//   walk #1   full reference walk from the FOV entry (bit-exact voxel
//             sequence), gathers image values for pushed elements and forms
//             p = sum w*x in walk order (projector.rs:135-141).  With TOF the
//             walk stops at the first step whose entry distance is already
//             past the Gaussian cutoff: every later weight is exactly 0.
//   walk #2   restarts from the state saved just before the first pushed
//             element and replays exactly the steps of walk #1, adding
//             w * (1/(p*ac)) to the backprojection (projector.rs:144-149).
// A LOR with a pushed index outside the image is dropped whole
// (projector.rs:103-118).  With TOF, walk #1 first crosses every boundary
// before the Gaussian window without computing weights (fast_forward:
// exact, a few operations per boundary).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.hpp"
#include "siddon_walk.hpp"

namespace petalo {

// Counter slots.
// CNT_SHIFTED: walks that start one voxel short of their entry (walk_setup);
// CNT_UNFOUND: generated sensitivity LORs that found no FOV-crossing pair of
// endpoints within kSensMaxAttempts draws.
// CNT_RUN_MISMATCH: LORs whose fill pass of k_bin_runs emitted a different
// number of runs than its count pass counted (never expected; the context
// build fails if it happens).
enum : int {
  CNT_LORS = 0, CNT_STEPS, CNT_NZ, CNT_SKIPPED, CNT_MISSED, CNT_RUNAWAY, CNT_SHIFTED, CNT_UNFOUND, CNT_RUN_MISMATCH,
  CNT_N
};

// Sum a per-lane value over the wavefront; lane 0 adds it to *dst.  Every
// lane of the wavefront must call it.
__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, uint32_t v) {
  unsigned long long s = v;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(dst, s);
}

// Statistics of a subset update's r = 1/(p*ac) values (those of LORs with
// tile runs), for the backprojection's choice of accumulator.  Workgroups of
// the reduction add theirs to one of kRSlots slots of four words (so that no
// address takes more than a few dozen atomics): the binary32 bits of the
// largest finite non-negative r (atomic max), non-zero where some r is
// negative or not finite, the binary64 bits of the sum of log2(r) over the
// finite positive r, and their count (the geometric mean, unlike the mean, is
// not carried by a few outliers).  Word kRSlots * 4 records the unit the
// backprojection's first work item chose (binary64 bits; 0: binary64 sums).
constexpr int kRSlots = 64;
constexpr int kRStatWords = kRSlots * 4 + 1;
constexpr int kRUnitWord = kRSlots * 4;

struct RStats {
  float rmax = 0.0f;  // largest finite non-negative r
  bool bad = false;   // some r negative or not finite
  double lsum = 0.0;  // sum of log2(r) over the finite positive ones
  uint32_t n = 0;     // their count
  __device__ __forceinline__ void note(float r) {
    const bool ok = r >= 0.0f && r <= 3.402823466e38f;
    bad |= !ok;
    if (ok) rmax = fmaxf(rmax, r);
    if (ok && r > 0.0f) {
      lsum += (double)__log2f(r);
      ++n;
    }
  }
  // combines the statistics of the wavefront's lanes into every lane's
  __device__ __forceinline__ void wave_combine() {
    uint32_t m = __float_as_uint(rmax), b = bad ? 1u : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
      b |= (uint32_t)__shfl_xor((int)b, off, 64);
      lsum += __shfl_xor(lsum, off, 64);
      n += (uint32_t)__shfl_xor((int)n, off, 64);
    }
    rmax = __uint_as_float(m);
    bad = b != 0;
  }
};

// Adds the workgroup's statistics to slot blockIdx.x % kRSlots of st (at
// most 16 wavefronts).  Every lane of the workgroup must call it.
__device__ __forceinline__ void block_r_stats(unsigned long long* st, RStats rs) {
  __shared__ float s_max[16];
  __shared__ uint32_t s_bad[16], s_n[16];
  __shared__ double s_lsum[16];
  rs.wave_combine();
  const uint32_t wave = threadIdx.x >> 6, waves = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_max[wave] = rs.rmax;
    s_bad[wave] = rs.bad ? 1u : 0u;
    s_lsum[wave] = rs.lsum;
    s_n[wave] = rs.n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.0f;
    uint32_t b = 0, n = 0;
    double lsum = 0.0;
    for (uint32_t k = 0; k < waves; ++k) {
      m = fmaxf(m, s_max[k]);
      b |= s_bad[k];
      lsum += s_lsum[k];
      n += s_n[k];
    }
    unsigned long long* slot = st + 4 * (blockIdx.x % kRSlots);
    if (m > 0.0f) atomicMax(slot, (unsigned long long)__float_as_uint(m));
    if (b) atomicOr(slot + 1, 1ull);
    if (lsum != 0.0) unsafeAtomicAdd(reinterpret_cast<double*>(slot + 2), lsum);
    if (n) atomicAdd(slot + 3, (unsigned long long)n);
  }
}

__device__ __forceinline__ Lor load_lor(const Lor* __restrict__ lors, uint32_t i) {
  const float4* p = reinterpret_cast<const float4*>(lors + i);
  const float4 a = p[0];
  const float4 b = p[1];
  return Lor{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

// Weight of the step about to be taken (siddon.rs:99-111).
template <bool TOF>
__host__ __device__ __forceinline__ float step_weight(const WalkS& w, const WalkC& c, const Step& s, const DevTof& t) {
  float wt = s.bnd - w.here;
  if (TOF) wt *= tof_factor(t, w.here, c.tof_peak);
  return wt;
}

// ---------------------------------------------------------------------------
// Plain MLEM projection: fused forward + back projection, f32 atomics.
template <bool TOF>
__global__ __launch_bounds__(256) void k_mlem_project(const Lor* __restrict__ lors, uint32_t n, DevFov f,
                                                      DevTof t, const float* __restrict__ img,
                                                      float* __restrict__ bp,
                                                      unsigned long long* __restrict__ counters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t skipped = 0;
  if (i < n) {
    const Lor l = load_lor(lors, i);
    WalkC c;
    WalkS w;
    if (walk_setup(l, f, c, w)) {
      // With every remaining count positive the walk provably ends within
      // nx+ny+nz-2 steps; otherwise it is walked in full against the cap.
      const bool bnd = bounded(w);
      // With TOF, skip exactly to the Gaussian window: no weight before it.
      const bool live = !(TOF && bnd) || fast_forward(w, c, tof_window_start(t, c.tof_peak));
      if (live) {
        const uint32_t cap = step_cap(f);
        WalkS start = w;
        uint32_t steps = 0, start_step = 0;
        bool started = false, bad = false;
        float p = 0.0f;
        for (;;) {
          if (!bnd && steps >= cap) { bad = true; break; }
          if (TOF && bnd && (w.here - c.tof_peak) >= t.cutoff) break;
          const Step s = next_boundary(w);
          const float wt = step_weight<TOF>(w, c, s, t);
          if (wt > 0.0f) {
            if (!started) { start = w; start_step = steps; started = true; }
            if ((uint32_t)w.index >= (uint32_t)f.nvox) { bad = true; break; }
            p += wt * img[w.index];
          }
          ++steps;
          if (advance(w, c, s)) break;
        }
        if (bad) {
          skipped = 1;
        } else if (started) {
          const float r = 1.0f / (p * l.ac);
          w = start;
          for (uint32_t k = start_step; k < steps; ++k) {
            const Step s = next_boundary(w);
            const float wt = step_weight<TOF>(w, c, s, t);
            if (wt > 0.0f) unsafeAtomicAdd(bp + w.index, wt * r);
            advance(w, c, s);
          }
        }
      }
    }
  }
  wave_add_u64(counters + CNT_SKIPPED, skipped);
}

// MLEM update (src/mlem.rs:45-52) fused with the reset of the
// backprojection: x = s > 0 ? x * (b * s) : 0.
__global__ __launch_bounds__(256) void k_update(float* __restrict__ x, float* __restrict__ bp,
                                                const float* __restrict__ s, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float b = bp[i];
  bp[i] = 0.0f;
  const float sj = s ? s[i] : 1.0f;
  x[i] = sj > 0.0f ? x[i] * (b * sj) : 0.0f;
}

__global__ void k_fill_f32(float* __restrict__ x, float v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_fill_f64(double* __restrict__ x, double v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_zero_u64(unsigned long long* __restrict__ x, int n) {
  if ((int)threadIdx.x < n) x[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------------------
// Full reference walk per LOR, counting only (no memory traffic besides the
// LOR): steps, pushed elements, skipped, missed and runaway LORs.
template <bool TOF>
__global__ __launch_bounds__(256) void k_walk_stats(const Lor* __restrict__ lors, uint32_t n, DevFov f, DevTof t,
                                                    unsigned long long* __restrict__ counters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t c_lor = 0, c_steps = 0, c_nz = 0, c_skip = 0, c_miss = 0, c_run = 0, c_shift = 0;
  if (i < n) {
    c_lor = 1;
    const Lor l = load_lor(lors, i);
    WalkC c;
    WalkS w;
    bool sh = false;
    if (!walk_setup(l, f, c, w, &sh)) {
      c_miss = 1;
    } else {
      c_shift = sh ? 1 : 0;
      const uint32_t cap = step_cap(f);
      bool bad = false;
      for (;;) {
        if (c_steps >= cap) { c_run = 1; bad = true; break; }
        ++c_steps;
        const Step s = next_boundary(w);
        const float wt = step_weight<TOF>(w, c, s, t);
        if (wt > 0.0f) {
          ++c_nz;
          if ((uint32_t)w.index >= (uint32_t)f.nvox) bad = true;
        }
        if (advance(w, c, s)) break;
      }
      c_skip = bad ? 1 : 0;
    }
  }
  wave_add_u64(counters + CNT_LORS, c_lor);
  wave_add_u64(counters + CNT_STEPS, c_steps);
  wave_add_u64(counters + CNT_NZ, c_nz);
  wave_add_u64(counters + CNT_SKIPPED, c_skip);
  wave_add_u64(counters + CNT_MISSED, c_miss);
  wave_add_u64(counters + CNT_RUNAWAY, c_run);
  wave_add_u64(counters + CNT_SHIFTED, c_shift);
}

// ---------------------------------------------------------------------------
// System-matrix rows (siddon.rs:69-73): element k of LOR i at [i*cap + k].
template <bool TOF>
__global__ __launch_bounds__(256) void k_rows(const Lor* __restrict__ lors, uint32_t n, DevFov f, DevTof t,
                                              uint32_t cap, int32_t* __restrict__ idx, float* __restrict__ wout,
                                              uint32_t* __restrict__ count, uint32_t* __restrict__ steps_out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Lor l = load_lor(lors, i);
  WalkC c;
  WalkS w;
  uint32_t m = 0, steps = 0;
  if (walk_setup(l, f, c, w)) {
    const uint32_t scap = step_cap(f);
    const size_t base = (size_t)i * cap;
    for (;;) {
      if (steps >= scap) { steps = 0xffffffffu; break; }
      ++steps;
      const Step s = next_boundary(w);
      const float wt = step_weight<TOF>(w, c, s, t);
      if (wt > 0.0f) {
        if (m < cap) {
          idx[base + m] = w.index;
          wout[base + m] = wt;
        }
        ++m;
      }
      if (advance(w, c, s)) break;
    }
  }
  count[i] = m;
  steps_out[i] = steps;
}

// ---------------------------------------------------------------------------
// Sensitivity-image LORs (continuous.rs:20-47): both ends uniform on the
// detector cylinder, z = L*(u - 1/2), theta = TAU*u; rejected until the LOR
// enters the FOV.  Attempt k of LOR i draws Philox block (i, k, 0, tag).  The
// reference draws until it finds one; here a LOR that finds none within
// kSensMaxAttempts draws (a FOV that the detector's LORs almost never cross)
// is counted in CNT_UNFOUND and the call fails.
constexpr uint32_t kSensTag = 0x5e45u;
constexpr uint32_t kSensMaxAttempts = 4096;

__device__ __forceinline__ bool sens_attempt(uint64_t seed, uint32_t i, uint32_t k, const DevFov& f, float L, float R,
                                             Lor& l);

__device__ __forceinline__ Lor sens_lor(uint64_t seed, uint32_t i, const DevFov& f, float L, float R,
                                        bool& found) {
  Lor l;
  for (uint32_t k = 0; k < kSensMaxAttempts; ++k)
    if (sens_attempt(seed, i, k, f, L, R, l)) {
      found = true;
      return l;
    }
  found = false;
  return Lor{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
}

// One draw of sens_lor: attempt k of LOR i.  Returns true when it enters the FOV.
__device__ __forceinline__ bool sens_attempt(uint64_t seed, uint32_t i, uint32_t k, const DevFov& f, float L, float R,
                                             Lor& l) {
  const U32x4 r = philox(seed, i, k, 0, kSensTag);
  const float z1 = L * (unit_f32(r.x) - 0.5f);
  const float t1 = kTau * unit_f32(r.y);
  const float z2 = L * (unit_f32(r.z) - 0.5f);
  const float t2 = kTau * unit_f32(r.w);
  float s1, c1, s2, c2;
  sincosf(t1, &s1, &c1);
  sincosf(t2, &s2, &c2);
  l = Lor{R * c1, R * s1, z1, R * c2, R * s2, z2, 0.0f, 1.0f};
  // most draws miss; a segment that passes the FOV's circumscribed cylinder,
  // or lies beyond its z range, by a clear margin (0.2% of the radius,
  // 0.1% of the half-length: far above the rounding of these few products
  // and of fov_entry's own arithmetic) misses the FOV, and fov_entry would
  // say so
  const float dx = l.x2 - l.x1, dy = l.y2 - l.y1;
  const float cr = l.x1 * dy - l.y1 * dx;  // (distance of the line from the z axis) * |(dx, dy)|
  if (cr * cr > (f.hw_x * f.hw_x + f.hw_y * f.hw_y) * 1.004f * (dx * dx + dy * dy)) return false;
  const float zg = f.hw_z * 1.001f;
  if ((z1 > zg && z2 > zg) || (z1 < -zg && z2 < -zg)) return false;
  float ex, ey, ez;
  return fov_entry(f, l.x1, l.y1, l.z1, l.x2, l.y2, l.z2, ex, ey, ez);
}

// LORs first .. first + n - 1 of the stream (each exactly sens_lor's: attempt
// k of LOR i draws block (i, k, 0, tag), the first hit is kept), written to
// out[0 .. n); LORs without a hit counted in counters[CNT_UNFOUND].  The
// rejection loop takes a geometric number of draws (5.4 on average for a
// 180 mm FOV in a 700 mm x 300 mm detector), so one LOR per lane would keep a
// wavefront for its slowest lane's draws; instead each wavefront takes
// kSensWaveLors consecutive LORs and a lane that has found its LOR takes the
// wavefront's next one, in lane order.  Launch with sens_lors_grid(n) blocks
// of 256.
constexpr uint32_t kSensWaveLors = 1024;
inline unsigned sens_lors_grid(size_t n) {
  const size_t waves = (n + kSensWaveLors - 1) / kSensWaveLors;
  return (unsigned)((waves + 3) / 4);
}

__global__ __launch_bounds__(256) void k_sens_lors(uint32_t n, uint32_t first, uint64_t seed, DevFov f, float L,
                                                   float R, Lor* __restrict__ out,
                                                   unsigned long long* __restrict__ counters) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t c0 = (uint64_t)wave * kSensWaveLors;
  if (c0 >= n) return;  // the whole wavefront
  const uint32_t c1 = (uint32_t)(c0 + kSensWaveLors < n ? c0 + kSensWaveLors : n);
  uint32_t i = (uint32_t)c0 + lane, k = 0, next = (uint32_t)c0 + 64u, unfound = 0;
  bool active = i < c1;
  while (__any(active)) {
    bool done = false;
    if (active) {
      Lor l;
      const bool hit = sens_attempt(seed, first + i, k, f, L, R, l);
      ++k;
      if (hit || k == kSensMaxAttempts) {
        if (!hit) {
          l = Lor{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
          ++unfound;
        }
        float4* p = reinterpret_cast<float4*>(out + i);
        p[0] = make_float4(l.x1, l.y1, l.z1, l.x2);
        p[1] = make_float4(l.y2, l.z2, l.dt, l.ac);
        done = true;
      }
    }
    const unsigned long long fin = __ballot(done);
    if (done) {
      i = next + (uint32_t)__popcll(fin & ((1ull << lane) - 1ull));
      k = 0;
      active = i < c1;
    }
    next += (uint32_t)__popcll(fin);
  }
  wave_add_u64(counters + CNT_UNFOUND, unfound);
}

// Sensitivity backprojection (project_one_lor_sens, projector.rs:72-74):
// forward walk on the attenuation FOV, p = sum w*mu, adapted to exp(-p);
// backward walk on the output FOV (TWO) or the same row, adding
// w * (1/exp(-p)) into binary64 accumulators.  mu == nullptr stands for an
// attenuation image of zeros (p = 0; the forward walk still counts).
template <bool GEN, bool TWO>
__global__ __launch_bounds__(256) void k_sens(const Lor* __restrict__ lors, uint32_t n, uint64_t seed, float L,
                                              float R, DevFov fa, const float* __restrict__ mu, DevFov fo,
                                              double* __restrict__ acc, unsigned long long* __restrict__ counters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t c_lor = 0, c_steps = 0, c_nz = 0, c_skip = 0, c_miss = 0, c_run = 0, c_shift = 0, c_unfound = 0;
  const DevTof notof = no_tof();
  if (i < n) {
    c_lor = 1;
    Lor l;
    if (GEN) {
      bool found;
      l = sens_lor(seed, i, fa, L, R, found);
      c_unfound = found ? 0u : 1u;
    } else {
      l = load_lor(lors, i);
    }
    WalkC c;
    WalkS w;
    bool sh = false;
    const bool hit = walk_setup(l, fa, c, w, &sh);
    c_shift = (hit && sh) ? 1u : 0u;
    bool bad = false;
    float p = 0.0f;
    WalkS start = w;
    uint32_t start_step = 0;
    bool started = false;
    if (!hit) {
      c_miss = 1;
    } else {
      const uint32_t cap = step_cap(fa);
      for (;;) {
        if (c_steps >= cap) { c_run = 1; bad = true; break; }
        const Step s = next_boundary(w);
        const float wt = step_weight<false>(w, c, s, notof);
        if (wt > 0.0f) {
          if (!started) { start = w; start_step = c_steps; started = true; }
          ++c_nz;
          if ((uint32_t)w.index >= (uint32_t)fa.nvox) { bad = true; break; }
          if (mu) p += wt * mu[w.index];  // mu == nullptr: no attenuation (p stays 0)
        }
        ++c_steps;
        if (advance(w, c, s)) break;
      }
    }
    const float r = 1.0f / expf(-p);
    if (!TWO) {
      if (bad) {
        c_skip = 1;
      } else if (started) {
        w = start;
        for (uint32_t k = start_step; k < c_steps; ++k) {
          const Step s = next_boundary(w);
          const float wt = step_weight<false>(w, c, s, notof);
          if (wt > 0.0f) unsafeAtomicAdd(acc + w.index, (double)(wt * r));
          advance(w, c, s);
        }
      }
    } else {
      WalkC cb;
      WalkS wb;
      if (!bad && walk_setup(l, fo, cb, wb)) {
        // check pass over the backward row, then the scatter pass
        const uint32_t cap = step_cap(fo);
        const WalkS b0 = wb;
        uint32_t nb = 0;
        for (;;) {
          if (nb >= cap) { bad = true; break; }
          const Step s = next_boundary(wb);
          const float wt = step_weight<false>(wb, cb, s, notof);
          ++nb;
          if (wt > 0.0f && (uint32_t)wb.index >= (uint32_t)fo.nvox) { bad = true; break; }
          if (advance(wb, cb, s)) break;
        }
        if (!bad) {
          wb = b0;
          for (uint32_t k = 0; k < nb; ++k) {
            const Step s = next_boundary(wb);
            const float wt = step_weight<false>(wb, cb, s, notof);
            if (wt > 0.0f) unsafeAtomicAdd(acc + wb.index, (double)(wt * r));
            advance(wb, cb, s);
          }
        }
      }
      c_skip = bad ? 1 : 0;
    }
  }
  wave_add_u64(counters + CNT_LORS, c_lor);
  wave_add_u64(counters + CNT_STEPS, c_steps);
  wave_add_u64(counters + CNT_NZ, c_nz);
  wave_add_u64(counters + CNT_SKIPPED, c_skip);
  wave_add_u64(counters + CNT_MISSED, c_miss);
  wave_add_u64(counters + CNT_RUNAWAY, c_run);
  wave_add_u64(counters + CNT_SHIFTED, c_shift);
  wave_add_u64(counters + CNT_UNFOUND, c_unfound);
}

// normalize (main.rs:77): out = (float)acc / (float)n.
__global__ void k_sens_finalize(const double* __restrict__ acc, float* __restrict__ out, float nl, size_t nv) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) out[i] = (float)acc[i] / nl;
}

}  // namespace petalo
