// walk64.hpp -- binary64 walks for the precision modes of a context.
//
//   PETALO_PRECISION_F64         the whole path in binary64: walk setup (FOV
//                                entry, TOF peak, index trackers, first
//                                boundaries), the walk, TOF weights, forward
//                                and back projection, image and update -- the
//                                reference's algorithm (src/projectors/
//                                siddon.rs:81-340, src/gauss.rs:31-60,
//                                src/projector.rs:33-153, src/mlem.rs:45-52)
//                                with every binary32 quantity widened;
//   PETALO_PRECISION_F32_WALK    the binary32 walk (the reference's) decides
//                                everything discrete -- hit, first voxel,
//                                start fraction in it (with the entry quirk
//                                of walk_setup's `shifted`), the crossed axis
//                                of every step, which steps push, where the
//                                walk ends -- and the weights, projections,
//                                image and update are binary64 along that
//                                step sequence: voxel lengths along the LOR
//                                from the binary64 direction, boundaries
//                                accumulated in binary64 from the binary32
//                                start fraction, the binary64 TOF peak and
//                                Gaussian.
// Against the binary32 path the second differs by rounding alone, against
// the first by the binary32 walk's discrete choices alone.  Both run one lane
// per LOR (two walks: forward, then backprojection from the first pushed
// step) with binary64 global atomics; they are reference modes for precision
// studies, not the fast path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "siddon_walk.hpp"

namespace petalo {

constexpr double kC64 = 0.299792458;          // mm/ps (units/src/lib.rs:156)
constexpr double kTau64 = 6.283185307179586;  // TAU
constexpr double kEntryEps64 = 1e-5;          // siddon.rs:134
constexpr double kTofFactor64 = 666.0;        // siddon.rs:107

struct DevFov64 {
  double hw_x, hw_y, hw_z;
  double vs_x, vs_y, vs_z;
  int32_t nx, ny, nz, nvox;
};

// FOV::new in binary64 (src/fov.rs:22-39): half_width = full/2,
// voxel_size = (half_width*2)/n.
inline DevFov64 make_fov64(const float full[3], const uint32_t n[3]) {
  DevFov64 f;
  f.hw_x = (double)full[0] / 2.0;
  f.hw_y = (double)full[1] / 2.0;
  f.hw_z = (double)full[2] / 2.0;
  f.vs_x = (f.hw_x * 2.0) / (double)n[0];
  f.vs_y = (f.hw_y * 2.0) / (double)n[1];
  f.vs_z = (f.hw_z * 2.0) / (double)n[2];
  f.nx = (int32_t)n[0];
  f.ny = (int32_t)n[1];
  f.nz = (int32_t)n[2];
  f.nvox = f.nx * f.ny * f.nz;
  return f;
}

struct DevTof64 {
  double sigma, cutoff, peak;
};

// Gaussian::new in binary64 (src/gauss.rs:31-45,58-60).
inline DevTof64 make_tof64(float sigma_ps, float cutoff) {
  DevTof64 t;
  t.sigma = (double)sigma_ps * kC64;
  t.peak = 1.0 / (t.sigma * sqrt(kTau64));
  t.cutoff = (double)cutoff * t.sigma;
  return t;
}

// Gaussian::call (src/gauss.rs:47-55) times 666 (siddon.rs:106-108).
__device__ __forceinline__ double tof_factor64(const DevTof64& t, double dx) {
  if (!(fabs(dx) < t.cutoff)) return 0.0;
  const double y = dx / t.sigma;
  return kTofFactor64 * (t.peak * exp(-0.5 * (y * y)));
}

__device__ __forceinline__ double norm3_64(double x, double y, double z) { return sqrt((x * x + y * y) + z * z); }

// Walk state in binary64.
struct Walk64 {
  double nb_x, nb_y, nb_z, here;
  double vl_x, vl_y, vl_z, tof_peak;
  int32_t index, d_x, d_y, d_z;
  int32_t r_x, r_y, r_z;
};

// One slab of the ray/box cast in binary64 (the restated ncollide3d cast).
__device__ __forceinline__ bool slab64(double o, double dir, double hw, double& tmin, double& tmax) {
  if (dir == 0.0) return !(o < -hw || o > hw);
  const double denom = 1.0 / dir;
  double a = (-hw - o) * denom;
  double b = (hw - o) * denom;
  if (a > b) {
    const double t = a;
    a = b;
    b = t;
  }
  tmin = fmax(tmin, a);
  tmax = fmin(tmax, b);
  return !(tmin > tmax);
}

// FOV::entry in binary64 on flipped endpoints.
__device__ __forceinline__ bool fov_entry64(const DevFov64& f, const double p1[3], const double p2[3], double e[3]) {
  const double dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
  const double len = norm3_64(dx, dy, dz);
  const double ux = dx / len, uy = dy / len, uz = dz / len;
  double tmin = 0.0, tmax = len;
  if (!slab64(p1[0], ux, f.hw_x, tmin, tmax)) return false;
  if (!slab64(p1[1], uy, f.hw_y, tmin, tmax)) return false;
  if (!slab64(p1[2], uz, f.hw_z, tmin, tmax)) return false;
  e[0] = p1[0] + ux * tmin;
  e[1] = p1[1] + uy * tmin;
  e[2] = p1[2] + uz * tmin;
  return true;
}

__device__ __forceinline__ int32_t floor_index64(double e) {
  const double fl = floor(e);
  if (!(fl > 0.0)) return 0;
  if (fl >= 18446744073709551616.0) return -1;
  if (fl < 2147483648.0) return (int32_t)fl;
  return (int32_t)(uint32_t)((unsigned long long)fl & 0xffffffffull);
}

// lor_fov_hit (siddon.rs:190-229) in binary64.
__device__ __forceinline__ bool walk_setup64(const Lor& l, const DevFov64& f, Walk64& w) {
  double p1[3] = {l.x1, l.y1, l.z1}, p2[3] = {l.x2, l.y2, l.z2};
  const bool fx = (p2[0] - p1[0]) < 0.0, fy = (p2[1] - p1[1]) < 0.0, fz = (p2[2] - p1[2]) < 0.0;
  if (fx) { p1[0] = -p1[0]; p2[0] = -p2[0]; }
  if (fy) { p1[1] = -p1[1]; p2[1] = -p2[1]; }
  if (fz) { p1[2] = -p1[2]; p2[2] = -p2[2]; }
  double e[3];
  if (!fov_entry64(f, p1, p2, e)) return false;
  const double half = norm3_64(p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]) / 2.0;
  const double shift = kC64 * (double)l.dt / 2.0;
  w.tof_peak = (half - shift) - norm3_64(e[0] - p1[0], e[1] - p1[1], e[2] - p1[2]);
  double vx = (e[0] + f.hw_x) / f.vs_x, vy = (e[1] + f.hw_y) / f.vs_y, vz = (e[2] + f.hw_z) / f.vs_z;
  if (fabs(vx) < kEntryEps64) vx = 0.0;
  if (fabs(vy) < kEntryEps64) vy = 0.0;
  if (fabs(vz) < kEntryEps64) vz = 0.0;
  const int32_t ix = floor_index64(vx), iy = floor_index64(vy), iz = floor_index64(vz);
  w.d_x = fx ? -1 : 1;
  w.d_y = fy ? -f.nx : f.nx;
  w.d_z = fz ? -(f.nx * f.ny) : f.nx * f.ny;
  w.r_x = f.nx - ix;
  w.r_y = f.ny - iy;
  w.r_z = f.nz - iz;
  const uint32_t jx = (uint32_t)(fx ? f.nx - 1 - ix : ix);
  const uint32_t jy = (uint32_t)(fy ? f.ny - 1 - iy : iy);
  const uint32_t jz = (uint32_t)(fz ? f.nz - 1 - iz : iz);
  w.index = (int32_t)(jx + (jy + jz * (uint32_t)f.ny) * (uint32_t)f.nx);
  const double dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
  const double len = norm3_64(dx, dy, dz);
  w.vl_x = f.vs_x / (dx / len);
  w.vl_y = f.vs_y / (dy / len);
  w.vl_z = f.vs_z / (dz / len);
  w.nb_x = (1.0 - (vx - floor(vx))) * w.vl_x;
  w.nb_y = (1.0 - (vy - floor(vy))) * w.vl_y;
  w.nb_z = (1.0 - (vz - floor(vz))) * w.vl_z;
  w.here = 0.0;
  return true;
}

// The binary64 arithmetic of a binary32 walk: lengths along the LOR from the
// binary64 direction of the same flipped endpoints, boundaries from the
// binary32 start fractions, the binary64 TOF peak from the binary64 entry
// point (the binary32 one where the binary64 cast misses a grazing LOR).
__device__ __forceinline__ void mixed_setup64(const Lor& l, const DevFov64& f, const WalkC& c, const WalkS& w32,
                                              const SetupExtra& x, Walk64& w) {
  double p1[3] = {l.x1, l.y1, l.z1}, p2[3] = {l.x2, l.y2, l.z2};
  if (c.d_x < 0) { p1[0] = -p1[0]; p2[0] = -p2[0]; }
  if (c.d_y < 0) { p1[1] = -p1[1]; p2[1] = -p2[1]; }
  if (c.d_z < 0) { p1[2] = -p1[2]; p2[2] = -p2[2]; }
  const double dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
  const double len = norm3_64(dx, dy, dz);
  w.vl_x = f.vs_x / (dx / len);
  w.vl_y = f.vs_y / (dy / len);
  w.vl_z = f.vs_z / (dz / len);
  w.nb_x = (1.0 - (double)x.frac_x) * w.vl_x;
  w.nb_y = (1.0 - (double)x.frac_y) * w.vl_y;
  w.nb_z = (1.0 - (double)x.frac_z) * w.vl_z;
  w.here = 0.0;
  double e[3];
  if (!fov_entry64(f, p1, p2, e)) {
    e[0] = x.ex;
    e[1] = x.ey;
    e[2] = x.ez;
  }
  const double half = norm3_64(p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]) / 2.0;
  w.tof_peak = (half - kC64 * (double)l.dt / 2.0) - norm3_64(e[0] - p1[0], e[1] - p1[1], e[2] - p1[2]);
  w.index = w32.index;
  w.d_x = c.d_x;
  w.d_y = c.d_y;
  w.d_z = c.d_z;
  w.r_x = w32.r_x;
  w.r_y = w32.r_y;
  w.r_z = w32.r_z;
}

// One step of the binary64 walk: argmin with ties x < y < z
// (geometry/src/vector.rs:150-158); returns the axis (0, 1, 2).
__device__ __forceinline__ int argmin64(const Walk64& w) {
  if (w.nb_x <= w.nb_y) return w.nb_x <= w.nb_z ? 0 : 2;
  return w.nb_y <= w.nb_z ? 1 : 2;
}

// Crosses axis a: returns true when the walk leaves the FOV.
__device__ __forceinline__ bool advance64(Walk64& w, int a) {
  const double bnd = a == 0 ? w.nb_x : (a == 1 ? w.nb_y : w.nb_z);
  w.here = bnd;
  if (a == 0) { w.nb_x += w.vl_x; w.index += w.d_x; return --w.r_x == 0; }
  if (a == 1) { w.nb_y += w.vl_y; w.index += w.d_y; return --w.r_y == 0; }
  w.nb_z += w.vl_z;
  w.index += w.d_z;
  return --w.r_z == 0;
}

__device__ __forceinline__ double bnd64(const Walk64& w, int a) { return a == 0 ? w.nb_x : (a == 1 ? w.nb_y : w.nb_z); }

// The binary64 weight of the step across axis a (siddon.rs:99-111).
template <bool TOF>
__device__ __forceinline__ double weight64(const Walk64& w, int a, const DevTof64& t) {
  double wt = bnd64(w, a) - w.here;
  if (TOF) wt *= tof_factor64(t, w.here - w.tof_peak);
  return wt;
}

// Both precision modes, one lane per LOR: walk #1 forms p = sum w * x in walk
// order (binary64), the LOR is dropped whole when a pushed index is outside
// the image or the walk runs into the step cap (projector.rs:103-118), then
// walk #2 replays from the first pushed step adding w * r, r = 1/(p * ac),
// into the binary64 backprojection.  MIXED: the binary32 walk (with its
// TOF window) decides the steps; the binary64 state follows it.
template <bool TOF, bool MIXED>
__global__ __launch_bounds__(256) void k_mlem_project64(const Lor* __restrict__ lors, uint32_t n, DevFov f32f,
                                                        DevTof t32, DevFov64 f, DevTof64 t,
                                                        const double* __restrict__ img, double* __restrict__ bp,
                                                        unsigned long long* __restrict__ counters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t skipped = 0;
  if (i < n) {
    const Lor l = load_lor(lors, i);
    WalkC c;
    WalkS w32;
    Walk64 w;
    bool hit;
    if (MIXED) {
      SetupExtra x;
      hit = walk_setup(l, f32f, c, w32, nullptr, &x);
      if (hit) mixed_setup64(l, f, c, w32, x, w);
    } else {
      hit = walk_setup64(l, f, w);
    }
    if (hit) {
      // the TOF exit below is taken only by walks that provably end (every
      // axis has voxels left); others are walked to the cap and dropped
      const bool bnd = MIXED ? bounded(w32) : (w.r_x > 0 && w.r_y > 0 && w.r_z > 0);
      const uint32_t cap = step_cap(f32f);
      Walk64 start = w;
      WalkS start32 = w32;
      uint32_t steps = 0, start_step = 0;
      bool started = false, bad = false;
      double p = 0.0;
      for (;;) {
        if (steps >= cap) { bad = true; break; }
        int a;
        bool push;
        double wt;
        if (MIXED) {
          if (TOF && bnd && (w32.here - c.tof_peak) >= t32.cutoff) break;  // every later binary32 weight is 0
          const Step s = next_boundary(w32);
          push = step_weight<TOF>(w32, c, s, t32) > 0.0f;
          a = s.ax ? 0 : (s.ay ? 1 : 2);
          wt = weight64<TOF>(w, a, t);
          if (push) {
            if (!started) { start = w; start32 = w32; start_step = steps; started = true; }
            if ((uint32_t)w.index >= (uint32_t)f.nvox) { bad = true; break; }
            p += wt * img[w.index];
          }
          ++steps;
          advance64(w, a);
          if (advance(w32, c, s)) break;
        } else {
          if (TOF && bnd && (w.here - w.tof_peak) >= t.cutoff) break;  // every later weight is 0
          a = argmin64(w);
          wt = weight64<TOF>(w, a, t);
          push = wt > 0.0;
          if (push) {
            if (!started) { start = w; start_step = steps; started = true; }
            if ((uint32_t)w.index >= (uint32_t)f.nvox) { bad = true; break; }
            p += wt * img[w.index];
          }
          ++steps;
          if (advance64(w, a)) break;
        }
      }
      if (bad) {
        skipped = 1;
      } else if (started) {
        const double r = 1.0 / (p * (double)l.ac);
        w = start;
        w32 = start32;
        for (uint32_t k = start_step; k < steps; ++k) {
          if (MIXED) {
            const Step s = next_boundary(w32);
            const bool push = step_weight<TOF>(w32, c, s, t32) > 0.0f;
            const int a = s.ax ? 0 : (s.ay ? 1 : 2);
            if (push) unsafeAtomicAdd(bp + w.index, weight64<TOF>(w, a, t) * r);
            advance64(w, a);
            advance(w32, c, s);
          } else {
            const int a = argmin64(w);
            const double wt = weight64<TOF>(w, a, t);
            if (wt > 0.0) unsafeAtomicAdd(bp + w.index, wt * r);
            advance64(w, a);
          }
        }
      }
    }
  }
  wave_add_u64(counters + CNT_SKIPPED, skipped);
}

// MLEM update in binary64 (src/mlem.rs:45-52), resetting the accumulator:
// x = s > 0 ? x * (b * s) : 0.
__global__ __launch_bounds__(256) void k_update64(double* __restrict__ x, double* __restrict__ bp,
                                                  const double* __restrict__ s, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double b = bp[i];
  bp[i] = 0.0;
  const double sj = s ? s[i] : 1.0;
  x[i] = sj > 0.0 ? x[i] * (b * sj) : 0.0;
}

// binary64 image -> binary32 (rounded to nearest).
__global__ void k_narrow(const double* __restrict__ x, float* __restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

__global__ void k_widen(const float* __restrict__ x, double* __restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}

}  // namespace petalo
