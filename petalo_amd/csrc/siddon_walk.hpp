// siddon_walk.hpp -- device-side Siddon setup and voxel walk for gfx950.
//
// Semantics follow the reference walk exactly (src/projectors/siddon.rs):
// the same binary32 operations in the same order, so that the voxel-index
// sequence of every LOR and every geometric weight is bit-identical to the
// reference's.  This translation unit must be compiled with
// -ffp-contract=off (no FMA contraction) and with hipcc's default correctly
// rounded binary32 division and square root.
//
// The state of a walk lives in scalars (no dynamically indexed arrays) so
// that it stays in VGPRs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>

namespace petalo {

// mm/ps, units/src/lib.rs:156 (binary32 rounds it to 0.29979247).
constexpr float kC = 0.299792458f;
// siddon.rs:134: entry-point components closer than this to 0 become 0.
constexpr float kEntryEps = 1e-5f;
// siddon.rs:106-108: factor applied to the TOF Gaussian.
constexpr float kTofFactor = 666.0f;
// std::f32::consts::TAU
constexpr float kTau = 6.2831855f;

struct DevFov {
  float hw_x, hw_y, hw_z;  // half widths
  float vs_x, vs_y, vs_z;  // voxel sizes
  int32_t nx, ny, nz;
  int32_t nvox;
};

struct DevTof {
  float sigma;      // mm (sigma_t * C)
  float cutoff;     // mm (cutoff * sigma)
  float peak;       // 1 / (sigma * sqrt(TAU))
  float rcp_sigma;  // 1 / sigma, correctly rounded
  int32_t on;
  // Replays of the tiled scheme: gauss(dx) * 666 = amp * exp2(-(dx * g_scale)^2)
  float g_scale;    // sqrt(log2(e) / 2) / sigma
  float amp;        // 666 * peak
};

// No TOF: the window test never applies.
__host__ __device__ inline DevTof no_tof() { return DevTof{1.0f, INFINITY, 0.0f, 1.0f, 0, 0.0f, 1.0f}; }

struct Lor {
  float x1, y1, z1, x2, y2, z2, dt, ac;
};

// Walk of one LOR, split into the part fixed at setup (WalkC) and the part
// that moves (WalkS).  Kernels hold WalkC in locals that are only read, so
// per-axis selects stay register selects (selecting between fields of a
// struct that lives in memory makes the compiler index it as an array).
struct WalkC {
  float vl_x, vl_y, vl_z;  // voxel size in LOR-distance units
  int32_t d_x, d_y, d_z;   // index change per axis
  float tof_peak;          // distance from the entry to the TOF peak
};

struct WalkS {
  float nb_x, nb_y, nb_z;  // distance along the LOR to the next boundary
  float here;              // distance travelled since the FOV entry
  int32_t index;           // current 1-D voxel index
  int32_t r_x, r_y, r_z;   // voxels remaining per axis
};

// nalgebra 0.30.1 Vector3::norm = sqrt((x*x + y*y) + z*z).
__host__ __device__ __forceinline__ float norm3(float x, float y, float z) {
  return sqrtf((x * x + y * y) + z * z);
}

// FOV::new (src/fov.rs:22-39).
inline DevFov make_fov(const float full[3], const uint32_t n[3]) {
  DevFov f;
  f.hw_x = full[0] / 2.0f;
  f.hw_y = full[1] / 2.0f;
  f.hw_z = full[2] / 2.0f;
  f.vs_x = (f.hw_x * 2.0f) / (float)n[0];
  f.vs_y = (f.hw_y * 2.0f) / (float)n[1];
  f.vs_z = (f.hw_z * 2.0f) / (float)n[2];
  f.nx = (int32_t)n[0];
  f.ny = (int32_t)n[1];
  f.nz = (int32_t)n[2];
  f.nvox = f.nx * f.ny * f.nz;
  return f;
}

// Gaussian::new (src/gauss.rs:31-45, 58-60).
inline DevTof make_tof(bool on, float sigma_ps, float cutoff) {
  DevTof t;
  t.on = on ? 1 : 0;
  t.sigma = sigma_ps * kC;
  const float root_two_pi = sqrtf(kTau);
  t.peak = 1.0f / (t.sigma * root_two_pi);
  t.cutoff = cutoff * t.sigma;
  t.rcp_sigma = 1.0f / t.sigma;
  t.g_scale = (float)(std::sqrt(0.5 * 1.4426950408889634) / (double)t.sigma);
  t.amp = kTofFactor * t.peak;
  return t;
}

// a / b correctly rounded, from the correctly rounded reciprocal of b:
// q0 = a*rcp is within one ulp of a/b, the residual a - q0*b is exact in one
// FMA, and q0 + residual*rcp rounds to the correctly rounded quotient
// (Markstein, IBM J. Res. Dev. 34 (1990) 111).  Checked exhaustively against
// binary32 division for every |a| in [2^-40, 2^12) with ten TOF widths.
// Below 2^-40 the quotient may differ in its last bit (the residual
// underflows), but there (a/b)^2 < 2^-76 and the Gaussian evaluates to
// exactly its peak either way.  Three instructions instead of the general
// division sequence.
__host__ __device__ __forceinline__ float div_rn(float a, float b, float rcp) {
  const float q0 = a * rcp;
  const float r = fmaf(-q0, b, a);
  return fmaf(r, rcp, q0);
}

// One slab of the ray/box cast (ncollide3d 0.33.0 AABB cast, called from
// FOV::entry, src/fov.rs:55-71).  Returns false on a miss.
__host__ __device__ __forceinline__ bool slab(float o, float dir, float hw, float& tmin, float& tmax) {
  if (dir == 0.0f) return !(o < -hw || o > hw);
  const float denom = 1.0f / dir;
  float a = (-hw - o) * denom;
  float b = (hw - o) * denom;
  if (a > b) {
    const float t = a;
    a = b;
    b = t;
  }
  tmin = fmaxf(tmin, a);
  tmax = fminf(tmax, b);
  return !(tmin > tmax);
}

// FOV::entry: ray from p1 along normalize(p2 - p1), max_toi = |p2 - p1|,
// solid cast; entry = p1 + dir * toi.
__host__ __device__ __forceinline__ bool fov_entry(const DevFov& f, float x1, float y1, float z1, float x2,
                                                   float y2, float z2, float& ex, float& ey, float& ez) {
  const float dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
  const float len = norm3(dx, dy, dz);
  const float ux = dx / len, uy = dy / len, uz = dz / len;
  float tmin = 0.0f, tmax = len;
  if (!slab(x1, ux, f.hw_x, tmin, tmax)) return false;
  if (!slab(y1, uy, f.hw_y, tmin, tmax)) return false;
  if (!slab(z1, uz, f.hw_z, tmin, tmax)) return false;
  ex = x1 + ux * tmin;
  ey = y1 + uy * tmin;
  ez = z1 + uz * tmin;
  return true;
}

// Rust `x.floor() as usize as i32`: saturating to [0, 2^64-1], then the low
// 32 bits.
__host__ __device__ __forceinline__ int32_t floor_index(float e) {
  const float fl = floorf(e);
  if (!(fl > 0.0f)) return 0;
  if (fl >= 18446744073709551616.0f) return -1;
  if (fl < 2147483648.0f) return (int32_t)fl;
  return (int32_t)(uint32_t)((unsigned long long)fl & 0xffffffffull);
}

// lor_fov_hit (siddon.rs:190-229) with flip_axes (:322-340),
// find_tof_peak (:156-163), find_entry_point (:138-153),
// index_trackers (:233-273), voxel_size (:179-185), first_boundaries
// (:166-174).  Returns false when the LOR misses the FOV.  *shifted (when
// given) is set when an entry coordinate in voxel units is still negative
// after the EPS snap: floor gives -1, which the saturating cast turns into
// voxel 0, so the walk starts one voxel short of its entry along that axis
// (its first step is a sliver) and every later index is one voxel further
// on -- the reference's own behaviour, kept; counted so that it can be
// reported.
// *extra (when given) receives the per-axis start fractions e - floor(e) and
// the entry point in flipped coordinates (for walks that replay these
// decisions in binary64).
struct SetupExtra {
  float frac_x, frac_y, frac_z;
  float ex, ey, ez;
};

__host__ __device__ __forceinline__ bool walk_setup(const Lor& l, const DevFov& f, WalkC& c, WalkS& w,
                                                    bool* shifted = nullptr, SetupExtra* extra = nullptr) {
  float x1 = l.x1, y1 = l.y1, z1 = l.z1, x2 = l.x2, y2 = l.y2, z2 = l.z2;
  const bool fx = (x2 - x1) < 0.0f, fy = (y2 - y1) < 0.0f, fz = (z2 - z1) < 0.0f;
  if (fx) { x1 = -x1; x2 = -x2; }
  if (fy) { y1 = -y1; y2 = -y2; }
  if (fz) { z1 = -z1; z2 = -z2; }

  float ex, ey, ez;
  if (!fov_entry(f, x1, y1, z1, x2, y2, z2, ex, ey, ez)) return false;

  const float half_len = norm3(x1 - x2, y1 - y2, z1 - z2) / 2.0f;
  const float tof_shift = kC * l.dt / 2.0f;
  const float p1_to_peak = half_len - tof_shift;
  const float p1_to_entry = norm3(ex - x1, ey - y1, ez - z1);
  c.tof_peak = p1_to_peak - p1_to_entry;

  float vx = (ex + f.hw_x) / f.vs_x;
  float vy = (ey + f.hw_y) / f.vs_y;
  float vz = (ez + f.hw_z) / f.vs_z;
  if (fabsf(vx) < kEntryEps) vx = 0.0f;
  if (fabsf(vy) < kEntryEps) vy = 0.0f;
  if (fabsf(vz) < kEntryEps) vz = 0.0f;
  if (shifted) *shifted = vx < 0.0f || vy < 0.0f || vz < 0.0f;

  const int32_t ix = floor_index(vx), iy = floor_index(vy), iz = floor_index(vz);
  c.d_x = fx ? -1 : 1;
  c.d_y = fy ? -f.nx : f.nx;
  c.d_z = fz ? -(f.nx * f.ny) : f.nx * f.ny;
  w.r_x = f.nx - ix;
  w.r_y = f.ny - iy;
  w.r_z = f.nz - iz;
  const uint32_t jx = (uint32_t)(fx ? f.nx - 1 - ix : ix);
  const uint32_t jy = (uint32_t)(fy ? f.ny - 1 - iy : iy);
  const uint32_t jz = (uint32_t)(fz ? f.nz - 1 - iz : iz);
  w.index = (int32_t)(jx + (jy + jz * (uint32_t)f.ny) * (uint32_t)f.nx);

  const float dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
  const float len = norm3(dx, dy, dz);
  c.vl_x = f.vs_x / (dx / len);
  c.vl_y = f.vs_y / (dy / len);
  c.vl_z = f.vs_z / (dz / len);

  const float frx = vx - floorf(vx), fry = vy - floorf(vy), frz = vz - floorf(vz);
  w.nb_x = (1.0f - frx) * c.vl_x;
  w.nb_y = (1.0f - fry) * c.vl_y;
  w.nb_z = (1.0f - frz) * c.vl_z;
  w.here = 0.0f;
  if (extra) *extra = SetupExtra{frx, fry, frz, ex, ey, ez};
  return true;
}

// Step cap: a legitimate walk ends within nx+ny+nz-2 steps.  The reference
// never ends an axis-parallel walk that starts exactly on the far face; such
// a walk is stopped here and its LOR dropped.
__host__ __device__ __forceinline__ uint32_t step_cap(const DevFov& f) {
  return (uint32_t)(2 * (f.nx + f.ny + f.nz) + 16);
}

// The boundary crossed next (Vector::argmin, geometry/src/vector.rs:150-158:
// ties go to x, then y) and its position, as one-hot axis flags so that the
// update below is branch-free select arithmetic (a switch on the axis makes
// the compiler spill the per-axis state to a scratch array).
struct Step {
  float bnd;
  bool ax, ay, az;
};

// The minimum of the three and equality tests against it pick the same
// axis as the comparison chain x<=y ? (x<=z ? x : z) : (y<=z ? y : z)
// whenever no NaN is present, and z when all three are NaN (p1 == p2).
// Mixed NaN boundaries do not arise from finite endpoints.
__host__ __device__ __forceinline__ Step next_boundary(const WalkS& w) {
  Step s;
  s.bnd = fminf(fminf(w.nb_x, w.nb_y), w.nb_z);
  s.ax = w.nb_x == s.bnd;
  s.ay = !s.ax && (w.nb_y == s.bnd);
  s.az = !s.ax && !s.ay;
  return s;
}

// exp(x) for the Gaussian, x <= 0.  On the device the hardware exp2 of
// x*log2(e) (a few ulp from the correctly rounded value; the reference's
// f32::exp is the host libm's, so TOF weights are compared within a
// tolerance in any case).
// exp(-z/2) for z = (dx/sigma)^2 >= 0.  On the device this is __expf(-0.5f*z)
// (v_exp_f32 of the argument times log2(e)) with the two constant factors
// merged: -0.5f*z is exact in binary32 wherever z is normal, so
// (-0.5f*z)*log2e and z*(-0.5f*log2e) round the same real number; where z is
// subnormal both arguments are tiny and the result is 1.
__host__ __device__ __forceinline__ float tof_gauss_exp(float z) {
#if defined(__HIP_DEVICE_COMPILE__)
  constexpr float kNegHalfLog2e = -0.5f * 1.44269504088896341f;
  return __builtin_amdgcn_exp2f(z * kNegHalfLog2e);
#else
  return expf(-0.5f * z);
#endif
}

// TOF Gaussian (src/gauss.rs:47-55) times the 666 mm factor (siddon.rs:106-108),
// for a voxel entry distance `here` known to lie inside the window
// |here - tof_peak| < cutoff.
__host__ __device__ __forceinline__ float tof_factor_in_window(const DevTof& t, float here, float tof_peak) {
  const float dx = here - tof_peak;
  const float y = div_rn(dx, t.sigma, t.rcp_sigma);
  const float z = y * y;
  return kTofFactor * (t.peak * tof_gauss_exp(z));
}

// The Gaussian's shape alone, exp(-(dx/sigma)^2 / 2) = exp2(-(dx * g_scale)^2),
// inside the window, for the replays of the tiled scheme: two multiplies and
// one exp2 instead of a correctly rounded division, a square and three
// multiplies; the 666 * peak factor is applied once per run.  Differs from
// tof_factor_in_window / (666 * peak) by a few ulp (the argument is at most
// 9 log2(e) / 2 inside a 3-sigma window).
__host__ __device__ __forceinline__ float tof_shape(const DevTof& t, float here, float tof_peak) {
  const float u = (here - tof_peak) * t.g_scale;
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_exp2f(-(u * u));
#else
  return std::exp2(-(u * u));
#endif
}

// The same, zero outside the window (strict cutoff, gauss.rs:47-55).
__host__ __device__ __forceinline__ float tof_factor(const DevTof& t, float here, float tof_peak) {
  const float dx = here - tof_peak;
  return fabsf(dx) < t.cutoff ? tof_factor_in_window(t, here, tof_peak) : 0.0f;
}

// Advance across the boundary of step s (siddon.rs:118-128): only the
// crossed axis moves.  Returns true when the walk has left the FOV (the
// crossed axis has no voxels left).
__host__ __device__ __forceinline__ bool advance(WalkS& w, const WalkC c, const Step s) {
  w.here = s.bnd;
  w.nb_x = s.ax ? w.nb_x + c.vl_x : w.nb_x;
  w.nb_y = s.ay ? w.nb_y + c.vl_y : w.nb_y;
  w.nb_z = s.az ? w.nb_z + c.vl_z : w.nb_z;
  w.index += s.ax ? +c.d_x : (s.ay ? +c.d_y : +c.d_z);
  w.r_x -= (int32_t)s.ax;
  w.r_y -= (int32_t)s.ay;
  w.r_z -= (int32_t)s.az;
  return (s.ax && w.r_x == 0) || (s.ay && w.r_y == 0) || (s.az && w.r_z == 0);
}

// Crossing without the per-axis remaining counts (for walks bounded by a
// step count, or by the packed counter below).
__host__ __device__ __forceinline__ void move(WalkS& w, const WalkC c, const Step s) {
  w.here = s.bnd;
  w.nb_x = s.ax ? w.nb_x + c.vl_x : w.nb_x;
  w.nb_y = s.ay ? w.nb_y + c.vl_y : w.nb_y;
  w.nb_z = s.az ? w.nb_z + c.vl_z : w.nb_z;
  w.index += s.ax ? +c.d_x : (s.ay ? +c.d_y : +c.d_z);
}

// Crossing of the next boundary bnd = min(nb_x, nb_y, nb_z), for the
// replays of runs (whose step count bounds them): the crossed axis is the one
// next_boundary picks (ties to x, then y; z when no boundary equals bnd) and
// only its boundary and the index move, as in move.  On the device each
// axis's two additions run under an execution mask of the lanes crossing
// that axis, instead of selects on all three axes for the boundaries and two
// for the index step: nine vector instructions per step instead of thirteen.
// The masks come from the same equality tests as next_boundary's, and exec
// is restored before the block ends.
__host__ __device__ __forceinline__ void cross(WalkS& w, const WalkC& c, float bnd) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t my, mz, sv;
  asm("v_cmp_eq_f32 vcc, %[nx], %[b]\n\t"
      "v_cmp_eq_f32 %[my], %[ny], %[b]\n\t"
      "s_mov_b64 %[sv], exec\n\t"
      "s_andn2_b64 %[my], %[my], vcc\n\t"
      "s_or_b64 %[mz], vcc, %[my]\n\t"
      "s_andn2_b64 %[mz], exec, %[mz]\n\t"
      "s_mov_b64 exec, vcc\n\t"
      "v_add_f32 %[nx], %[nx], %[vx]\n\t"
      "v_add_u32 %[i], %[i], %[dx]\n\t"
      "s_mov_b64 exec, %[my]\n\t"
      "v_add_f32 %[ny], %[ny], %[vy]\n\t"
      "v_add_u32 %[i], %[i], %[dy]\n\t"
      "s_mov_b64 exec, %[mz]\n\t"
      "v_add_f32 %[nz], %[nz], %[vz]\n\t"
      "v_add_u32 %[i], %[i], %[dz]\n\t"
      "s_mov_b64 exec, %[sv]"
      : [nx] "+v"(w.nb_x), [ny] "+v"(w.nb_y), [nz] "+v"(w.nb_z), [i] "+v"(w.index), [my] "=&s"(my),
        [mz] "=&s"(mz), [sv] "=&s"(sv)
      : [b] "v"(bnd), [vx] "v"(c.vl_x), [vy] "v"(c.vl_y), [vz] "v"(c.vl_z), [dx] "v"(c.d_x), [dy] "v"(c.d_y),
        [dz] "v"(c.d_z)
      : "vcc");
  w.here = bnd;
#else
  Step s;
  s.bnd = bnd;
  s.ax = w.nb_x == bnd;
  s.ay = !s.ax && (w.nb_y == bnd);
  s.az = !s.ax && !s.ay;
  move(w, c, s);
#endif
}

// The three remaining counts of a bounded walk packed in one word: field d
// (bits 10d .. 10d+8) holds r_d - 1, bit 10d+9 is a guard that is set.
// Crossing axis d subtracts 1 << 10d; the crossing that exhausts the axis
// (field 0) borrows the guard, and the walk has left the FOV exactly when a
// guard is clear.  Valid while every r_d <= 512.
constexpr uint32_t kPackGuards = 0x200u | (0x200u << 10) | (0x200u << 20);
constexpr int32_t kPackMax = 512;

__host__ __device__ __forceinline__ uint32_t pack_remaining(const WalkS& w) {
  return ((uint32_t)(w.r_x - 1) | 0x200u) | (((uint32_t)(w.r_y - 1) | 0x200u) << 10) |
         (((uint32_t)(w.r_z - 1) | 0x200u) << 20);
}

// Crosses step s; returns true when the walk has left the FOV.
__host__ __device__ __forceinline__ bool advance_packed(WalkS& w, const WalkC c, const Step s, uint32_t& rem) {
  move(w, c, s);
  rem -= s.ax ? 1u : (s.ay ? (1u << 10) : (1u << 20));
  return (rem & kPackGuards) != kPackGuards;
}

// v after m additions of d, each rounded to binary32 (round to nearest,
// ties to even): the value the walk's boundary of one axis reaches after m
// crossings of that axis, nb, nb + vl, ..., computed in O(binades crossed)
// instead of O(m), bit for bit.  For v, d > 0 (+inf allowed).  While v lies
// in the binade [2^e, 2^(e+1)) with ulp u = 2^(e-23), write v = V u (V an
// integer, 2^23 <= V < 2^24) and d / u = q + r (q an integer, 0 <= r < 1,
// exact: a power-of-two scaling).  If V + q <= 2^24 - 1, the exact sum
// (V + q + r) u lies below 2^(e+1), where binary32 values are the multiples
// of u, so it rounds to (V + Q) u with Q = q + (r > 1/2) -- for r != 1/2.
// r == 1/2 rounds to the even neighbour: after one hardware addition V is
// even, and from an even V each addition lands on V + q + (q odd), even
// again, so Q = q + (q & 1) from then on.  So k additions jump V to V + kQ
// while the start of the last one, V + (k-1)Q, still satisfies the bound;
// the addition that leaves the binade is done by the hardware in the same
// round, and the next binade repeats the argument.  Values below 8 d (the
// start of a walk, where binades hold few crossings) take hardware additions
// first, which also keeps d / u below 2^24.  The integers V, q, Q and k are
// carried in binary32 (all below 2^24, so exact) and combined by fused
// multiply-adds.
__host__ __device__ __forceinline__ float repeat_add(float v, float d, uint32_t m) {
  if (m == 0u) return v;
  if (!(v < INFINITY && d < INFINITY)) return v + d;
  // below 8 d a binade holds few crossings (the start of a walk): hardware
  // additions, at most 16, so that the rounds below cross few binades
  for (int j = 0; j < 16 && m > 0u && v * 0.125f < d; ++j) {
    v = v + d;
    --m;
  }
  while (m > 0) {
    if (!(v >= d)) {
      v = v + d;
      --m;
      continue;
    }
    uint32_t bits;
    memcpy(&bits, &v, 4);
    const int32_t e = (int32_t)(bits >> 23) - 127;  // v >= d > 0 is normal here
    float Vf = (float)((bits & 0x7fffffu) | 0x800000u);
    const float t = ldexpf(d, 23 - e);  // d / u, exact (or below 1/2 when it underflows)
    const float qf = floorf(t);
    const float r = t - qf;
    float Qf = qf + (r > 0.5f ? 1.0f : 0.0f);
    if (r == 0.5f) {  // a tie: one hardware addition makes V even
      v = v + d;
      --m;
      uint32_t b2;
      memcpy(&b2, &v, 4);
      if (m == 0u || (int32_t)(b2 >> 23) - 127 != e) continue;  // done, or left the binade
      Vf = (float)((b2 & 0x7fffffu) | 0x800000u);
      Qf = qf + (float)((uint32_t)qf & 1u);
    }
    // integers below 2^24 are exact in binary32, and so is every sum or fused
    // product-sum here that stays at or below 2^24; above, the comparisons
    // with 2^24 - 1 still come out right (the rounded value stays above it)
    if (Vf + qf > 16777215.0f) {  // the next addition leaves the binade
      v = v + d;
      --m;
      continue;
    }
    float kf = (float)m;
    if (fmaf(kf - 1.0f, Qf, Vf + qf) > 16777215.0f) {
      // k - 1 = floor((2^24 - 1 - q - V) / Q): the rounded quotient is at
      // most one above it, and never below
      const float num = (16777215.0f - qf) - Vf;
      float kk = floorf(num / Qf);
      if (fmaf(kk, Qf, -num) > 0.0f) kk -= 1.0f;
      kf = kk + 1.0f;
    }
    v = ldexpf(fmaf(kf, Qf, Vf), e - 23);  // V + kQ <= 2^24: exact
    m -= (uint32_t)kf;
    if (m > 0u) {  // the addition that leaves the binade
      v = v + d;
      --m;
    }
  }
  return v;
}

// True when every remaining count is positive: then each step decrements
// one of them and the walk ends within r_x + r_y + r_z - 2 steps, below the
// step cap.
__host__ __device__ __forceinline__ bool bounded(const WalkS& w) { return w.r_x > 0 && w.r_y > 0 && w.r_z > 0; }

// Start of the range of entry distances that may carry TOF weight: any
// step entering its voxel before this distance has |here - peak| >= cutoff
// after binary32 rounding (the margin covers the rounding of here - peak
// for distances below ~8 m), so its weight is exactly 0.
__host__ __device__ __forceinline__ float tof_window_start(const DevTof& t, float tof_peak) {
  return (tof_peak - t.cutoff) - (1e-3f + 2e-6f * (fabsf(tof_peak) + t.cutoff));
}

// Cross, without computing weights, every boundary strictly before T.
// Each axis's boundary positions are the same repeated binary32 additions
// the walk performs, so the state reached is exactly the walk's state after
// its last step whose boundary lies before T (boundaries are crossed in
// position order, so all of them come first).  Returns false when the walk
// leaves the FOV before T (crossing the last boundary of an axis), in which
// case it pushes no element at or after T.  Requires bounded(w).
__host__ __device__ __forceinline__ bool fast_forward(WalkS& w, const WalkC& c, float T) {
  // Each axis on its own: count its boundaries before T (bounded by its
  // remaining voxels, so an axis at -inf cannot loop forever) and keep the
  // last one crossed.
  float nx = w.nb_x, ny = w.nb_y, nz = w.nb_z;
  float hx = w.here, hy = w.here, hz = w.here;
  int32_t kx = 0, ky = 0, kz = 0;
  while (nx < T && kx < w.r_x) { hx = nx; nx += c.vl_x; ++kx; }
  while (ny < T && ky < w.r_y) { hy = ny; ny += c.vl_y; ++ky; }
  while (nz < T && kz < w.r_z) { hz = nz; nz += c.vl_z; ++kz; }
  // Crossing the last boundary of an axis ends the walk.
  if (kx == w.r_x || ky == w.r_y || kz == w.r_z) return false;
  w.nb_x = nx;
  w.nb_y = ny;
  w.nb_z = nz;
  w.here = fmaxf(fmaxf(hx, hy), hz);
  w.r_x -= kx;
  w.r_y -= ky;
  w.r_z -= kz;
  w.index += kx * c.d_x + ky * c.d_y + kz * c.d_z;
  return true;
}

}  // namespace petalo
