// ingest.hpp -- LOR ingestion: cuts, energy and position smearing,
// scattergram binning and conversion of reco_info/lors records into MLEM
// LORs.  Host and device functions; binary32 throughout, in the reference's
// operation order.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "philox.hpp"
#include "siddon_walk.hpp"

namespace petalo {

// One reco_info/lors record (Hdf5Lor, src/io/hdf5.rs:282-297).
struct H5Lor {
  float dt;                  // ns
  float x1, y1, z1, x2, y2, z2;  // mm
  float q1, q2;              // charges
  float e1, e2;              // keV
};

// [input] bounds: min <= v < max, each side optional (config/mlem.rs:141-146).
struct BoundsF {
  bool has_min, has_max;
  float min, max;
  __host__ __device__ bool contains(float v) const { return (!has_min || min <= v) && (!has_max || max > v); }
};

// Element geometry derived from (r_min, dr, dz, da) (discrete.rs:170-187).
struct DevDiscretize {
  float dr, dz;
  float n_radial;     // (r_min + dr/2) / dr
  float d_azimuthal;  // 2 pi / round(2 pi (r_min + dr/2) / da)
  int32_t adjust;     // 0 none, 1 centre, 2 random, 3 random z and phi
};

struct DevIngest {
  BoundsF charge, energy;
  float fwhm;
  float pre_cut;      // 511 * (1 - 1.6 fwhm) (hdf5.rs:108)
  bool has_z_max;
  float z_max;
  DevDiscretize d;
  uint64_t seed;
};

constexpr float kTwoPi = 6.2831855f;  // units::TWOPI, std::f32::consts::TAU

__host__ __device__ inline DevDiscretize make_discretize(float r_min, float dr, float dz, float da, int32_t adjust) {
  DevDiscretize d;
  d.dr = dr;
  d.dz = dz;
  const float r = r_min + dr / 2.0f;
  d.n_radial = r / dr;
  const float mid_circumference = kTwoPi * r;
  const float n_azimuthal = roundf(mid_circumference / da);
  d.d_azimuthal = kTwoPi / n_azimuthal;
  d.adjust = adjust;
  return d;
}

// smear(x) = x - 0.5 + U[0, 1) (discrete.rs:213).
__host__ __device__ inline float smear_unit(float x, float u) { return x - 0.5f + u; }

// Discretize::make_adjust_fn (discrete.rs:78-118) with the uniforms u[0..2]
// (z, phi, r) of this end.
__host__ __device__ inline void adjust_point(const DevDiscretize& d, float& x, float& y, float& z, const float u[3]) {
  if (d.adjust == 0) return;
  // cell_indices (discrete.rs:60-66)
  const float original_phi = atan2f(y, x);
  const int32_t n_phi = (int32_t)roundf(original_phi / d.d_azimuthal);
  const int32_t n_z = (int32_t)roundf(z / d.dz);
  float r, phi;
  if (d.adjust == 1) {
    r = d.n_radial * d.dr;
    z = (float)n_z * d.dz;
    phi = (float)n_phi * d.d_azimuthal;
  } else if (d.adjust == 2) {
    z = smear_unit((float)n_z, u[0]) * d.dz;
    phi = smear_unit((float)n_phi, u[1]) * d.d_azimuthal;
    r = smear_unit(d.n_radial, u[2]) * d.dr;
  } else {
    z = smear_unit((float)n_z, u[0]) * d.dz;
    phi = smear_unit((float)n_phi, u[1]) * d.d_azimuthal;
    r = hypotf(x, y);
  }
  x = r * cosf(phi);
  y = r * sinf(phi);
}

constexpr uint32_t kIngestTagEnergy = 0xe6e7u;
constexpr uint32_t kIngestTagPosition = 0x9051u;

// Standard normal from two uniforms in [0, 1) (Box-Muller; 1 - u1 is in (0, 1]).
__host__ __device__ inline float std_normal(float u1, float u2) {
  return sqrtf(-2.0f * logf(1.0f - u1)) * cosf(kTwoPi * u2);
}

// Energy smearing of make_smear_energy (hdf5.rs:230-238): Normal(e, fwhm e / 2.35).
__host__ __device__ inline float smear_energy(float e, float fwhm, float z) {
  const float f = fwhm * e;
  const float sigma = f / 2.35f;
  return e + sigma * z;
}

// Cuts and smearing of record `index` (hdf5.rs:113-127, 240-265); false when
// the record is cut.
__host__ __device__ inline bool ingest_record(const DevIngest& p, uint64_t index, H5Lor& l) {
  if (p.has_z_max && !(fabsf(l.z1) < p.z_max && fabsf(l.z2) < p.z_max)) return false;
  if (!(p.charge.contains(l.q1) && p.charge.contains(l.q2))) return false;
  if (!(p.pre_cut < l.e1 && p.pre_cut < l.e2)) return false;
  const uint32_t lo = (uint32_t)index, hi = (uint32_t)(index >> 32);
  const U32x4 re = philox(p.seed, lo, hi, 0, kIngestTagEnergy);
  l.e1 = smear_energy(l.e1, p.fwhm, std_normal(unit_f32(re.x), unit_f32(re.y)));
  l.e2 = smear_energy(l.e2, p.fwhm, std_normal(unit_f32(re.z), unit_f32(re.w)));
  if (!(p.energy.contains(l.e1) && p.energy.contains(l.e2))) return false;
  const U32x4 a = philox(p.seed, lo, hi, 1, kIngestTagPosition);
  const U32x4 b = philox(p.seed, lo, hi, 2, kIngestTagPosition);
  const float u1[3] = {unit_f32(a.x), unit_f32(a.y), unit_f32(b.x)};
  const float u2[3] = {unit_f32(a.z), unit_f32(a.w), unit_f32(b.y)};
  adjust_point(p.d, l.x1, l.y1, l.z1, u1);
  adjust_point(p.d, l.x2, l.y2, l.z2, u2);
  return true;
}

// ---------------------------------------------------------------------------
// Detector elements (Discretize::centre_all_elements, discrete.rs:150-171):
// n_azimuthal x (2 n_half_axial + 1) element centres, azimuth outer, axial
// inner, at radius n_radial * dr.
struct DevElements {
  DevDiscretize d;
  float r;
  int32_t n_half_axial;
  uint32_t n_axial, n_azimuthal;
  uint64_t n_points;
};

__host__ __device__ inline DevElements make_elements(float r_min, float dr, float dz, float da, float length,
                                                     int32_t adjust) {
  DevElements e;
  e.d = make_discretize(r_min, dr, dz, da, adjust);
  e.r = e.d.n_radial * dr;
  const float half_length = length / 2.0f;
  e.n_half_axial = (int32_t)(uint32_t)roundf(half_length / dz);
  e.n_axial = 2u * (uint32_t)e.n_half_axial + 1u;
  const float rr = r_min + dr / 2.0f;
  e.n_azimuthal = (uint32_t)roundf(kTwoPi * rr / da);
  e.n_points = (uint64_t)e.n_axial * e.n_azimuthal;
  return e;
}

__host__ __device__ inline void element_centre(const DevElements& e, uint64_t idx, float& x, float& y, float& z) {
  const uint32_t az = (uint32_t)(idx / e.n_axial);
  const int32_t n_z = (int32_t)(idx % e.n_axial) - e.n_half_axial;
  const float phi = e.d.d_azimuthal * (float)az;
  x = e.r * cosf(phi);
  y = e.r * sinf(phi);
  z = (float)n_z * e.d.dz;
}

// Pair t of make_lors (discrete.rs make_lors, bin/make_sensitivity_image/
// discrete.rs:43-62): for i in 0..N, for k in i..N, t counts (i, k); the LOR
// runs from element k to element i, both ends adjusted.
__host__ __device__ inline void pair_of(uint64_t n, uint64_t t, uint64_t& i, uint64_t& k) {
  auto row_start = [n](uint64_t r) { return r * n - (r * (r - 1)) / 2; };  // r*N - r(r-1)/2
  const double b = 2.0 * (double)n + 1.0;
  double g = (b - sqrt(b * b - 8.0 * (double)t)) / 2.0;
  uint64_t r = g < 0.0 ? 0 : (uint64_t)g;
  if (r >= n) r = n - 1;
  while (r + 1 < n && row_start(r + 1) <= t) ++r;
  while (r > 0 && row_start(r) > t) --r;
  i = r;
  k = r + (t - row_start(r));
}

constexpr uint32_t kPairTag = 0x9a1bu;

__host__ __device__ inline Lor pair_lor(const DevElements& e, uint64_t seed, uint64_t t) {
  uint64_t i, k;
  pair_of(e.n_points, t, i, k);
  float px, py, pz, qx, qy, qz;
  element_centre(e, k, px, py, pz);
  element_centre(e, i, qx, qy, qz);
  const uint32_t lo = (uint32_t)t, hi = (uint32_t)(t >> 32);
  const U32x4 a = philox(seed, lo, hi, 1, kPairTag);
  const U32x4 b = philox(seed, lo, hi, 2, kPairTag);
  const float up[3] = {unit_f32(a.x), unit_f32(a.y), unit_f32(b.x)};
  const float uq[3] = {unit_f32(a.z), unit_f32(a.w), unit_f32(b.y)};
  adjust_point(e.d, px, py, pz, up);
  adjust_point(e.d, qx, qy, qz, uq);
  return Lor{px, py, pz, qx, qy, qz, 0.0f, 1.0f};
}

// ---------------------------------------------------------------------------
// Scattergram (lorogram.rs): five axes over LORs, counts of trues and of
// scatters.  Axis semantics follow ndhistogram 0.9.0 (Cargo.lock:1733-1735):
// Uniform(n, low, high) has an underflow bin 0, bins 1..n and an overflow
// bin n+1, with the bin of x at 1 + floor((x - low) / (high - low) * n);
// UniformCyclic(n, low, high) wraps x into [low, high) first and has no
// flow bins; a NaN coordinate falls in no bin.
struct Axis {
  uint32_t n;      // bins (without flow bins)
  bool cyclic;
  float low, high;
  __host__ __device__ uint32_t stride_bins() const { return cyclic ? n : n + 2; }
  // bin index, or -1 for none
  __host__ __device__ int64_t index(float x) const {
    const float width = high - low;
    if (cyclic) {
      float f = (x - low) / width;
      if (!(f == f)) return -1;
      f = f - floorf(f);
      int64_t i = (int64_t)(f * (float)n);
      if (i >= (int64_t)n) i = n - 1;
      return i;
    }
    const float f = (x - low) / width;
    if (!(f == f)) return -1;
    if (f < 0.0f) return 0;
    if (f >= 1.0f) return (int64_t)n + 1;
    int64_t i = (int64_t)(f * (float)n);
    if (i >= (int64_t)n) i = n - 1;
    return i + 1;
  }
};

struct DevScattergram {
  Axis phi, z, dz, r, dt;
  uint64_t bins;  // product of the axes' stride_bins
};

// LOR coordinates of the scattergram axes (lorogram.rs:166-221), from the
// record's points in mm.  The reference's dt axis is indexed by the LOR's
// distance from the z axis (lorogram.rs:155-160); kept so.
__host__ __device__ inline int64_t scattergram_bin(const DevScattergram& s, const H5Lor& l) {
  const float dx = l.x2 - l.x1, dy = l.y2 - l.y1;
  const float num = dx * l.y1 - dy * l.x1;
  const float den = sqrtf(dx * dx + dy * dy);
  const float r_signed = num / den;
  const float r = fabsf(num) / den;
  float phi = atan2f(dy, dx);
  if (r_signed < 0.0f) phi = phi + 3.14159265358979323846f;
  const float zm = (l.z1 + l.z2) / 2.0f;
  const float dz = fabsf(l.z1 - l.z2);
  const int64_t i0 = s.phi.index(phi), i1 = s.z.index(zm), i2 = s.dz.index(dz), i3 = s.r.index(r),
                i4 = s.dt.index(r);
  if (i0 < 0 || i1 < 0 || i2 < 0 || i3 < 0 || i4 < 0) return -1;
  return (((i0 * s.z.stride_bins() + i1) * s.dz.stride_bins() + i2) * s.r.stride_bins() + i3) *
             s.dt.stride_bins() + i4;
}

// (scatters + trues) / trues, FLT_MAX for an empty bin (lorogram.rs:65-74).
__host__ __device__ inline float scattergram_value(uint32_t trues, uint32_t scatters) {
  if (trues > 0) {
    const float t = (float)trues;
    return ((float)scatters + t) / t;
  }
  return 3.40282347e+38f;
}

}  // namespace petalo
