// fom_oracle.cpp -- CPU restatement of the figures of merit (TEST
// INFRASTRUCTURE ONLY, part of liboracle.so; see petalo_oracle.h).
//
//   ROI membership      src/fom.rs:34-63 (binary32 as written)
//   values_inside_roi   src/fom.rs:81-88 (voxel index order)
//   mean, mu_and_sigma  src/fom.rs:276-288 (sequential sums)
//   Image::foms, crc    src/fom.rs:426-461
//   phantom analyses    src/bin/foms.rs:42-250
//
// Every statistic is computed twice: in binary32 in the reference's order
// (T = float) and in binary64 (T = double), each voxel value widened first.
#include <cmath>
#include <cstdint>
#include <vector>

#include "petalo_oracle.h"

namespace {

struct Grid {
  float hw[3], vs[3];
  uint32_t n[3];
};

Grid grid_of(const float full[3], const uint32_t n[3]) {
  Grid g;
  for (int a = 0; a < 3; ++a) {
    g.hw[a] = full[a] / 2.0f;             // fov.rs:27
    g.vs[a] = (g.hw[a] * 2.0f) / (float)n[a];  // fov.rs:28-30
    g.n[a] = n[a];
  }
  return g;
}

// FOV::voxel_centre1 (fov.rs:42-53): index1_to_3 with x fastest.
void centre1(const Grid& g, size_t i, float p[3]) {
  const size_t ix = i % g.n[0], iy = (i / g.n[0]) % g.n[1], iz = i / ((size_t)g.n[0] * g.n[1]);
  const size_t idx[3] = {ix, iy, iz};
  for (int a = 0; a < 3; ++a) p[a] = ((float)idx[a] + 0.5f) * g.vs[a] - g.hw[a];
}

// ROI::contains_fn (fom.rs:34-63).
bool roi_contains(const orc_roi_t& roi, const float p[3]) {
  const float r = roi.r_mm;
  switch (roi.kind) {
    case 0: {  // Sphere((cx, cy, cz), r)
      const float x = p[0] - roi.c[0], y = p[1] - roi.c[1], z = p[2] - roi.c[2];
      return x * x + y * y + z * z < r * r;
    }
    case 1: {  // CylinderX((cy, cz), r)
      const float y = p[1] - roi.c[0], z = p[2] - roi.c[1];
      return y * y + z * z < r * r;
    }
    case 2: {  // CylinderY((cx, cz), r)
      const float x = p[0] - roi.c[0], z = p[2] - roi.c[1];
      return x * x + z * z < r * r;
    }
    case 3: {  // CylinderZ((cx, cy), r)
      const float x = p[0] - roi.c[0], y = p[1] - roi.c[1];
      return x * x + y * y < r * r;
    }
    default: {  // DiscZ((cx, cy, z), r)
      const float x = p[0] - roi.c[0], y = p[1] - roi.c[1];
      return roi.c[2] == p[2] && x * x + y * y < r * r;
    }
  }
}

// Image::values_inside_roi (fom.rs:81-88).
std::vector<float> values_inside(const Grid& g, const float* img, const orc_roi_t& roi) {
  std::vector<float> out;
  const size_t nv = (size_t)g.n[0] * g.n[1] * g.n[2];
  float p[3];
  for (size_t i = 0; i < nv; ++i) {
    centre1(g, i, p);
    if (roi_contains(roi, p)) out.push_back(img[i]);
  }
  return out;
}

// fom::mean (fom.rs:276-278): reduce(a + b) / len; false when empty.
template <class T, class V> bool mean_of(const std::vector<V>& v, T& m) {
  if (v.empty()) return false;
  T s = (T)v[0];
  for (size_t k = 1; k < v.size(); ++k) s = s + (T)v[k];
  m = s / (T)v.size();
  return true;
}

// fom::mu_and_sigma (fom.rs:280-288).
template <class T, class V> bool mu_sigma(const std::vector<V>& v, T& mu, T& sigma) {
  if (!mean_of(v, mu)) return false;
  T s = (T)0;
  for (const V& x : v) {
    const T d = (T)x - mu;
    s = s + d * d;
  }
  sigma = std::sqrt(s / (T)v.size());
  return true;
}

// fom::crc (fom.rs:452-461).
template <class T> T crc(T roi_measured, T roi_activity, T bgd_measured, T bgd_activity) {
  return (T)100 * (roi_activity > bgd_activity
                       ? ((roi_measured / bgd_measured) - (T)1) / ((roi_activity / bgd_activity) - (T)1)
                       : (T)1 - (roi_measured / bgd_measured));
}

// fom::position_to_index / index_to_position (fom.rs:115-123).
uint64_t pos_to_index(float x, float hw, float vs) {
  const float q = (x + hw) / vs;
  if (!(q > 0.0f)) return 0;  // saturating `as usize`
  if (q >= 18446744073709551616.0f) return UINT64_MAX;
  return (uint64_t)q;
}
float index_to_pos(uint64_t i, float hw, float vs) { return ((float)i + 0.5f) * vs - hw; }
float closest_centre(float x, float hw, float vs) { return index_to_pos(pos_to_index(x, hw, vs), hw, vs); }

orc_roi_t disc(float x, float y, float z, float r) {
  orc_roi_t d;
  d.kind = 4;
  d.c[0] = x;
  d.c[1] = y;
  d.c[2] = z;
  d.r_mm = r;
  return d;
}

template <class T>
int phantom_foms(const Grid& g, const float* img, int phantom, float foms[24], float* aocs, size_t cap,
                 size_t* n_aocs) {
  // nema7_foms / jaszczak_foms (foms.rs:89-211)
  struct S { int pos; float d; float a; };
  std::vector<S> spheres;
  std::vector<std::pair<float, float>> xys;
  std::vector<float> zs;
  float ring_r, sphere_z;
  if (phantom == 0) {
    ring_r = 114.4f / 2.0f;
    sphere_z = 0.0f;
    spheres = {{1, 10.0f, 4.0f}, {2, 13.0f, 4.0f}, {3, 17.0f, 4.0f}, {4, 22.0f, 4.0f}, {5, 28.0f, 4.0f},
               {0, 37.0f, 4.0f}};
    xys = {{0.0f, -86.0f},  {0.0f, -82.0f},  {60.0f, -82.0f}, {-60.0f, -82.0f}, {-100.0f, -62.0f},
           {100.0f, -62.0f}, {-110.0f, -20.0f}, {110.0f, -20.0f}, {-85.0f, 40.0f},  {85.0f, 40.0f},
           {-52.0f, 72.0f}, {52.0f, 72.0f},   {0.0f, 82.0f}};
    zs = {-20.0f, -10.0f, 0.0f, 10.0f, 20.0f};
  } else {
    ring_r = 54.0f;
    sphere_z = 34.0f;
    spheres = {{0, 9.5f, 4.0f}, {1, 12.7f, 4.0f}, {2, 15.9f, 4.0f}, {3, 19.1f, 4.0f}, {4, 25.4f, 4.0f},
               {5, 31.8f, 4.0f}};
    xys = {{73.4f, -32.0f}, {67.4f, 35.2f}, {1.5f, 77.8f}, {-65.0f, 38.2f}, {-69.6f, -32.4f}, {-2.4f, -78.4f},
           {0.0f, 0.0f}};
    zs = {5.0f, 15.0f, 25.0f, 35.0f, 45.0f, 55.0f, 65.0f, 75.0f};
  }
  const T bg_activity = (T)1;
  const float hw = g.hw[2], vs = g.vs[2];
  // sphere_foms (foms.rs:42-87)
  for (float& z : zs) z = closest_centre(z, hw, vs);
  const float fz = closest_centre(sphere_z, hw, vs);
  T crc37 = (T)NAN;
  bool have37 = false;
  for (size_t k = 0; k < spheres.size(); ++k) {
    // sphere() (foms.rs:214-218)
    const float radians = 6.2831855f * (float)spheres[k].pos / 6.0f;
    const float x = ring_r * std::cos(radians), y = ring_r * std::sin(radians), r = spheres[k].d / 2.0f;
    // contrast_and_variability (foms.rs:222-250)
    T sm, ss;
    if (!mu_sigma<T>(values_inside(g, img, disc(x, y, fz, r)), sm, ss)) return -1;
    std::vector<T> bg_means;
    for (const auto& xy : xys)
      for (float z : zs) {
        T m;
        if (!mean_of<T>(values_inside(g, img, disc(xy.first, xy.second, z, r)), m)) return -1;
        bg_means.push_back(m);
      }
    T bg_mean = (T)NAN, bg_sd = (T)NAN;
    mu_sigma<T>(bg_means, bg_mean, bg_sd);
    const T act = (T)spheres[k].a;
    const T c = crc<T>(sm, act, bg_mean, bg_activity);
    T snr = (T)100 * (sm - bg_mean) / ss;
    if (!(act > bg_activity)) snr = -snr;
    foms[4 * k + 0] = r;
    foms[4 * k + 1] = (float)c;
    foms[4 * k + 2] = (float)((T)100 * bg_sd / bg_mean);
    foms[4 * k + 3] = (float)snr;
    if (r == 37.0f / 2.0f) crc37 = c, have37 = true;
  }
  size_t n = 0;
  if (phantom == 0) {
    // 7.4.2 (foms.rs:136-170)
    const float hi_limit = 70.0f - 30.0f, lo_limit = 70.0f - 180.0f + 30.0f;
    const float hi_c = closest_centre(hi_limit, hw, vs), lo_c = closest_centre(lo_limit, hw, vs);
    const float hi = hi_c <= hi_limit ? hi_c : closest_centre(hi_c - vs, hw, vs);
    const float lo = lo_c >= lo_limit ? lo_c : closest_centre(lo_c + vs, hw, vs);
    const uint64_t hi_i = pos_to_index(hi, hw, vs), lo_i = pos_to_index(lo, hw, vs);
    if (!have37) return -1;
    // lung voxels: CylinderZ((0, 0), 15 mm); each slice: DiscZ((0, 0, z_i), 15 mm)
    const orc_roi_t lung = {3, {0.0f, 0.0f, 0.0f}, 15.0f};
    for (uint64_t i = lo_i; i <= hi_i && lo_i <= hi_i; ++i) {
      const orc_roi_t d = disc(0.0f, 0.0f, index_to_pos(i, hw, vs), 30.0f / 2.0f);
      std::vector<float> v;
      const size_t nv = (size_t)g.n[0] * g.n[1] * g.n[2];
      float p[3];
      for (size_t j = 0; j < nv; ++j) {
        centre1(g, j, p);
        if (roi_contains(lung, p) && roi_contains(d, p)) v.push_back(img[j]);
      }
      T m;
      if (!mean_of<T>(v, m)) return -1;
      if (n < cap && aocs) aocs[n] = (float)((T)100 * m / crc37);
      ++n;
      if (i == UINT64_MAX) break;
    }
  }
  if (n_aocs) *n_aocs = n;
  return 0;
}

}  // namespace

extern "C" {

int orc_roi_stats(const float full_mm[3], const uint32_t n[3], const float* image, const orc_roi_t* rois,
                  size_t n_rois, int f64, uint64_t* counts, double* means, double* sigmas) {
  const Grid g = grid_of(full_mm, n);
  for (size_t k = 0; k < n_rois; ++k) {
    const std::vector<float> v = values_inside(g, image, rois[k]);
    counts[k] = v.size();
    if (v.empty()) {
      means[k] = sigmas[k] = NAN;
    } else if (f64) {
      mu_sigma<double>(v, means[k], sigmas[k]);
    } else {
      float m, s;
      mu_sigma<float>(v, m, s);
      means[k] = m;
      sigmas[k] = s;
    }
  }
  return 0;
}

int orc_image_foms(const float full_mm[3], const uint32_t n[3], const float* image, const orc_roi_t* rois,
                   const float* activities, size_t n_rois, const orc_roi_t* bg_rois, size_t n_bg,
                   float bg_activity, int f64, float* crcs, float* snrs, float* bg_measured) {
  // Image::foms (fom.rs:426-450)
  const Grid g = grid_of(full_mm, n);
  auto run = [&](auto zero) -> int {
    using T = decltype(zero);
    T bg = (T)0;
    for (size_t k = 0; k < n_bg; ++k) {
      T m;
      if (!mean_of<T>(values_inside(g, image, bg_rois[k]), m)) return -1;
      bg = bg + m;
    }
    bg = bg / (T)n_bg;
    for (size_t k = 0; k < n_rois; ++k) {
      T mu, sigma;
      if (!mu_sigma<T>(values_inside(g, image, rois[k]), mu, sigma)) return -1;
      crcs[k] = (float)crc<T>(mu, (T)activities[k], bg, (T)bg_activity);
      snrs[k] = (float)((mu - bg) / sigma);
    }
    if (bg_measured) *bg_measured = (float)bg;
    return 0;
  };
  return f64 ? run(0.0) : run(0.0f);
}

int orc_phantom_foms(const float full_mm[3], const uint32_t n[3], const float* image, int phantom, int f64,
                     float foms[24], float* aocs, size_t aoc_capacity, size_t* n_aocs) {
  const Grid g = grid_of(full_mm, n);
  return f64 ? phantom_foms<double>(g, image, phantom, foms, aocs, aoc_capacity, n_aocs)
             : phantom_foms<float>(g, image, phantom, foms, aocs, aoc_capacity, n_aocs);
}

}  // extern "C"
