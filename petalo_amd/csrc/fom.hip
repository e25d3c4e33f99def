// fom.hip -- figures of merit on reconstructed images (src/fom.rs,
// src/bin/foms.rs): region statistics on the device, the phantom analyses on
// the host from those statistics.
//
// A region's voxels are those whose centre passes the region's binary32 test
// (fom.rs:34-63).  The host bounds each region by a box of voxels, one voxel
// wider than the region on each side; the device tests every voxel of the
// box, so membership is decided only by the exact test.  One workgroup per
// region: count and sum, a block reduction, then the sum of squared
// deviations from the mean (two passes, as fom::mu_and_sigma, fom.rs:280-288),
// in binary64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "hip_util.hpp"
#include "siddon_walk.hpp"

using namespace petalo;

namespace {

struct RoiBox {
  int32_t kind;
  float c0, c1, c2, r;
  int32_t lo[3];
  int32_t n[3];  // box extent in voxels; any 0 = empty box
};

// Voxel centre along one axis (fov.rs:42-48), binary32.
__device__ inline float centre_of(int32_t i, float vs, float hw) { return ((float)i + 0.5f) * vs - hw; }

// ROI::contains_fn (fom.rs:34-63), binary32 in the reference's order.
__device__ inline bool contains(const RoiBox& b, float x, float y, float z) {
  const float rr = b.r * b.r;
  switch (b.kind) {
    case PETALO_ROI_SPHERE: {
      const float dx = x - b.c0, dy = y - b.c1, dz = z - b.c2;
      return dx * dx + dy * dy + dz * dz < rr;
    }
    case PETALO_ROI_CYLINDER_X: {
      const float dy = y - b.c0, dz = z - b.c1;
      return dy * dy + dz * dz < rr;
    }
    case PETALO_ROI_CYLINDER_Y: {
      const float dx = x - b.c0, dz = z - b.c1;
      return dx * dx + dz * dz < rr;
    }
    case PETALO_ROI_CYLINDER_Z: {
      const float dx = x - b.c0, dy = y - b.c1;
      return dx * dx + dy * dy < rr;
    }
    default: {  // PETALO_ROI_DISC_Z
      const float dx = x - b.c0, dy = y - b.c1;
      return b.c2 == z && dx * dx + dy * dy < rr;
    }
  }
}

constexpr int kRoiThreads = 256;

template <class T> __device__ inline T block_sum(T v, T* lds) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int s = kRoiThreads / 2; s > 0; s >>= 1) {
    if (t < s) lds[t] += lds[t + s];
    __syncthreads();
  }
  const T r = lds[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kRoiThreads) void k_roi_stats(const float* __restrict__ img, DevFov f,
                                                         const RoiBox* __restrict__ boxes,
                                                         petalo_roi_stats_t* __restrict__ out) {
  __shared__ double lds_d[kRoiThreads];
  __shared__ unsigned long long lds_u[kRoiThreads];
  const RoiBox b = boxes[blockIdx.x];
  const uint32_t n0 = (uint32_t)b.n[0], n01 = (uint32_t)b.n[0] * (uint32_t)b.n[1];
  const uint32_t total = n01 * (uint32_t)b.n[2];
  // pass 1: count and sum
  double s = 0.0;
  unsigned long long c = 0;
  for (uint32_t t = threadIdx.x; t < total; t += kRoiThreads) {
    const int32_t iz = b.lo[2] + (int32_t)(t / n01);
    const uint32_t rem = t % n01;
    const int32_t iy = b.lo[1] + (int32_t)(rem / n0), ix = b.lo[0] + (int32_t)(rem % n0);
    if (contains(b, centre_of(ix, f.vs_x, f.hw_x), centre_of(iy, f.vs_y, f.hw_y), centre_of(iz, f.vs_z, f.hw_z))) {
      s += (double)img[ix + (iy + iz * f.ny) * f.nx];
      ++c;
    }
  }
  const unsigned long long count = block_sum(c, lds_u);
  const double sum = block_sum(s, lds_d);
  const double mean = count ? sum / (double)count : __builtin_nan("");
  // pass 2: squared deviations from the mean
  double m2 = 0.0;
  if (count) {
    for (uint32_t t = threadIdx.x; t < total; t += kRoiThreads) {
      const int32_t iz = b.lo[2] + (int32_t)(t / n01);
      const uint32_t rem = t % n01;
      const int32_t iy = b.lo[1] + (int32_t)(rem / n0), ix = b.lo[0] + (int32_t)(rem % n0);
      if (contains(b, centre_of(ix, f.vs_x, f.hw_x), centre_of(iy, f.vs_y, f.hw_y),
                   centre_of(iz, f.vs_z, f.hw_z))) {
        const double d = (double)img[ix + (iy + iz * f.ny) * f.nx] - mean;
        m2 += d * d;
      }
    }
  }
  const double ss = block_sum(m2, lds_d);
  if (threadIdx.x == 0) {
    petalo_roi_stats_t o;
    o.count = count;
    o.mean = mean;
    o.sigma = count ? std::sqrt(ss / (double)count) : __builtin_nan("");
    out[blockIdx.x] = o;
  }
}

// Voxel range [lo, lo + n) along one axis that holds every voxel whose centre
// can lie within r of c, one voxel of margin on each side; clamped to the grid.
void axis_span(double c, double r, double hw, double vs, int32_t nvox, int32_t& lo, int32_t& n) {
  double a = std::floor((c - r + hw) / vs) - 1.0, b = std::floor((c + r + hw) / vs) + 1.0;
  a = std::max(a, 0.0);
  b = std::min(b, (double)nvox - 1.0);
  if (!(a <= b)) {
    lo = 0;
    n = 0;
    return;
  }
  lo = (int32_t)a;
  n = (int32_t)(b - a) + 1;
}

RoiBox make_box(const petalo_roi_t& roi, const DevFov& f) {
  RoiBox b{};
  b.kind = roi.kind;
  b.c0 = roi.c[0];
  b.c1 = roi.c[1];
  b.c2 = roi.c[2];
  b.r = roi.r_mm;
  const double r = std::fabs((double)roi.r_mm);
  const double hw[3] = {f.hw_x, f.hw_y, f.hw_z}, vs[3] = {f.vs_x, f.vs_y, f.vs_z};
  const int32_t nv[3] = {f.nx, f.ny, f.nz};
  // centre and radius per axis; an axis the region does not bound spans the grid
  bool bounded[3] = {true, true, true};
  double c[3] = {roi.c[0], roi.c[1], roi.c[2]}, rad[3] = {r, r, r};
  switch (roi.kind) {
    case PETALO_ROI_SPHERE: break;
    case PETALO_ROI_CYLINDER_X: bounded[0] = false, c[1] = roi.c[0], c[2] = roi.c[1]; break;
    case PETALO_ROI_CYLINDER_Y: bounded[1] = false, c[0] = roi.c[0], c[2] = roi.c[1]; break;
    case PETALO_ROI_CYLINDER_Z: bounded[2] = false; break;
    default: rad[2] = 0.0; break;  // DISC_Z: the slice(s) whose centre can equal c2
  }
  for (int a = 0; a < 3; ++a) {
    if (bounded[a]) {
      axis_span(c[a], rad[a], hw[a], vs[a], nv[a], b.lo[a], b.n[a]);
    } else {
      b.lo[a] = 0;
      b.n[a] = nv[a];
    }
  }
  if (!b.n[0] || !b.n[1] || !b.n[2]) b.n[0] = b.n[1] = b.n[2] = 0;
  return b;
}

void check_rois(const petalo_roi_t* rois, size_t n) {
  for (size_t k = 0; k < n; ++k) {
    const petalo_roi_t& r = rois[k];
    if (r.kind < PETALO_ROI_SPHERE || r.kind > PETALO_ROI_DISC_Z)
      throw std::invalid_argument("region " + std::to_string(k) + ": unknown kind");
    if (!std::isfinite(r.r_mm) || !std::isfinite(r.c[0]) || !std::isfinite(r.c[1]) || !std::isfinite(r.c[2]))
      throw std::invalid_argument("region " + std::to_string(k) + ": centre and radius must be finite");
  }
}

// Statistics of every region of one image (host pointer), one launch.
std::vector<petalo_roi_stats_t> region_stats(const petalo_fov_t& fov, const float* image,
                                             const std::vector<petalo_roi_t>& rois) {
  std::vector<petalo_roi_stats_t> out(rois.size());
  if (rois.empty()) return out;
  check_rois(rois.data(), rois.size());
  require_device();
  const DevFov f = make_fov(fov.full_size_mm, fov.n);
  std::vector<RoiBox> boxes(rois.size());
  for (size_t k = 0; k < rois.size(); ++k) boxes[k] = make_box(rois[k], f);
  const size_t nv = (size_t)f.nx * f.ny * f.nz;
  hipStream_t s;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  DevBuf<float> dimg(nv);
  DevBuf<RoiBox> dbox(boxes.size());
  DevBuf<petalo_roi_stats_t> dout(rois.size());
  HIP_CHECK(hipMemcpyAsync(dimg.p, image, nv * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dbox.p, boxes.data(), boxes.size() * sizeof(RoiBox), hipMemcpyHostToDevice, s));
  for (size_t first = 0; first < rois.size(); first += 65535) {
    const unsigned g = (unsigned)std::min<size_t>(65535, rois.size() - first);
    k_roi_stats<<<g, kRoiThreads, 0, s>>>(dimg.p, f, dbox.p + first, dout.p + first);
    HIP_CHECK(hipGetLastError());
  }
  HIP_CHECK(hipMemcpyAsync(out.data(), dout.p, out.size() * sizeof(petalo_roi_stats_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return out;
}

// fom::crc (fom.rs:452-461): percent; hot when the region is hotter than the
// background, cold otherwise.
double crc(double roi_measured, double roi_activity, double bgd_measured, double bgd_activity) {
  return 100.0 * (roi_activity > bgd_activity ? ((roi_measured / bgd_measured) - 1.0) /
                                                     ((roi_activity / bgd_activity) - 1.0)
                                               : 1.0 - (roi_measured / bgd_measured));
}

// fom::mu_and_sigma (fom.rs:280-288) of a list of values.
void mu_and_sigma(const std::vector<double>& v, double& mu, double& sigma) {
  double s = 0.0;
  for (double x : v) s += x;
  mu = s / (double)v.size();
  double m2 = 0.0;
  for (double x : v) m2 += (x - mu) * (x - mu);
  sigma = std::sqrt(m2 / (double)v.size());
}

petalo_roi_t disc_z(float x, float y, float z, float r) { return petalo_roi_t{PETALO_ROI_DISC_Z, {x, y, z}, r}; }

// fom::position_to_index (fom.rs:115-118): binary32 ratio, saturating cast.
uint64_t position_to_index(float position, float half_width, float voxel_size) {
  const float q = (position + half_width) / voxel_size;
  if (!(q > 0.0f)) return 0;
  if (q >= 18446744073709551616.0f) return std::numeric_limits<uint64_t>::max();
  return (uint64_t)q;
}

// fom::index_to_position (fom.rs:120-123).
float index_to_position(uint64_t index, float half_width, float voxel_size) {
  return ((float)index + 0.5f) * voxel_size - half_width;
}

float nearest_centre(float x, float hw, float vs) { return index_to_position(position_to_index(x, hw, vs), hw, vs); }

struct PhantomSpec {
  float ring_r, sphere_z;
  struct S { uint16_t pos; float diameter, activity; } spheres[6];
  std::vector<float> bg_x, bg_y, bg_z;
  float bg_activity;
};

// nema7_foms / jaszczak_foms constants (foms.rs:89-134, 174-200).
PhantomSpec phantom_spec(int32_t phantom) {
  PhantomSpec p;
  p.bg_activity = 1.0f;
  if (phantom == PETALO_PHANTOM_NEMA7) {
    p.ring_r = 114.4f / 2.0f;
    p.sphere_z = 0.0f;
    const PhantomSpec::S s[6] = {{1, 10.0f, 4.0f}, {2, 13.0f, 4.0f}, {3, 17.0f, 4.0f},
                                 {4, 22.0f, 4.0f}, {5, 28.0f, 4.0f}, {0, 37.0f, 4.0f}};
    std::copy(s, s + 6, p.spheres);
    p.bg_x = {0.0f, 0.0f, 60.0f, -60.0f, -100.0f, 100.0f, -110.0f, 110.0f, -85.0f, 85.0f, -52.0f, 52.0f, 0.0f};
    p.bg_y = {-86.0f, -82.0f, -82.0f, -82.0f, -62.0f, -62.0f, -20.0f, -20.0f, 40.0f, 40.0f, 72.0f, 72.0f, 82.0f};
    p.bg_z = {-20.0f, -10.0f, 0.0f, 10.0f, 20.0f};
  } else {
    p.ring_r = 54.0f;
    p.sphere_z = 34.0f;
    const PhantomSpec::S s[6] = {{0, 9.5f, 4.0f},  {1, 12.7f, 4.0f}, {2, 15.9f, 4.0f},
                                 {3, 19.1f, 4.0f}, {4, 25.4f, 4.0f}, {5, 31.8f, 4.0f}};
    std::copy(s, s + 6, p.spheres);
    p.bg_x = {73.4f, 67.4f, 1.5f, -65.0f, -69.6f, -2.4f, 0.0f};
    p.bg_y = {-32.0f, 35.2f, 77.8f, 38.2f, -32.4f, -78.4f, 0.0f};
    p.bg_z = {5.0f, 15.0f, 25.0f, 35.0f, 45.0f, 55.0f, 65.0f, 75.0f};
  }
  return p;
}

void require_voxels(const petalo_roi_stats_t& st, const char* what) {
  if (st.count == 0) throw std::invalid_argument(std::string(what) + " contains no voxel centre");
}

}  // namespace

extern "C" {

int petalo_roi_stats(const petalo_fov_t* fov, const float* image, const petalo_roi_t* rois, size_t n_rois,
                     petalo_roi_stats_t* out) {
  return guarded([&] {
    std::string why;
    if (!fov_ok(fov, why)) return fail(PETALO_ERR_ARG, why);
    if (n_rois && (!image || !rois || !out)) return fail(PETALO_ERR_ARG, "NULL argument");
    if (n_rois == 0) return PETALO_OK;
    const std::vector<petalo_roi_t> v(rois, rois + n_rois);
    const auto st = region_stats(*fov, image, v);
    std::copy(st.begin(), st.end(), out);
    return PETALO_OK;
  });
}

int petalo_image_foms(const petalo_fov_t* fov, const float* image, const petalo_roi_t* rois,
                      const float* activities, size_t n_rois, const petalo_roi_t* background_rois,
                      size_t n_background, float background_activity, float* crcs_out, float* snrs_out,
                      float* background_measured_out) {
  return guarded([&] {
    std::string why;
    if (!fov_ok(fov, why)) return fail(PETALO_ERR_ARG, why);
    if (!image || n_background == 0 || !background_rois) return fail(PETALO_ERR_ARG, "no background regions");
    if (n_rois && (!rois || !activities || !crcs_out || !snrs_out)) return fail(PETALO_ERR_ARG, "NULL argument");
    std::vector<petalo_roi_t> all(background_rois, background_rois + n_background);
    all.insert(all.end(), rois, rois + n_rois);
    const auto st = region_stats(*fov, image, all);
    // background_measured = sum of background means / their number (fom.rs:428-431)
    double bg = 0.0;
    for (size_t k = 0; k < n_background; ++k) {
      require_voxels(st[k], ("background region " + std::to_string(k)).c_str());
      bg += st[k].mean;
    }
    bg /= (double)n_background;
    for (size_t k = 0; k < n_rois; ++k) {
      const petalo_roi_stats_t& r = st[n_background + k];
      require_voxels(r, ("region " + std::to_string(k)).c_str());
      crcs_out[k] = (float)crc(r.mean, activities[k], bg, background_activity);
      snrs_out[k] = (float)((r.mean - bg) / r.sigma);  // fom.rs:440
    }
    if (background_measured_out) *background_measured_out = (float)bg;
    return PETALO_OK;
  });
}

int petalo_phantom_foms(const petalo_fov_t* fov, const float* image, int32_t phantom, petalo_fom_t foms_out[6],
                        float* aocs_out, size_t aoc_capacity, size_t* n_aocs_out) {
  return guarded([&] {
    std::string why;
    if (!fov_ok(fov, why)) return fail(PETALO_ERR_ARG, why);
    if (!image || !foms_out) return fail(PETALO_ERR_ARG, "NULL argument");
    if (phantom != PETALO_PHANTOM_NEMA7 && phantom != PETALO_PHANTOM_JASZCZAK)
      return fail(PETALO_ERR_ARG, "unknown phantom");
    const PhantomSpec p = phantom_spec(phantom);
    const DevFov f = make_fov(fov->full_size_mm, fov->n);
    const float hw = f.hw_z, vs = f.vs_z;
    // sphere_foms (foms.rs:42-87): regions z-aligned with slice centres
    const float fz = nearest_centre(p.sphere_z, hw, vs);
    std::vector<float> bg_z;
    for (float z : p.bg_z) bg_z.push_back(nearest_centre(z, hw, vs));
    std::vector<petalo_roi_t> rois;
    float sx[6], sy[6], sr[6];
    for (int k = 0; k < 6; ++k) {
      // sphere(): the k-th of six angular positions (foms.rs:214-218)
      const float radians = 6.2831855f * (float)p.spheres[k].pos / 6.0f;
      sx[k] = p.ring_r * std::cos(radians);
      sy[k] = p.ring_r * std::sin(radians);
      sr[k] = p.spheres[k].diameter / 2.0f;
      rois.push_back(disc_z(sx[k], sy[k], fz, sr[k]));
      // background discs of this sphere's radius: xy outer, z inner (foms.rs:237-241)
      for (size_t b = 0; b < p.bg_x.size(); ++b)
        for (float z : bg_z) rois.push_back(disc_z(p.bg_x[b], p.bg_y[b], z, sr[k]));
    }
    const size_t per_sphere = 1 + p.bg_x.size() * bg_z.size();
    // NEMA7 7.4.2: lung insert slices at least 30 mm from the phantom's ends (foms.rs:136-170)
    uint64_t lo_index = 1, hi_index = 0;
    if (phantom == PETALO_PHANTOM_NEMA7) {
      const float hi_limit = 70.0f - 30.0f, lo_limit = 70.0f - 180.0f + 30.0f;
      const float hi_c = nearest_centre(hi_limit, hw, vs), lo_c = nearest_centre(lo_limit, hw, vs);
      const float hi = hi_c <= hi_limit ? hi_c : nearest_centre(hi_c - vs, hw, vs);
      const float lo = lo_c >= lo_limit ? lo_c : nearest_centre(lo_c + vs, hw, vs);
      hi_index = position_to_index(hi, hw, vs);
      lo_index = position_to_index(lo, hw, vs);
      if (lo_index <= hi_index && hi_index - lo_index >= (1u << 24))
        return fail(PETALO_ERR_ARG, "too many lung-insert slices");
      for (uint64_t i = lo_index; i <= hi_index && lo_index <= hi_index; ++i)
        rois.push_back(disc_z(0.0f, 0.0f, index_to_position(i, hw, vs), 30.0f / 2.0f));
    }
    const auto st = region_stats(*fov, image, rois);
    double crc37 = std::numeric_limits<double>::quiet_NaN();
    for (int k = 0; k < 6; ++k) {
      // contrast_and_variability (foms.rs:222-250)
      const petalo_roi_stats_t& sph = st[k * per_sphere];
      require_voxels(sph, "sphere region");
      std::vector<double> bg_means;
      for (size_t j = 1; j < per_sphere; ++j) {
        require_voxels(st[k * per_sphere + j], "background region");
        bg_means.push_back(st[k * per_sphere + j].mean);
      }
      double bg_mean, bg_sd;
      mu_and_sigma(bg_means, bg_mean, bg_sd);
      const double a = p.spheres[k].activity;
      const double c = crc(sph.mean, a, bg_mean, p.bg_activity);
      double snr = 100.0 * (sph.mean - bg_mean) / sph.sigma;
      if (!(a > p.bg_activity)) snr = -snr;
      foms_out[k] = petalo_fom_t{sr[k], (float)c, (float)(100.0 * bg_sd / bg_mean), (float)snr};
      if (sr[k] == 37.0f / 2.0f) crc37 = c;  // the reference keeps the 37 mm sphere's crc as "bg_37"
    }
    size_t n_aoc = 0;
    if (phantom == PETALO_PHANTOM_NEMA7 && lo_index <= hi_index) {
      n_aoc = (size_t)(hi_index - lo_index + 1);
      for (size_t i = 0; i < n_aoc; ++i) {
        const petalo_roi_stats_t& r = st[6 * per_sphere + i];
        require_voxels(r, "lung-insert region");
        if (i < aoc_capacity && aocs_out) aocs_out[i] = (float)(100.0 * r.mean / crc37);
      }
    }
    if (n_aocs_out) *n_aocs_out = n_aoc;
    return PETALO_OK;
  });
}

}  // extern "C"
