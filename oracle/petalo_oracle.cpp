// petalo_oracle.cpp -- CPU oracle (TEST INFRASTRUCTURE ONLY; see petalo_oracle.h).
//
// Built with -ffp-contract=off and without fast-math so that every binary32
// operation is a single correctly rounded IEEE operation, as in the Rust
// reference (rustc never contracts a*b+c into an FMA).
//
// Parity status of each function (what pins it; the tests live in tests/):
//   fov_new / voxel_centre ... reference KATs src/fov.rs:100-115 (exact).
//   fov_entry ................ binary64 slab closed form + hand_picked KATs
//                              (siddon.rs:381-433).  The binary32 operation
//                              order of ncollide3d 0.33.0's ray/AABB cast is
//                              restated from its published algorithm; bitwise
//                              toi for oblique LORs is "parity unpinned"
//                              (the reference cannot be built here).
//   walk (row) ............... hand_picked KATs (exact voxel sequence, total
//                              length within 1 ulp), sum-of-weights proptest
//                              invariant (siddon.rs:440-490), independent
//                              classic Siddon and brute-force voxel clipping
//                              in binary64.
//   gauss .................... closed-form normal pdf, strict cutoff.
//   mlem update .............. exact invariant sum_j x+_j/s_j = N_projected,
//                              OSEM partition KAT, FP64 twin agreement.
//   sensitivity .............. mu = 0: s = (1/N) sum w; uniform mu: per-LOR
//                              exp(mu * chord) closed form.

#include "petalo_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

namespace {

// ---- constants (units/src/lib.rs:36-47,156; siddon.rs:106-108,134) --------
template <class T> struct K;
template <> struct K<float> {
  static constexpr float C = 0.299792458f;     // mm/ps, units/src/lib.rs:156
  static constexpr float TAU = 6.2831855f;     // std::f32::consts::TAU
  static constexpr float EPS = 1e-5f;          // siddon.rs:134
  static constexpr float TOF_FACTOR = 666.0f;  // siddon.rs:107
};
template <> struct K<double> {
  static constexpr double C = 0.299792458;
  static constexpr double TAU = 6.283185307179586;
  static constexpr double EPS = 1e-5;
  static constexpr double TOF_FACTOR = 666.0;
};

inline float fmax_(float a, float b) { return std::fmax(a, b); }
inline double fmax_(double a, double b) { return std::fmax(a, b); }
inline float fmin_(float a, float b) { return std::fmin(a, b); }
inline double fmin_(double a, double b) { return std::fmin(a, b); }

// nalgebra 0.30.1 Vector3::norm: sqrt of the unrolled dot (a + b) + c.
template <class T> inline T norm3(T x, T y, T z) { return std::sqrt((x * x + y * y) + z * z); }

template <class T> struct Fov {
  T hw[3];  // half_width
  T vs[3];  // voxel_size
  int64_t n[3];
  int64_t nvox() const { return n[0] * n[1] * n[2]; }
};

// src/fov.rs:22-39: half_width = full/2; voxel_size = (half_width*2)/n.
template <class T> Fov<T> fov_new(const float full[3], const uint32_t n[3]) {
  Fov<T> f;
  for (int i = 0; i < 3; ++i) {
    f.hw[i] = T(full[i]) / T(2);
    f.n[i] = n[i];
  }
  for (int i = 0; i < 3; ++i) f.vs[i] = (f.hw[i] * T(2)) / T(n[i]);
  return f;
}

// Entry-point arithmetic used by fov_entry (orc_set_entry_variant):
//   0  the restated ncollide3d slab cast: entry = p1 + dir * toi per component
//      (the reference's FOV::entry as far as its third-party arithmetic is
//      known; the default);
//   1  the same, then the coordinate of the face that set toi replaced by that
//      face's exact position (-half_width: after flip_axes every direction
//      component is >= 0, so a LOR enters through a lower face).  This is the
//      entry an exact cast gives; it never lands outside the box, so no walk
//      starts shifted (see Hit::shifted).  It bounds how much the unpinned
//      binary32 order of the cast can move an image.
int g_entry_variant = 0;

// src/fov.rs:55-71 -> ncollide3d 0.33.0 Cuboid::toi_with_ray (solid = true),
// i.e. the AABB slab cast with tmin = 0, tmax = max_toi; identity isometry.
template <class T> bool fov_entry(const Fov<T>& f, const T p1[3], const T p2[3], T out[3]) {
  const T d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const T len = norm3(d[0], d[1], d[2]);
  const T dir[3] = {d[0] / len, d[1] / len, d[2] / len};  // normalize: v / norm
  T tmin = T(0);
  T tmax = len;  // max_toi = (p2 - p1).norm()
  int face = -1;  // axis whose slab set tmin (-1: p1 inside the box)
  for (int i = 0; i < 3; ++i) {
    if (dir[i] == T(0)) {
      if (p1[i] < -f.hw[i] || p1[i] > f.hw[i]) return false;
    } else {
      const T denom = T(1) / dir[i];
      T a = (-f.hw[i] - p1[i]) * denom;
      T b = (f.hw[i] - p1[i]) * denom;
      if (a > b) std::swap(a, b);
      if (a > tmin) face = i;
      tmin = fmax_(tmin, a);
      tmax = fmin_(tmax, b);
      if (tmin > tmax) return false;
    }
  }
  for (int i = 0; i < 3; ++i) out[i] = p1[i] + dir[i] * tmin;
  if (g_entry_variant == 1 && face >= 0) out[face] = -f.hw[face];
  return true;
}

template <class T> struct Gauss {
  T sigma, cutoff, peak;
};

// src/gauss.rs:31-45,58-60: sigma_x = sigma_t * C; peak = 1/(sigma*sqrt(TAU)).
template <class T> Gauss<T> gauss_new(T sigma_ps, T cutoff) {
  Gauss<T> g;
  g.sigma = sigma_ps * K<T>::C;
  const T root_two_pi = std::sqrt(K<T>::TAU);
  g.peak = T(1) / (g.sigma * root_two_pi);
  g.cutoff = cutoff * g.sigma;
  return g;
}

// src/gauss.rs:47-55 (strict cutoff).
template <class T> T gauss_call(const Gauss<T>& g, T dx) {
  if (std::fabs(dx) < g.cutoff) {
    const T y = dx / g.sigma;
    const T z = y * y;
    return g.peak * std::exp(T(-0.5) * z);
  }
  return T(0);
}

// Rust `x.floor() as usize` (saturating) then `as i32` (low 32 bits).
template <class T> inline int32_t floor_usize_i32(T x) {
  const T fl = std::floor(x);
  uint64_t u;
  if (!(fl > T(0))) u = 0;  // negative, zero or NaN
  else if (fl >= T(18446744073709551616.0)) u = UINT64_MAX;
  else u = (uint64_t)fl;
  return (int32_t)(uint32_t)(u & 0xffffffffu);
}

template <class T> struct Hit {
  T next[3];
  T vs[3];
  int32_t index;
  int32_t delta[3];
  int32_t remaining[3];
  T tof_peak;
  T frac[3];     // e - floor(e): how far into its first voxel the walk starts, per axis
  T entry[3];    // entry point (flipped coordinates)
  bool shifted;  // an entry coordinate e is negative after the EPS snap: floor(e) = -1
                 // saturates to voxel 0 (siddon.rs:312-313), so the walk starts
                 // one voxel short of the entry along that axis and every index
                 // after its first (sliver) step is one voxel further along
};

// siddon.rs:190-229 lor_fov_hit and helpers :138-185, :233-273, :322-340.
template <class T> bool lor_fov_hit(const orc_lor_t& lor, const Fov<T>& f, Hit<T>& h) {
  T p1[3], p2[3];
  bool flipped[3];
  for (int i = 0; i < 3; ++i) {
    p1[i] = T(lor.p1[i]);
    p2[i] = T(lor.p2[i]);
  }
  // flip_axes :322-340 (direction computed before any flip)
  const T dir0[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  for (int i = 0; i < 3; ++i) {
    flipped[i] = dir0[i] < T(0);
    if (flipped[i]) {
      p1[i] = -p1[i];
      p2[i] = -p2[i];
    }
  }
  T entry[3];
  if (!fov_entry(f, p1, p2, entry)) return false;

  // find_tof_peak :156-163
  const T half_lor_length = norm3(p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]) / T(2);
  const T tof_shift = K<T>::C * T(lor.dt_ps) / T(2);
  const T p1_to_peak = half_lor_length - tof_shift;
  const T p1_to_entry = norm3(entry[0] - p1[0], entry[1] - p1[1], entry[2] - p1[2]);
  h.tof_peak = p1_to_peak - p1_to_entry;

  // find_entry_point :138-153
  T e[3];
  h.shifted = false;
  for (int i = 0; i < 3; ++i) {
    h.entry[i] = entry[i];
    e[i] = (entry[i] + f.hw[i]) / f.vs[i];
    if (std::fabs(e[i]) < K<T>::EPS) e[i] = T(0);
    if (e[i] < T(0)) h.shifted = true;
  }

  // index_trackers :233-273
  int32_t ii[3], nn[3];
  for (int i = 0; i < 3; ++i) {
    ii[i] = floor_usize_i32(e[i]);
    nn[i] = (int32_t)f.n[i];
  }
  h.delta[0] = flipped[0] ? -1 : 1;
  h.delta[1] = flipped[1] ? -nn[0] : nn[0];
  h.delta[2] = flipped[2] ? -(nn[0] * nn[1]) : nn[0] * nn[1];
  for (int i = 0; i < 3; ++i) h.remaining[i] = nn[i] - ii[i];
  int32_t jj[3];
  for (int i = 0; i < 3; ++i) jj[i] = flipped[i] ? nn[i] - 1 - ii[i] : ii[i];
  h.index = (int32_t)((uint32_t)jj[0] + ((uint32_t)jj[1] + (uint32_t)jj[2] * (uint32_t)nn[1]) * (uint32_t)nn[0]);

  // voxel_size :179-185 (recomputes normalize(p2 - p1) on flipped points)
  {
    const T d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    const T len = norm3(d[0], d[1], d[2]);
    for (int i = 0; i < 3; ++i) h.vs[i] = f.vs[i] / (d[i] / len);
  }
  // first_boundaries :166-174
  for (int i = 0; i < 3; ++i) {
    const T frac = e[i] - std::floor(e[i]);
    h.frac[i] = frac;
    h.next[i] = (T(1) - frac) * h.vs[i];
  }
  return true;
}

template <class T> struct Row {
  std::vector<int64_t> idx;
  std::vector<T> w;
  std::vector<uint8_t> axis;  // crossed axis of every walk step (filled on request)
  std::vector<uint8_t> push;  // whether the step pushed an element (filled on request)
  uint64_t steps = 0;
  bool runaway = false;
  bool shifted = false;
  bool record = false;        // fill axis/push
  void clear() {
    idx.clear();
    w.clear();
    axis.clear();
    push.clear();
    steps = 0;
    runaway = false;
    shifted = false;
  }
};

// Step cap.  A legitimate walk ends within nx+ny+nz-2 steps; the reference
// walk never ends for an axis-parallel LOR starting exactly on the FOV's far
// face.  Both this oracle and the HIP kernels stop such walks here and drop
// the LOR.
inline uint64_t step_cap(const int64_t n[3]) { return (uint64_t)(2 * (n[0] + n[1] + n[2]) + 16); }

// vector.rs:150-158 argmin: ties go to x, then y.
template <class T> inline int argmin3(const T v[3]) {
  if (v[0] <= v[1]) return v[0] <= v[2] ? 0 : 2;
  return v[1] <= v[2] ? 1 : 2;
}

// siddon.rs:81-130 update_smatrix_row_inner.
template <class T> void walk(const Hit<T>& hit, const Gauss<T>* g, const int64_t n[3], Row<T>& row) {
  T next[3] = {hit.next[0], hit.next[1], hit.next[2]};
  int32_t index = hit.index;
  int32_t remaining[3] = {hit.remaining[0], hit.remaining[1], hit.remaining[2]};
  T here = T(0);
  const uint64_t cap = step_cap(n);
  for (;;) {
    if (row.steps >= cap) {
      row.runaway = true;
      break;
    }
    ++row.steps;
    const int d = argmin3(next);
    const T bnd = next[d];
    T weight = bnd - here;
    if (g) {
      const T gv = gauss_call(*g, here - hit.tof_peak);
      weight *= K<T>::TOF_FACTOR * gv;
    }
    if (weight > T(0)) {
      row.idx.push_back((int64_t)index);
      row.w.push_back(weight);
    }
    if (row.record) {
      row.axis.push_back((uint8_t)d);
      row.push.push_back(weight > T(0) ? 1 : 0);
    }
    here = bnd;
    next[d] += hit.vs[d];
    index = (int32_t)((uint32_t)index + (uint32_t)hit.delta[d]);
    remaining[d] -= 1;
    if (remaining[d] == 0) break;
  }
}

template <class T> bool make_row(const orc_lor_t& lor, const Fov<T>& f, const Gauss<T>* g, Row<T>& row) {
  row.clear();
  Hit<T> h;
  if (!lor_fov_hit(lor, f, h)) return false;
  row.shifted = h.shifted;
  walk(h, g, f.n, row);
  return true;
}

// "binary32 walk decisions, binary64 arithmetic": the binary32 walk (the
// reference's) decides everything discrete -- whether the LOR hits, the first
// voxel, how far into it the walk starts (the binary32 fraction frac, which
// carries the entry quirk of Hit::shifted), the crossed axis of every step,
// which steps push an element and when the walk ends -- and the weights are
// then evaluated in binary64 along that same step sequence: binary64 voxel
// lengths along the LOR, boundaries accumulated in binary64 from
// (1 - frac) * length, the binary64 TOF peak and Gaussian.  Images from these
// rows differ from the binary32 oracle's by rounding alone and from the
// binary64 oracle's by the walk's discrete choices alone.
bool make_row_mixed(const orc_lor_t& lor, const Fov<float>& f32, const Fov<double>& f64, const Gauss<double>* g,
                    const Gauss<float>* g32, Row<double>& row, Row<float>& scratch) {
  row.clear();
  Hit<float> h;
  if (!lor_fov_hit(lor, f32, h)) return false;
  scratch.clear();
  scratch.record = true;
  walk(h, g32, f32.n, scratch);
  row.steps = scratch.steps;
  row.runaway = scratch.runaway;
  row.shifted = h.shifted;
  // binary64 geometry of the flipped LOR
  double p1[3], p2[3];
  for (int i = 0; i < 3; ++i) {
    p1[i] = lor.p1[i];
    p2[i] = lor.p2[i];
    if (h.delta[i] < 0) {  // flipped in binary32 (flip_axes); the same here
      p1[i] = -p1[i];
      p2[i] = -p2[i];
    }
  }
  const double d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double len = norm3(d[0], d[1], d[2]);
  double vl[3], next[3];
  for (int i = 0; i < 3; ++i) {
    vl[i] = f64.vs[i] / (d[i] / len);
    next[i] = (1.0 - (double)h.frac[i]) * vl[i];
  }
  double entry[3];
  if (!fov_entry(f64, p1, p2, entry))
    for (int i = 0; i < 3; ++i) entry[i] = h.entry[i];
  const double half = norm3(p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]) / 2.0;
  const double tof_peak = (half - K<double>::C * (double)lor.dt_ps / 2.0) -
                          norm3(entry[0] - p1[0], entry[1] - p1[1], entry[2] - p1[2]);
  double here = 0.0;
  int32_t index = h.index;
  for (size_t k = 0; k < scratch.axis.size(); ++k) {
    const int a = scratch.axis[k];
    const double bnd = next[a];
    double w = bnd - here;
    if (g) w *= K<double>::TOF_FACTOR * gauss_call(*g, here - tof_peak);
    if (scratch.push[k]) {
      row.idx.push_back((int64_t)index);
      row.w.push_back(w);
    }
    here = bnd;
    next[a] += vl[a];
    index = (int32_t)((uint32_t)index + (uint32_t)h.delta[a]);
  }
  return true;
}

// The binary32 FOV and Gaussian behind a mixed-precision projection.
struct Mixed {
  Fov<float> f;
  const Gauss<float>* g;
};

// projector.rs:104-116: `index as usize >= len` (negative i32 -> huge usize).
template <class T> bool row_in_range(const Row<T>& r, int64_t len) {
  for (int64_t i : r.idx)
    if (i < 0 || i >= len) return false;
  return true;
}

template <class T> void count(orc_stats_t& s, const Row<T>& r, bool hit) {
  s.n_lors += 1;
  s.n_steps += r.steps;
  s.n_nz += r.idx.size();
  if (!hit) s.n_missed += 1;
  if (r.runaway) s.n_runaway += 1;
  if (hit && r.shifted) s.n_shifted += 1;
}

template <class T> T mlem_adapt(T p, const orc_lor_t& lor) { return p * T(lor.additive_correction); }
template <class T> T sens_adapt(T p, const orc_lor_t&) { return std::exp(-p); }

// project_lors (projector.rs:33-63) with 10 000-LOR jobs (mlem.rs:39,57-66):
// job c goes to thread c % n_threads; each thread folds into a private
// backprojection; the private buffers are summed in thread order.
// With mx (T = double only), rows come from make_row_mixed.
template <class T, class Adapt>
void project_lors(const Fov<T>& ffwd, const Fov<T>* fbck_opt, const Gauss<T>* g, const orc_lor_t* lors,
                  size_t n_lors, const T* image, int n_threads, Adapt adapt, std::vector<T>& bp_out,
                  orc_stats_t& stats, const Mixed* mx = nullptr) {
  const Fov<T>& fbck = fbck_opt ? *fbck_opt : ffwd;
  const size_t chunk = 10000;
  const size_t n_chunks = (n_lors + chunk - 1) / chunk;
  if (n_threads < 1) n_threads = 1;
  if ((size_t)n_threads > n_chunks) n_threads = (int)std::max<size_t>(1, n_chunks);
  const int64_t fwd_len = ffwd.nvox(), bck_len = fbck.nvox();
  std::vector<std::vector<T>> bps(n_threads, std::vector<T>((size_t)bck_len, T(0)));
  std::vector<orc_stats_t> st(n_threads);
  for (auto& s : st) std::memset(&s, 0, sizeof s);
  auto work = [&](int t) {
    Row<T> rf, rb;
    Row<float> scratch;
    std::vector<T>& bp = bps[t];
    for (size_t c = (size_t)t; c < n_chunks; c += (size_t)n_threads) {
      const size_t lo = c * chunk, hi = std::min(n_lors, lo + chunk);
      for (size_t i = lo; i < hi; ++i) {
        const orc_lor_t& lor = lors[i];
        bool hit;
        if constexpr (std::is_same<T, double>::value) {
          hit = mx ? make_row_mixed(lor, mx->f, ffwd, g, mx->g, rf, scratch) : make_row(lor, ffwd, g, rf);
        } else {
          hit = make_row(lor, ffwd, g, rf);
        }
        count(st[t], rf, hit);
        const Row<T>* rbp = &rf;
        if (fbck_opt) {
          make_row(lor, fbck, g, rb);
          rbp = &rb;
        }
        bool ok = !rf.runaway && !rbp->runaway && row_in_range(*rbp, bck_len);
        if (ok && fbck_opt) ok = row_in_range(rf, fwd_len);
        if (!ok) {
          st[t].n_skipped += 1;
          continue;
        }
        // forward_project :135-141
        T p = T(0);
        for (size_t k = 0; k < rf.idx.size(); ++k) p += rf.w[k] * image[rf.idx[k]];
        const T a = adapt(p, lor);
        // back_project :144-149
        const T r = T(1) / a;
        for (size_t k = 0; k < rbp->idx.size(); ++k) bp[rbp->idx[k]] += rbp->w[k] * r;
      }
    }
  };
  if (n_threads == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
  }
  bp_out = std::move(bps[0]);
  for (int t = 1; t < n_threads; ++t)
    for (int64_t j = 0; j < bck_len; ++j) bp_out[j] = bp_out[j] + bps[t][j];
  for (const auto& s : st) {
    stats.n_lors += s.n_lors;
    stats.n_steps += s.n_steps;
    stats.n_nz += s.n_nz;
    stats.n_skipped += s.n_skipped;
    stats.n_missed += s.n_missed;
    stats.n_runaway += s.n_runaway;
    stats.n_shifted += s.n_shifted;
  }
}

template <class T>
int mlem_update(const float full[3], const uint32_t n[3], const orc_tof_t* tof, const orc_lor_t* lors,
                size_t n_lors, const T* sens, T* image, int n_threads, orc_stats_t* stats, bool mixed = false) {
  const Fov<T> f = fov_new<T>(full, n);
  Gauss<T> g;
  if (tof) g = gauss_new<T>(T(tof->sigma_ps), T(tof->cutoff));
  Gauss<float> g32;
  if (tof) g32 = gauss_new<float>(tof->sigma_ps, tof->cutoff);
  const Mixed mx{fov_new<float>(full, n), tof ? &g32 : nullptr};
  orc_stats_t st;
  std::memset(&st, 0, sizeof st);
  std::vector<T> bp;
  project_lors<T>(f, nullptr, tof ? &g : nullptr, lors, n_lors, image, n_threads, mlem_adapt<T>, bp, st,
                  mixed ? &mx : nullptr);
  // apply_sensitivity_image mlem.rs:45-52
  const int64_t nv = f.nvox();
  for (int64_t j = 0; j < nv; ++j) {
    const T s = sens ? sens[j] : T(1);
    if (s > T(0)) image[j] *= bp[j] * s;
    else image[j] = T(0);
  }
  if (stats) *stats = st;
  return 0;
}

template <class T>
int sensitivity(const float att_full[3], const uint32_t att_n[3], const float* mu, const float* out_full,
                const uint32_t* out_n, const orc_lor_t* lors, size_t n_lors, int n_threads, T* out,
                orc_stats_t* stats) {
  const Fov<T> fa = fov_new<T>(att_full, att_n);
  Fov<T> fo;
  const bool two = out_full && out_n;
  if (two) fo = fov_new<T>(out_full, out_n);
  const int64_t na = fa.nvox();
  std::vector<T> muT((size_t)na);
  for (int64_t j = 0; j < na; ++j) muT[j] = T(mu[j]);
  orc_stats_t st;
  std::memset(&st, 0, sizeof st);
  std::vector<T> bp;
  project_lors<T>(fa, two ? &fo : nullptr, nullptr, lors, n_lors, muT.data(), n_threads, sens_adapt<T>, bp, st);
  // normalize main.rs:77: e /= n as f32
  const T nl = T(n_lors);
  for (size_t j = 0; j < bp.size(); ++j) out[j] = bp[j] / nl;
  if (stats) *stats = st;
  return 0;
}

template <class T>
int64_t row_impl(const float full_mm[3], const uint32_t n[3], const orc_tof_t* tof, const orc_lor_t* lor,
                        int64_t* idx, T* w, int64_t cap, uint64_t* steps) {
  const Fov<T> f = fov_new<T>(full_mm, n);
  Gauss<T> g;
  if (tof) g = gauss_new<T>(T(tof->sigma_ps), T(tof->cutoff));
  Row<T> r;
  make_row(*lor, f, tof ? &g : nullptr, r);
  const int64_t m = (int64_t)r.idx.size();
  for (int64_t k = 0; k < m && k < cap; ++k) {
    idx[k] = r.idx[k];
    w[k] = r.w[k];
  }
  if (steps) *steps = r.steps;
  return r.runaway ? -1 - m : m;
}

}  // namespace

// ============================ C ABI ========================================
extern "C" {

void orc_fov_derive_f32(const float full_mm[3], const uint32_t n[3], float half_width[3], float voxel_size[3]) {
  const Fov<float> f = fov_new<float>(full_mm, n);
  for (int i = 0; i < 3; ++i) {
    half_width[i] = f.hw[i];
    voxel_size[i] = f.vs[i];
  }
}

int orc_fov_entry_f32(const float full_mm[3], const uint32_t n[3], const float p1[3], const float p2[3],
                      float entry[3]) {
  const Fov<float> f = fov_new<float>(full_mm, n);
  return fov_entry(f, p1, p2, entry) ? 1 : 0;
}

float orc_gauss_f32(float sigma_ps, float cutoff, float dx_mm) {
  return gauss_call(gauss_new<float>(sigma_ps, cutoff), dx_mm);
}

double orc_gauss_f64(double sigma_ps, double cutoff, double dx_mm) {
  return gauss_call(gauss_new<double>(sigma_ps, cutoff), dx_mm);
}

int64_t orc_row_f32(const float full_mm[3], const uint32_t n[3], const orc_tof_t* tof, const orc_lor_t* lor,
                    int64_t* idx, float* w, int64_t cap, uint64_t* steps) {
  return row_impl<float>(full_mm, n, tof, lor, idx, w, cap, steps);
}

int64_t orc_row_f64(const float full_mm[3], const uint32_t n[3], const orc_tof_t* tof, const orc_lor_t* lor,
                    int64_t* idx, double* w, int64_t cap, uint64_t* steps) {
  return row_impl<double>(full_mm, n, tof, lor, idx, w, cap, steps);
}

// Classic Siddon (R. L. Siddon, Med. Phys. 12 (1985) 252): parametric line
// P(a) = p1 + a (p2 - p1), a in [0,1]; merge the a-values of all plane
// crossings inside [a_min, a_max]; each consecutive pair is one voxel,
// identified from the segment midpoint.  Written independently of the
// reference's incremental walk.
int64_t orc_classic_siddon_f64(const double full_mm[3], const uint32_t n[3], const double p1[3],
                               const double p2[3], int64_t* idx, double* len, int64_t cap) {
  double hw[3], vs[3], d[3];
  for (int i = 0; i < 3; ++i) {
    hw[i] = full_mm[i] / 2;
    vs[i] = full_mm[i] / n[i];
    d[i] = p2[i] - p1[i];
  }
  double amin = 0, amax = 1;
  for (int i = 0; i < 3; ++i) {
    if (d[i] == 0) {
      if (p1[i] < -hw[i] || p1[i] > hw[i]) return 0;
      continue;
    }
    double a = (-hw[i] - p1[i]) / d[i], b = (hw[i] - p1[i]) / d[i];
    if (a > b) std::swap(a, b);
    amin = std::max(amin, a);
    amax = std::min(amax, b);
  }
  if (amin >= amax) return 0;
  std::vector<double> as{amin, amax};
  for (int i = 0; i < 3; ++i) {
    if (d[i] == 0) continue;
    for (uint32_t k = 0; k <= n[i]; ++k) {
      const double x = -hw[i] + k * vs[i];
      const double a = (x - p1[i]) / d[i];
      if (a > amin && a < amax) as.push_back(a);
    }
  }
  std::sort(as.begin(), as.end());
  const double L = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  int64_t m = 0;
  for (size_t k = 0; k + 1 < as.size(); ++k) {
    const double a0 = as[k], a1 = as[k + 1];
    if (!(a1 > a0)) continue;
    const double am = 0.5 * (a0 + a1);
    int64_t ij[3];
    bool in = true;
    for (int i = 0; i < 3; ++i) {
      const double x = p1[i] + am * d[i];
      ij[i] = (int64_t)std::floor((x + hw[i]) / vs[i]);
      if (ij[i] < 0 || ij[i] >= (int64_t)n[i]) in = false;
    }
    if (!in) continue;
    if (m < cap) {
      idx[m] = ij[0] + (ij[1] + ij[2] * (int64_t)n[1]) * (int64_t)n[0];
      len[m] = (a1 - a0) * L;
    }
    ++m;
  }
  return m;
}

static double clip_len(const double lo[3], const double hi[3], const double p1[3], const double p2[3]) {
  double d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  double a0 = 0, a1 = 1;
  for (int i = 0; i < 3; ++i) {
    if (d[i] == 0) {
      if (p1[i] < lo[i] || p1[i] > hi[i]) return 0;
      continue;
    }
    double a = (lo[i] - p1[i]) / d[i], b = (hi[i] - p1[i]) / d[i];
    if (a > b) std::swap(a, b);
    a0 = std::max(a0, a);
    a1 = std::min(a1, b);
  }
  if (a1 <= a0) return 0;
  return (a1 - a0) * std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

void orc_bruteforce_lengths_f64(const double full_mm[3], const uint32_t n[3], const double p1[3],
                                const double p2[3], double* out) {
  double hw[3], vs[3];
  for (int i = 0; i < 3; ++i) {
    hw[i] = full_mm[i] / 2;
    vs[i] = full_mm[i] / n[i];
  }
  for (uint32_t z = 0; z < n[2]; ++z)
    for (uint32_t y = 0; y < n[1]; ++y)
      for (uint32_t x = 0; x < n[0]; ++x) {
        const uint32_t ix[3] = {x, y, z};
        double lo[3], hi[3];
        for (int i = 0; i < 3; ++i) {
          lo[i] = -hw[i] + ix[i] * vs[i];
          hi[i] = -hw[i] + (ix[i] + 1) * vs[i];
        }
        out[x + (y + (size_t)z * n[1]) * n[0]] = clip_len(lo, hi, p1, p2);
      }
}

double orc_chord_f64(const double full_mm[3], const double p1[3], const double p2[3]) {
  double lo[3], hi[3];
  for (int i = 0; i < 3; ++i) {
    lo[i] = -full_mm[i] / 2;
    hi[i] = full_mm[i] / 2;
  }
  return clip_len(lo, hi, p1, p2);
}

int orc_mlem_update_f32(const float full_mm[3], const uint32_t n[3], const orc_tof_t* tof, const orc_lor_t* lors,
                        size_t n_lors, const float* sensitivity, float* image, int n_threads, orc_stats_t* stats) {
  return mlem_update<float>(full_mm, n, tof, lors, n_lors, sensitivity, image, n_threads, stats);
}

int orc_mlem_update_f64(const float full_mm[3], const uint32_t n[3], const orc_tof_t* tof, const orc_lor_t* lors,
                        size_t n_lors, const double* sensitivity, double* image, int n_threads, orc_stats_t* stats) {
  return mlem_update<double>(full_mm, n, tof, lors, n_lors, sensitivity, image, n_threads, stats);
}

int orc_mlem_update_mixed(const float full_mm[3], const uint32_t n[3], const orc_tof_t* tof, const orc_lor_t* lors,
                          size_t n_lors, const double* sensitivity, double* image, int n_threads,
                          orc_stats_t* stats) {
  return mlem_update<double>(full_mm, n, tof, lors, n_lors, sensitivity, image, n_threads, stats, true);
}

int64_t orc_row_mixed(const float full_mm[3], const uint32_t n[3], const orc_tof_t* tof, const orc_lor_t* lor,
                      int64_t* idx, double* w, int64_t cap, uint64_t* steps) {
  const Fov<float> f32 = fov_new<float>(full_mm, n);
  const Fov<double> f64 = fov_new<double>(full_mm, n);
  Gauss<double> g;
  Gauss<float> g32;
  if (tof) {
    g = gauss_new<double>(tof->sigma_ps, tof->cutoff);
    g32 = gauss_new<float>(tof->sigma_ps, tof->cutoff);
  }
  Row<double> r;
  Row<float> scratch;
  make_row_mixed(*lor, f32, f64, tof ? &g : nullptr, tof ? &g32 : nullptr, r, scratch);
  const int64_t m = (int64_t)r.idx.size();
  for (int64_t k = 0; k < m && k < cap; ++k) {
    idx[k] = r.idx[k];
    w[k] = r.w[k];
  }
  if (steps) *steps = r.steps;
  return r.runaway ? -1 - m : m;
}

// find_lors / random_point_on_cylinder (continuous.rs:20-47) as the library
// draws them: attempt k of LOR i takes Philox4x32-10 block (i, k, 0, 0x5e45)
// under key (seed lo, seed hi); uniforms are the top 24 bits / 2^24;
// p1 = (R cos t1, R sin t1, z1), p2 likewise, z = L (u - 1/2), t = TAU u,
// with (z1, t1, z2, t2) from the block's four words; a pair is kept when
// FOV::entry finds the box (unflipped endpoints, as find_lors calls it).
int64_t orc_sensitivity_lors(const float full_mm[3], const uint32_t n[3], size_t n_lors, uint64_t seed, float L,
                             float D, orc_lor_t* out, uint32_t* attempts) {
  const Fov<float> f = fov_new<float>(full_mm, n);
  const float R = D / 2.0f;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  int64_t found_all = 0;
  for (size_t i = 0; i < n_lors; ++i) {
    bool found = false;
    for (uint32_t k = 0; k < 4096 && !found; ++k) {
      const uint32_t ctr[4] = {(uint32_t)i, k, 0u, 0x5e45u};
      uint32_t r[4];
      orc_philox4x32_10(ctr, key, r);
      auto u = [](uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); };
      const float z1 = L * (u(r[0]) - 0.5f), t1 = K<float>::TAU * u(r[1]);
      const float z2 = L * (u(r[2]) - 0.5f), t2 = K<float>::TAU * u(r[3]);
      const float p1[3] = {R * std::cos(t1), R * std::sin(t1), z1};
      const float p2[3] = {R * std::cos(t2), R * std::sin(t2), z2};
      float e[3];
      if (fov_entry(f, p1, p2, e)) {
        out[i] = orc_lor_t{{p1[0], p1[1], p1[2]}, {p2[0], p2[1], p2[2]}, 0.0f, 1.0f};
        if (attempts) attempts[i] = k + 1;
        found = true;
      }
    }
    if (!found) {
      out[i] = orc_lor_t{{0, 0, 0}, {0, 0, 0}, 0.0f, 1.0f};
      if (attempts) attempts[i] = 0;
    } else {
      ++found_all;
    }
  }
  return found_all;
}

int orc_set_entry_variant(int variant) {
  if (variant < 0 || variant > 1) return -1;
  g_entry_variant = variant;
  return 0;
}

int orc_entry_shifted_f32(const float full_mm[3], const uint32_t n[3], const orc_lor_t* lor) {
  const Fov<float> f = fov_new<float>(full_mm, n);
  Hit<float> h;
  if (!lor_fov_hit(*lor, f, h)) return -1;
  return h.shifted ? 1 : 0;
}

int orc_backproject_mlem_f32(const float full_mm[3], const uint32_t n[3], const orc_tof_t* tof,
                             const orc_lor_t* lors, size_t n_lors, const float* image, float* bp, int n_threads,
                             orc_stats_t* stats) {
  const Fov<float> f = fov_new<float>(full_mm, n);
  Gauss<float> g;
  if (tof) g = gauss_new<float>(tof->sigma_ps, tof->cutoff);
  orc_stats_t st;
  std::memset(&st, 0, sizeof st);
  std::vector<float> b;
  project_lors<float>(f, nullptr, tof ? &g : nullptr, lors, n_lors, image, n_threads, mlem_adapt<float>, b, st);
  std::copy(b.begin(), b.end(), bp);
  if (stats) *stats = st;
  return 0;
}

int orc_sensitivity_f32(const float att_full_mm[3], const uint32_t att_n[3], const float* mu,
                        const float* out_full_mm, const uint32_t* out_n, const orc_lor_t* lors, size_t n_lors,
                        int n_threads, float* out, orc_stats_t* stats) {
  return sensitivity<float>(att_full_mm, att_n, mu, out_full_mm, out_n, lors, n_lors, n_threads, out, stats);
}

int orc_sensitivity_f64(const float att_full_mm[3], const uint32_t att_n[3], const float* mu,
                        const float* out_full_mm, const uint32_t* out_n, const orc_lor_t* lors, size_t n_lors,
                        int n_threads, double* out, orc_stats_t* stats) {
  return sensitivity<double>(att_full_mm, att_n, mu, out_full_mm, out_n, lors, n_lors, n_threads, out, stats);
}

// main.rs:11-17 turns the CLI value (cm^2/g) into an AreaPerMass in base
// units (mm, kg); main.rs:80-94 divides it by (1/mm)/(kg/m^3).
float orc_rho_to_mu_factor_f32(float rho_to_mu) {
  const float g = 0.001f;  // kg(0.001)
  const float cm = 10.0f;  // mm(10.0)
  const float rho_unit = g / (cm * cm * cm);
  const float mu_unit = 1.0f / cm;
  const float apm = rho_to_mu * (mu_unit / rho_unit);
  const float kg = 1.0f;
  const float m = 1000.0f;
  const float rho_unit2 = kg / (m * m * m);
  const float mu_unit2 = 1.0f / 1.0f;
  return apm / (mu_unit2 / rho_unit2);
}

int orc_walk_stats_f32(const float full_mm[3], const uint32_t n[3], const orc_tof_t* tof, const orc_lor_t* lors,
                       size_t n_lors, int n_threads, orc_stats_t* stats) {
  const Fov<float> f = fov_new<float>(full_mm, n);
  Gauss<float> g;
  if (tof) g = gauss_new<float>(tof->sigma_ps, tof->cutoff);
  if (n_threads < 1) n_threads = 1;
  std::vector<orc_stats_t> st(n_threads);
  for (auto& s : st) std::memset(&s, 0, sizeof s);
  auto work = [&](int t) {
    Row<float> r;
    for (size_t i = (size_t)t; i < n_lors; i += (size_t)n_threads) {
      bool hit = make_row(lors[i], f, tof ? &g : nullptr, r);
      count(st[t], r, hit);
      if (r.runaway || !row_in_range(r, f.nvox())) st[t].n_skipped += 1;
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < n_threads; ++t) th.emplace_back(work, t);
  for (auto& x : th) x.join();
  std::memset(stats, 0, sizeof *stats);
  for (const auto& s : st) {
    stats->n_lors += s.n_lors;
    stats->n_steps += s.n_steps;
    stats->n_nz += s.n_nz;
    stats->n_skipped += s.n_skipped;
    stats->n_missed += s.n_missed;
    stats->n_runaway += s.n_runaway;
    stats->n_shifted += s.n_shifted;
  }
  return 0;
}


// ---------------------------------------------------------------------------
// LOR ingestion

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  // Salmon, Moraes, Dror, Shaw (SC'11): 10 rounds of the Philox S-box with
  // multipliers 0xD2511F53, 0xCD9E8D57 and Weyl key increments 0x9E3779B9,
  // 0xBB67AE85.
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    const uint64_t a = (uint64_t)0xD2511F53u * c0;
    const uint64_t b = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)b;
    const uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)a;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

}  // extern "C"

namespace {

struct Rec { float dt, x1, y1, z1, x2, y2, z2, q1, q2, e1, e2; };

float uni(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

void draw(uint64_t seed, uint64_t rec, uint32_t block, uint32_t tag, float u[4]) {
  const uint32_t ctr[4] = {(uint32_t)rec, (uint32_t)(rec >> 32), block, tag};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32_10(ctr, key, o);
  for (int k = 0; k < 4; ++k) u[k] = uni(o[k]);
}

bool in_bounds(int32_t has_min, float mn, int32_t has_max, float mx, float v) {
  // config/mlem.rs:141-146
  return (!has_min || mn <= v) && (!has_max || mx > v);
}

// discrete.rs:170-187 and 78-118, binary32 in the reference's order
struct Disc { float dr, dz, n_radial, d_azimuthal; int adjust; };

Disc disc_of(const orc_ingest_t& p) {
  const float tau = 6.2831855f;
  Disc d;
  d.dr = p.dr_mm;
  d.dz = p.dz_mm;
  const float r = p.r_min_mm + (p.dr_mm / 2.0f);
  d.n_radial = r / p.dr_mm;
  const float mid = tau * r;
  const float n_az = std::round(mid / p.da_mm);
  d.d_azimuthal = tau / n_az;
  d.adjust = p.adjust;
  return d;
}

void adjust_end(const Disc& d, float& x, float& y, float& z, float uz, float uphi, float ur) {
  if (d.adjust == 0) return;
  const int32_t n_phi = (int32_t)std::round(std::atan2(y, x) / d.d_azimuthal);
  const int32_t n_z = (int32_t)std::round(z / d.dz);
  float r, phi;
  if (d.adjust == 1) {
    r = d.n_radial * d.dr;
    z = (float)n_z * d.dz;
    phi = (float)n_phi * d.d_azimuthal;
  } else {
    z = (((float)n_z - 0.5f) + uz) * d.dz;
    phi = (((float)n_phi - 0.5f) + uphi) * d.d_azimuthal;
    r = d.adjust == 2 ? ((d.n_radial - 0.5f) + ur) * d.dr : std::hypot(x, y);
  }
  x = r * std::cos(phi);
  y = r * std::sin(phi);
}

// ndhistogram 0.9.0 axes as restated in the library's header
struct Ax {
  uint32_t n; bool cyclic; float low, high;
  uint64_t stride() const { return cyclic ? n : (uint64_t)n + 2; }
  int64_t index(float x) const {
    const float f = (x - low) / (high - low);
    if (std::isnan(f)) return -1;
    if (cyclic) {
      const float g = f - std::floor(f);
      int64_t i = (int64_t)(g * (float)n);
      return i >= (int64_t)n ? (int64_t)n - 1 : i;
    }
    if (f < 0.0f) return 0;
    if (f >= 1.0f) return (int64_t)n + 1;
    int64_t i = (int64_t)(f * (float)n);
    return (i >= (int64_t)n ? (int64_t)n - 1 : i) + 1;
  }
};

}  // namespace

extern "C" {

int64_t orc_discrete_lors(float r_min, float dr, float dz, float da, float length, int adjust, uint64_t seed,
                          orc_lor_t* out, size_t capacity, uint64_t* n_elements) {
  orc_ingest_t p{};
  p.r_min_mm = r_min;
  p.dr_mm = dr;
  p.dz_mm = dz;
  p.da_mm = da;
  p.adjust = adjust;
  const Disc d = disc_of(p);
  // centre_all_elements (discrete.rs:150-171)
  const float tau = 6.2831855f;
  const float rr = r_min + (dr / 2.0f);
  const uint32_t n_az = (uint32_t)std::round((tau * rr) / da);
  const float r = d.n_radial * dr;
  const int32_t n_half = (int32_t)(uint32_t)std::round((length / 2.0f) / dz);
  std::vector<std::array<float, 3>> pts;
  for (uint32_t a = 0; a < n_az; ++a) {
    const float phi = d.d_azimuthal * (float)a;
    for (int32_t nz = -n_half; nz <= n_half; ++nz) pts.push_back({r * std::cos(phi), r * std::sin(phi), (float)nz * dz});
  }
  const uint64_t n = pts.size();
  if (n_elements) *n_elements = n;
  uint64_t t = 0;
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t k = i; k < n; ++k, ++t) {
      if (t >= capacity) continue;
      float pe[3] = {pts[k][0], pts[k][1], pts[k][2]}, qe[3] = {pts[i][0], pts[i][1], pts[i][2]};
      float a4[4], b4[4];
      draw(seed, t, 1, 0x9a1bu, a4);
      draw(seed, t, 2, 0x9a1bu, b4);
      adjust_end(d, pe[0], pe[1], pe[2], a4[0], a4[1], b4[0]);
      adjust_end(d, qe[0], qe[1], qe[2], a4[2], a4[3], b4[1]);
      out[t] = orc_lor_t{{pe[0], pe[1], pe[2]}, {qe[0], qe[1], qe[2]}, 0.0f, 1.0f};
    }
  return (int64_t)t;
}

int64_t orc_ingest_f32(const float* records, size_t n_records, const orc_ingest_t* pp,
                       const orc_sgram_spec_t* sp, orc_lor_t* out, uint64_t* n_read) {
  const orc_ingest_t& p = *pp;
  const uint64_t lo = p.has_events_min ? std::min<uint64_t>(p.events_min, n_records) : 0;
  const uint64_t hi = p.has_events_max ? std::min<uint64_t>(p.events_max, n_records) : n_records;
  const uint64_t n = hi > lo ? hi - lo : 0;
  if (n_read) *n_read = n;
  const Disc d = disc_of(p);
  // hdf5.rs:108: 511 * (1 - 1.6 fwhm)
  const float pre = 511.0f * (1.0f - 1.6f * p.smear_energy_fwhm);
  std::vector<Rec> kept;
  for (uint64_t k = 0; k < n; ++k) {
    const uint64_t i = lo + k;
    Rec r;
    std::memcpy(&r, records + 11 * i, sizeof r);
    if (p.has_z_max && !(std::fabs(r.z1) < p.z_max_mm && std::fabs(r.z2) < p.z_max_mm)) continue;
    if (!(in_bounds(p.has_charge_min, p.charge_min, p.has_charge_max, p.charge_max, r.q1) &&
          in_bounds(p.has_charge_min, p.charge_min, p.has_charge_max, p.charge_max, r.q2))) continue;
    if (!(pre < r.e1 && pre < r.e2)) continue;
    float ue[4];
    draw(p.seed, i, 0, 0xe6e7u, ue);
    // hdf5.rs:230-238: Normal(e, fwhm*e/2.35); standard normal by Box-Muller
    auto smear_e = [&](float e, float u1, float u2) {
      const float zn = std::sqrt(-2.0f * std::log(1.0f - u1)) * std::cos(6.2831855f * u2);
      const float f = p.smear_energy_fwhm * e;
      const float sigma = f / 2.35f;
      return e + sigma * zn;
    };
    r.e1 = smear_e(r.e1, ue[0], ue[1]);
    r.e2 = smear_e(r.e2, ue[2], ue[3]);
    if (!(in_bounds(p.has_energy_min, p.energy_min, p.has_energy_max, p.energy_max, r.e1) &&
          in_bounds(p.has_energy_min, p.energy_min, p.has_energy_max, p.energy_max, r.e2))) continue;
    float a[4], b[4];
    draw(p.seed, i, 1, 0x9051u, a);
    draw(p.seed, i, 2, 0x9051u, b);
    adjust_end(d, r.x1, r.y1, r.z1, a[0], a[1], b[0]);
    adjust_end(d, r.x2, r.y2, r.z2, a[2], a[3], b[1]);
    kept.push_back(r);
  }
  // scattergram (lorogram.rs, build_scattergram.rs:60-158)
  bool sg = false;
  Ax ax[5];
  if (sp) {
    const bool hz = sp->z_bins || sp->z_length_mm > 0, hdz = sp->dz_bins || sp->dz_max_mm > 0,
               hr = sp->r_bins || sp->r_max_mm > 0, hdt = sp->dt_bins || sp->dt_max_ps > 0;
    sg = sp->phi_bins || hz || hdz || hr || hdt;
    const float l = 66666666.6f, t = 66666666.6f;
    auto nb = [](uint32_t v, bool on) { return on ? (v ? v : 30u) : 1u; };
    const float zl = hz ? (sp->z_length_mm > 0 ? sp->z_length_mm : 200.0f) : l;
    ax[0] = Ax{sp->phi_bins ? sp->phi_bins : 1u, true, 0.0f, 6.2831855f};
    ax[1] = Ax{nb(sp->z_bins, hz), false, -(zl / 2.0f), zl / 2.0f};
    ax[2] = Ax{nb(sp->dz_bins, hdz), false, 0.0f, hdz ? (sp->dz_max_mm > 0 ? sp->dz_max_mm : 1000.0f) : l};
    ax[3] = Ax{nb(sp->r_bins, hr), false, 0.0f, hr ? (sp->r_max_mm > 0 ? sp->r_max_mm : 30.0f) : l};
    const float tm = hdt ? (sp->dt_max_ps > 0 ? sp->dt_max_ps : 1700.0f) : t;
    ax[4] = Ax{nb(sp->dt_bins, hdt), false, -tm, tm};
  }
  auto bin_of = [&](const Rec& r) -> int64_t {
    // lorogram.rs:166-221 (the dt axis is indexed by the distance from the axis)
    const float dx = r.x2 - r.x1, dy = r.y2 - r.y1;
    const float num = dx * r.y1 - dy * r.x1;
    const float den = std::sqrt(dx * dx + dy * dy);
    float phi = std::atan2(dy, dx);
    if (num / den < 0.0f) phi = phi + 3.1415927f;
    const float dist = std::fabs(num) / den;
    const float c[5] = {phi, (r.z1 + r.z2) / 2.0f, std::fabs(r.z1 - r.z2), dist, dist};
    int64_t b = 0;
    for (int k = 0; k < 5; ++k) {
      const int64_t i = ax[k].index(c[k]);
      if (i < 0) return -1;
      b = b * (int64_t)ax[k].stride() + i;
    }
    return b;
  };
  std::unordered_map<int64_t, std::pair<uint64_t, uint64_t>> hist;  // bin -> (trues, scatters)
  if (sg)
    for (const Rec& r : kept) {
      const int64_t b = bin_of(r);
      if (b < 0) continue;
      auto& h = hist[b];
      if (std::fmin(r.e1, r.e2) < 510.0f) ++h.second; else ++h.first;
    }
  for (size_t k = 0; k < kept.size(); ++k) {
    const Rec& r = kept[k];
    float ac = 1.0f;
    if (sg) {
      const int64_t b = bin_of(r);
      const auto it = b < 0 ? hist.end() : hist.find(b);
      const uint64_t t = it == hist.end() ? 0 : it->second.first;
      const uint64_t s = it == hist.end() ? 0 : it->second.second;
      ac = t > 0 ? ((float)s + (float)t) / (float)t : 3.40282347e+38f;  // lorogram.rs:65-74
    }
    // LOR::from(Hdf5Lor) (hdf5.rs:298-320): ns -> ps
    out[k] = orc_lor_t{{r.x1, r.y1, r.z1}, {r.x2, r.y2, r.z2}, r.dt * 1000.0f, ac};
  }
  return (int64_t)kept.size();
}
}  // extern "C"
