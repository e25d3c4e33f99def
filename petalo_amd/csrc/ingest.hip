// ingest.hip -- C ABI of the LOR ingestion (petalo_ingest_lors): records ->
// cuts and smearing -> compaction -> scattergram -> MLEM LORs, on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "hip_util.hpp"
#include "ingest.hpp"
#include "siddon_walk.hpp"

using namespace petalo;

namespace {

// Cuts and smearing, one lane per record of the event range.
__global__ __launch_bounds__(256) void k_ingest(const H5Lor* __restrict__ in, uint64_t first, uint32_t n, DevIngest p,
                                                H5Lor* __restrict__ out, uint32_t* __restrict__ keep) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  H5Lor l = in[i];
  const bool k = ingest_record(p, first + i, l);
  out[i] = l;
  keep[i] = k ? 1u : 0u;
}

// Kept records to consecutive positions (pos = exclusive scan of keep).
__global__ __launch_bounds__(256) void k_compact(const H5Lor* __restrict__ in, const uint32_t* __restrict__ keep,
                                                 const uint32_t* __restrict__ pos, uint32_t n,
                                                 H5Lor* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && keep[i]) out[pos[i]] = in[i];
}

// Trues (min E >= 510 keV) and scatters into the scattergram (hdf5.rs:82-86).
__global__ __launch_bounds__(256) void k_sgram_fill(const H5Lor* __restrict__ l, uint32_t n, DevScattergram s,
                                                    uint32_t* __restrict__ trues, uint32_t* __restrict__ scatters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const H5Lor x = l[i];
  const int64_t b = scattergram_bin(s, x);
  if (b < 0) return;
  const bool scatter = fminf(x.e1, x.e2) < 510.0f;
  atomicAdd(scatter ? scatters + b : trues + b, 1u);
}

// LOR::from(Hdf5Lor) (hdf5.rs:298-320): dt ns -> ps; additive correction
// from the scattergram when there is one.
__global__ __launch_bounds__(256) void k_to_lor(const H5Lor* __restrict__ l, uint32_t n, DevScattergram s,
                                                const uint32_t* __restrict__ trues,
                                                const uint32_t* __restrict__ scatters, Lor* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const H5Lor x = l[i];
  float ac = 1.0f;
  if (trues) {
    const int64_t b = scattergram_bin(s, x);
    ac = b < 0 ? scattergram_value(0u, 0u) : scattergram_value(trues[b], scatters[b]);
  }
  out[i] = Lor{x.x1, x.y1, x.z1, x.x2, x.y2, x.z2, x.dt * 1000.0f, ac};
}

// BuildScattergram::build (build_scattergram.rs:60-158); false: none requested.
bool make_scattergram(const petalo_scattergram_spec_t* sp, DevScattergram& s) {
  if (!sp) return false;
  const bool phi = sp->phi_bins > 0;
  const bool z = sp->z_bins > 0 || sp->z_length_mm > 0.0f;
  const bool dz = sp->dz_bins > 0 || sp->dz_max_mm > 0.0f;
  const bool r = sp->r_bins > 0 || sp->r_max_mm > 0.0f;
  const bool dt = sp->dt_bins > 0 || sp->dt_max_ps > 0.0f;
  if (!(phi || z || dz || r || dt)) return false;
  constexpr uint32_t kDefaultBins = 30;
  const float l = 66666666.6f, t = 66666666.6f;  // the builder's "unset" ranges
  auto bins = [&](uint32_t b) { return b > 0 ? b : kDefaultBins; };
  auto range = [](float v, float dflt) { return v > 0.0f ? v : dflt; };
  s.phi = Axis{phi ? sp->phi_bins : 1u, true, 0.0f, kTwoPi};
  const float zl = z ? range(sp->z_length_mm, 200.0f) : l;
  s.z = Axis{z ? bins(sp->z_bins) : 1u, false, -(zl / 2.0f), zl / 2.0f};
  s.dz = Axis{dz ? bins(sp->dz_bins) : 1u, false, 0.0f, dz ? range(sp->dz_max_mm, 1000.0f) : l};
  s.r = Axis{r ? bins(sp->r_bins) : 1u, false, 0.0f, r ? range(sp->r_max_mm, 30.0f) : l};
  const float tm = dt ? range(sp->dt_max_ps, 1700.0f) : t;
  s.dt = Axis{dt ? bins(sp->dt_bins) : 1u, false, -tm, tm};
  s.bins = (uint64_t)s.phi.stride_bins() * s.z.stride_bins() * s.dz.stride_bins() * s.r.stride_bins() *
           s.dt.stride_bins();
  return true;
}

DevIngest make_ingest(const petalo_ingest_t& p) {
  DevIngest d;
  d.charge = BoundsF{p.has_charge_min != 0, p.has_charge_max != 0, p.charge_min, p.charge_max};
  d.energy = BoundsF{p.has_energy_min != 0, p.has_energy_max != 0, p.energy_min, p.energy_max};
  d.fwhm = p.smear_energy_fwhm;
  d.pre_cut = 511.0f * (1.0f - 1.6f * p.smear_energy_fwhm);
  d.has_z_max = p.has_z_max != 0;
  d.z_max = p.z_max_mm;
  d.d = make_discretize(p.r_min_mm, p.dr_mm, p.dz_mm, p.da_mm, p.adjust);
  d.seed = p.seed;
  return d;
}

}  // namespace

extern "C" {

int petalo_ingest_lors(const float* records, size_t n_records, const petalo_ingest_t* params,
                       const petalo_scattergram_spec_t* scatter, petalo_lor_t* out, size_t capacity,
                       size_t* n_out, uint64_t* n_read_out) {
  return guarded([&] {
    if (!params || !n_out || (n_records && !records)) return fail(PETALO_ERR_ARG, "NULL argument");
    const petalo_ingest_t& p = *params;
    if (p.adjust < PETALO_ADJUST_NONE || p.adjust > PETALO_ADJUST_RANDOM_Z_PHI)
      return fail(PETALO_ERR_ARG, "unknown element adjustment");
    if (p.adjust != PETALO_ADJUST_NONE &&
        !(p.dr_mm > 0.0f && p.dz_mm > 0.0f && p.da_mm > 0.0f && p.r_min_mm >= 0.0f))
      return fail(PETALO_ERR_ARG, "element discretization needs positive dr, dz, da");
    if (!(p.smear_energy_fwhm >= 0.0f)) return fail(PETALO_ERR_ARG, "energy FWHM must be >= 0");
    // the event range [min, max) clipped to the records (hdf5.rs:36-66, 103-106)
    const uint64_t lo = p.has_events_min ? std::min<uint64_t>(p.events_min, n_records) : 0;
    const uint64_t hi = p.has_events_max ? std::min<uint64_t>(p.events_max, n_records) : n_records;
    const uint64_t n = hi > lo ? hi - lo : 0;
    if (n >= (1ull << 31)) return fail(PETALO_ERR_ARG, "2^31 or more records in the event range");
    if (n_read_out) *n_read_out = n;
    DevScattergram sg{};
    const bool have_sg = make_scattergram(scatter, sg);
    if (have_sg && sg.bins >= (1ull << 31)) return fail(PETALO_ERR_ARG, "scattergram has 2^31 or more bins");
    *n_out = 0;
    if (n == 0) return PETALO_OK;
    require_device();
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    const uint32_t m = (uint32_t)n;
    DevBuf<H5Lor> rec(m), smeared(m), kept(m);
    DevBuf<uint32_t> keep(m + 1), pos(m + 1);
    upload_h2d(rec.p, records + lo * 11, (size_t)m * sizeof(H5Lor), s);
    k_ingest<<<grid_for(m), kBlock, 0, s>>>(rec.p, lo, m, make_ingest(p), smeared.p, keep.p);
    HIP_CHECK(hipMemsetAsync(keep.p + m, 0, 4, s));
    size_t tmp_bytes = 0;
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, keep.p, pos.p, (int)m + 1, s));
    DevBuf<unsigned char> tmp(std::max<size_t>(tmp_bytes, 1));
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, keep.p, pos.p, (int)m + 1, s));
    k_compact<<<grid_for(m), kBlock, 0, s>>>(smeared.p, keep.p, pos.p, m, kept.p);
    HIP_CHECK(hipGetLastError());
    uint32_t n_kept = 0;
    HIP_CHECK(hipMemcpyAsync(&n_kept, pos.p + m, 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (n_kept > capacity) return fail(PETALO_ERR_SIZE, "output capacity too small");
    if (!out && n_kept) return fail(PETALO_ERR_ARG, "out is NULL");
    DevBuf<uint32_t> trues, scatters;
    if (have_sg && n_kept) {
      trues.alloc(sg.bins);
      scatters.alloc(sg.bins);
      HIP_CHECK(hipMemsetAsync(trues.p, 0, sg.bins * 4, s));
      HIP_CHECK(hipMemsetAsync(scatters.p, 0, sg.bins * 4, s));
      k_sgram_fill<<<grid_for(n_kept), kBlock, 0, s>>>(kept.p, n_kept, sg, trues.p, scatters.p);
    }
    DevBuf<Lor> lors(std::max<uint32_t>(n_kept, 1));
    if (n_kept) {
      k_to_lor<<<grid_for(n_kept), kBlock, 0, s>>>(kept.p, n_kept, sg, have_sg ? trues.p : nullptr,
                                                   have_sg ? scatters.p : nullptr, lors.p);
      HIP_CHECK(hipGetLastError());
      download_d2h(out, lors.p, (size_t)n_kept * sizeof(Lor), s);
    }
    HIP_CHECK(hipStreamSynchronize(s));
    *n_out = n_kept;
    return PETALO_OK;
  });
}

}  // extern "C"
