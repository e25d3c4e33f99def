// capi.hip -- C ABI of the MI355X reconstruction hot path (include/petalo_hip.h).
//
// Device work: the kernels in kernels.hpp.  Host work here: argument checks,
// device memory, the OSEM schedule (src/mlem.rs:70-96) and the LOR ordering
// applied inside each subset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/petalo_hip.h"
#include "capi_internal.hpp"
#include "hip_util.hpp"
#include "kernels.hpp"
#include "tiles.hpp"
#include "walk64.hpp"
#include "ingest.hpp"

#include <hipcub/hipcub.hpp>

using namespace petalo;

namespace petalo {

thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

}  // namespace petalo

namespace {

DevFov to_dev(const petalo_fov_t& f) { return make_fov(f.full_size_mm, f.n); }

uint64_t nvox(const petalo_fov_t& f) { return (uint64_t)f.n[0] * f.n[1] * f.n[2]; }

DevTof to_dev(const petalo_tof_t* t) {
  if (!t || !t->enabled) return no_tof();
  return make_tof(true, t->sigma_ps, t->cutoff);
}

// Tiled scheme: number of work items per subset update (several rounds of
// resident workgroups) and the least work per item, in replayed steps.
constexpr uint64_t kWorkItems = 2048;
constexpr uint64_t kItemStepsPerVoxel = 16;
constexpr uint64_t kMinItemSteps = 8192;
// Largest (tile, LOR block) table of the partial layout; past it the
// partials stay in sorted order.
// LDS entries of k_lor_reduce_win (80 KiB: two workgroups per CU); larger
// LOR-block windows are gathered from memory.
constexpr uint32_t kReduceWindow = 20480;
// Timed subset updates whose kernel timings a context keeps.
constexpr size_t kTimingRing = 256;
// Largest tile (LDS entries, row padding included): its binary64 copy fills
// the 160 KiB of LDS of a CU.
// LDS entries of the largest tile: binary64 tiles fill the 160 KB of LDS but
// for 512 bytes that the tile kernels' own shared variables may take.
constexpr int32_t kMaxTileVoxels = 20416;
// Default tile budget: two workgroups' binary64 tiles per CU.
constexpr int32_t kScatterEntries = 10240;

// LOR order inside a subset, computed on the device: by the image tile that
// holds the LOR's TOF peak (its midpoint without TOF: where its weighted
// voxels cluster), then by the azimuth of its direction (folded to [0, pi)),
// its signed transaxial distance from the axis and its axial midpoint.  LORs
// close in this order are near-parallel neighbours: their walks cross the
// same voxels, so a wavefront's gathers and LDS adds share cache lines and its
// lanes take similar numbers of steps.  Key: tile (bits 48-63, saturated),
// azimuth (38-47), distance (20-37, +-400 mm), midpoint (0-19, +-400 mm).
struct OrderParams {
  double full[3];
  uint32_t n[3];
  int32_t tx, ty, tz, ntx, nty;
  int32_t tof;
};

__device__ __forceinline__ uint64_t quantise(double v, double lo, double hi, int bits) {
  double t = (v - lo) / (hi - lo);
  t = fmin(fmax(t, 0.0), 1.0 - 1e-12);
  return (uint64_t)(t * (double)(1ull << bits));
}

__global__ void k_order_keys(const Lor* __restrict__ lors, uint32_t n, OrderParams p, uint64_t* __restrict__ keys,
                             uint32_t* __restrict__ idx) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const Lor l = lors[k];
  const double d[3] = {(double)l.x2 - l.x1, (double)l.y2 - l.y1, (double)l.z2 - l.z1};
  const double p1[3] = {l.x1, l.y1, l.z1};
  double t = 0.5;
  const double L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (p.tof && L > 0) t = fmin(fmax(0.5 - 0.299792458 * (double)l.dt / (2.0 * L), 0.0), 1.0);
  int32_t v[3];
  for (int a = 0; a < 3; ++a) {
    const double c = p1[a] + t * d[a];
    const double vs = p.full[a] / p.n[a];
    const double x = floor((c + 0.5 * p.full[a]) / vs);
    v[a] = (int32_t)fmin(fmax(x, 0.0), (double)p.n[a] - 1);
  }
  const uint64_t tile = (uint64_t)(v[0] / p.tx) + (uint64_t)p.ntx * ((uint64_t)(v[1] / p.ty) +
                                                                     (uint64_t)p.nty * (uint64_t)(v[2] / p.tz));
  double phi = atan2(d[1], d[0]);
  double ux = d[0], uy = d[1];
  if (phi < 0) {
    phi += M_PI;
    ux = -ux;
    uy = -uy;
  }
  if (phi >= M_PI) phi -= M_PI;
  const double ln = sqrt(ux * ux + uy * uy);
  const double s = ln > 0 ? (p1[0] * uy - p1[1] * ux) / ln : 0.0;  // signed distance
  const double zm = 0.5 * ((double)l.z1 + l.z2);
  keys[k] = (min(tile, (uint64_t)0xffff) << 48) | (quantise(phi, 0.0, M_PI, 10) << 38) |
            (quantise(s, -400.0, 400.0, 18) << 20) | quantise(zm, -400.0, 400.0, 20);
  idx[k] = k;
}

__global__ void k_gather_lors(const Lor* __restrict__ in, const uint32_t* __restrict__ idx, uint32_t n,
                              Lor* __restrict__ out) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) out[r] = in[idx[r]];
}

__global__ void k_lor_ac(const Lor* __restrict__ lors, size_t n, float* __restrict__ ac) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ac[i] = lors[i].ac;
}

// The context's LORs on the device: uploaded once, then (sort) each subset's
// LORs reordered as above; the remainder past the last subset is kept as is.
}  // namespace

namespace petalo {

// Host transfers through pinned staging (declared in capi_internal.hpp): a
// process-wide ring of two pinned buffers, one transfer at a time.
constexpr size_t kStageChunk = 16u << 20;  // bytes per pinned buffer
constexpr size_t kStageMin = 32u << 20;    // smaller copies are plain
struct StageRing {
  std::mutex mu;
  void* pin[2] = {nullptr, nullptr};
};
// never freed: kept for the process (and not released by static destructors,
// which may run after the HIP runtime has shut down)
StageRing* const g_stage = new StageRing();

struct StageEvents {
  hipEvent_t e[2] = {nullptr, nullptr};
  StageEvents() {
    for (hipEvent_t& x : e) HIP_CHECK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  }
  ~StageEvents() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

// memcpy of n bytes split over up to 8 host threads (the caller's included).
void parallel_copy(char* to, const char* from, size_t n) {
  const unsigned hc = std::thread::hardware_concurrency();
  const unsigned n_threads = std::max(1u, std::min(8u, hc ? hc : 1u));
  const size_t per = (n + n_threads - 1) / n_threads;
  std::vector<std::thread> th;
  for (unsigned k = 1; k < n_threads; ++k) {
    const size_t a = std::min(n, (size_t)k * per), e = std::min(n, a + per);
    if (a < e) th.emplace_back([=] { std::memcpy(to + a, from + a, e - a); });
  }
  std::memcpy(to, from, std::min(n, per));
  for (std::thread& t : th) t.join();
}

void upload_h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
  std::unique_lock<std::mutex> lk(g_stage->mu, std::try_to_lock);
  if (bytes < kStageMin || !lk.owns_lock()) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return;
  }
  void** pin = g_stage->pin;
  for (int b = 0; b < 2; ++b)
    if (!pin[b]) HIP_CHECK(hipHostMalloc(&pin[b], kStageChunk, hipHostMallocDefault));
  StageEvents ev;
  bool pending[2] = {false, false};
  size_t off = 0;
  int b = 0;
  try {
    while (off < bytes) {
      const size_t n = std::min(kStageChunk, bytes - off);
      if (pending[b]) HIP_CHECK(hipEventSynchronize(ev.e[b]));  // the buffer's previous DMA is done
      parallel_copy(static_cast<char*>(pin[b]), static_cast<const char*>(src) + off, n);
      HIP_CHECK(hipMemcpyAsync(static_cast<char*>(dst) + off, pin[b], n, hipMemcpyHostToDevice, st));
      HIP_CHECK(hipEventRecord(ev.e[b], st));
      pending[b] = true;
      off += n;
      b ^= 1;
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);  // no DMA may still read a buffer the next transfer refills
    throw;
  }
  HIP_CHECK(hipStreamSynchronize(st));
}

void download_d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
  std::unique_lock<std::mutex> lk(g_stage->mu, std::try_to_lock);
  if (bytes < kStageMin || !lk.owns_lock()) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return;
  }
  void** pin = g_stage->pin;
  for (int b = 0; b < 2; ++b)
    if (!pin[b]) HIP_CHECK(hipHostMalloc(&pin[b], kStageChunk, hipHostMallocDefault));
  StageEvents ev;
  auto dma = [&](size_t off, int b) {
    const size_t n = std::min(kStageChunk, bytes - off);
    HIP_CHECK(hipMemcpyAsync(pin[b], static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipEventRecord(ev.e[b], st));
  };
  try {
    dma(0, 0);
    int b = 0;
    for (size_t off = 0; off < bytes; off += kStageChunk, b ^= 1) {
      // the other buffer was copied out in the previous round: refill it
      if (off + kStageChunk < bytes) dma(off + kStageChunk, b ^ 1);
      HIP_CHECK(hipEventSynchronize(ev.e[b]));
      parallel_copy(static_cast<char*>(dst) + off, static_cast<const char*>(pin[b]),
                    std::min(kStageChunk, bytes - off));
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);  // no DMA may still write a buffer the next transfer reads
    throw;
  }
}

}  // namespace petalo

namespace {

void place_lors(DevBuf<Lor>& out, hipStream_t st, const petalo_lor_t* lors, size_t n_lors, size_t n_subsets,
                size_t m, bool tof, bool sort, const petalo_fov_t& fov, const TileGeom& g) {
  out.alloc(std::max<size_t>(n_lors, 1));
  if (!n_lors) return;
  if (!sort || m < 2) {
    upload_h2d(out.p, lors, n_lors * sizeof(Lor), st);
    return;
  }
  DevBuf<Lor> raw(n_lors);
  upload_h2d(raw.p, lors, n_lors * sizeof(Lor), st);
  OrderParams p;
  for (int a = 0; a < 3; ++a) {
    p.full[a] = fov.full_size_mm[a];
    p.n[a] = fov.n[a];
  }
  p.tx = g.tx;
  p.ty = g.ty;
  p.tz = g.tz;
  p.ntx = g.ntx;
  p.nty = g.nty;
  p.tof = tof ? 1 : 0;
  DevBuf<uint64_t> keys(m), keys2(m);
  DevBuf<uint32_t> idx(m), idx2(m);
  size_t tmp_bytes = 0;
  HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys.p, keys2.p, idx.p, idx2.p, (int)m, 0, 64, st));
  DevBuf<unsigned char> tmp(std::max<size_t>(tmp_bytes, 1));
  for (size_t s = 0; s < n_subsets; ++s) {
    const Lor* src = raw.p + s * m;
    k_order_keys<<<grid_for(m), kBlock, 0, st>>>(src, (uint32_t)m, p, keys.p, idx.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, keys.p, keys2.p, idx.p, idx2.p, (int)m, 0, 64, st));
    k_gather_lors<<<grid_for(m), kBlock, 0, st>>>(src, idx2.p, (uint32_t)m, out.p + s * m);
    HIP_CHECK(hipGetLastError());
  }
  const size_t done = n_subsets * m;
  if (done < n_lors)
    HIP_CHECK(hipMemcpyAsync(out.p + done, raw.p + done, (n_lors - done) * sizeof(Lor), hipMemcpyDeviceToDevice,
                             st));
  HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace

// ============================================================================
struct petalo_mlem {
  int device = 0;
  hipStream_t stream = nullptr;
  petalo_fov_t fov{};
  DevFov dfov{};
  DevTof dtof{};
  bool tof_on = false;
  size_t n_lors = 0;
  size_t n_subsets = 1;
  size_t set_size = 0;
  int scheme = PETALO_SCHEME_ATOMIC;
  DevBuf<Lor> lors;
  DevBuf<float> image, bp, sens;
  DevBuf<unsigned long long> counters;
  DevBuf<unsigned long long> rstat;  // r statistics of the tiled update (kRStatWords)
  petalo_osem_t osem{};
  bool timing = false;
  // timing: per timed update, events at its start, forward end, scatter end,
  // update end and forward tile kernel end, in a ring of kTimingRing sets
  // (created when timing is first switched on) so that updates need not wait
  // for each other's timings; ev is the set of the last timed update
  std::vector<std::array<hipEvent_t, 5>> evr;
  hipEvent_t* ev = nullptr;
  size_t ev_next = 0;   // set of the next timed update
  size_t ev_count = 0;  // timed updates since the last drain (at most the ring)
  bool have_timing = false;
  // tiled scheme
  TileGeom tiles{};
  std::vector<DevBuf<Run>> runs;     // per subset, sorted by tile
  std::vector<DevBuf<int4>> work;    // per subset: (tile, first run, end run, 0)
  std::vector<uint64_t> n_runs;      // per subset
  std::vector<DevBuf<uint32_t>> run_off;    // per subset: LOR-major run offsets (n + 1)
  std::vector<DevBuf<uint32_t>> run_pos;    // per subset: sorted position of each LOR-major run
  std::vector<DevBuf<uint4>> pos4;          // per subset: the first four of them per LOR (k_pack_pos4) ...
  std::vector<DevBuf<uint2>> pos2;          // ... or as 16-bit window offsets (k_pack_pos2)
  std::vector<DevBuf<uint32_t>> lpos;       // per subset: partial position of each sorted run
  std::vector<uint32_t> n_blocks;           // per subset: LOR blocks of 2^kWinShift of the partial layout
  std::vector<DevBuf<uint32_t>> win;        // per subset: partial window of each LOR block (n_blocks + 1)
  std::vector<uint32_t> max_win;            // per subset: largest window
  std::vector<DevBuf<uint32_t>> irregular;  // per subset: LORs with unbounded walks
  DevBuf<float> partial;             // per-run partial projections (largest subset)
  DevBuf<float> rbuf;                // 1/(p*ac) per LOR of the current subset
  DevBuf<float> ac;                  // additive corrections, in context LOR order
  DevBuf<double> bp64;               // binary64 backprojection
  std::vector<std::vector<uint64_t>> item_steps;  // per subset: replayed steps of each work item
  uint64_t item_target = 0;          // replayed steps per work item; 0: the automatic rule (cut_items)
  std::vector<std::vector<uint64_t>> tile_start;  // per subset: first sorted run of every tile (ntiles + 1)
  // binary64 precision modes (PETALO_PRECISION_F64 / _F32_WALK)
  int precision = PETALO_PRECISION_F32;
  DevFov64 dfov64{};
  DevTof64 dtof64{};
  DevBuf<double> image64, sens64;
  // execution
  bool own_stream = true;            // false inside a batch (the batch's stream)
  bool use_graphs = true;            // replay each subset update as a captured HIP graph
  std::vector<hipGraphExec_t> graphs;  // per subset, captured on first use
};

namespace {

// The binary64 modes: one lane per LOR, binary64 walk/weights/image.
void launch_project64(petalo_mlem* c, size_t lo, size_t count) {
  if (count == 0) return;
  const Lor* L = c->lors.p + lo;
  const unsigned g = grid_for(count);
  const bool mixed = c->precision == PETALO_PRECISION_F32_WALK;
  hipStream_t st = c->stream;
  const uint32_t n = (uint32_t)count;
  if (c->tof_on) {
    if (mixed)
      k_mlem_project64<true, true><<<g, kBlock, 0, st>>>(L, n, c->dfov, c->dtof, c->dfov64, c->dtof64,
                                                         c->image64.p, c->bp64.p, c->counters.p);
    else
      k_mlem_project64<true, false><<<g, kBlock, 0, st>>>(L, n, c->dfov, c->dtof, c->dfov64, c->dtof64,
                                                          c->image64.p, c->bp64.p, c->counters.p);
  } else {
    if (mixed)
      k_mlem_project64<false, true><<<g, kBlock, 0, st>>>(L, n, c->dfov, c->dtof, c->dfov64, c->dtof64,
                                                          c->image64.p, c->bp64.p, c->counters.p);
    else
      k_mlem_project64<false, false><<<g, kBlock, 0, st>>>(L, n, c->dfov, c->dtof, c->dfov64, c->dtof64,
                                                           c->image64.p, c->bp64.p, c->counters.p);
  }
  HIP_CHECK(hipGetLastError());
}

void launch_project(petalo_mlem* c, size_t lo, size_t count) {
  if (count == 0) return;
  const Lor* L = c->lors.p + lo;
  const unsigned g = grid_for(count);
  if (c->tof_on)
    k_mlem_project<true><<<g, kBlock, 0, c->stream>>>(L, (uint32_t)count, c->dfov, c->dtof, c->image.p, c->bp.p,
                                                       c->counters.p);
  else
    k_mlem_project<false><<<g, kBlock, 0, c->stream>>>(L, (uint32_t)count, c->dfov, c->dtof, c->image.p, c->bp.p,
                                                        c->counters.p);
  HIP_CHECK(hipGetLastError());
}

// Splits n voxels into equal tiles of at most `target` voxels.
int32_t even_split(uint32_t n, int32_t target) {
  const int32_t parts = (int32_t)((n + target - 1) / target);
  return (int32_t)((n + parts - 1) / parts);
}

// Tile shape for an image.  The LDS row pitch is odd (tx or tx + 1): with an
// even pitch every voxel of a column falls on the same LDS bank, and the
// binary64 atomics of lanes working on nearby parallel LORs collide.  By
// default the tile is about 40 x 40 voxels transaxially, cut evenly from the
// image, and as deep axially as keeps its binary64 copy within
// kScatterEntries, so two 1024-lane workgroups fit a CU's LDS.
TileGeom tile_shape(const petalo_fov_t& fov, const int32_t shape[3]) {
  TileGeom g;
  g.tx = shape[0] > 0 ? shape[0] : even_split(fov.n[0], 40);
  g.ty = shape[1] > 0 ? shape[1] : even_split(fov.n[1], 40);
  g.px = g.tx | 1;
  g.tz = shape[2] > 0 ? shape[2] : even_split(fov.n[2], std::max(1, kScatterEntries / (g.px * g.ty)));
  g.ntx = (int32_t)((fov.n[0] + g.tx - 1) / g.tx);
  g.nty = (int32_t)((fov.n[1] + g.ty - 1) / g.ty);
  g.ntz = (int32_t)((fov.n[2] + g.tz - 1) / g.tz);
  g.ntiles = g.ntx * g.nty * g.ntz;
  g.tvox = g.px * g.ty * g.tz;
  return g;
}

TileGeom make_tiles(const petalo_fov_t& fov, const int32_t shape[3]) {
  const TileGeom g = tile_shape(fov, shape);
  if (g.ntiles >= 65536) throw std::runtime_error("image too large for the tiled scheme (65536 tiles or more)");
  if (g.tvox > kMaxTileVoxels) throw std::runtime_error("tile over 20416 LDS entries (row pitch included)");
  if (g.tx + g.ty + g.tz > (int32_t)kRunMaxSteps) throw std::runtime_error("tile too long for the run record");
  return g;
}

// Partial-projection layout.  The runs are replayed in (tile, block of
// 2^kLorBlockShift LORs, length bucket) order; their partials are stored by
// (block of 2^kWinShift LORs, tile) group instead, each group keeping its
// runs' replay order.  A wavefront's runs share a tile and mostly a window
// block, so the forward kernel's stores stay nearly contiguous, while the
// partials of one window block -- all that the reduction of its LORs
// gathers, about 4 runs x 4096 LORs x 4 B -- lie together and fit in LDS.

// wkey[j] = (window block << tile_bits) | tile of sorted run j.
__global__ void k_window_keys(const Run* __restrict__ unsorted, const uint64_t* __restrict__ sorted_keys,
                              const uint32_t* __restrict__ perm, uint64_t n, uint32_t tile_bits,
                              uint64_t* __restrict__ wkey) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint64_t blk = unsorted[perm[j]].lor >> kWinShift;
  wkey[j] = (blk << tile_bits) | (sorted_keys[j] >> 32);
}

// pos[perm[j]] = j.
__global__ void k_invert_perm(const uint32_t* __restrict__ perm, uint64_t n, uint32_t* __restrict__ pos) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) pos[perm[j]] = (uint32_t)j;
}

// win[b] = first stored position of window block b (UINT32_MAX where the
// block has no runs; the caller fills those in).
__global__ void k_window_starts(const uint64_t* __restrict__ wkey, uint64_t n, uint32_t tile_bits,
                                uint32_t n_blocks, uint32_t* __restrict__ win) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint64_t b = wkey[j] >> tile_bits;
  if (b < n_blocks && (j == 0 || (wkey[j - 1] >> tile_bits) != b)) win[b] = (uint32_t)j;
}

// Puts the runs in sorted order, records where each LOR-major run's partial
// goes (lpos of its sorted index) and writes a (tile << 12) | length summary
// of the sorted runs for the host.
__global__ void k_place_runs(const Run* __restrict__ in, const uint64_t* __restrict__ sorted_keys,
                             const uint32_t* __restrict__ perm, uint64_t n, Run* __restrict__ out,
                             uint32_t* __restrict__ pos, uint32_t* __restrict__ summary,
                             const uint32_t* __restrict__ lpos) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t src = perm[j];
  const Run ru = in[src];
  out[j] = ru;
  pos[src] = lpos[j];
  if (summary) summary[j] = ((uint32_t)(sorted_keys[j] >> 32) << 12) | run_len(ru.code);
}

// Irregular LORs of a subset (count 0xffffffff from the binning count pass):
// flag[i] = 1 and their count set to 0; counts[n] = 0 for the exclusive scan.
__global__ void k_flag_irregular(uint32_t* __restrict__ counts, uint32_t n, unsigned char* __restrict__ flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == n) counts[n] = 0;
  if (i >= n) return;
  const bool irr = counts[i] == 0xffffffffu;
  flag[i] = irr ? 1 : 0;
  if (irr) counts[i] = 0;
}

// *sum += counts[0 .. n), in 64 bits (zeroed by the caller).  Each lane sums
// a grid-strided share and each wavefront adds once, so that the one address
// takes a few thousand atomics rather than one per 64 counts.  Launch with
// k_sum_grid(n) blocks of kBlock.
inline unsigned k_sum_grid(size_t n) { return std::min<unsigned>(grid_for(n), 1024u); }
__global__ void k_sum_u32(const uint32_t* __restrict__ counts, uint32_t n, unsigned long long* __restrict__ sum) {
  unsigned long long s = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    s += counts[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(sum, s);
}

// steps[j] = length of sorted run j (then scanned into a running total).
__global__ void k_run_steps(const Run* __restrict__ runs, uint64_t n, uint64_t* __restrict__ steps) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) steps[j] = run_len(runs[j].code);
}

// out[k] = the running total of steps before sorted run at[k] (0 at 0).
__global__ void k_steps_before(const uint64_t* __restrict__ incl, const uint64_t* __restrict__ at, uint32_t n,
                               uint64_t* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = at[k] == 0 ? 0ull : incl[at[k] - 1];
}

// First sorted position of every tile; keys hold the tile from bit tshift
// (values past ntiles, which a well-formed key never holds, are ignored).
__global__ void k_tile_starts(const uint64_t* __restrict__ keys, uint64_t n, uint32_t tshift, uint32_t ntiles,
                              uint64_t* __restrict__ start) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint64_t t = keys[j] >> tshift;
  if (t < ntiles && (j == 0 || (keys[j - 1] >> tshift) != t)) start[t] = j;
}

// One tile's share of the work items: its sorted runs [a, b), the running
// step total before a, its steps, its item count and its first item.
struct TileCut {
  uint64_t a, b, base, steps;
  uint32_t n, first, tile, pad;
};

// Work item k: tile t's chunk m (k = first + m) ends after the first run j
// whose steps since the tile's start reach (m + 1)/n of the tile's (binary
// search over the inclusive running total incl; the last chunk at b); a
// chunk left empty (two boundaries on one run) has first == end.
__global__ void k_cut_items(const uint64_t* __restrict__ incl, const TileCut* __restrict__ tc, uint32_t n_tiles,
                            uint32_t n_items, int4* __restrict__ items, uint64_t* __restrict__ steps) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_items) return;
  uint32_t lo = 0, hi = n_tiles;  // the tile whose items hold k
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (tc[mid].first <= k) lo = mid; else hi = mid;
  }
  const TileCut t = tc[lo];
  const uint32_t m = k - t.first;
  auto end_of = [&](uint32_t q) -> uint64_t {  // one past the last run of chunk q
    if (q + 1 >= t.n) return t.b;
    const double goal = (double)(q + 1) * (double)t.steps;
    uint64_t l = t.a, h = t.b - 1;  // the first j in [a, b) with (incl[j] - base) n >= goal
    while (l < h) {
      const uint64_t mid = l + (h - l) / 2;
      if ((double)(incl[mid] - t.base) * (double)t.n >= goal) h = mid; else l = mid + 1;
    }
    return l + 1;
  };
  const uint64_t e = end_of(m);
  const uint64_t a = m == 0 ? t.a : end_of(m - 1);
  const uint64_t b = e > a ? e : a;
  items[k] = make_int4((int)t.tile, (int)a, (int)b, 0);
  steps[k] = b > a ? incl[b - 1] - (a == 0 ? 0ull : incl[a - 1]) : 0ull;
}

__global__ void k_iota(uint32_t* __restrict__ x, uint64_t n) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] = (uint32_t)j;
}

inline unsigned grid64(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void zero_counters(unsigned long long* p, hipStream_t s);

// A device buffer that only grows.
template <class T> struct Grow {
  DevBuf<T> b;
  T* ensure(size_t n) {
    if (n > b.n) b.alloc(n + n / 4);
    return b.p;
  }
};

// Temporary device buffers of a context's setup (build_runs, cut_items),
// reused across its subsets: allocating and freeing a dozen buffers per
// subset cost more than the setup's kernels when the subsets are small.
struct SetupScratch {
  Grow<unsigned char> flag, tmp, wtmp;
  Grow<unsigned long long> sum, chk;
  Grow<uint32_t> n_sel, perm, perm2, order;
  Grow<Run> unsorted;
  Grow<uint64_t> keys, keys2, wkey, wkey2, dstart, incl, at, before, dst;
  Grow<TileCut> dtc;
  Grow<int4> dit;
};

// Work items of one subset: chunks of one tile's runs holding about equal
// numbers of replayed steps, largest first so that the last workgroups to
// start are the short ones.  With c->item_target the chunks hold about that
// many steps (and at least kMinItemSteps); automatically, kWorkItems in
// total, and, since each item loads (forward) and flushes (scatter) its
// whole tile, at least kItemStepsPerVoxel steps per LDS entry of the tile.
// Cut on the device from a running total of the sorted runs' lengths
// (k_cut_items); the host orders the few thousand items.
void cut_items(petalo_mlem* c, size_t sub, SetupScratch* scratch = nullptr) {
  SetupScratch local;
  SetupScratch& S = scratch ? *scratch : local;
  const TileGeom& g = c->tiles;
  hipStream_t st = c->stream;
  const uint64_t total = c->n_runs[sub];
  const std::vector<uint64_t>& starts = c->tile_start[sub];
  c->item_steps[sub].clear();
  std::vector<int4> items;
  if (total) {
    uint64_t* incl = S.incl.ensure(total);
    k_run_steps<<<grid64(total), kBlock, 0, st>>>(c->runs[sub].p, total, incl);
    HIP_CHECK(hipGetLastError());
    size_t tb = 0;
    HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tb, incl, incl, (int)total, st));
    unsigned char* tmp = S.tmp.ensure(std::max<size_t>(tb, 1));
    HIP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, tb, incl, incl, (int)total, st));
    // running totals at the tile starts
    const size_t nb = (size_t)g.ntiles + 1;
    uint64_t* at = S.at.ensure(nb);
    uint64_t* before = S.before.ensure(nb);
    HIP_CHECK(hipMemcpyAsync(at, starts.data(), nb * 8, hipMemcpyHostToDevice, st));
    k_steps_before<<<grid_for(nb), kBlock, 0, st>>>(incl, at, (uint32_t)nb, before);
    HIP_CHECK(hipGetLastError());
    std::vector<uint64_t> bs(nb);
    HIP_CHECK(hipMemcpyAsync(bs.data(), before, nb * 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    const uint64_t total_steps = bs[g.ntiles];
    const uint64_t target =
        c->item_target ? std::max<uint64_t>(kMinItemSteps, c->item_target)
                       : std::max<uint64_t>(std::max<uint64_t>(kMinItemSteps, total_steps / kWorkItems + 1),
                                            kItemStepsPerVoxel * (uint64_t)g.tvox);
    std::vector<TileCut> tc;
    uint32_t n_items = 0;
    for (int32_t t = 0; t < g.ntiles; ++t) {
      if (starts[t] == starts[t + 1]) continue;
      const uint64_t steps = bs[t + 1] - bs[t];
      const uint32_t n = (uint32_t)std::max<uint64_t>(1, (steps + target - 1) / target);
      tc.push_back(TileCut{starts[t], starts[t + 1], bs[t], steps, n, n_items, (uint32_t)t, 0});
      n_items += n;
    }
    if (n_items) {
      TileCut* dtc = S.dtc.ensure(tc.size());
      int4* dit = S.dit.ensure(n_items);
      uint64_t* dst = S.dst.ensure(n_items);
      HIP_CHECK(hipMemcpyAsync(dtc, tc.data(), tc.size() * sizeof(TileCut), hipMemcpyHostToDevice, st));
      k_cut_items<<<grid_for(n_items), kBlock, 0, st>>>(incl, dtc, (uint32_t)tc.size(), n_items, dit, dst);
      HIP_CHECK(hipGetLastError());
      std::vector<int4> hit(n_items);
      std::vector<uint64_t> hst(n_items);
      HIP_CHECK(hipMemcpyAsync(hit.data(), dit, n_items * sizeof(int4), hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipMemcpyAsync(hst.data(), dst, n_items * 8, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      std::vector<std::pair<uint64_t, int4>> sized;
      for (uint32_t k = 0; k < n_items; ++k)
        if (hit[k].z > hit[k].y) sized.push_back({hst[k], hit[k]});
      std::stable_sort(sized.begin(), sized.end(), [](const std::pair<uint64_t, int4>& u,
                                                      const std::pair<uint64_t, int4>& v) { return u.first > v.first; });
      items.reserve(sized.size());
      for (auto& it : sized) {
        items.push_back(it.second);
        c->item_steps[sub].push_back(it.first);
      }
    }
  }
  c->work[sub].alloc(std::max<size_t>(items.size(), 1));
  if (!items.empty())
    HIP_CHECK(hipMemcpyAsync(c->work[sub].p, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice, st));
  c->work[sub].n = items.size();
  HIP_CHECK(hipStreamSynchronize(st));
}

// Cuts the subset's LOR walks into per-tile runs, sorted by tile, and the
// per-workgroup work items over them; lists the LORs with unbounded walks.
void build_runs(petalo_mlem* c, size_t sub, size_t lo, size_t count, SetupScratch* scratch = nullptr) {
  SetupScratch local;
  SetupScratch& S = scratch ? *scratch : local;
  const TileGeom& g = c->tiles;
  hipStream_t st = c->stream;
  uint64_t total = 0;
  size_t n_irr = 0;
  c->run_off[sub].alloc(count + 1);
  c->irregular[sub].alloc(std::max<size_t>(count, 1));
  if (count) {
    // count pass; LORs with unbounded walks listed (in order) and given no
    // runs; the run offsets scanned on the device
    unsigned char* flag = S.flag.ensure(count);
    unsigned long long* sum = S.sum.ensure(1);
    uint32_t* n_sel = S.n_sel.ensure(1);
    const Lor* L = c->lors.p + lo;
    uint32_t* nr = c->run_off[sub].p;  // counts, then scanned in place
    if (c->tof_on)
      k_bin_runs<true, false><<<bin_grid(count), kBlock, 0, st>>>(L, (uint32_t)count, c->dfov, c->dtof, g, nr,
                                                                  nullptr, nullptr, nullptr, 0u);
    else
      k_bin_runs<false, false><<<bin_grid(count), kBlock, 0, st>>>(L, (uint32_t)count, c->dfov, c->dtof, g, nr,
                                                                   nullptr, nullptr, nullptr, 0u);
    k_flag_irregular<<<grid_for(count + 1), kBlock, 0, st>>>(nr, (uint32_t)count, flag);
    HIP_CHECK(hipMemsetAsync(sum, 0, 8, st));
    k_sum_u32<<<k_sum_grid(count), kBlock, 0, st>>>(nr, (uint32_t)count, sum);
    HIP_CHECK(hipGetLastError());
    size_t sel_bytes = 0, scan_bytes = 0;
    hipcub::CountingInputIterator<uint32_t> idx(0);
    HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, sel_bytes, idx, flag, c->irregular[sub].p, n_sel,
                                            (int)count, st));
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, nr, nr, (int)count + 1, st));
    unsigned char* tmp = S.tmp.ensure(std::max<size_t>(std::max(sel_bytes, scan_bytes), 1));
    HIP_CHECK(hipcub::DeviceSelect::Flagged(tmp, sel_bytes, idx, flag, c->irregular[sub].p, n_sel,
                                            (int)count, st));
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, scan_bytes, nr, nr, (int)count + 1, st));
    unsigned long long h_sum = 0;
    uint32_t h_sel = 0;
    HIP_CHECK(hipMemcpyAsync(&h_sum, sum, 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(&h_sel, n_sel, 4, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));  // (also before the temporaries are freed)
    if (h_sum >= (1ull << 31)) throw std::runtime_error("2^31 or more tile runs in one subset");
    total = h_sum;
    n_irr = h_sel;
  } else {
    HIP_CHECK(hipMemsetAsync(c->run_off[sub].p, 0, 4, st));
  }
  c->irregular[sub].n = n_irr;
  c->n_blocks[sub] = (uint32_t)((count + (1u << kWinShift) - 1) >> kWinShift);
  c->win[sub].alloc(c->n_blocks[sub] + 1);
  c->max_win[sub] = 0;
  HIP_CHECK(hipMemsetAsync(c->win[sub].p, 0, ((size_t)c->n_blocks[sub] + 1) * 4, st));  // empty windows
  c->runs[sub].alloc(std::max<uint64_t>(total, 1));
  c->run_pos[sub].alloc(std::max<uint64_t>(total, 1));
  std::vector<uint64_t> starts(g.ntiles + 1, UINT64_MAX);
  if (total) {
    const Lor* L = c->lors.p + lo;
    Run* unsorted = S.unsorted.ensure(total);
    uint64_t* keys = S.keys.ensure(total);
    uint64_t* keys2 = S.keys2.ensure(total);
    uint32_t* perm = S.perm.ensure(total);
    uint32_t* perm2 = S.perm2.ensure(total);
    unsigned long long* chk = S.chk.ensure(CNT_N);
    zero_counters(chk, st);
    if (c->tof_on)
      k_bin_runs<true, true><<<bin_grid(count), kBlock, 0, st>>>(L, (uint32_t)count, c->dfov, c->dtof, g, nullptr,
                                                                 c->run_off[sub].p, unsorted, keys, kLorBlockShift,
                                                                 chk);
    else
      k_bin_runs<false, true><<<bin_grid(count), kBlock, 0, st>>>(L, (uint32_t)count, c->dfov, c->dtof, g, nullptr,
                                                                  c->run_off[sub].p, unsorted, keys, kLorBlockShift,
                                                                  chk);
    HIP_CHECK(hipGetLastError());
    {
      unsigned long long h[CNT_N];
      HIP_CHECK(hipMemcpyAsync(h, chk, sizeof h, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      if (h[CNT_RUN_MISMATCH])
        throw std::runtime_error("tile binning: fill and count passes disagree on " + std::to_string(h[CNT_RUN_MISMATCH]) +
                                 " LORs");
    }
    k_iota<<<grid64(total), kBlock, 0, st>>>(perm, total);
    int end_bit = 33;
    while ((1 << (end_bit - 32)) < g.ntiles) ++end_bit;
    size_t tmp_bytes = 0;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys2, perm, perm2, (int)total, 0,
                                                 end_bit, st));
    unsigned char* tmp = S.tmp.ensure(std::max<size_t>(tmp_bytes, 1));
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys2, perm, perm2, (int)total, 0,
                                                 end_bit, st));
    // partial layout by (window block, tile): a stable sort of the sorted
    // runs by that key gives each run's stored position
    int tile_bits = 1;
    while ((1 << tile_bits) < g.ntiles) ++tile_bits;
    int blk_bits = 1;
    while ((1ull << blk_bits) < c->n_blocks[sub]) ++blk_bits;
    c->lpos[sub].alloc(total);
    std::vector<uint32_t> hw(c->n_blocks[sub] + 1, UINT32_MAX);
    {
      uint64_t* wkey = S.wkey.ensure(total);
      uint64_t* wkey2 = S.wkey2.ensure(total);
      uint32_t* order = S.order.ensure(total);
      k_window_keys<<<grid64(total), kBlock, 0, st>>>(unsorted, keys2, perm2, total, (uint32_t)tile_bits,
                                                      wkey);
      k_iota<<<grid64(total), kBlock, 0, st>>>(perm, total);
      HIP_CHECK(hipGetLastError());
      size_t wbytes = 0;
      HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, wbytes, wkey, wkey2, perm, order, (int)total, 0,
                                                   tile_bits + blk_bits, st));
      unsigned char* wtmp = S.wtmp.ensure(std::max<size_t>(wbytes, 1));
      HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(wtmp, wbytes, wkey, wkey2, perm, order, (int)total, 0,
                                                   tile_bits + blk_bits, st));
      k_invert_perm<<<grid64(total), kBlock, 0, st>>>(order, total, c->lpos[sub].p);
      HIP_CHECK(hipMemcpyAsync(c->win[sub].p, hw.data(), hw.size() * 4, hipMemcpyHostToDevice, st));
      k_window_starts<<<grid64(total), kBlock, 0, st>>>(wkey2, total, (uint32_t)tile_bits, c->n_blocks[sub],
                                                        c->win[sub].p);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(hw.data(), c->win[sub].p, hw.size() * 4, hipMemcpyDeviceToHost, st));
      // (also waits for the temporaries' users: they are freed at scope end)
      HIP_CHECK(hipStreamSynchronize(st));
    }
    // blocks without runs start where the next one does
    hw[c->n_blocks[sub]] = (uint32_t)total;
    for (size_t b = c->n_blocks[sub]; b-- > 0;)
      if (hw[b] == UINT32_MAX) hw[b] = hw[b + 1];
    uint32_t mw = 0;
    for (size_t b = 0; b < c->n_blocks[sub]; ++b) mw = std::max(mw, hw[b + 1] - hw[b]);
    c->max_win[sub] = mw;
    HIP_CHECK(hipMemcpyAsync(c->win[sub].p, hw.data(), hw.size() * 4, hipMemcpyHostToDevice, st));
    k_place_runs<<<grid64(total), kBlock, 0, st>>>(unsorted, keys2, perm2, total, c->runs[sub].p,
                                                   c->run_pos[sub].p, nullptr, c->lpos[sub].p);
    // first sorted run of every tile (UINT64_MAX where a tile has none)
    uint64_t* dstart = S.dstart.ensure(g.ntiles);
    HIP_CHECK(hipMemsetAsync(dstart, 0xff, (size_t)g.ntiles * 8, st));
    k_tile_starts<<<grid64(total), kBlock, 0, st>>>(keys2, total, 32u, (uint32_t)g.ntiles, dstart);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(starts.data(), dstart, (size_t)g.ntiles * 8, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
  }
  starts[g.ntiles] = total;
  for (int k = g.ntiles - 1; k >= 0; --k)
    if (starts[k] == UINT64_MAX) starts[k] = starts[k + 1];
  // the reduction's first four positions per LOR: 16-bit window offsets
  // where every window allows them, else full positions
  if (c->max_win[sub] <= kMaxCompactWindow) {
    c->pos4[sub].alloc(0);
    c->pos2[sub].alloc(std::max<size_t>(count, 1));
    if (count)
      k_pack_pos2<<<grid_for(count), kBlock, 0, st>>>(c->run_off[sub].p, c->run_pos[sub].p, (uint32_t)count,
                                                      c->win[sub].p, c->pos2[sub].p);
  } else {
    c->pos2[sub].alloc(0);
    c->pos4[sub].alloc(std::max<size_t>(count, 1));
    if (count)
      k_pack_pos4<<<grid_for(count), kBlock, 0, st>>>(c->run_off[sub].p, c->run_pos[sub].p, (uint32_t)count,
                                                      c->pos4[sub].p);
  }
  HIP_CHECK(hipGetLastError());
  c->n_runs[sub] = total;
  c->tile_start[sub] = std::move(starts);
  cut_items(c, sub, &S);
}

// 1/(p*ac) per LOR from the partials of the forward kernel, then the
// one-lane path for the subset's unbounded walks.
void launch_reduce(petalo_mlem* c, size_t sub, size_t lo, size_t count) {
  if (count == 0) return;
  const Lor* L = c->lors.p + lo;
  hipStream_t st = c->stream;
  const uint32_t lds = std::min(c->max_win[sub], kReduceWindow);
  if (c->pos2[sub].p)
    k_lor_reduce_win<false, true><<<c->n_blocks[sub], kWinThreads, (size_t)lds * sizeof(float), st>>>(
        c->ac.p + lo, (uint32_t)count, c->win[sub].p, lds, c->run_off[sub].p, c->run_pos[sub].p, c->pos4[sub].p,
        c->pos2[sub].p, c->partial.p, c->rbuf.p, c->rstat.p);
  else
    k_lor_reduce_win<false, false><<<c->n_blocks[sub], kWinThreads, (size_t)lds * sizeof(float), st>>>(
        c->ac.p + lo, (uint32_t)count, c->win[sub].p, lds, c->run_off[sub].p, c->run_pos[sub].p, c->pos4[sub].p,
        c->pos2[sub].p, c->partial.p, c->rbuf.p, c->rstat.p);
  HIP_CHECK(hipGetLastError());
  const size_t n_irr = c->irregular[sub].n;
  if (n_irr) {
    if (c->tof_on)
      k_fwd_irregular<true><<<grid_for(n_irr), kBlock, 0, st>>>(L, c->irregular[sub].p, (uint32_t)n_irr, c->dfov,
                                                                c->dtof, c->image.p, c->bp64.p, c->counters.p);
    else
      k_fwd_irregular<false><<<grid_for(n_irr), kBlock, 0, st>>>(L, c->irregular[sub].p, (uint32_t)n_irr, c->dfov,
                                                                 c->dtof, c->image.p, c->bp64.p, c->counters.p);
    HIP_CHECK(hipGetLastError());
  }
}

void launch_tiled(petalo_mlem* c, size_t sub, size_t lo, size_t count, bool timing) {
  if (count == 0) return;
  hipStream_t st = c->stream;
  const size_t nw = c->work[sub].n;
  k_zero_u64<<<1, 320, 0, st>>>(c->rstat.p, kRStatWords);  // the reduction's r statistics
  if (nw) {
    const size_t lds = ((size_t)c->tiles.tvox + 1) * sizeof(float);
    if (c->tof_on)
      k_fwd_tile<true><<<(unsigned)nw, 1024, lds, st>>>(c->runs[sub].p, c->work[sub].p, c->image.p, c->dfov,
                                                        c->dtof, c->tiles, c->partial.p, nullptr,
                                                        c->lpos[sub].p);
    else
      k_fwd_tile<false><<<(unsigned)nw, 1024, lds, st>>>(c->runs[sub].p, c->work[sub].p, c->image.p, c->dfov,
                                                         c->dtof, c->tiles, c->partial.p, nullptr,
                                                         c->lpos[sub].p);
    HIP_CHECK(hipGetLastError());
  }
  if (timing) HIP_CHECK(hipEventRecord(c->ev[4], st));
  launch_reduce(c, sub, lo, count);
  if (timing) HIP_CHECK(hipEventRecord(c->ev[1], st));
  if (nw) {
    const size_t lds = (size_t)c->tiles.tvox * sizeof(double);
    const int4* sw = c->work[sub].p;
    const size_t nws = nw;
    unsigned long long* rs = c->rstat.p;
    if (c->tof_on)
      k_scatter<true><<<(unsigned)nws, 1024, lds, st>>>(c->runs[sub].p, sw, c->rbuf.p, c->dfov, c->dtof,
                                                       c->tiles, c->bp64.p, nullptr, rs);
    else
      k_scatter<false><<<(unsigned)nws, 1024, lds, st>>>(c->runs[sub].p, sw, c->rbuf.p, c->dfov,
                                                        c->dtof, c->tiles, c->bp64.p, nullptr, rs);
    HIP_CHECK(hipGetLastError());
  }
}

void fill_f32(float* p, float v, size_t n, hipStream_t s) {
  if (!n) return;
  k_fill_f32<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
  HIP_CHECK(hipGetLastError());
}

void fill_f64(double* p, double v, size_t n, hipStream_t s) {
  if (!n) return;
  k_fill_f64<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
  HIP_CHECK(hipGetLastError());
}

void zero_counters(unsigned long long* p, hipStream_t s) {
  k_zero_u64<<<1, 64, 0, s>>>(p, CNT_N);
  HIP_CHECK(hipGetLastError());
}

// The tile kernels take their LDS tile as dynamic shared memory; allow up to
// the largest tile on the current device, for every context (a lower limit
// set for one context would break another with larger tiles).
void ensure_lds_attributes() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  static std::mutex mu;
  static std::vector<bool> done;
  std::lock_guard<std::mutex> lock(mu);
  if ((size_t)dev < done.size() && done[dev]) return;
  const int lds = kMaxTileVoxels * (int)sizeof(double);
  const void* fns[] = {(const void*)k_scatter<true>,        (const void*)k_scatter<false>,
                       (const void*)k_fwd_tile<true>,       (const void*)k_fwd_tile<false>,
                       (const void*)k_scatter_batch<true>,  (const void*)k_scatter_batch<false>,
                       (const void*)k_fwd_tile_batch<true>, (const void*)k_fwd_tile_batch<false>,
                       (const void*)k_lor_reduce_win<false, true>, (const void*)k_lor_reduce_win<false, false>,
                       (const void*)k_lor_reduce_win_batch, (const void*)k_batch_persistent,
                       (const void*)k_scatter_unit32};
  for (const void* f : fns) HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  if (done.size() <= (size_t)dev) done.resize(dev + 1, false);
  done[dev] = true;
}

// Everything one subset update launches, on the context's stream.  With
// timing, HIP events bracket the phases (not while a graph is captured).
void enqueue_update(petalo_mlem* c, bool timing) {
  const size_t nv = nvox(c->fov);
  const size_t sub = c->osem.subset - 1;
  const size_t lo = sub * c->set_size;  // Osem::subset (mlem.rs:90-95)
  zero_counters(c->counters.p, c->stream);
  if (timing) {
    c->ev = c->evr[c->ev_next].data();
    c->ev_next = (c->ev_next + 1) % c->evr.size();
    c->ev_count = std::min(c->ev_count + 1, c->evr.size());
    HIP_CHECK(hipEventRecord(c->ev[0], c->stream));
  }
  if (c->precision != PETALO_PRECISION_F32) {
    launch_project64(c, lo, c->set_size);
    if (timing) {
      HIP_CHECK(hipEventRecord(c->ev[4], c->stream));
      HIP_CHECK(hipEventRecord(c->ev[1], c->stream));
      HIP_CHECK(hipEventRecord(c->ev[2], c->stream));
    }
    k_update64<<<grid_for(nv), kBlock, 0, c->stream>>>(c->image64.p, c->bp64.p, c->sens64.p, (uint32_t)nv);
  } else if (c->scheme == PETALO_SCHEME_TILE) {
    launch_tiled(c, sub, lo, c->set_size, timing);
    if (timing) HIP_CHECK(hipEventRecord(c->ev[2], c->stream));
    k_update_f64<<<grid_for(nv), kBlock, 0, c->stream>>>(c->image.p, c->bp64.p, c->sens.p, (uint32_t)nv);
  } else {
    launch_project(c, lo, c->set_size);
    if (timing) {
      HIP_CHECK(hipEventRecord(c->ev[4], c->stream));
      HIP_CHECK(hipEventRecord(c->ev[1], c->stream));
      HIP_CHECK(hipEventRecord(c->ev[2], c->stream));
    }
    k_update<<<grid_for(nv), kBlock, 0, c->stream>>>(c->image.p, c->bp.p, c->sens.p, (uint32_t)nv);
  }
  HIP_CHECK(hipGetLastError());
  if (timing) HIP_CHECK(hipEventRecord(c->ev[3], c->stream));
}

// The subset update as a HIP graph: captured from the stream on first use
// (kernel nodes only), then launched whole.
void launch_update_graph(petalo_mlem* c, hipStream_t st, std::vector<hipGraphExec_t>& graphs, size_t sub,
                         const std::function<void()>& enqueue) {
  if (graphs.size() <= sub) graphs.resize(sub + 1, nullptr);
  if (!graphs[sub]) {
    hipGraph_t g = nullptr;
    HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_CHECK(hipStreamEndCapture(st, &g));
    hipGraphExec_t e = nullptr;
    const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_CHECK(err);
    graphs[sub] = e;
  }
  (void)c;
  HIP_CHECK(hipGraphLaunch(graphs[sub], st));
}

}  // namespace

extern "C" {

const char* petalo_last_error(void) { return petalo::g_last_error.c_str(); }

int petalo_version(void) { return 100; }

int petalo_device_count(int* count) {
  if (!count) return fail(PETALO_ERR_ARG, "count is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return PETALO_OK;
}

int petalo_device_synchronize(int device) {
  return guarded([&] {
    require_device();
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipDeviceSynchronize());
    return PETALO_OK;
  });
}

}  // extern "C"

namespace {

// Argument checks of petalo_mlem_create; "" when the arguments are usable.
std::string check_mlem_args(const petalo_fov_t* fov, const petalo_tof_t* tof, const petalo_lor_t* lors,
                            size_t n_lors, size_t n_subsets, const petalo_mlem_options_t& opt) {
  std::string why;
  if (!fov_ok(fov, why)) return why;
  if (n_subsets == 0) return "n_subsets must be at least 1";
  if (n_lors && !lors) return "lors is NULL";
  if (n_lors >= (1ull << 32)) return "more than 2^32-1 LORs";
  if (tof && tof->enabled && !(tof->sigma_ps > 0.0f)) return "TOF sigma must be positive";
  if (opt.scheme != PETALO_SCHEME_AUTO && opt.scheme != PETALO_SCHEME_ATOMIC && opt.scheme != PETALO_SCHEME_TILE)
    return "unsupported backprojection scheme";
  for (int a = 0; a < 3; ++a)
    if (opt.tile[a] < 0) return "negative tile dimension";
  if (opt.precision < PETALO_PRECISION_F32 || opt.precision > PETALO_PRECISION_F32_WALK) return "unknown precision";
  if (opt.precision != PETALO_PRECISION_F32 && opt.autotune) return "autotune applies to the binary32 path only";
  if (opt.item_steps < 0) return "negative item_steps";
  return "";
}

// Builds a context on the current device; its stream is `shared` when given
// (a batch's), else its own.  Throws on failure.
petalo_mlem* create_context(const petalo_fov_t* fov, const petalo_tof_t* tof, const petalo_lor_t* lors,
                            size_t n_lors, const float* sensitivity, size_t n_subsets,
                            const petalo_mlem_options_t& opt, hipStream_t shared) {
    auto* c = new petalo_mlem();
    try {
      c->device = opt.device;
      c->item_target = (uint64_t)opt.item_steps;
      c->fov = *fov;
      c->dfov = to_dev(*fov);
      c->dtof = to_dev(tof);
      c->tof_on = tof && tof->enabled;
      c->n_lors = n_lors;
      c->n_subsets = n_subsets;
      c->set_size = n_lors / n_subsets;  // remainder LORs ignored (mlem.rs:91)
      c->precision = opt.precision;
      // the binary64 modes are one lane per LOR with global atomics
      c->scheme = (opt.scheme == PETALO_SCHEME_ATOMIC || opt.precision != PETALO_PRECISION_F32) ? PETALO_SCHEME_ATOMIC
                                                                                                  : PETALO_SCHEME_TILE;
      c->osem = petalo_osem_t{n_subsets, 1, 1};
      if (shared) {
        c->stream = shared;
        c->own_stream = false;
      } else {
        HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      }
      const size_t nv = nvox(*fov);

      // LORs: reorder inside each subset (subset membership is unchanged).
      place_lors(c->lors, c->stream, lors, n_lors, n_subsets, c->set_size, c->tof_on, opt.sort_lors != 0, *fov,
                 tile_shape(*fov, opt.tile));

      c->image.alloc(nv);
      c->bp.alloc(nv);
      c->counters.alloc(CNT_N);
      c->rstat.alloc(kRStatWords);
      k_zero_u64<<<1, 320, 0, c->stream>>>(c->rstat.p, kRStatWords);
      fill_f32(c->image.p, 1.0f, nv, c->stream);  // Image::ones (mlem.rs:17)
      fill_f32(c->bp.p, 0.0f, nv, c->stream);
      zero_counters(c->counters.p, c->stream);
      if (sensitivity) {
        c->sens.alloc(nv);
        HIP_CHECK(hipMemcpy(c->sens.p, sensitivity, nv * sizeof(float), hipMemcpyHostToDevice));
      }
      if (c->precision != PETALO_PRECISION_F32) {
        c->dfov64 = make_fov64(fov->full_size_mm, fov->n);
        c->dtof64 = c->tof_on ? make_tof64(tof->sigma_ps, tof->cutoff) : DevTof64{1.0, 0.0, 0.0};
        c->image64.alloc(nv);
        c->bp64.alloc(nv);
        fill_f64(c->image64.p, 1.0, nv, c->stream);  // Image::ones (mlem.rs:17)
        fill_f64(c->bp64.p, 0.0, nv, c->stream);
        if (sensitivity) {
          c->sens64.alloc(nv);
          k_widen<<<grid_for(nv), kBlock, 0, c->stream>>>(c->sens.p, c->sens64.p, nv);
          HIP_CHECK(hipGetLastError());
        }
      } else if (c->scheme == PETALO_SCHEME_TILE) {
        c->tiles = make_tiles(*fov, opt.tile);
        ensure_lds_attributes();
        c->runs.resize(n_subsets);
        c->item_steps.resize(n_subsets);
        c->work.resize(n_subsets);
        c->run_off.resize(n_subsets);
        c->run_pos.resize(n_subsets);
        c->pos4.resize(n_subsets);
        c->pos2.resize(n_subsets);
        c->lpos.resize(n_subsets);
        c->n_blocks.resize(n_subsets, 0);
        c->win.resize(n_subsets);
        c->max_win.resize(n_subsets, 0);
        c->irregular.resize(n_subsets);
        c->n_runs.assign(n_subsets, 0);
        c->tile_start.resize(n_subsets);
        c->rbuf.alloc(std::max<size_t>(c->set_size, 1));
        c->ac.alloc(std::max<size_t>(n_lors, 1));
        if (n_lors) {
          k_lor_ac<<<(unsigned)((n_lors + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(c->lors.p, n_lors, c->ac.p);
          HIP_CHECK(hipGetLastError());
        }
        c->bp64.alloc(nv);
        fill_f64(c->bp64.p, 0.0, nv, c->stream);
        uint64_t max_runs = 1;
        SetupScratch scratch;
        for (size_t sub = 0; sub < n_subsets; ++sub) {
          build_runs(c, sub, sub * c->set_size, c->set_size, &scratch);
          max_runs = std::max<uint64_t>(max_runs, c->n_runs[sub]);
        }
        c->partial.alloc(max_runs);
      }
      HIP_CHECK(hipStreamSynchronize(c->stream));
    } catch (...) {
      petalo_mlem_destroy(c);
      throw;
    }
    return c;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// The autotuner: times one subset update of each candidate on up to
// kTuneLors LORs of the first subset (one warm-up update, then three) and
// returns the options of the fastest, keeping the default tile shape unless
// another candidate is at least kTuneMargin faster (the candidates often lie
// within the timing noise of each other).  Candidates: the default tile, the
// plain atomic scheme and tiles of about 24, 32, 48 and 64 voxels across,
// each with its default depth; shapes the image does not admit are skipped.
constexpr size_t kTuneLors = 1u << 21;
constexpr double kTuneMargin = 0.03;

petalo_mlem_options_t autotune(const petalo_fov_t* fov, const petalo_tof_t* tof, const petalo_lor_t* lors,
                               size_t n_lors, const float* sensitivity, size_t n_subsets,
                               petalo_mlem_options_t opt) {
  opt.autotune = 0;
  const size_t m = std::min(n_lors / n_subsets, kTuneLors);
  petalo_mlem_options_t best = opt;
  best.scheme = PETALO_SCHEME_TILE;
  if (m == 0) return best;
  std::vector<petalo_mlem_options_t> cands;
  std::vector<std::array<int32_t, 3>> seen;
  for (int32_t w : {0, 24, 32, 48, 64}) {
    petalo_mlem_options_t t = opt;
    t.scheme = PETALO_SCHEME_TILE;
    for (int k = 0; k < 2; ++k) t.tile[k] = w ? even_split(fov->n[k], w) : 0;  // 0: the default
    t.tile[2] = 0;  // default depth for this cross-section
    const TileGeom g = tile_shape(*fov, t.tile);
    const std::array<int32_t, 3> shape{g.tx, g.ty, g.tz};
    if (std::find(seen.begin(), seen.end(), shape) != seen.end()) continue;
    if (g.ntiles >= 65536 || g.tvox > kMaxTileVoxels || g.tx + g.ty + g.tz > (int32_t)kRunMaxSteps) continue;
    seen.push_back(shape);
    cands.push_back(t);
    if (w == 0) {
      petalo_mlem_options_t a = opt;
      a.scheme = PETALO_SCHEME_ATOMIC;
      cands.push_back(a);
    }
  }
  double best_ms = 1e300;
  for (const petalo_mlem_options_t& cand : cands) {
    petalo_mlem* c = nullptr;
    try {
      c = create_context(fov, tof, lors, m, sensitivity, 1, cand, nullptr);
    } catch (const HipError&) {
      throw;  // a device failure is not a property of the candidate
    } catch (const std::exception&) {
      continue;  // e.g. out of host memory for this candidate's runs: skip it
    }
    double ms = 1e300;
    try {
      petalo_osem_t id;
      auto step = [&] {
        if (petalo_mlem_next(c, nullptr, &id) != PETALO_OK) throw std::runtime_error(petalo_last_error());
      };
      step();
      HIP_CHECK(hipStreamSynchronize(c->stream));
      const auto t0 = std::chrono::steady_clock::now();
      for (int k = 0; k < 3; ++k) step();
      HIP_CHECK(hipStreamSynchronize(c->stream));
      ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / 3.0;
    } catch (...) {
      petalo_mlem_destroy(c);
      throw;
    }
    petalo_mlem_destroy(c);
    // the first candidate is the default; a later one must beat the best by the margin
    if (&cand == &cands[0] || ms < best_ms * (1.0 - kTuneMargin)) {
      best_ms = ms;
      best = cand;
    }
  }
  return best;
}

// The autotuner's second dimension: the work granularity of a tiled context,
// timed on the context itself.  Candidates: the
// automatic rule of cut_items and targets of 2^14 .. 2^20 replayed steps per
// work item, those that cut the first subset into a number of items no
// earlier candidate gave; each re-cuts (cut_items) the items of the timed subsets (the
// first min(n_subsets, 4)), runs them once untimed and three times timed
// (the best counts), and the automatic rule stays unless another is at least
// kTuneMargin faster.  The
// items of every subset are then cut with the choice, the image, counters
// and subset position reset.
void tune_items(petalo_mlem* c) {
  if (c->scheme != PETALO_SCHEME_TILE) return;
  const size_t nt = std::min<size_t>(c->n_subsets, 4);
  const bool graphs = c->use_graphs;
  c->use_graphs = false;
  auto recut = [&](uint64_t target, size_t upto) {
    c->item_target = target;
    for (size_t sub = 0; sub < upto; ++sub) cut_items(c, sub);
  };
  auto updates = [&]() {
    c->osem = petalo_osem_t{c->n_subsets, 1, 1};
    petalo_osem_t id;
    for (size_t k = 0; k < nt; ++k)
      if (petalo_mlem_next(c, nullptr, &id) != PETALO_OK) throw std::runtime_error(petalo_last_error());
    HIP_CHECK(hipStreamSynchronize(c->stream));
  };
  std::vector<size_t> seen;
  uint64_t best = 0;
  double best_ms = 1e300;
  for (uint64_t target : {(uint64_t)0, (uint64_t)1 << 14, (uint64_t)1 << 15, (uint64_t)1 << 16, (uint64_t)1 << 17,
                          (uint64_t)1 << 18, (uint64_t)1 << 19, (uint64_t)1 << 20}) {
    recut(target, nt);
    const size_t n_items = c->work[0].n;
    if (std::find(seen.begin(), seen.end(), n_items) != seen.end()) continue;
    seen.push_back(n_items);
    updates();  // warm-up
    double ms = 1e300;  // the best of three: a few updates time noisily
    for (int rep = 0; rep < 3; ++rep) {
      const auto t0 = std::chrono::steady_clock::now();
      updates();
      ms = std::min(ms, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    if (target == 0 || ms < best_ms * (1.0 - kTuneMargin)) {
      best_ms = ms;
      best = target;
    }
  }
  recut(best, c->n_subsets);
  const size_t nv = nvox(c->fov);
  fill_f32(c->image.p, 1.0f, nv, c->stream);  // Image::ones (mlem.rs:17)
  fill_f64(c->bp64.p, 0.0, nv, c->stream);
  zero_counters(c->counters.p, c->stream);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->osem = petalo_osem_t{c->n_subsets, 1, 1};
  c->use_graphs = graphs;
}

}  // namespace

extern "C" {

int petalo_mlem_create(petalo_mlem_t** out, const petalo_fov_t* fov, const petalo_tof_t* tof,
                       const petalo_lor_t* lors, size_t n_lors, const float* sensitivity, size_t n_subsets,
                       const petalo_mlem_options_t* options) {
  return guarded([&] {
    if (!out) return fail(PETALO_ERR_ARG, "out is NULL");
    *out = nullptr;
    petalo_mlem_options_t opt{PETALO_SCHEME_AUTO, 1, 0, {0, 0, 0}, 0, PETALO_PRECISION_F32, 0};
    if (options) opt = *options;
    const std::string why = check_mlem_args(fov, tof, lors, n_lors, n_subsets, opt);
    if (!why.empty()) return fail(PETALO_ERR_ARG, why);
    require_device();
    HIP_CHECK(hipSetDevice(opt.device));
    const bool tune = opt.scheme == PETALO_SCHEME_AUTO && opt.autotune;
    if (tune) opt = autotune(fov, tof, lors, n_lors, sensitivity, n_subsets, opt);
    petalo_mlem* c = create_context(fov, tof, lors, n_lors, sensitivity, n_subsets, opt, nullptr);
    try {
      if (tune && opt.item_steps == 0) tune_items(c);
    } catch (...) {
      petalo_mlem_destroy(c);
      throw;
    }
    *out = c;
    return PETALO_OK;
  });
}

int petalo_mlem_next(petalo_mlem_t* c, float* image_out, petalo_osem_t* id_out) {
  return guarded([&] {
    if (!c) return fail(PETALO_ERR_ARG, "context is NULL");
    HIP_CHECK(hipSetDevice(c->device));
    const size_t nv = nvox(c->fov);
    if (c->timing || !c->use_graphs) {
      enqueue_update(c, c->timing);
      c->have_timing = c->timing;
    } else {
      launch_update_graph(c, c->stream, c->graphs, c->osem.subset - 1, [&] { enqueue_update(c, false); });
      c->have_timing = false;
    }
    if (id_out) *id_out = c->osem;
    // Osem::advance (mlem.rs:82-88)
    c->osem.subset += 1;
    if (c->osem.subset > c->osem.n_subsets) {
      c->osem.subset = 1;
      c->osem.iteration += 1;
    }
    if (image_out) {
      if (c->precision != PETALO_PRECISION_F32) {
        k_narrow<<<grid_for(nv), kBlock, 0, c->stream>>>(c->image64.p, c->image.p, nv);
        HIP_CHECK(hipGetLastError());
      }
      HIP_CHECK(hipMemcpyAsync(image_out, c->image.p, nv * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    return PETALO_OK;
  });
}

int petalo_mlem_get_image(petalo_mlem_t* c, float* out) {
  return guarded([&] {
    if (!c || !out) return fail(PETALO_ERR_ARG, "NULL argument");
    HIP_CHECK(hipSetDevice(c->device));
    const size_t nv = nvox(c->fov);
    if (c->precision != PETALO_PRECISION_F32) {
      k_narrow<<<grid_for(nv), kBlock, 0, c->stream>>>(c->image64.p, c->image.p, nv);
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipMemcpyAsync(out, c->image.p, nv * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return PETALO_OK;
  });
}

int petalo_mlem_get_image_f64(petalo_mlem_t* c, double* out) {
  return guarded([&] {
    if (!c || !out) return fail(PETALO_ERR_ARG, "NULL argument");
    if (c->precision == PETALO_PRECISION_F32) return fail(PETALO_ERR_ARG, "not a binary64 context");
    HIP_CHECK(hipSetDevice(c->device));
    HIP_CHECK(hipMemcpyAsync(out, c->image64.p, nvox(c->fov) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return PETALO_OK;
  });
}

int petalo_mlem_set_image(petalo_mlem_t* c, const float* img) {
  return guarded([&] {
    if (!c || !img) return fail(PETALO_ERR_ARG, "NULL argument");
    HIP_CHECK(hipSetDevice(c->device));
    const size_t nv = nvox(c->fov);
    HIP_CHECK(hipMemcpyAsync(c->image.p, img, nv * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (c->precision != PETALO_PRECISION_F32) {
      k_widen<<<grid_for(nv), kBlock, 0, c->stream>>>(c->image.p, c->image64.p, nv);
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return PETALO_OK;
  });
}

int petalo_mlem_walk_stats(petalo_mlem_t* c, size_t subset, petalo_stats_t* out) {
  return guarded([&] {
    if (!c || !out) return fail(PETALO_ERR_ARG, "NULL argument");
    if (subset > c->n_subsets) return fail(PETALO_ERR_ARG, "subset out of range");
    HIP_CHECK(hipSetDevice(c->device));
    DevBuf<unsigned long long> cnt(CNT_N);
    zero_counters(cnt.p, c->stream);
    const size_t lo = subset ? (subset - 1) * c->set_size : 0;
    const size_t m = subset ? c->set_size : c->n_lors;
    if (m) {
      if (c->tof_on)
        k_walk_stats<true><<<grid_for(m), kBlock, 0, c->stream>>>(c->lors.p + lo, (uint32_t)m, c->dfov, c->dtof,
                                                                  cnt.p);
      else
        k_walk_stats<false><<<grid_for(m), kBlock, 0, c->stream>>>(c->lors.p + lo, (uint32_t)m, c->dfov, c->dtof,
                                                                   cnt.p);
      HIP_CHECK(hipGetLastError());
    }
    unsigned long long h[CNT_N];
    HIP_CHECK(hipMemcpyAsync(h, cnt.p, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    *out = petalo_stats_t{h[CNT_LORS], h[CNT_STEPS], h[CNT_NZ], h[CNT_SKIPPED], h[CNT_MISSED], h[CNT_RUNAWAY],
                          h[CNT_SHIFTED]};
    return PETALO_OK;
  });
}

int petalo_mlem_last_skipped(petalo_mlem_t* c, uint64_t* skipped) {
  return guarded([&] {
    if (!c || !skipped) return fail(PETALO_ERR_ARG, "NULL argument");
    HIP_CHECK(hipSetDevice(c->device));
    unsigned long long h[CNT_N];
    HIP_CHECK(hipMemcpyAsync(h, c->counters.p, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    *skipped = h[CNT_SKIPPED];
    return PETALO_OK;
  });
}

int petalo_mlem_accumulator(petalo_mlem_t* c, double* unit, double* r_max, double* r_geomean) {
  return guarded([&] {
    if (!c || !unit) return fail(PETALO_ERR_ARG, "NULL argument");
    *unit = 0.0;
    if (r_max) *r_max = NAN;
    if (r_geomean) *r_geomean = NAN;
    if (!c->rstat.p) return PETALO_OK;  // not a tiled context
    HIP_CHECK(hipSetDevice(c->device));
    std::vector<unsigned long long> h(kRStatWords);
    HIP_CHECK(hipMemcpyAsync(h.data(), c->rstat.p, kRStatWords * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    std::memcpy(unit, &h[kRUnitWord], 8);
    float m = 0.0f;
    double lsum = 0.0, n = 0.0;
    for (int k = 0; k < kRSlots; ++k) {
      uint32_t mb = (uint32_t)h[4 * k];
      float mk;
      std::memcpy(&mk, &mb, 4);
      m = std::max(m, mk);
      double lk;
      std::memcpy(&lk, &h[4 * k + 2], 8);
      lsum += lk;
      n += (double)h[4 * k + 3];
    }
    if (n > 0) {
      if (r_max) *r_max = m;
      if (r_geomean) *r_geomean = std::exp2(lsum / n);
    }
    return PETALO_OK;
  });
}

int petalo_mlem_set_timing(petalo_mlem_t* c, int on) {
  return guarded([&] {
    if (!c) return fail(PETALO_ERR_ARG, "context is NULL");
    if (on && c->evr.empty()) {
      HIP_CHECK(hipSetDevice(c->device));
      c->evr.resize(kTimingRing);
      for (auto& set : c->evr) set.fill(nullptr);
      for (auto& set : c->evr)
        for (auto& e : set) HIP_CHECK(hipEventCreate(&e));
    }
    c->timing = on != 0;
    c->have_timing = false;
    c->ev_count = 0;
    return PETALO_OK;
  });
}

int petalo_mlem_set_graphs(petalo_mlem_t* c, int on) {
  if (!c) return fail(PETALO_ERR_ARG, "context is NULL");
  c->use_graphs = on != 0;
  return PETALO_OK;
}

int petalo_mlem_last_timing(petalo_mlem_t* c, double* project_ms, double* update_ms) {
  return guarded([&] {
    if (!c) return fail(PETALO_ERR_ARG, "context is NULL");
    if (!c->have_timing) return fail(PETALO_ERR_ARG, "timing is off or no update has run");
    HIP_CHECK(hipSetDevice(c->device));
    HIP_CHECK(hipEventSynchronize(c->ev[3]));
    float a = 0, b = 0;
    HIP_CHECK(hipEventElapsedTime(&a, c->ev[0], c->ev[2]));
    HIP_CHECK(hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
    if (project_ms) *project_ms = a;
    if (update_ms) *update_ms = b;
    return PETALO_OK;
  });
}

int petalo_mlem_last_phase_timing(petalo_mlem_t* c, double* forward_ms, double* scatter_ms, double* update_ms) {
  return guarded([&] {
    if (!c) return fail(PETALO_ERR_ARG, "context is NULL");
    if (!c->have_timing) return fail(PETALO_ERR_ARG, "timing is off or no update has run");
    HIP_CHECK(hipSetDevice(c->device));
    HIP_CHECK(hipEventSynchronize(c->ev[3]));
    float a = 0, b = 0, d = 0;
    HIP_CHECK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIP_CHECK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    HIP_CHECK(hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
    if (forward_ms) *forward_ms = a;
    if (scatter_ms) *scatter_ms = b;
    if (update_ms) *update_ms = d;
    return PETALO_OK;
  });
}

int petalo_mlem_last_kernel_timing(petalo_mlem_t* c, double* ms, size_t n) {
  return guarded([&] {
    if (!c || (!ms && n)) return fail(PETALO_ERR_ARG, "NULL argument");
    if (!c->have_timing) return fail(PETALO_ERR_ARG, "timing is off or no update has run");
    HIP_CHECK(hipSetDevice(c->device));
    HIP_CHECK(hipEventSynchronize(c->ev[3]));
    float t[PETALO_N_TIMES] = {0, 0, 0, 0};
    HIP_CHECK(hipEventElapsedTime(&t[PETALO_TIME_FORWARD_TILES], c->ev[0], c->ev[4]));
    HIP_CHECK(hipEventElapsedTime(&t[PETALO_TIME_FORWARD_REDUCE], c->ev[4], c->ev[1]));
    HIP_CHECK(hipEventElapsedTime(&t[PETALO_TIME_SCATTER], c->ev[1], c->ev[2]));
    HIP_CHECK(hipEventElapsedTime(&t[PETALO_TIME_UPDATE], c->ev[2], c->ev[3]));
    for (size_t k = 0; k < n && k < PETALO_N_TIMES; ++k) ms[k] = t[k];
    return PETALO_OK;
  });
}

int petalo_mlem_drain_kernel_timings(petalo_mlem_t* c, double* ms, size_t max_updates, size_t* n_updates) {
  return guarded([&] {
    if (!c || !n_updates || (!ms && max_updates)) return fail(PETALO_ERR_ARG, "NULL argument");
    HIP_CHECK(hipSetDevice(c->device));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    const size_t n = std::min(c->ev_count, max_updates);
    const size_t ring = c->evr.size();
    for (size_t u = 0; u < n; ++u) {
      // the n most recent timed updates, oldest first
      const hipEvent_t* e = c->evr[(c->ev_next + ring - n + u) % ring].data();
      float t[PETALO_N_TIMES] = {0, 0, 0, 0};
      HIP_CHECK(hipEventElapsedTime(&t[PETALO_TIME_FORWARD_TILES], e[0], e[4]));
      HIP_CHECK(hipEventElapsedTime(&t[PETALO_TIME_FORWARD_REDUCE], e[4], e[1]));
      HIP_CHECK(hipEventElapsedTime(&t[PETALO_TIME_SCATTER], e[1], e[2]));
      HIP_CHECK(hipEventElapsedTime(&t[PETALO_TIME_UPDATE], e[2], e[3]));
      for (size_t k = 0; k < PETALO_N_TIMES; ++k) ms[u * PETALO_N_TIMES + k] = t[k];
    }
    *n_updates = n;
    c->ev_count = 0;
    return PETALO_OK;
  });
}

int petalo_mlem_info(petalo_mlem_t* c, int32_t* scheme, uint64_t* n_runs_total) {
  if (!c) return fail(PETALO_ERR_ARG, "context is NULL");
  if (scheme) *scheme = c->scheme;
  if (n_runs_total) {
    uint64_t t = 0;
    for (uint64_t v : c->n_runs) t += v;
    *n_runs_total = t;
  }
  return PETALO_OK;
}

int petalo_mlem_tile_shape(petalo_mlem_t* c, int32_t tile[3]) {
  if (!c || !tile) return fail(PETALO_ERR_ARG, "NULL argument");
  const bool t = c->scheme == PETALO_SCHEME_TILE;
  tile[0] = t ? c->tiles.tx : 0;
  tile[1] = t ? c->tiles.ty : 0;
  tile[2] = t ? c->tiles.tz : 0;
  return PETALO_OK;
}

int petalo_mlem_work_items(petalo_mlem_t* c, uint64_t* item_steps, uint64_t* n_items_total) {
  if (!c) return fail(PETALO_ERR_ARG, "context is NULL");
  if (item_steps) *item_steps = c->item_target;
  if (n_items_total) {
    uint64_t t = 0;
    for (const auto& w : c->work) t += w.n;
    *n_items_total = t;
  }
  return PETALO_OK;
}

int petalo_mlem_destroy(petalo_mlem_t* c) {
  if (!c) return PETALO_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& set : c->evr)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e);
  for (auto& g : c->graphs)
    if (g) (void)hipGraphExecDestroy(g);
  c->graphs.clear();
  c->lors.release();
  c->image.release();
  c->bp.release();
  c->sens.release();
  c->counters.release();
  c->runs.clear();
  c->work.clear();
  c->run_off.clear();
  c->run_pos.clear();
  c->pos4.clear();
  c->pos2.clear();
  c->lpos.clear();
  c->win.clear();
  c->irregular.clear();
  c->partial.release();
  c->ac.release();
  c->rbuf.release();
  c->bp64.release();
  c->image64.release();
  c->sens64.release();
  if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PETALO_OK;
}

// ---------------------------------------------------------------------------
// Batched reconstructions
}  // extern "C"

struct petalo_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<petalo_mlem*> ctx;
  size_t n_subsets = 1;
  petalo_osem_t osem{};
  std::vector<DevBuf<ReconDesc>> descs;   // per subset: one entry per reconstruction
  std::vector<DevBuf<int4>> items[2];     // [TOF][subset]: (tile, first, end, reconstruction)
  std::vector<DevBuf<ReduceDesc>> reduce; // per subset
  std::vector<unsigned> reduce_blocks;    // per subset: window blocks of all reconstructions
  std::vector<uint32_t> reduce_lds;       // per subset: LDS entries of k_lor_reduce_win_batch
  DevBuf<UpdateDesc> update;
  unsigned update_blocks = 0;
  size_t lds_f32 = 0, lds_f64 = 0;        // largest tile over the reconstructions
  bool use_graphs = true;
  std::vector<hipGraphExec_t> graphs;     // per subset
  bool timing = false, have_timing = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
  uint64_t item_target = 0;               // replayed steps per merged item; 0: the automatic rule
  // persistent schedule (k_batch_persistent): per subset its task list and
  // the tasks per (reconstruction, phase); one sync buffer for all
  bool persistent = false;
  std::vector<DevBuf<int4>> ptasks;
  std::vector<DevBuf<uint32_t>> pneed;
  std::vector<size_t> plds;               // per subset: dynamic LDS bytes of the launch
  std::vector<uint32_t> preduce_lds;      // per subset: reduction window entries staged in LDS
  DevBuf<uint32_t> psync;
  unsigned pgrid = 0;                     // resident workgroups of the launch
};

namespace {

// Work items of a whole batch per subset update.
constexpr uint64_t kBatchItems = 3072;
// Reduction window entries the persistent launch stages in LDS (79 KB).
constexpr uint32_t kPersistReduceEntries = 79 * 256;

// Zeroes the persistent schedule's sync words but the sticky error.
__global__ void k_zero_sync(uint32_t* __restrict__ sync, uint32_t n) {
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x)
    if (k != kSyncSticky) sync[k] = 0u;
}

// The rare unbounded walks of every reconstruction, one lane each.
void batch_irregular(petalo_batch* b, size_t sub) {
  for (petalo_mlem* c : b->ctx) {
    const size_t n_irr = c->irregular[sub].n;
    if (!n_irr) continue;
    const Lor* L = c->lors.p + sub * c->set_size;
    if (c->tof_on)
      k_fwd_irregular<true><<<grid_for(n_irr), kBlock, 0, b->stream>>>(L, c->irregular[sub].p, (uint32_t)n_irr,
                                                                       c->dfov, c->dtof, c->image.p, c->bp64.p,
                                                                       c->counters.p);
    else
      k_fwd_irregular<false><<<grid_for(n_irr), kBlock, 0, b->stream>>>(L, c->irregular[sub].p, (uint32_t)n_irr,
                                                                        c->dfov, c->dtof, c->image.p, c->bp64.p,
                                                                        c->counters.p);
  }
  HIP_CHECK(hipGetLastError());
}

void batch_enqueue(petalo_batch* b, size_t sub) {
  hipStream_t st = b->stream;
  const uint32_t n = (uint32_t)b->ctx.size();
  k_zero_counters_batch<<<1, 256, 0, st>>>(b->update.p, n, CNT_N);
  if (b->persistent) {
    k_zero_sync<<<1, 256, 0, st>>>(b->psync.p, kSyncDone + 3 * n);
    batch_irregular(b, sub);
    const size_t nt = b->ptasks[sub].n;
    if (nt) {
      const unsigned grid = (unsigned)std::min<size_t>(nt, b->pgrid);
      k_batch_persistent<<<grid, 1024, b->plds[sub], st>>>(b->ptasks[sub].p, (uint32_t)nt, b->descs[sub].p,
                                                          b->reduce[sub].p, b->update.p, b->pneed[sub].p,
                                                          b->psync.p, b->preduce_lds[sub]);
      HIP_CHECK(hipGetLastError());
    }
    return;
  }
  const ReconDesc* rd = b->descs[sub].p;
  if (b->items[1][sub].n)
    k_fwd_tile_batch<true><<<(unsigned)b->items[1][sub].n, 1024, b->lds_f32, st>>>(rd, b->items[1][sub].p);
  if (b->items[0][sub].n)
    k_fwd_tile_batch<false><<<(unsigned)b->items[0][sub].n, 1024, b->lds_f32, st>>>(rd, b->items[0][sub].p);
  HIP_CHECK(hipGetLastError());
  if (b->reduce_blocks[sub])
    k_lor_reduce_win_batch<<<b->reduce_blocks[sub], kWinThreads, (size_t)b->reduce_lds[sub] * sizeof(float), st>>>(
        b->reduce[sub].p, n, b->reduce_lds[sub]);
  HIP_CHECK(hipGetLastError());
  batch_irregular(b, sub);
  if (b->items[1][sub].n)
    k_scatter_batch<true><<<(unsigned)b->items[1][sub].n, 1024, b->lds_f64, st>>>(rd, b->items[1][sub].p);
  if (b->items[0][sub].n)
    k_scatter_batch<false><<<(unsigned)b->items[0][sub].n, 1024, b->lds_f64, st>>>(rd, b->items[0][sub].p);
  HIP_CHECK(hipGetLastError());
  k_update_batch<<<b->update_blocks, kBlock, 0, st>>>(b->update.p, n);
  HIP_CHECK(hipGetLastError());
}

}  // namespace

namespace {

// Descriptors and work items of one subset update of a batch: every
// reconstruction's items of that subset (as its context cut them), adjacent
// items of one tile merged up to item_target replayed steps (0: the batch's
// share, total / kBatchItems), so that small reconstructions do not pay a
// tile load and flush per few runs; then largest first over the whole batch.
void batch_build_subset(petalo_batch* b, size_t sub, uint64_t item_target) {
  const size_t n = b->ctx.size();
  std::vector<ReconDesc> d(n);
  std::vector<ReduceDesc> rd(n);
  uint32_t rblocks = 0;
  // (reconstruction, tile, first, end) with steps, from each context's items
  struct Item { int r, tile, a, e; uint64_t steps; bool tof; };
  std::vector<Item> all;
  uint64_t total_steps = 0;
  for (size_t r = 0; r < n; ++r) {
    petalo_mlem* c = b->ctx[r];
    const uint32_t m = (uint32_t)c->set_size;
    d[r] = ReconDesc{c->runs[sub].p, c->image.p, c->partial.p, c->rbuf.p, c->bp64.p, c->lpos[sub].p,
                     c->rstat.p, c->dfov, c->dtof, c->tiles};
    rd[r] = ReduceDesc{c->ac.p + sub * c->set_size, c->run_off[sub].p, c->run_pos[sub].p, c->pos4[sub].p,
                       c->partial.p, c->rbuf.p, c->win[sub].p, c->pos2[sub].p, c->rstat.p, m,
                       rblocks};
    rblocks += c->n_blocks[sub];
    b->reduce_lds[sub] = std::max(b->reduce_lds[sub], std::min(c->max_win[sub], kReduceWindow));
    std::vector<int4> h(c->work[sub].n);
    if (!h.empty())
      HIP_CHECK(hipMemcpy(h.data(), c->work[sub].p, h.size() * sizeof(int4), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < h.size(); ++k) {
      all.push_back(Item{(int)r, h[k].x, h[k].y, h[k].z, c->item_steps[sub][k], c->tof_on});
      total_steps += c->item_steps[sub][k];
    }
  }
  const uint64_t target = item_target ? std::max<uint64_t>(kMinItemSteps, item_target)
                                      : std::max<uint64_t>(kMinItemSteps, total_steps / kBatchItems + 1);
  std::sort(all.begin(), all.end(), [](const Item& u, const Item& v) {
    return u.r != v.r ? u.r < v.r : (u.tile != v.tile ? u.tile < v.tile : u.a < v.a);
  });
  std::vector<Item> merged;
  for (const Item& x : all) {
    if (!merged.empty()) {
      Item& m = merged.back();
      if (m.r == x.r && m.tile == x.tile && m.e == x.a && m.steps + x.steps <= target) {
        m.e = x.e;
        m.steps += x.steps;
        continue;
      }
    }
    merged.push_back(x);
  }
  std::stable_sort(merged.begin(), merged.end(), [](const Item& u, const Item& v) { return u.steps > v.steps; });
  // the persistent schedule's tasks: reconstructions by decreasing work, each
  // one's forward items (largest first), then its reduction blocks, its
  // backprojection items and its update chunks, phase after phase
  {
    std::vector<uint64_t> rsteps(n, 0);
    for (const Item& x : merged) rsteps[x.r] += x.steps;
    std::vector<size_t> order(n);
    for (size_t r = 0; r < n; ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](size_t u, size_t v) { return rsteps[u] > rsteps[v]; });
    std::vector<std::vector<int4>> per(n);
    for (const Item& x : merged) per[x.r].push_back(make_int4(x.tile, x.a, x.e, x.tof ? 4 : 0));
    std::vector<int4> tasks;
    std::vector<uint32_t> need(3 * n, 0u);
    auto head = [](uint32_t kind, size_t r, bool tof) { return (int)(kind | (tof ? 4u : 0u) | ((uint32_t)r << 3)); };
    for (size_t r : order)
      for (const int4& w : per[r]) {
        tasks.push_back(make_int4(head(kTaskForward, r, w.w != 0), w.x, w.y, w.z));
        ++need[3 * r + kTaskForward];
      }
    for (size_t r : order)
      for (uint32_t blk = 0; blk < b->ctx[r]->n_blocks[sub]; ++blk) {
        tasks.push_back(make_int4(head(kTaskReduce, r, false), (int)blk, 0, 0));
        ++need[3 * r + kTaskReduce];
      }
    for (size_t r : order)
      for (const int4& w : per[r]) {
        tasks.push_back(make_int4(head(kTaskScatter, r, w.w != 0), w.x, w.y, w.z));
        ++need[3 * r + kTaskScatter];
      }
    constexpr uint32_t kUpdateChunk = 16384;  // voxels per update task
    for (size_t r : order) {
      const uint32_t nv = (uint32_t)nvox(b->ctx[r]->fov);
      for (uint32_t v = 0; v < nv; v += kUpdateChunk)
        tasks.push_back(make_int4(head(kTaskUpdate, r, false), (int)v, (int)std::min(nv, v + kUpdateChunk), 0));
    }
    if (n >= (1u << 29)) throw std::runtime_error("too many reconstructions for the persistent schedule");
    b->ptasks[sub].alloc(std::max<size_t>(tasks.size(), 1));
    if (!tasks.empty())
      HIP_CHECK(hipMemcpy(b->ptasks[sub].p, tasks.data(), tasks.size() * sizeof(int4), hipMemcpyHostToDevice));
    b->ptasks[sub].n = tasks.size();
    b->pneed[sub].alloc(3 * n);
    HIP_CHECK(hipMemcpy(b->pneed[sub].p, need.data(), need.size() * 4, hipMemcpyHostToDevice));
  }
  b->descs[sub].alloc(n);
  HIP_CHECK(hipMemcpy(b->descs[sub].p, d.data(), n * sizeof(ReconDesc), hipMemcpyHostToDevice));
  b->reduce[sub].alloc(n);
  HIP_CHECK(hipMemcpy(b->reduce[sub].p, rd.data(), n * sizeof(ReduceDesc), hipMemcpyHostToDevice));
  b->reduce_blocks[sub] = rblocks;
  for (int t = 0; t < 2; ++t) {
    std::vector<int4> w;
    for (const Item& x : merged)
      if (x.tof == (t == 1)) w.push_back(make_int4(x.tile, x.a, x.e, x.r));
    b->items[t][sub].alloc(std::max<size_t>(w.size(), 1));
    if (!w.empty())
      HIP_CHECK(hipMemcpy(b->items[t][sub].p, w.data(), w.size() * sizeof(int4), hipMemcpyHostToDevice));
    b->items[t][sub].n = w.size();
  }
  // the persistent launch stages reduction windows of up to 79 KB (larger
  // ones are gathered from memory), so that two workgroups fit a CU's LDS
  // with the kernel's own shared variables
  b->preduce_lds[sub] = std::min<uint32_t>(b->reduce_lds[sub], kPersistReduceEntries);
  b->plds[sub] = std::max(std::max(b->lds_f32, b->lds_f64), (size_t)b->preduce_lds[sub] * sizeof(float));
}

// The batch's work granularity, tuned as tune_items does for one context:
// candidates the automatic rule and 2^14 .. 2^20 replayed steps per item,
// applied to every reconstruction's cut and to the merge; the first min(n_subsets, 4) subset updates of the whole
// batch run once untimed and three times timed per candidate (the best
// counts); the automatic rule
// stays unless another is at least kTuneMargin faster.  Then every subset is
// cut with the choice and the images, counters and subset positions reset.
void tune_batch(petalo_batch* b) {
  const size_t nt = std::min<size_t>(b->n_subsets, 4);
  const bool graphs = b->use_graphs;
  b->use_graphs = false;
  auto recut = [&](uint64_t target, size_t upto) {
    for (petalo_mlem* c : b->ctx) {
      c->item_target = target;
      for (size_t sub = 0; sub < upto; ++sub) cut_items(c, sub);
    }
    for (size_t sub = 0; sub < upto; ++sub) batch_build_subset(b, sub, target);
    b->item_target = target;
  };
  auto reset_position = [&] {
    b->osem = petalo_osem_t{b->n_subsets, 1, 1};
    for (petalo_mlem* c : b->ctx) c->osem = b->osem;
  };
  auto updates = [&] {
    reset_position();
    petalo_osem_t id;
    for (size_t k = 0; k < nt; ++k)
      if (petalo_batch_next(b, &id) != PETALO_OK) throw std::runtime_error(petalo_last_error());
    HIP_CHECK(hipStreamSynchronize(b->stream));
  };
  std::vector<size_t> seen;
  uint64_t best = 0;
  double best_ms = 1e300;
  for (uint64_t target : {(uint64_t)0, (uint64_t)1 << 14, (uint64_t)1 << 15, (uint64_t)1 << 16, (uint64_t)1 << 17,
                          (uint64_t)1 << 18, (uint64_t)1 << 19, (uint64_t)1 << 20}) {
    recut(target, nt);
    const size_t n_items = b->items[0][0].n + b->items[1][0].n;
    if (std::find(seen.begin(), seen.end(), n_items) != seen.end()) continue;
    seen.push_back(n_items);
    updates();  // warm-up
    double ms = 1e300;  // the best of three: a few updates time noisily
    for (int rep = 0; rep < 3; ++rep) {
      const auto t0 = std::chrono::steady_clock::now();
      updates();
      ms = std::min(ms, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    if (target == 0 || ms < best_ms * (1.0 - kTuneMargin)) {
      best_ms = ms;
      best = target;
    }
  }
  recut(best, b->n_subsets);
  for (petalo_mlem* c : b->ctx) {
    const size_t nv = nvox(c->fov);
    fill_f32(c->image.p, 1.0f, nv, b->stream);  // Image::ones (mlem.rs:17)
    fill_f64(c->bp64.p, 0.0, nv, b->stream);
    zero_counters(c->counters.p, b->stream);
  }
  HIP_CHECK(hipStreamSynchronize(b->stream));
  reset_position();
  b->use_graphs = graphs;
}

}  // namespace

extern "C" {

int petalo_batch_destroy(petalo_batch_t* b);

int petalo_batch_create(petalo_batch_t** out, size_t n, const petalo_fov_t* fovs, const petalo_tof_t* tofs,
                        const petalo_lor_t* const* lors, const size_t* n_lors, const float* const* sens,
                        size_t n_subsets, const petalo_mlem_options_t* options) {
  return guarded([&] {
    if (!out) return fail(PETALO_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n == 0) return fail(PETALO_ERR_ARG, "a batch needs at least one reconstruction");
    if (!fovs || !lors || !n_lors) return fail(PETALO_ERR_ARG, "NULL argument");
    petalo_mlem_options_t opt{PETALO_SCHEME_TILE, 1, 0, {0, 0, 0}, 0, PETALO_PRECISION_F32, 0};
    if (options) opt = *options;
    if (opt.scheme == PETALO_SCHEME_ATOMIC) return fail(PETALO_ERR_ARG, "batches run the tiled scheme");
    if (opt.precision != PETALO_PRECISION_F32) return fail(PETALO_ERR_ARG, "batches run the binary32 path");
    // autotune: the tile shapes stay as given or defaulted; the work granularity is tuned
    const bool tune = opt.autotune != 0 && opt.item_steps == 0;
    opt.autotune = 0;  // the contexts themselves are not tuned
    opt.scheme = PETALO_SCHEME_TILE;
    for (size_t r = 0; r < n; ++r) {
      const std::string why = check_mlem_args(&fovs[r], tofs ? &tofs[r] : nullptr, lors[r], n_lors[r],
                                              n_subsets, opt);
      if (!why.empty()) return fail(PETALO_ERR_ARG, "reconstruction " + std::to_string(r) + ": " + why);
    }
    require_device();
    HIP_CHECK(hipSetDevice(opt.device));
    auto* b = new petalo_batch();
    try {
      b->device = opt.device;
      b->n_subsets = n_subsets;
      b->item_target = (uint64_t)opt.item_steps;
      b->osem = petalo_osem_t{n_subsets, 1, 1};
      HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
      for (auto& e : b->ev) HIP_CHECK(hipEventCreate(&e));
      for (size_t r = 0; r < n; ++r)
        b->ctx.push_back(create_context(&fovs[r], tofs ? &tofs[r] : nullptr, lors[r], n_lors[r],
                                        sens ? sens[r] : nullptr, n_subsets, opt, b->stream));
      for (petalo_mlem* c : b->ctx) {
        b->lds_f32 = std::max(b->lds_f32, ((size_t)c->tiles.tvox + 1) * sizeof(float));
        b->lds_f64 = std::max(b->lds_f64, (size_t)c->tiles.tvox * sizeof(double));
      }
      b->descs.resize(n_subsets);
      b->ptasks.resize(n_subsets);
      b->pneed.resize(n_subsets);
      b->plds.assign(n_subsets, 0);
      b->preduce_lds.assign(n_subsets, 0);
      b->psync.alloc(kSyncDone + 3 * n);
      HIP_CHECK(hipMemset(b->psync.p, 0, (kSyncDone + 3 * n) * sizeof(uint32_t)));
      {
        // resident workgroups of the persistent launch (its largest LDS is
        // set per subset; two 1024-lane workgroups per CU at most)
        int cus = 0;
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
        b->pgrid = (unsigned)std::max(1, 2 * cus);
      }
      b->reduce.resize(n_subsets);
      b->reduce_blocks.assign(n_subsets, 0);
      b->reduce_lds.assign(n_subsets, 0);
      for (int t = 0; t < 2; ++t) b->items[t].resize(n_subsets);
      {
        std::vector<UpdateDesc> u(n);
        uint32_t blocks = 0;
        for (size_t r = 0; r < n; ++r) {
          petalo_mlem* c = b->ctx[r];
          const uint32_t nv = (uint32_t)nvox(c->fov);
          u[r] = UpdateDesc{c->image.p, c->bp64.p, c->sens.p, c->counters.p, c->rstat.p, nv, blocks};
          blocks += grid_for(nv);
        }
        b->update.alloc(n);
        HIP_CHECK(hipMemcpy(b->update.p, u.data(), n * sizeof(UpdateDesc), hipMemcpyHostToDevice));
        b->update_blocks = blocks;
      }
      for (size_t sub = 0; sub < n_subsets; ++sub) batch_build_subset(b, sub, (uint64_t)opt.item_steps);
      if (tune) tune_batch(b);
      HIP_CHECK(hipStreamSynchronize(b->stream));
    } catch (...) {
      petalo_batch_destroy(b);
      throw;
    }
    *out = b;
    return PETALO_OK;
  });
}

int petalo_batch_work_items(petalo_batch_t* b, uint64_t* item_steps, uint64_t* n_items_total) {
  if (!b) return fail(PETALO_ERR_ARG, "batch is NULL");
  if (item_steps) *item_steps = b->item_target;
  if (n_items_total) {
    uint64_t t = 0;
    for (int k = 0; k < 2; ++k)
      for (const auto& w : b->items[k]) t += w.n;
    *n_items_total = t;
  }
  return PETALO_OK;
}

int petalo_batch_next(petalo_batch_t* b, petalo_osem_t* id_out) {
  return guarded([&] {
    if (!b) return fail(PETALO_ERR_ARG, "batch is NULL");
    HIP_CHECK(hipSetDevice(b->device));
    const size_t sub = b->osem.subset - 1;
    if (b->timing) HIP_CHECK(hipEventRecord(b->ev[0], b->stream));
    if (b->use_graphs)
      launch_update_graph(nullptr, b->stream, b->graphs, sub, [&] { batch_enqueue(b, sub); });
    else
      batch_enqueue(b, sub);
    if (b->timing) HIP_CHECK(hipEventRecord(b->ev[1], b->stream));
    b->have_timing = b->timing;
    if (id_out) *id_out = b->osem;
    auto advance = [](petalo_osem_t& o) {  // Osem::advance (mlem.rs:82-88)
      o.subset += 1;
      if (o.subset > o.n_subsets) {
        o.subset = 1;
        o.iteration += 1;
      }
    };
    advance(b->osem);
    for (petalo_mlem* c : b->ctx) advance(c->osem);
    return PETALO_OK;
  });
}

int petalo_batch_get_image(petalo_batch_t* b, size_t r, float* out) {
  return guarded([&] {
    if (!b || !out) return fail(PETALO_ERR_ARG, "NULL argument");
    if (r >= b->ctx.size()) return fail(PETALO_ERR_ARG, "reconstruction index out of range");
    HIP_CHECK(hipSetDevice(b->device));
    petalo_mlem* c = b->ctx[r];
    HIP_CHECK(hipMemcpyAsync(out, c->image.p, nvox(c->fov) * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    uint32_t sticky = 0;
    HIP_CHECK(hipMemcpyAsync(&sticky, b->psync.p + kSyncSticky, 4, hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    if (sticky)
      return fail(PETALO_ERR_HIP, "a persistent-schedule update waited past its time limit: its images are not valid");
    return PETALO_OK;
  });
}

int petalo_batch_set_options(petalo_batch_t* b, int timing, int use_graphs) {
  if (!b) return fail(PETALO_ERR_ARG, "batch is NULL");
  b->timing = timing != 0;
  b->use_graphs = use_graphs != 0;
  b->have_timing = false;
  return PETALO_OK;
}

int petalo_batch_set_schedule(petalo_batch_t* b, int persistent) {
  return guarded([&] {
    if (!b) return fail(PETALO_ERR_ARG, "batch is NULL");
    HIP_CHECK(hipSetDevice(b->device));
    if ((persistent != 0) != b->persistent) {
      HIP_CHECK(hipStreamSynchronize(b->stream));
      for (auto& g : b->graphs)
        if (g) {
          (void)hipGraphExecDestroy(g);
          g = nullptr;
        }
      b->persistent = persistent != 0;
    }
    return PETALO_OK;
  });
}

int petalo_batch_last_ms(petalo_batch_t* b, double* ms) {
  return guarded([&] {
    if (!b || !ms) return fail(PETALO_ERR_ARG, "NULL argument");
    if (!b->have_timing) return fail(PETALO_ERR_ARG, "timing is off or no update has run");
    HIP_CHECK(hipSetDevice(b->device));
    HIP_CHECK(hipEventSynchronize(b->ev[1]));
    float t = 0;
    HIP_CHECK(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
    *ms = t;
    return PETALO_OK;
  });
}

int petalo_batch_destroy(petalo_batch_t* b) {
  if (!b) return PETALO_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& g : b->graphs)
    if (g) (void)hipGraphExecDestroy(g);
  for (petalo_mlem* c : b->ctx) petalo_mlem_destroy(c);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  b->descs.clear();
  b->reduce.clear();
  b->update.release();
  for (auto& v : b->items) v.clear();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PETALO_OK;
}

// ---------------------------------------------------------------------------
}  // extern "C"

namespace {
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
}  // namespace

extern "C" {

int petalo_copy_probe(size_t bytes, int iters, double* ms_per_copy) {
  return guarded([&] {
    if (bytes < 16 || iters < 1) return fail(PETALO_ERR_ARG, "need at least 16 bytes and 1 iteration");
    require_device();
    const size_t n = bytes / 16;
    DevBuf<float4> a(n), b(n);
    HIP_CHECK(hipMemset(a.p, 0, n * 16));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    k_copy_probe<<<2048, 256>>>(a.p, b.p, n);
    HIP_CHECK(hipEventRecord(e0));
    for (int k = 0; k < iters; ++k) k_copy_probe<<<2048, 256>>>(a.p, b.p, n);
    HIP_CHECK(hipEventRecord(e1));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_copy) *ms_per_copy = ms / iters;
    return PETALO_OK;
  });
}

int petalo_system_matrix_rows(const petalo_fov_t* fov, const petalo_tof_t* tof, const petalo_lor_t* lors,
                              size_t n_lors, size_t cap, int32_t* idx, float* w, uint32_t* count,
                              uint32_t* steps) {
  return guarded([&] {
    std::string why;
    if (!fov_ok(fov, why)) return fail(PETALO_ERR_ARG, why);
    if (n_lors && (!lors || !count || !steps || (cap && (!idx || !w)))) return fail(PETALO_ERR_ARG, "NULL argument");
    if (n_lors >= (1ull << 32) || cap >= (1ull << 32)) return fail(PETALO_ERR_ARG, "too many LORs or too large cap");
    if (n_lors == 0) return PETALO_OK;
    require_device();
    const DevFov f = to_dev(*fov);
    const DevTof t = to_dev(tof);
    DevBuf<Lor> dl(n_lors);
    DevBuf<int32_t> di(std::max<size_t>(n_lors * cap, 1));
    DevBuf<float> dw(std::max<size_t>(n_lors * cap, 1));
    DevBuf<uint32_t> dc(n_lors), ds(n_lors);
    upload_h2d(dl.p, lors, n_lors * sizeof(Lor), nullptr);
    if (t.on)
      k_rows<true><<<grid_for(n_lors), kBlock>>>(dl.p, (uint32_t)n_lors, f, t, (uint32_t)cap, di.p, dw.p, dc.p, ds.p);
    else
      k_rows<false><<<grid_for(n_lors), kBlock>>>(dl.p, (uint32_t)n_lors, f, t, (uint32_t)cap, di.p, dw.p, dc.p, ds.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    if (cap) {
      download_d2h(idx, di.p, n_lors * cap * sizeof(int32_t), nullptr);
      download_d2h(w, dw.p, n_lors * cap * sizeof(float), nullptr);
    }
    download_d2h(count, dc.p, n_lors * sizeof(uint32_t), nullptr);
    download_d2h(steps, ds.p, n_lors * sizeof(uint32_t), nullptr);
    return PETALO_OK;
  });
}

// ---------------------------------------------------------------------------
int petalo_sensitivity_lors(const petalo_fov_t* fov, size_t n_lors, uint64_t seed, float L, float D,
                            petalo_lor_t* out) {
  return guarded([&] {
    std::string why;
    if (!fov_ok(fov, why)) return fail(PETALO_ERR_ARG, why);
    if (n_lors && !out) return fail(PETALO_ERR_ARG, "out is NULL");
    if (n_lors >= (1ull << 32)) return fail(PETALO_ERR_ARG, "more than 2^32-1 LORs");
    if (!(L > 0.0f) || !(D > 0.0f)) return fail(PETALO_ERR_ARG, "detector length and diameter must be positive");
    if (n_lors == 0) return PETALO_OK;
    require_device();
    DevBuf<Lor> d(n_lors);
    DevBuf<unsigned long long> cnt(CNT_N);
    zero_counters(cnt.p, nullptr);
    k_sens_lors<<<sens_lors_grid(n_lors), kBlock>>>((uint32_t)n_lors, 0u, seed, to_dev(*fov), L, D / 2.0f, d.p,
                                                    cnt.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(out, d.p, n_lors * sizeof(Lor), hipMemcpyDeviceToHost));
    unsigned long long h[CNT_N];
    HIP_CHECK(hipMemcpy(h, cnt.p, sizeof h, hipMemcpyDeviceToHost));
    if (h[CNT_UNFOUND])
      return fail(PETALO_ERR_ARG, std::to_string(h[CNT_UNFOUND]) + " sensitivity LORs found no line through the FOV in " +
                                      std::to_string(kSensMaxAttempts) + " draws (FOV outside the detector's reach)");
    return PETALO_OK;
  });
}

}  // extern "C"

namespace {

// Sensitivity image on the tiled scheme, LORs in batches of kSensBatch.
constexpr uint32_t kSensBatch = 1u << 25;
// Work item size of the sensitivity kernels, in runs: each item clears and
// flushes its whole tile, so items are larger than the reconstruction's.
constexpr uint32_t kSensItemRuns = 16384;
// Most length classes of a sensitivity batch's segment binning.
constexpr uint32_t kSensClasses = 4;

// Unbounded walks (count 0xffffffff): copied out for the one-lane kernel,
// their count set to 0.  counts[n] is set to 0 for the exclusive scan.
// Sort key of a sensitivity LOR: about the number of steps of its walk (the
// crossings of each axis before the first axis runs out, from the boundary
// arithmetic; an estimate, for ordering only), so that the binning kernel's
// lanes, which take several LORs each, get walks of similar lengths.  The
// estimate >> shift, at most 255: one 8-bit radix pass sorts the batch.
__global__ void k_walk_len_keys(const Lor* __restrict__ lors, uint32_t n, DevFov f, uint32_t shift,
                                uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Lor l = load_lor(lors, i);
  WalkC c;
  WalkS w;
  float est = 0.0f;
  if (walk_setup(l, f, c, w) && bounded(w)) {
    const float nb[3] = {w.nb_x, w.nb_y, w.nb_z}, vl[3] = {c.vl_x, c.vl_y, c.vl_z};
    const int32_t r[3] = {w.r_x, w.r_y, w.r_z};
    float d = 3.0e38f;
    for (int a = 0; a < 3; ++a) d = fminf(d, nb[a] + (float)(r[a] - 1) * vl[a]);
    for (int a = 0; a < 3; ++a) {
      float q = (d - nb[a]) / vl[a];     // NaN or negative for an axis that is never crossed
      q = q >= 0.0f ? fminf(q + 1.0f, (float)r[a]) : 0.0f;
      est += q;
    }
  }
  keys[i] = min((uint32_t)fminf(est, 65535.0f) >> shift, 255u);
  idx[i] = i;
}

__global__ void k_gather_lors_by(const Lor* __restrict__ in, const uint32_t* __restrict__ idx, uint32_t n,
                                 Lor* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = in[idx[j]];
}

// irr holds cap LORs; *n_irr counts every unbounded walk taken, also those
// past cap (the caller then takes the batch's again, k_retake_irregular).
__global__ void k_take_irregular(uint32_t* __restrict__ counts, uint32_t n, const Lor* __restrict__ lors,
                                 const uint32_t* __restrict__ order, Lor* __restrict__ irr, uint32_t cap,
                                 uint32_t* __restrict__ n_irr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == n) counts[n] = 0;
  if (i >= n || counts[i] != 0xffffffffu) return;
  counts[i] = 0;
  const uint32_t at = atomicAdd(n_irr, 1u);
  if (at < cap) irr[at] = lors[order ? order[i] : i];
}

// The unbounded walks of LORs 0 .. n - 1 (lors[order[i]]) found again from
// the walk setup, as the binning kernels flag them, into irr (room for cap).
__global__ void k_retake_irregular(const Lor* __restrict__ lors, const uint32_t* __restrict__ order, uint32_t n,
                                   DevFov f, Lor* __restrict__ irr, uint32_t cap, uint32_t* __restrict__ n_irr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Lor l = load_lor(lors, order ? order[i] : i);
  WalkC c;
  WalkS w;
  if (!walk_setup(l, f, c, w) || bounded(w)) return;
  const uint32_t at = atomicAdd(n_irr, 1u);
  if (at < cap) irr[at] = l;
}



// Tile counts up to which a sensitivity batch is binned straight into tile
// order (k_bin_runs_seg keeps a cursor and the end of each (tile, length
// class) segment in LDS); above it the runs are binned LOR-major and sorted
// by tile.
constexpr uint32_t kSegMaxTiles = 2048;

// Work items of a sensitivity batch binned by k_bin_runs_seg: tile t's runs
// are [soff[t * ng], soff[(t + 1) * ng]) (soff[ntiles * ng] = the total),
// cut into chunks of `per` -- the full chunks of all tiles first, then each
// tile's remainder, so that the largest items start first -- followed by
// empty items (first == end) up to cap, which their workgroups skip.  One
// workgroup of 1024 lanes; lane q takes a contiguous range of tiles.
__global__ __launch_bounds__(1024) void k_seg_items(const uint32_t* __restrict__ soff, uint32_t ng, uint32_t ntiles,
                                                    uint32_t per, uint32_t cap, int4* __restrict__ items) {
  __shared__ uint32_t full_s[1024], rem_s[1024];
  const uint32_t q = threadIdx.x, span = (ntiles + 1023u) / 1024u;
  const uint32_t t0 = min(q * span, ntiles), t1 = min(t0 + span, ntiles);
  uint32_t nf = 0, nr = 0;
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t c = soff[(size_t)(t + 1) * ng] - soff[(size_t)t * ng];
    nf += c / per;
    nr += c % per ? 1u : 0u;
  }
  full_s[q] = nf;
  rem_s[q] = nr;
  __syncthreads();
  for (uint32_t off = 1; off < 1024u; off <<= 1) {  // inclusive scans
    const uint32_t a = q >= off ? full_s[q - off] : 0u, b = q >= off ? rem_s[q - off] : 0u;
    __syncthreads();
    full_s[q] += a;
    rem_s[q] += b;
    __syncthreads();
  }
  const uint32_t n_full = full_s[1023], n_items = n_full + rem_s[1023];
  uint32_t fi = full_s[q] - nf, ri = n_full + rem_s[q] - nr;
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t a = soff[(size_t)t * ng], e = soff[(size_t)(t + 1) * ng];
    uint32_t x = a;
    for (; e - x >= per; x += per, ++fi)
      if (fi < cap) items[fi] = make_int4((int)t, (int)x, (int)(x + per), 0);
    if (x < e) {
      if (ri < cap) items[ri] = make_int4((int)t, (int)x, (int)e, 0);
      ++ri;
    }
  }
  for (uint32_t k = n_items + q; k < cap; k += 1024u) items[k] = make_int4(0, 0, 0, 0);
}

// Sensitivity image (one FOV) with the tiled kernels: per batch, bin the
// LORs into runs in tile order -- straight into (tile, workgroup) segments
// sized by a counting pass, or, above kSegMaxTiles tiles, LOR-major and then
// sorted by tile (the runs read through the permutation) -- then (attenuation
// present) forward project mu over the runs, reduce to 1/exp(-p) per LOR, and
// backproject;
// with mu == 0 everywhere every LOR has p = 0 exactly and r = 1, so the
// forward half is skipped.  Unbounded walks go to the one-lane kernel.
// produce(dst, first, count, stream) writes LORs first .. first + count - 1.
// by_length: bin each batch in order of walk length (LORs in random order:
// the binning lanes then walk similar lengths); false keeps the source's
// order (neighbouring element pairs share cache lines and LDS banks).
// want_stats: count every step and pushed step (the counting pass then walks
// every step instead of jumping from tile to tile).
using LorSource = std::function<void(Lor*, uint64_t, uint32_t, hipStream_t)>;


// Device buffers and stream of the tiled sensitivity image, kept per device
// between calls: the batch buffers of 2^25 LORs and the runs of a batch
// come to about 10 GB at 120^3, and allocating and freeing them took about
// 4 ms of a 50 ms call.  A call leases its device's set (SensLease), or a fresh one
// when a concurrent call holds it; a set given back when the slot is taken,
// or after an error, is freed.  petalo_release_workspaces frees the kept
// sets.
struct SensWork {
  hipStream_t stream = nullptr;
  Grow<float> dmu, dout, rbuf, partial;
  Grow<double> acc;
  Grow<unsigned long long> cnt, rstat;
  Grow<Lor> dl, lb, lraw, irr;
  Grow<uint32_t> counts, offs, n_irr, seg, soff, lk, lk2, li, li2, perm, perm2, pos;
  Grow<uint4> pos4;
  Grow<unsigned char> lsort_tmp, scan_tmp, sort_tmp;
  Grow<uint64_t> dstart, keys, keys2;
  Grow<Run> unsorted;
  Grow<int4> work;
  SensWork() { HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)); }
  ~SensWork() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  SensWork(const SensWork&) = delete;
  SensWork& operator=(const SensWork&) = delete;
};

constexpr int kSensDevices = 64;
std::mutex g_sens_mu;
// never destroyed: a set still kept when the process exits is left to the
// driver, as freeing device memory from static destructors may run after the
// HIP runtime has shut down
std::unique_ptr<SensWork>* const g_sens_work = new std::unique_ptr<SensWork>[kSensDevices];

struct SensLease {
  int dev = -1;
  std::unique_ptr<SensWork> w;
  SensLease() {
    int d = 0;
    HIP_CHECK(hipGetDevice(&d));
    if (d >= 0 && d < kSensDevices) {
      std::lock_guard<std::mutex> lk(g_sens_mu);
      dev = d;
      w = std::move(g_sens_work[d]);
    }
    if (!w) w.reset(new SensWork());
  }
  ~SensLease() {
    // after an error the set's stream may hold failed work: drop it
    if (dev < 0 || std::uncaught_exceptions() > 0) return;
    if (hipStreamSynchronize(w->stream) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(g_sens_mu);
    if (!g_sens_work[dev]) g_sens_work[dev] = std::move(w);
  }
  SensLease(const SensLease&) = delete;
  SensLease& operator=(const SensLease&) = delete;
};

// Work items of a sensitivity batch, built on the device from the tile
// starts of the sorted runs (UINT64_MAX where a tile has none): every tile's
// runs in chunks of `per` -- the full chunks of all tiles first, then each
// tile's remainder, so that the largest items start first -- and empty items
// (first == end) up to cap, which their workgroups skip.  One workgroup;
// thread 0 walks the tiles (a few hundred to a few thousand).
__global__ void k_sens_items(const uint64_t* __restrict__ start, uint32_t ntiles, uint64_t total, uint32_t per,
                             uint32_t cap, int4* __restrict__ items) {
  __shared__ uint32_t n_items;
  if (threadIdx.x == 0) {
    uint32_t n = 0;
    for (int pass = 0; pass < 2; ++pass) {
      int32_t pend = -1;  // the last non-empty tile seen, whose end is the next one's start
      uint64_t pend_start = 0;
      for (uint32_t t = 0; t <= ntiles; ++t) {
        const uint64_t st = t < ntiles ? start[t] : total;
        if (t < ntiles && st == UINT64_MAX) continue;
        if (pend >= 0) {
          const uint64_t a = pend_start, e = st, full = a + (e - a) / per * per;
          if (pass == 0) {
            for (uint64_t x = a; x < full && n < cap; x += per) items[n++] = make_int4(pend, (int)x, (int)(x + per), 0);
          } else if (full < e && n < cap) {
            items[n++] = make_int4(pend, (int)full, (int)e, 0);
          }
        }
        pend = (int32_t)t;
        pend_start = st;
      }
    }
    n_items = n;
  }
  __syncthreads();
  for (uint32_t k = n_items + threadIdx.x; k < cap; k += blockDim.x) items[k] = make_int4(0, 0, 0, 0);
}

// Whether every value of mu is zero (either sign); scanned in blocks, so a
// non-zero image ends the scan early.
bool all_zero(const float* mu, size_t n) {
  const size_t kBlk = 4096;
  for (size_t a = 0; a < n; a += kBlk) {
    const size_t e = std::min(n, a + kBlk);
    uint32_t any = 0;
    for (size_t k = a; k < e; ++k) {
      uint32_t b;
      std::memcpy(&b, mu + k, 4);
      any |= b & 0x7fffffffu;
    }
    if (any) return false;
  }
  return true;
}

// mu_zero: the attenuation image (dmu on the device) is zero everywhere.
void sens_tiled(SensWork& ws, const petalo_fov_t& fov, const DevFov& f, bool mu_zero, const float* dmu,
                uint64_t n_lors, const LorSource& produce, double* acc, unsigned long long* cnt, bool by_length,
                bool want_stats) {
  hipStream_t s = ws.stream;
  // the default transaxial tile, as deep as a 64-bit LDS tile allows (the
  // unit-weight scatter's 32-bit tile then still fits two workgroups per CU,
  // the 64-bit one one): deeper tiles cut the runs per LOR, and the binning
  // scales with them (for 120^3: 40 x 40 x 12, 4.1 runs per LOR against 5.3
  // at the reconstruction's depth 6)
  const int32_t no_shape[3] = {0, 0, 0};
  const TileGeom g0 = tile_shape(fov, no_shape);
  const int32_t sens_shape[3] = {0, 0, even_split(fov.n[2], std::max(1, kMaxTileVoxels / (g0.px * g0.ty)))};
  const TileGeom g = make_tiles(fov, sens_shape);
  // tile order straight from the binning (segments), or LOR-major runs sorted by tile
  const bool segs = (uint32_t)g.ntiles <= kSegMaxTiles;
  const uint32_t nt = (uint32_t)g.ntiles;
  // length classes of the segment binning: as many as fit 8 bytes per (tile,
  // class) in 32 KB of the fill pass's LDS, at most kSensClasses
  RunClasses rc;
  if (segs) {
    while (rc.n < kSensClasses && (size_t)nt * rc.n * 2 <= 4096) rc.n *= 2;
    rc.scale = (double)rc.n / (double)(g.tx + g.ty + g.tz);
  }
  const uint32_t ncls = rc.n;
  // without attenuation every r is 1: the scatter's tiles sum in 32-bit fixed
  // point, a step adding round(w * 2^e) units with e such that an item's
  // kSensItemRuns runs of at most 1.25 voxel diagonals each stay below 2^32
  bool unit32 = false;
  float unit32_s = 0.0f;
  double unit32_unit = 0.0;
  if (mu_zero) {
    int lg = 0;
    while ((1u << lg) < kSensItemRuns) ++lg;
    const float diag = std::sqrt(f.vs_x * f.vs_x + f.vs_y * f.vs_y + f.vs_z * f.vs_z);
    const float bound = 1.25f * diag;
    int k = 0;
    std::frexp(bound, &k);  // bound < 2^k
    const int e = (32 - lg) - k;
    if (bound > 0.0f && std::isfinite(bound) && e > -120 && e < 120) {
      unit32 = true;
      unit32_s = std::ldexp(1.0f, e);
      unit32_unit = std::ldexp(1.0, -e);
    }
  }
  // walk lengths are below nx + ny + nz steps: the shift that brings them under 256
  uint32_t len_shift = 0;
  while ((((uint64_t)fov.n[0] + fov.n[1] + fov.n[2]) >> len_shift) > 255u) ++len_shift;
  const uint32_t shift = mu_zero ? 32u : kLorBlockShift;
  const uint32_t tshift = mu_zero ? 12u : 32u;
  int tile_bits = 1;
  while ((1 << tile_bits) < g.ntiles) ++tile_bits;
  const int end_bit = (int)tshift + tile_bits;
  const DevTof notof = no_tof();
  ensure_lds_attributes();
  const size_t bmax = (size_t)std::min<uint64_t>(n_lors, kSensBatch);
  const size_t ng_max = bin_grid(bmax);
  const size_t n_seg = segs ? (size_t)nt * ncls * ng_max + 1 : 0;
  // buffers of the whole call: batch-sized ones now, run-sized ones grown
  // to the largest batch (no allocation, and one wait, per batch)
  // by length, the segment binning reads lraw through the sort's order and needs no copy in lb
  Lor* lb = ws.lb.ensure(by_length && segs ? 0 : bmax);
  Lor* lraw = ws.lraw.ensure(by_length ? bmax : 0);
  // unbounded walks of several batches (a handful per 2^25 random LORs):
  // grown when a batch brings more than the room left
  Lor* irr = ws.irr.ensure(1u << 16);
  size_t irr_cap = ws.irr.b.n;
  uint32_t* counts = ws.counts.ensure(bmax + 1);
  uint32_t* offs = ws.offs.ensure(bmax + 1);
  uint32_t* n_irr = ws.n_irr.ensure(1);
  uint32_t* seg = ws.seg.ensure(n_seg);
  uint32_t* soff = ws.soff.ensure(n_seg);
  uint32_t* lk = ws.lk.ensure(by_length ? bmax : 0);
  uint32_t* lk2 = ws.lk2.ensure(by_length ? bmax : 0);
  uint32_t* li = ws.li.ensure(by_length ? bmax : 0);
  uint32_t* li2 = ws.li2.ensure(by_length ? bmax : 0);
  float* rbuf = ws.rbuf.ensure(mu_zero ? 0 : bmax);
  uint4* pos4 = ws.pos4.ensure(mu_zero ? 0 : bmax);
  unsigned long long* rstat = ws.rstat.ensure(kRStatWords);  // each batch's r statistics
  size_t lsort_bytes = 0, scan_bytes = 0, seg_scan_bytes = 0;
  if (by_length)
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, lsort_bytes, lk, lk2, li, li2, (int)bmax, 0, 8, s));
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, counts, offs, (int)bmax + 1, s));
  if (segs) {
    if (n_seg >= (1ull << 31)) throw std::runtime_error("too many (workgroup, tile) segments in a sensitivity batch");
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, seg_scan_bytes, seg, soff, (int)n_seg, s));
    scan_bytes = std::max(scan_bytes, seg_scan_bytes);
  }
  unsigned char* lsort_tmp = ws.lsort_tmp.ensure(std::max<size_t>(lsort_bytes, 1));
  unsigned char* scan_tmp = ws.scan_tmp.ensure(std::max<size_t>(scan_bytes, 1));
  uint64_t* dstart = ws.dstart.ensure(segs ? 0 : g.ntiles);
  uint32_t irr_pending = 0;  // unbounded walks held in irr
  for (uint64_t first = 0; first < n_lors; first += kSensBatch) {
    const uint32_t b = (uint32_t)std::min<uint64_t>(kSensBatch, n_lors - first);
    const uint32_t ng = bin_grid(b);
    produce(by_length ? lraw : lb, first, b, s);
    HIP_CHECK(hipGetLastError());
    if (by_length) {
      // the batch in order of walk length (the image is a sum over LORs)
      k_walk_len_keys<<<grid_for(b), kBlock, 0, s>>>(lraw, b, f, len_shift, lk, li);
      HIP_CHECK(hipGetLastError());
      size_t sb = lsort_bytes;
      HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(lsort_tmp, sb, lk, lk2, li, li2, (int)b, 0, 8, s));
      // the segment binning reads the batch through the order; the sorted
      // binning walks a reordered copy
      if (!segs) k_gather_lors_by<<<grid_for(b), kBlock, 0, s>>>(lraw, li2, b, lb);
      HIP_CHECK(hipGetLastError());
    }
    const Lor* src_lors = (by_length && segs) ? lraw : lb;
    const uint32_t* lor_order = (by_length && segs) ? li2 : nullptr;
    const size_t nsb = (size_t)nt * ncls * ng + 1;  // this batch's segments and the closing zero
    if (segs) {
      HIP_CHECK(hipMemsetAsync(seg + (nsb - 1), 0, 4, s));
      const size_t lds = (size_t)nt * ncls * 4;
      if (want_stats)
        k_bin_runs_seg<false, true><<<ng, kBlock, lds, s>>>(src_lors, lor_order, b, f, g, rc, counts, seg, nullptr,
                                                            nullptr, nullptr, nullptr, cnt);
      else
        k_bin_runs_seg<false, false><<<ng, kBlock, lds, s>>>(src_lors, lor_order, b, f, g, rc, counts, seg, nullptr,
                                                             nullptr, nullptr, nullptr, cnt);
    } else {
      k_bin_runs<false, false><<<ng, kBlock, 0, s>>>(lb, b, f, notof, g, counts, nullptr, nullptr, nullptr,
                                                      shift, cnt);
    }
    HIP_CHECK(hipGetLastError());
    // unbounded walks collect in irr over the batches (the one-lane kernel
    // costs its longest walk whatever their number) and are projected after
    // the last batch, or when a batch brings more than irr has room for
    if (irr_pending == 0) HIP_CHECK(hipMemsetAsync(n_irr, 0, 4, s));
    k_take_irregular<<<grid_for((size_t)b + 1), kBlock, 0, s>>>(counts, b, src_lors, lor_order, irr,
                                                                 (uint32_t)irr_cap, n_irr);
    HIP_CHECK(hipGetLastError());
    uint32_t total = 0, h_irr = 0;
    if (!segs || !mu_zero) {
      size_t tb = scan_bytes;
      HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb, counts, offs, (int)b + 1, s));
    }
    if (segs) {
      size_t tb = scan_bytes;
      HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb, seg, soff, (int)nsb, s));
      HIP_CHECK(hipMemcpyAsync(&total, soff + (nsb - 1), 4, hipMemcpyDeviceToHost, s));
    } else {
      HIP_CHECK(hipMemcpyAsync(&total, offs + b, 4, hipMemcpyDeviceToHost, s));
    }
    HIP_CHECK(hipMemcpyAsync(&h_irr, n_irr, 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));  // the batch's run count sizes the run buffers
    if (total >= (1u << 31)) throw std::runtime_error("2^31 or more tile runs in one sensitivity batch");
    if (total) {
      const uint32_t cap = total / kSensItemRuns + nt + 1;
      int4* items = ws.work.ensure(cap);
      Run* un = ws.unsorted.ensure(total);
      // sorted position of every LOR-major run (attenuation present)
      uint32_t* ps = mu_zero ? nullptr : ws.pos.ensure(total);
      uint32_t* order = nullptr;  // runs in tile order are un[order[j]]; nullptr: un[j]
      if (segs) {
        k_bin_runs_seg<true, false><<<ng, kBlock, (size_t)nt * ncls * 8, s>>>(src_lors, lor_order, b, f, g, rc, nullptr,
                                                                              seg, soff, offs, un, ps, cnt);
        // tile t's runs: segments [t * ncls * ng, (t + 1) * ncls * ng)
        k_seg_items<<<1, 1024, 0, s>>>(soff, ncls * ng, nt, kSensItemRuns, cap, items);
        HIP_CHECK(hipGetLastError());
      } else {
        uint64_t* k1 = ws.keys.ensure(total);
        uint64_t* k2 = ws.keys2.ensure(total);
        uint32_t* p1 = ws.perm.ensure(total);
        uint32_t* p2 = ws.perm2.ensure(total);
        k_bin_runs<false, true><<<ng, kBlock, 0, s>>>(lb, b, f, notof, g, nullptr, offs, un, k1, shift, cnt);
        k_iota<<<grid64(total), kBlock, 0, s>>>(p1, total);
        HIP_CHECK(hipGetLastError());
        size_t sort_bytes = 0;
        HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k1, k2, p1, p2, (int)total, 0, end_bit, s));
        unsigned char* stmp = ws.sort_tmp.ensure(std::max<size_t>(sort_bytes, 1));
        HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(stmp, sort_bytes, k1, k2, p1, p2, (int)total, 0, end_bit, s));
        if (ps) k_invert_perm<<<grid64(total), kBlock, 0, s>>>(p2, total, ps);
        HIP_CHECK(hipMemsetAsync(dstart, 0xff, (size_t)g.ntiles * 8, s));
        k_tile_starts<<<grid64(total), kBlock, 0, s>>>(k2, total, tshift, nt, dstart);
        k_sens_items<<<1, 256, 0, s>>>(dstart, nt, total, kSensItemRuns, cap, items);
        HIP_CHECK(hipGetLastError());
        order = p2;
      }
      if (!mu_zero) {
        float* part = ws.partial.ensure(total);
        k_fwd_tile<false><<<cap, 1024, ((size_t)g.tvox + 1) * sizeof(float), s>>>(un, items, dmu, f, notof, g, part,
                                                                                   order, nullptr);
        k_pack_pos4<<<grid_for(b), kBlock, 0, s>>>(offs, ps, b, pos4);
        // this batch's r statistics, for the backprojection's accumulator
        k_zero_u64<<<1, 320, 0, s>>>(rstat, kRStatWords);
        k_lor_reduce<true><<<reduce_grid(b), kBlock, 0, s>>>(nullptr, b, offs, ps, pos4, part, rbuf,
                                                             rstat);
        HIP_CHECK(hipGetLastError());
      }
      if (unit32)
        k_scatter_unit32<<<cap, 1024, (size_t)g.tvox * sizeof(uint32_t), s>>>(un, items, f, g, acc, order, unit32_s,
                                                                              unit32_unit);
      else
        k_scatter<false><<<cap, 1024, (size_t)g.tvox * sizeof(double), s>>>(un, items, mu_zero ? nullptr : rbuf, f,
                                                                            notof, g, acc, order, rstat);
      HIP_CHECK(hipGetLastError());
    }
    if (h_irr > irr_cap) {
      // irr overflowed: project the earlier batches' walks (intact below
      // irr_pending), make room for this batch's and take them again
      if (irr_pending) {
        k_sens<false, false><<<grid_for(irr_pending), kBlock, 0, s>>>(irr, irr_pending, 0, 0.0f, 0.0f, f,
                                                                       mu_zero ? nullptr : dmu, f, acc, cnt);
        HIP_CHECK(hipGetLastError());
      }
      const uint32_t mine = h_irr - irr_pending;
      HIP_CHECK(hipStreamSynchronize(s));  // before irr is reallocated
      irr = ws.irr.ensure(mine);
      irr_cap = ws.irr.b.n;
      n_irr = ws.n_irr.ensure(1);
      HIP_CHECK(hipMemsetAsync(n_irr, 0, 4, s));
      k_retake_irregular<<<grid_for(b), kBlock, 0, s>>>(src_lors, lor_order, b, f, irr, (uint32_t)irr_cap, n_irr);
      HIP_CHECK(hipGetLastError());
      uint32_t again = 0;
      HIP_CHECK(hipMemcpyAsync(&again, n_irr, 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      if (again != mine) throw std::runtime_error("sensitivity: unbounded walks counted differently when taken again");
      h_irr = mine;
    }
    irr_pending = h_irr;  // the running count
  }
  if (irr_pending) {
    k_sens<false, false><<<grid_for(irr_pending), kBlock, 0, s>>>(irr, irr_pending, 0, 0.0f, 0.0f, f,
                                                                   mu_zero ? nullptr : dmu, f, acc, cnt);
    HIP_CHECK(hipGetLastError());
  }
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

__global__ __launch_bounds__(256) void k_pair_lors(uint64_t first, uint32_t n, DevElements e, uint64_t seed,
                                                   Lor* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Lor l = pair_lor(e, seed, first + i);
  float4* p = reinterpret_cast<float4*>(out + i);
  p[0] = make_float4(l.x1, l.y1, l.z1, l.x2);
  p[1] = make_float4(l.y2, l.z2, l.dt, l.ac);
}

}  // namespace

extern "C" {

int petalo_make_sensitivity_image_discrete(const petalo_fov_t* att_fov, const float* mu,
                                           const petalo_fov_t* out_fov, float detector_length_mm, float r_min_mm,
                                           float dr_mm, float dz_mm, float da_mm, int32_t adjust, uint64_t seed,
                                           float* out, petalo_stats_t* stats) {
  return guarded([&] {
    std::string why;
    if (!fov_ok(att_fov, why)) return fail(PETALO_ERR_ARG, why);
    if (out_fov && !fov_ok(out_fov, why)) return fail(PETALO_ERR_ARG, why);
    if (!mu || !out) return fail(PETALO_ERR_ARG, "NULL argument");
    if (adjust < PETALO_ADJUST_NONE || adjust > PETALO_ADJUST_RANDOM_Z_PHI)
      return fail(PETALO_ERR_ARG, "unknown element adjustment");
    if (!(dr_mm > 0.0f && dz_mm > 0.0f && da_mm > 0.0f && r_min_mm >= 0.0f && detector_length_mm >= 0.0f))
      return fail(PETALO_ERR_ARG, "element discretization needs positive dr, dz, da");
    const DevElements el = make_elements(r_min_mm, dr_mm, dz_mm, da_mm, detector_length_mm, adjust);
    const uint64_t npts = el.n_points;
    if (npts < 2 || npts > (1ull << 31)) return fail(PETALO_ERR_ARG, "element count out of range");
    const uint64_t n_pairs = npts * (npts + 1) / 2;  // i <= k, self pairs included (discrete.rs:55)
    require_device();
    const DevFov fa = to_dev(*att_fov);
    const DevFov fo = to_dev(out_fov ? *out_fov : *att_fov);
    const size_t na = nvox(*att_fov), no = nvox(out_fov ? *out_fov : *att_fov);
    SensLease lease;  // the device's kept buffers and stream
    SensWork& ws = *lease.w;
    hipStream_t s = ws.stream;
    float* dmu = ws.dmu.ensure(na);
    float* dout = ws.dout.ensure(no);
    double* acc = ws.acc.ensure(no);
    unsigned long long* cnt = ws.cnt.ensure(CNT_N);
    // an image of zeros is cleared on the device rather than copied
    const bool mu_zero = all_zero(mu, na);
    if (mu_zero)
      HIP_CHECK(hipMemsetAsync(dmu, 0, na * sizeof(float), s));
    else
      HIP_CHECK(hipMemcpyAsync(dmu, mu, na * sizeof(float), hipMemcpyHostToDevice, s));
    fill_f64(acc, 0.0, no, s);
    zero_counters(cnt, s);
    LorSource src = [&](Lor* dst, uint64_t first, uint32_t n, hipStream_t st) {
      k_pair_lors<<<grid_for(n), kBlock, 0, st>>>(first, n, el, seed, dst);
    };
    if (!out_fov) {
      sens_tiled(ws, *att_fov, fa, mu_zero, dmu, n_pairs, src, acc, cnt, false, stats != nullptr);
    } else {
      // two grids: one lane per LOR, binary64 global atomics
      DevBuf<Lor> lb((size_t)std::min<uint64_t>(n_pairs, kSensBatch));
      for (uint64_t first = 0; first < n_pairs; first += kSensBatch) {
        const uint32_t b = (uint32_t)std::min<uint64_t>(kSensBatch, n_pairs - first);
        src(lb.p, first, b, s);
        k_sens<false, true><<<grid_for(b), kBlock, 0, s>>>(lb.p, b, 0, 0.0f, 0.0f, fa, dmu, fo, acc, cnt);
        HIP_CHECK(hipGetLastError());
      }
    }
    // normalised by n (n - 1) / 2 over the element count (discrete.rs:16-20)
    const double norm = (double)npts * (double)(npts - 1) / 2.0;
    k_sens_finalize<<<grid_for(no), kBlock, 0, s>>>(acc, dout, (float)norm, no);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, dout, no * sizeof(float), hipMemcpyDeviceToHost, s));
    unsigned long long h[CNT_N];
    HIP_CHECK(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (h[CNT_RUN_MISMATCH])
      throw std::runtime_error("tile binning: fill and count passes disagree on " + std::to_string(h[CNT_RUN_MISMATCH]) +
                               " LORs");
    if (stats)
      *stats = petalo_stats_t{n_pairs, h[CNT_STEPS], h[CNT_NZ], h[CNT_SKIPPED], h[CNT_MISSED], h[CNT_RUNAWAY],
                              h[CNT_SHIFTED]};
    return PETALO_OK;
  });
}

int petalo_make_sensitivity_image(const petalo_fov_t* att_fov, const float* mu, const petalo_fov_t* out_fov,
                                  const petalo_lor_t* lors, size_t n_lors, uint64_t seed, float L, float D,
                                  float* out, petalo_stats_t* stats) {
  return guarded([&] {
    std::string why;
    if (!fov_ok(att_fov, why)) return fail(PETALO_ERR_ARG, why);
    if (out_fov && !fov_ok(out_fov, why)) return fail(PETALO_ERR_ARG, why);
    if (!mu || !out) return fail(PETALO_ERR_ARG, "NULL argument");
    if (n_lors >= (1ull << 32)) return fail(PETALO_ERR_ARG, "more than 2^32-1 LORs");
    if (!lors && (!(L > 0.0f) || !(D > 0.0f)))
      return fail(PETALO_ERR_ARG, "detector length and diameter must be positive");
    require_device();
    const DevFov fa = to_dev(*att_fov);
    const DevFov fo = to_dev(out_fov ? *out_fov : *att_fov);
    const size_t na = nvox(*att_fov), no = nvox(out_fov ? *out_fov : *att_fov);
    SensLease lease;  // the device's kept buffers and stream
    SensWork& ws = *lease.w;
    hipStream_t s = ws.stream;
    float* dmu = ws.dmu.ensure(na);
    float* dout = ws.dout.ensure(no);
    double* acc = ws.acc.ensure(no);
    unsigned long long* cnt = ws.cnt.ensure(CNT_N);
    DevBuf<Lor> dl;
    // an image of zeros is cleared on the device rather than copied
    const bool mu_zero = all_zero(mu, na);
    if (mu_zero)
      HIP_CHECK(hipMemsetAsync(dmu, 0, na * sizeof(float), s));
    else
      HIP_CHECK(hipMemcpyAsync(dmu, mu, na * sizeof(float), hipMemcpyHostToDevice, s));
    if (lors && n_lors) {
      dl.alloc(n_lors);
      upload_h2d(dl.p, lors, n_lors * sizeof(Lor), s);
    }
    fill_f64(acc, 0.0, no, s);
    zero_counters(cnt, s);
    const bool two = out_fov != nullptr;
    if (n_lors) {
      const unsigned g = grid_for(n_lors);
      const uint32_t n = (uint32_t)n_lors;
      const float R = D / 2.0f;
      if (!two) {
        const float Rr = R;
        LorSource src = lors ? LorSource([&](Lor* dst, uint64_t first, uint32_t cnt_, hipStream_t st) {
          HIP_CHECK(hipMemcpyAsync(dst, dl.p + first, (size_t)cnt_ * sizeof(Lor), hipMemcpyDeviceToDevice, st));
        })
                             : LorSource([&](Lor* dst, uint64_t first, uint32_t cnt_, hipStream_t st) {
          k_sens_lors<<<sens_lors_grid(cnt_), kBlock, 0, st>>>(cnt_, (uint32_t)first, seed, fa, L, Rr, dst, cnt);
        });
        sens_tiled(ws, *att_fov, fa, mu_zero, dmu, n_lors, src, acc, cnt, true, stats != nullptr);
      } else {
        // two FOVs: one lane per LOR, binary64 global atomics
        if (lors) k_sens<false, true><<<g, kBlock, 0, s>>>(dl.p, n, seed, L, R, fa, dmu, fo, acc, cnt);
        else      k_sens<true, true><<<g, kBlock, 0, s>>>(nullptr, n, seed, L, R, fa, dmu, fo, acc, cnt);
      }
      HIP_CHECK(hipGetLastError());
    }
    k_sens_finalize<<<grid_for(no), kBlock, 0, s>>>(acc, dout, (float)n_lors, no);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, dout, no * sizeof(float), hipMemcpyDeviceToHost, s));
    unsigned long long h[CNT_N];
    HIP_CHECK(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (h[CNT_RUN_MISMATCH])
      throw std::runtime_error("tile binning: fill and count passes disagree on " + std::to_string(h[CNT_RUN_MISMATCH]) +
                               " LORs");
    if (h[CNT_UNFOUND])
      return fail(PETALO_ERR_ARG, std::to_string(h[CNT_UNFOUND]) + " sensitivity LORs found no line through the FOV in " +
                                      std::to_string(kSensMaxAttempts) + " draws (FOV outside the detector's reach)");
    if (stats)
      *stats = petalo_stats_t{(uint64_t)n_lors, h[CNT_STEPS], h[CNT_NZ], h[CNT_SKIPPED], h[CNT_MISSED],
                              h[CNT_RUNAWAY], h[CNT_SHIFTED]};
    return PETALO_OK;
  });
}

int petalo_release_workspaces(void) {
  return guarded([&] {
    std::unique_ptr<SensWork> kept[kSensDevices];
    {
      std::lock_guard<std::mutex> lk(g_sens_mu);
      for (int d = 0; d < kSensDevices; ++d) kept[d] = std::move(g_sens_work[d]);
    }
    int cur = 0;
    HIP_CHECK(hipGetDevice(&cur));
    for (int d = 0; d < kSensDevices; ++d) {
      if (!kept[d]) continue;
      HIP_CHECK(hipSetDevice(d));  // each set is freed on its own device
      kept[d].reset();
    }
    HIP_CHECK(hipSetDevice(cur));
    return PETALO_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The sensitivity image without any allocation, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops):
// petalo_sensitivity_image_noalloc.
//
// It is petalo_make_sensitivity_image for caller-given LORs on one FOV, with
// device pointers in and out and the caller's stream: the one-lane kernel
// k_sens (forward walk over mu, p -> exp(-p), backprojection of w / exp(-p)
// with binary64 atomics) and k_sens_finalize (division by n).  Its scratch --
// the binary64 accumulator and the work counters -- is the file-scope array
// `scratch` below, at most kNoAllocMaxVoxels voxels; every device that runs
// it holds its own copy until the process ends.  Host code never uses the
// array's name as an address: on a device's first call the address is looked
// up with hipGetSymbolAddress and kept per device.  Calls on one device share
// the array and clear it themselves (fill kernels, no memset nodes), so calls
// that overlap on one device race.
namespace {

// 2^21 voxels (a 128^3 image); the counters follow the accumulator.
constexpr size_t kNoAllocMaxVoxels = size_t(1) << 21;
constexpr size_t kCounterOffset = kNoAllocMaxVoxels * sizeof(double);
constexpr size_t kScratchBytes = kCounterOffset + 256;
static_assert(CNT_N * sizeof(unsigned long long) <= 256, "counters fit the scratch tail");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device, looked up once

// The scratch array's device address on the current device.
void* scratch_address(int dev) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" int petalo_sensitivity_image_noalloc(const petalo_fov_t* fov, const float* mu, const petalo_lor_t* lors,
                                                size_t n_lors, float* out, void* stream) {
  return guarded([&] {
    std::string why;
    if (!fov_ok(fov, why)) return fail(PETALO_ERR_ARG, why);
    if (!mu || !out || (n_lors && !lors)) return fail(PETALO_ERR_ARG, "NULL argument");
    const size_t nv = (size_t)fov->n[0] * fov->n[1] * fov->n[2];
    if (nv > kNoAllocMaxVoxels)
      return fail(PETALO_ERR_SIZE, "image of " + std::to_string(nv) + " voxels exceeds the " +
                                       std::to_string(kNoAllocMaxVoxels) +
                                       " voxels of petalo_sensitivity_image_noalloc's scratch");
    if (n_lors >= (1ull << 32)) return fail(PETALO_ERR_ARG, "more than 2^32-1 LORs");
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PETALO_ERR_ARG, "device ordinal beyond the supported 64");
    unsigned char* base = static_cast<unsigned char*>(scratch_address(dev));
    double* acc = reinterpret_cast<double*>(base);
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(base + kCounterOffset);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const DevFov f = make_fov(fov->full_size_mm, fov->n);
    k_fill_f64<<<grid_for(nv), kBlock, 0, s>>>(acc, 0.0, nv);
    k_zero_u64<<<1, 64, 0, s>>>(cnt, CNT_N);
    if (n_lors)
      k_sens<false, false><<<grid_for(n_lors), kBlock, 0, s>>>(reinterpret_cast<const Lor*>(lors), (uint32_t)n_lors, 0,
                                                               0.0f, 0.0f, f, mu, f, acc, cnt);
    k_sens_finalize<<<grid_for(nv), kBlock, 0, s>>>(acc, out, (float)n_lors, nv);
    HIP_CHECK(hipGetLastError());
    return PETALO_OK;
  });
}
