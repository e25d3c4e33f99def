// tiles.hpp -- tiled backprojection for gfx950.
//
// Global float atomics cost one L2 request per distinct 128-byte line per
// wavefront instruction (about 2e10 requests/s on MI355X), so scattering one
// lane per LOR into the image is bound by them.  binary64 adds into LDS run
// at about 1.1e12/s.  The backprojection is therefore split:
//
//   once per context   k_bin_runs walks every LOR exactly (fast-forwarded to
//                      the TOF window) and cuts its pushed steps into runs,
//                      one per image tile the walk crosses (a 6-connected walk
//                      is monotone in every flipped axis, so it visits each
//                      box in one contiguous run).  A run records the walk
//                      state at its first pushed step and its length.  Runs
//                      are sorted by tile.
//   every update       k_fwd_tile replays the runs against an LDS copy of
//                      each tile of the image and writes per-run partial
//                      projections; k_lor_reduce sums them per LOR into
//                      r = 1/(p*ac).  k_scatter gives each workgroup a chunk of
//                      one tile's runs: it replays every run from its recorded
//                      state, computing each weight anew, adds w*r into a
//                      binary64 LDS copy of the tile, and flushes the tile with
//                      lane-coalesced binary64 global atomics.
//
// LORs whose walk is not provably bounded (an axis with no voxels left at
// the entry) take the exact one-lane path with global atomics
// (k_fwd_irregular).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "siddon_walk.hpp"

namespace petalo {

// Tile shape: nominal dimensions in voxels (edge tiles are cut by the FOV).
struct TileGeom {
  int32_t tx, ty, tz;     // nominal tile dimensions
  int32_t ntx, nty, ntz;  // tiles per axis
  int32_t ntiles;
  int32_t px;             // LDS row pitch in entries (>= tx; an odd pitch spreads a
                          // column of voxels over the LDS banks)
  int32_t tvox;           // px*ty*tz: LDS entries per tile
};

// One run: the walk state at the first pushed step of a LOR inside a tile
// (40 bytes).
struct __align__(8) Run {
  float nb_x, nb_y, nb_z, here;
  float vl_x, vl_y, vl_z, tof_peak;
  uint32_t code;  // bits 0-11: steps in the run; bits 12-14: axis flips (x, y, z);
                  // bits 15-31: tile-local LDS index of the first voxel (x fastest)
  uint32_t lor;   // index of the LOR within its subset
};

constexpr uint32_t kRunMaxSteps = 0xfffu;
constexpr uint32_t kRunMaxLocal = 0x1ffffu;
__host__ __device__ __forceinline__ uint32_t run_len(uint32_t code) { return code & 0xfffu; }
__host__ __device__ __forceinline__ uint32_t run_flips(uint32_t code) { return (code >> 12) & 7u; }
__host__ __device__ __forceinline__ uint32_t run_local(uint32_t code) { return code >> 15; }

// Per-step voxel coordinates (original axes) tracked by the binning walk.
struct Coord {
  int32_t x, y, z;
  int32_t sx, sy, sz;  // +1 or -1 per axis
};

__host__ __device__ __forceinline__ int32_t tile_of(const TileGeom& g, const Coord& q) {
  return q.x / g.tx + g.ntx * (q.y / g.ty + g.nty * (q.z / g.tz));
}

__host__ __device__ __forceinline__ uint32_t local_of(const TileGeom& g, const Coord& q) {
  return (uint32_t)((q.x % g.tx) + g.px * ((q.y % g.ty) + g.ty * (q.z % g.tz)));
}

// Walk #1 with TOF window fast-forward, as in k_mlem_project.  Calls
// on_step(w, c, s, weight, coord) for every walked step; bounded walks only.
template <bool TOF, class F>
__host__ __device__ __forceinline__ void walk_pushes(const Lor& l, const DevFov& f, const DevTof& t, F&& on_step) {
  WalkC c;
  WalkS w;
  if (!walk_setup(l, f, c, w)) return;
  if (!bounded(w)) return;
  if (TOF && !fast_forward(w, c, tof_window_start(t, c.tof_peak))) return;
  Coord q;
  {
    const int32_t nxy = f.nx * f.ny;
    q.z = w.index / nxy;
    const int32_t rem = w.index - q.z * nxy;
    q.y = rem / f.nx;
    q.x = rem - q.y * f.nx;
    q.sx = c.d_x < 0 ? -1 : 1;
    q.sy = c.d_y < 0 ? -1 : 1;
    q.sz = c.d_z < 0 ? -1 : 1;
  }
  for (;;) {
    if (TOF && (w.here - c.tof_peak) >= t.cutoff) break;
    const Step s = next_boundary(w);
    const float wt = step_weight<TOF>(w, c, s, t);
    on_step(w, c, wt, q);
    q.x += s.ax ? q.sx : 0;
    q.y += s.ay ? q.sy : 0;
    q.z += s.az ? q.sz : 0;
    if (advance(w, c, s)) break;
  }
}

// Cuts the pushed steps of LOR i into runs, one per tile crossed; calls
// emit(run, tile) for each and returns their number.
template <bool TOF, class Emit>
__host__ __device__ __forceinline__ uint32_t bin_lor(const Lor& l, uint32_t i, const DevFov& f, const DevTof& t,
                                                     const TileGeom& g, Emit&& emit) {
  uint32_t count = 0;
  int32_t cur = -1;        // tile of the open run
  uint32_t steps_in = 0;   // steps since the open run's first push
  uint32_t last_push = 0;  // steps_in at the last push of the open run
  Run open{};
  auto close = [&]() {
    if (cur < 0) return;
    open.code = (open.code & ~kRunMaxSteps) | (last_push + 1);
    emit(open, (uint32_t)cur);
    ++count;
    cur = -1;
  };
  walk_pushes<TOF>(l, f, t, [&](const WalkS& w, const WalkC& c, float wt, const Coord& q) {
    const int32_t tl = tile_of(g, q);
    if (cur >= 0) {
      ++steps_in;
      // a run ends where the walk leaves its tile, or when its length field is full
      if (tl != cur || steps_in >= kRunMaxSteps) close();
    }
    if (wt > 0.0f) {
      if (cur < 0) {
        cur = tl;
        steps_in = 0;
        open.nb_x = w.nb_x;
        open.nb_y = w.nb_y;
        open.nb_z = w.nb_z;
        open.here = w.here;
        open.vl_x = c.vl_x;
        open.vl_y = c.vl_y;
        open.vl_z = c.vl_z;
        open.tof_peak = c.tof_peak;
        const uint32_t flips = (q.sx < 0 ? 1u : 0u) | (q.sy < 0 ? 2u : 0u) | (q.sz < 0 ? 4u : 0u);
        open.code = (local_of(g, q) << 15) | (flips << 12);
        open.lor = i;
      }
      last_push = steps_in;
    }
  });
  close();
  return count;
}

// Binning walk: FILL = false counts runs per LOR into n_runs[i]; FILL = true
// writes them at runs[offset[i] ...] (LOR-major) with their sort keys:
// tile, then block of 2^lor_shift consecutive LORs (LORs are ordered
// spatially, so a block's 1/(p*ac) values stay cache-resident while its runs
// in a tile are replayed), then longest run first (lanes of a wavefront get
// runs of similar length).
template <bool TOF, bool FILL>
__global__ __launch_bounds__(256) void k_bin_runs(const Lor* __restrict__ lors, uint32_t n, DevFov f, DevTof t,
                                                  TileGeom g, uint32_t* __restrict__ n_runs,
                                                  const uint32_t* __restrict__ offset, Run* __restrict__ runs,
                                                  uint64_t* __restrict__ keys, uint32_t lor_shift) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Lor l = load_lor(lors, i);
  const uint32_t base = FILL ? offset[i] : 0;
  uint32_t k = 0;
  const uint32_t count = bin_lor<TOF>(l, i, f, t, g, [&](const Run& ru, uint32_t tile) {
    if (FILL) {
      runs[base + k] = ru;
      keys[base + k] = ((uint64_t)tile << 32) | ((uint64_t)(i >> lor_shift) << 12) | (kRunMaxSteps - run_len(ru.code));
    }
    ++k;
  });
  if (!FILL) {
    // unbounded walks are projected by the one-lane kernel instead
    WalkC c;
    WalkS w;
    n_runs[i] = (walk_setup(l, f, c, w) && !bounded(w)) ? 0xffffffffu : count;
  }
}

// Replays a run from its recorded state, calling add(local_index, w * r)
// for every pushed step.  The weights are recomputed with the same binary32
// operations as walk #1.
template <bool TOF, class Add>
__host__ __device__ __forceinline__ void replay_run(const Run& ru, float r, const DevTof& t, const TileGeom& g,
                                                    Add&& add) {
  const int32_t dlx = 1, dly = g.px, dlz = g.px * g.ty;
  WalkS w;
  w.nb_x = ru.nb_x;
  w.nb_y = ru.nb_y;
  w.nb_z = ru.nb_z;
  w.here = ru.here;
  WalkC c;
  c.vl_x = ru.vl_x;
  c.vl_y = ru.vl_y;
  c.vl_z = ru.vl_z;
  c.tof_peak = ru.tof_peak;
  const uint32_t flips = run_flips(ru.code);
  c.d_x = (flips & 1u) ? -dlx : dlx;
  c.d_y = (flips & 2u) ? -dly : dly;
  c.d_z = (flips & 4u) ? -dlz : dlz;
  w.index = (int32_t)run_local(ru.code);
  const uint32_t m = run_len(ru.code);  // the run length bounds the replay
  for (uint32_t k = 0; k < m; ++k) {
    const Step s = next_boundary(w);
    // Every step of a run lies inside the TOF window, so the window test of
    // step_weight is skipped: the run spans the walk's first to last pushed
    // step in the tile, pushed steps lie inside the window, `here` does not
    // decrease along the walk and the window is an interval of here - peak.
    float wt = s.bnd - w.here;
    if (TOF) wt *= tof_factor_in_window(t, w.here, c.tof_peak);
    if (wt > 0.0f) add((uint32_t)w.index, wt * r);
    move(w, c, s);
  }
}

// Forward, tiled: one workgroup per work item copies its tile of the image
// into LDS and replays its runs, writing each run's partial projection
// (sum of w*x over the run's pushed steps, in walk order) at the run's
// sorted position (coalesced; writing them LOR-major instead scatters 4-byte
// stores over the whole array and costs more than the gather in
// k_lor_reduce).  Each lane loads its next run record before replaying the
// current one, so the record's latency hides behind the replay (the replay
// itself touches only LDS).
template <bool TOF>
__global__ __launch_bounds__(1024) void k_fwd_tile(const Run* __restrict__ runs, const int4* __restrict__ work,
                                                   const float* __restrict__ img, DevFov f, DevTof t, TileGeom g,
                                                   float* __restrict__ partial) {
  extern __shared__ float timg[];
  const int4 wi = work[blockIdx.x];
  const int32_t tl = wi.x;
  const int32_t tzi = tl / (g.ntx * g.nty);
  const int32_t tyi = (tl / g.ntx) % g.nty;
  const int32_t txi = tl % g.ntx;
  const int32_t x0 = txi * g.tx, y0 = tyi * g.ty, z0 = tzi * g.tz;
  for (int32_t k = threadIdx.x; k < g.tvox; k += blockDim.x) {
    const int32_t lx = k % g.px, ly = (k / g.px) % g.ty, lz = k / (g.px * g.ty);
    const int32_t x = x0 + lx, y = y0 + ly, z = z0 + lz;
    timg[k] = (lx < g.tx && x < f.nx && y < f.ny && z < f.nz) ? img[x + f.nx * (y + f.ny * z)] : 0.0f;
  }
  __syncthreads();
  int32_t j = wi.y + (int32_t)threadIdx.x;
  if (j >= wi.z) return;
  const int32_t last = wi.z - 1;
  const int32_t stride = (int32_t)blockDim.x;
  Run cur = runs[j];
  for (;;) {
    const int32_t jn = j + stride;
    const Run nxt = runs[jn < last ? jn : last];  // in bounds; unused past the end
    float p = 0.0f;
    replay_run<TOF>(cur, 1.0f, t, g, [&](uint32_t li, float wt) { p += wt * timg[li]; });
    partial[j] = p;
    if (jn > last) break;
    cur = nxt;
    j = jn;
  }
}

// r = 1/(p*ac) per LOR with p the sum of its runs' partials in walk order;
// run_pos[run_off[i] ...] are the sorted positions of LOR i's runs (LORs
// without runs get 1/(0*ac), as a row without elements does).
//
// Each lane takes kReduceLors LORs, 64 apart (loads stay coalesced across
// the wavefront), and issues the loads of all of them together: the first
// four runs of every LOR in one batch, any further runs in a loop.  Each sum
// keeps its order.  Workgroups land on the eight XCDs round-robin by index;
// XCD x takes the x-th eighth of the LORs, so each L2 caches the partials of
// one eighth.  Launch with reduce_grid(n) workgroups of 256.
constexpr uint32_t kReduceLors = 4;
inline unsigned reduce_grid(size_t n) {
  const size_t share = (n + 7) / 8;
  const size_t per_group = 256 * kReduceLors;
  return (unsigned)(8 * ((share + per_group - 1) / per_group));
}
__global__ __launch_bounds__(256) void k_lor_reduce(const float* __restrict__ ac, uint32_t n,
                                                    const uint32_t* __restrict__ run_off,
                                                    const uint32_t* __restrict__ run_pos,
                                                    const float* __restrict__ partial, float* __restrict__ r_out) {
  const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3;
  const uint32_t share = (n + 7u) / 8u;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t first = slot * blockDim.x * kReduceLors + wave * 64u * kReduceLors + lane;
  uint32_t idx[kReduceLors], b[kReduceLors], e[kReduceLors];
  bool on[kReduceLors];
#pragma unroll
  for (uint32_t u = 0; u < kReduceLors; ++u) {
    const uint32_t local = first + 64u * u;
    idx[u] = xcd * share + local;
    on[u] = local < share && idx[u] < n;
    b[u] = on[u] ? run_off[idx[u]] : 0u;
    e[u] = on[u] ? run_off[idx[u] + 1] : 0u;
  }
  uint32_t q[kReduceLors][4];
#pragma unroll
  for (uint32_t u = 0; u < kReduceLors; ++u)
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) q[u][k] = b[u] + k < e[u] ? run_pos[b[u] + k] : 0u;
  float v[kReduceLors][4];
#pragma unroll
  for (uint32_t u = 0; u < kReduceLors; ++u)
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) v[u][k] = b[u] + k < e[u] ? partial[q[u][k]] : 0.0f;
#pragma unroll
  for (uint32_t u = 0; u < kReduceLors; ++u) {
    if (!on[u]) continue;
    float p = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (b[u] + k < e[u]) p += v[u][k];
    for (uint32_t k = b[u] + 4; k < e[u]; ++k) p += partial[run_pos[k]];
    r_out[idx[u]] = 1.0f / (p * ac[idx[u]]);
  }
}

// LORs whose walk is not provably bounded: exact one-lane projection and
// backprojection with global binary64 atomics; dropped whole on a pushed
// index outside the image (projector.rs:103-118).
template <bool TOF>
__global__ __launch_bounds__(256) void k_fwd_irregular(const Lor* __restrict__ lors, const uint32_t* __restrict__ list,
                                                       uint32_t n, DevFov f, DevTof t, const float* __restrict__ img,
                                                       double* __restrict__ bp, unsigned long long* __restrict__ counters) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t skipped = 0;
  if (j < n) {
    const Lor l = load_lor(lors, list[j]);
    WalkC c;
    WalkS w;
    if (walk_setup(l, f, c, w)) {
      const uint32_t cap = step_cap(f);
      WalkS start = w;
      uint32_t steps = 0, start_step = 0;
      bool started = false, bad = false;
      float p = 0.0f;
      for (;;) {
        if (steps >= cap) { bad = true; break; }
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
          if (wt > 0.0f) unsafeAtomicAdd(bp + w.index, (double)(wt * r));
          advance(w, c, s);
        }
      }
    }
  }
  wave_add_u64(counters + CNT_SKIPPED, skipped);
}

// Backprojection, tiled: one workgroup per work item (tile, first run, end
// run).  Lanes keep two runs in flight: the record of the run after next and
// the 1/(p*ac) of the next run load while the current run is replayed.
template <bool TOF>
__global__ __launch_bounds__(1024) void k_scatter(const Run* __restrict__ runs, const int4* __restrict__ work,
                                                  const float* __restrict__ r_in, DevFov f, DevTof t, TileGeom g,
                                                  double* __restrict__ bp) {
  extern __shared__ double tile[];
  const int4 wi = work[blockIdx.x];
  const int32_t tl = wi.x;
  for (int32_t k = threadIdx.x; k < g.tvox; k += blockDim.x) tile[k] = 0.0;
  __syncthreads();
  int32_t j = wi.y + (int32_t)threadIdx.x;
  if (j < wi.z) {
    const int32_t last = wi.z - 1;
    const int32_t stride = (int32_t)blockDim.x;
    const int32_t j1 = j + stride;
    Run cur = runs[j];
    Run nxt = runs[j1 < last ? j1 : last];
    float r = r_in[cur.lor];
    for (;;) {
      const int32_t jn = j + stride;
      const int32_t jnn = jn + stride;
      const float rn = r_in[nxt.lor];
      const Run nn = runs[jnn < last ? jnn : last];
      replay_run<TOF>(cur, r, t, g, [&](uint32_t li, float v) { atomicAdd(tile + li, (double)v); });
      if (jn > last) break;
      cur = nxt;
      nxt = nn;
      r = rn;
      j = jn;
    }
  }
  __syncthreads();
  // flush the tile: consecutive lanes take consecutive x, so each wavefront
  // atomic touches few cache lines
  const int32_t tzi = tl / (g.ntx * g.nty);
  const int32_t tyi = (tl / g.ntx) % g.nty;
  const int32_t txi = tl % g.ntx;
  const int32_t x0 = txi * g.tx, y0 = tyi * g.ty, z0 = tzi * g.tz;
  for (int32_t k = threadIdx.x; k < g.tvox; k += blockDim.x) {
    const double v = tile[k];
    if (v == 0.0) continue;
    const int32_t lx = k % g.px, ly = (k / g.px) % g.ty, lz = k / (g.px * g.ty);
    const int32_t x = x0 + lx, y = y0 + ly, z = z0 + lz;
    if (lx < g.tx && x < f.nx && y < f.ny && z < f.nz) unsafeAtomicAdd(bp + x + f.nx * (y + f.ny * z), v);
  }
}

// MLEM update with the binary64 backprojection, which is reset:
// x = s > 0 ? x * (b * s) : 0 with b rounded to binary32 (mlem.rs:45-52).
__global__ __launch_bounds__(256) void k_update_f64(float* __restrict__ x, double* __restrict__ bp,
                                                    const float* __restrict__ s, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float b = (float)bp[i];
  bp[i] = 0.0;
  const float sj = s ? s[i] : 1.0f;
  x[i] = sj > 0.0f ? x[i] * (b * sj) : 0.0f;
}

}  // namespace petalo
