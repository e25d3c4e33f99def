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

// A lane's walk over the LDS entries k = first, first + stride, ... of a
// tile with their local coordinates (k = lx + px * (ly + ty * lz)), advanced
// by carries instead of a division per entry.
struct TileCursor {
  int32_t k, lx, ly, lz;
  int32_t sx, sy, sz;  // stride = sx + px * (sy + ty * sz), sx < px, sy < ty
  __device__ __forceinline__ TileCursor(const TileGeom& g, int32_t first, int32_t stride) {
    k = first;
    lx = first % g.px;
    ly = (first / g.px) % g.ty;
    lz = first / (g.px * g.ty);
    sx = stride % g.px;
    sy = (stride / g.px) % g.ty;
    sz = stride / (g.px * g.ty);
  }
  __device__ __forceinline__ void next(const TileGeom& g, int32_t stride) {
    k += stride;
    lx += sx;
    const int32_t cx = lx >= g.px ? 1 : 0;
    lx -= cx ? g.px : 0;
    ly += sy + cx;
    const int32_t cy = ly >= g.ty ? 1 : 0;
    ly -= cy ? g.ty : 0;
    lz += sz + cy;
  }
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

// Position of the binning walk relative to the tiles, tracked per step
// without division: per axis the voxel steps left in the current tile in the
// walk's direction (cx, cy, cz), and the tile id.
struct Coord {
  int32_t cx, cy, cz;     // steps along the axis until the walk leaves the tile
  int32_t tile;
  int32_t sx, sy, sz;     // +1 or -1 per axis (walk direction, original axes)
};

// Tile-local LDS index of the current voxel.
__host__ __device__ __forceinline__ uint32_t local_of(const TileGeom& g, const Coord& q) {
  const int32_t lx = q.sx > 0 ? g.tx - q.cx : q.cx - 1;
  const int32_t ly = q.sy > 0 ? g.ty - q.cy : q.cy - 1;
  const int32_t lz = q.sz > 0 ? g.tz - q.cz : q.cz - 1;
  return (uint32_t)(lx + g.px * (ly + g.ty * lz));
}

// Moves q one voxel along the crossed axis of s (exactly one of ax, ay, az),
// with selects only: the walk's lanes take different axes.
__host__ __device__ __forceinline__ void coord_step(Coord& q, const TileGeom& g, const Step& s) {
  const int32_t left = (s.ax ? q.cx : (s.ay ? q.cy : q.cz)) - 1;
  const bool cross = left == 0;
  const int32_t reload = s.ax ? g.tx : (s.ay ? g.ty : g.tz);
  const int32_t next = cross ? reload : left;
  q.cx = s.ax ? next : q.cx;
  q.cy = s.ay ? next : q.cy;
  q.cz = s.az ? next : q.cz;
  const int32_t dt = s.ax ? q.sx : (s.ay ? q.sy * g.ntx : q.sz * g.ntx * g.nty);
  q.tile += cross ? dt : 0;
}

// Walk #1 with TOF window fast-forward, as in k_mlem_project.  Calls
// on_step(w, c, weight, coord) for every walked step; bounded walks only.
template <bool TOF, class F>
__host__ __device__ __forceinline__ void walk_pushes(const Lor& l, const DevFov& f, const DevTof& t,
                                                     const TileGeom& g, F&& on_step) {
  WalkC c;
  WalkS w;
  if (!walk_setup(l, f, c, w)) return;
  if (!bounded(w)) return;
  if (TOF && !fast_forward(w, c, tof_window_start(t, c.tof_peak))) return;
  Coord q;
  {
    const int32_t nxy = f.nx * f.ny;
    const int32_t z = w.index / nxy;
    const int32_t rem = w.index - z * nxy;
    const int32_t y = rem / f.nx;
    const int32_t x = rem - y * f.nx;
    q.sx = c.d_x < 0 ? -1 : 1;
    q.sy = c.d_y < 0 ? -1 : 1;
    q.sz = c.d_z < 0 ? -1 : 1;
    q.cx = q.sx > 0 ? g.tx - x % g.tx : x % g.tx + 1;
    q.cy = q.sy > 0 ? g.ty - y % g.ty : y % g.ty + 1;
    q.cz = q.sz > 0 ? g.tz - z % g.tz : z % g.tz + 1;
    q.tile = x / g.tx + g.ntx * (y / g.ty + g.nty * (z / g.tz));
  }
  for (;;) {
    if (TOF && (w.here - c.tof_peak) >= t.cutoff) break;
    const Step s = next_boundary(w);
    const float wt = step_weight<TOF>(w, c, s, t);
    on_step(w, c, wt, q);
    coord_step(q, g, s);
    if (advance(w, c, s)) break;
  }
}

// Cuts the pushed steps of LOR i into runs, one per tile crossed; calls
// emit(run, tile) for each and returns their number.
template <bool TOF, class Emit>
__host__ __device__ __forceinline__ uint32_t bin_lor(const Lor& l, uint32_t i, const DevFov& f, const DevTof& t,
                                                     const TileGeom& g, Emit&& emit, uint32_t* n_steps = nullptr,
                                                     uint32_t* n_nz = nullptr) {
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
  walk_pushes<TOF>(l, f, t, g, [&](const WalkS& w, const WalkC& c, float wt, const Coord& q) {
    if (n_steps) ++*n_steps;
    if (n_nz && wt > 0.0f) ++*n_nz;
    const int32_t tl = q.tile;
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

// Length classes of runs, by which the sensitivity batches order each
// tile's runs (so that a replaying wavefront's lanes get runs of about one
// length): n classes of an estimated step count s, class min(n - 1, s *
// scale) with scale = n / (tx + ty + tz).  The count pass (tile_runs) and the
// fill pass (bin_lor_jump) must put every run in the same class, so s comes
// from values both compute alike: the binary64 estimates of the crossings
// that enter (h) and leave (b) the tile, s = (b - h) * (1/vl_x + 1/vl_y +
// 1/vl_z); a walk both step through (degenerate LORs) uses its run's exact
// length.
struct RunClasses {
  uint32_t n = 1;
  double scale = 0.0;
  __host__ __device__ __forceinline__ uint32_t of_steps(double st) const {
    const double c = st * scale;
    return c >= 1.0 ? (c < (double)(n - 1) ? (uint32_t)c : n - 1) : 0u;
  }
  __host__ __device__ __forceinline__ uint32_t of_crossings(double h, double b, double inv_vl_sum) const {
    return of_steps((b - h) * inv_vl_sum);
  }
};

__host__ __device__ __forceinline__ double inv_vl_sum(const WalkC& c) {
  return (double)(1.0f / c.vl_x) + (double)(1.0f / c.vl_y) + (double)(1.0f / c.vl_z);
}

// The tiles in which bin_lor<false> emits a run for LOR i, in walk order,
// without walking its steps: calls on_run(tile) for each and returns their
// number.  The walk leaves a tile at the first, in (position, axis x < y <
// z) order, of the three axes' tile-leaving crossings, and the FOV at the
// first of their FOV-leaving ones; each of these is one axis's boundary
// after a known number m of crossings of that axis.  Its binary32 value
// (m additions of vl to nb, each rounded) lies within m 2^-23 b of the
// binary64 estimate b = nb + m vl: every addition rounds by at most half an
// ulp, 2^-24 of a value at most the result.  The order of the three is
// taken from the estimates where their error intervals are disjoint, and
// from the exact values (repeat_add) where they are not.  The walk's
// positions never decrease and its step lengths are the differences of
// consecutive positions, so a tile has a pushed step (weight > 0, no TOF)
// exactly when the crossing that leaves it lies beyond the one that
// entered it; that test too is decided on the estimates where it can be.
// Walks whose boundaries are not positive (degenerate LORs) are stepped
// through by bin_lor.
__host__ __device__ __forceinline__ double crossing_estimate(float nb, float vl, uint32_t m, double& err) {
  const double b = (double)nb + (double)m * (double)vl;
  err = b < (double)INFINITY ? (double)m * 0x1p-23 * b : 0.0;
  return b;
}

template <class OnRun>
__host__ __device__ __forceinline__ uint32_t tile_runs(const Lor& l, uint32_t i, const DevFov& f, const TileGeom& g,
                                                       const RunClasses& rc, OnRun&& on_run) {
  WalkC c;
  WalkS w;
  if (!walk_setup(l, f, c, w) || !bounded(w)) return 0;
  if (!(w.nb_x > 0.0f && w.nb_y > 0.0f && w.nb_z > 0.0f && c.vl_x > 0.0f && c.vl_y > 0.0f && c.vl_z > 0.0f))
    return bin_lor<false>(l, i, f, no_tof(), g, [&](const Run& ru, uint32_t tl) {
      on_run(tl, rc.of_steps((double)run_len(ru.code)));
    });
  const double ivs = inv_vl_sum(c);
  const float nb[3] = {w.nb_x, w.nb_y, w.nb_z}, vl[3] = {c.vl_x, c.vl_y, c.vl_z};
  int32_t tile;
  uint32_t idx[3], last[3];
  bool out[3];
  double b[3], err[3];
  {
    const int32_t nxy = f.nx * f.ny;
    const int32_t z = w.index / nxy;
    const int32_t rem = w.index - z * nxy;
    const int32_t y = rem / f.nx;
    const int32_t x = rem - y * f.nx;
    const bool px = c.d_x > 0, py = c.d_y > 0, pz = c.d_z > 0;
    // crossings of each axis until the walk leaves the tile along it
    const uint32_t cc[3] = {(uint32_t)(px ? g.tx - x % g.tx : x % g.tx + 1),
                            (uint32_t)(py ? g.ty - y % g.ty : y % g.ty + 1),
                            (uint32_t)(pz ? g.tz - z % g.tz : z % g.tz + 1)};
    tile = x / g.tx + g.ntx * (y / g.ty + g.nty * (z / g.tz));
    const int32_t rr[3] = {w.r_x, w.r_y, w.r_z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      last[a] = (uint32_t)rr[a] - 1u;  // index of the crossing that leaves the FOV
      idx[a] = cc[a] - 1u < last[a] ? cc[a] - 1u : last[a];
      out[a] = cc[a] - 1u >= last[a];
      b[a] = crossing_estimate(nb[a], vl[a], idx[a], err[a]);
    }
  }
  const uint32_t step[3] = {(uint32_t)g.tx, (uint32_t)g.ty, (uint32_t)g.tz};
  const int32_t dtile[3] = {c.d_x > 0 ? 1 : -1, (c.d_y > 0 ? 1 : -1) * g.ntx, (c.d_z > 0 ? 1 : -1) * g.ntx * g.nty};
  double h = 0.0, herr = 0.0;  // the crossing that entered the tile (the walk starts at here = 0)
  bool hstart = true;          // no crossing yet; else its axis's nb, vl and index give its exact value
  float hnb = 0.0f, hvl = 0.0f;
  uint32_t hidx = 0;
  uint32_t count = 0;
  for (;;) {
    int e = 0;
    if (b[1] < b[e]) e = 1;
    if (b[2] < b[e]) e = 2;
    bool sure = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (a != e) sure = sure && b[e] + err[e] < b[a] - err[a];
    float Be = 0.0f;
    bool exact = false;
    if (!sure) {
      float B[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) B[a] = repeat_add(nb[a], vl[a], idx[a]);
      e = 0;
      if (B[1] < B[e]) e = 1;
      if (B[2] < B[e]) e = 2;
      Be = e == 0 ? B[0] : (e == 1 ? B[1] : B[2]);
      exact = true;
    }
    uint32_t ie = idx[0], le = last[0], se = step[0];
    bool oe = out[0];
    float nbe = nb[0], vle = vl[0];
    int32_t de = dtile[0];
    double bee = b[0], eee = err[0];
#pragma unroll
    for (int a = 1; a < 3; ++a)
      if (a == e) {
        ie = idx[a];
        le = last[a];
        se = step[a];
        oe = out[a];
        nbe = nb[a];
        vle = vl[a];
        de = dtile[a];
        bee = b[a];
        eee = err[a];
      }
    bool push = bee - eee > h + herr;
    if (!push) {
      if (!exact) Be = repeat_add(nbe, vle, ie);
      const float hx = hstart ? w.here : repeat_add(hnb, hvl, hidx);
      push = Be > hx;
    }
    if (push) {
      on_run((uint32_t)tile, rc.of_crossings(h, bee, ivs));
      ++count;
    }
    if (oe) break;
    h = bee;
    herr = eee;
    hstart = false;
    hnb = nbe;
    hvl = vle;
    hidx = ie;
    tile += de;
    const uint32_t nx = ie + se < le ? ie + se : le;
    double en;
    const double bn = crossing_estimate(nbe, vle, nx, en);
    const bool on = ie + se >= le;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (a == e) {
        idx[a] = nx;
        out[a] = on;
        b[a] = bn;
        err[a] = en;
      }
  }
  return count;
}

// bin_lor without TOF by tile jumps (the fill pass's form of tile_runs):
// emits exactly bin_lor<false>'s runs and adds its step count, without
// generating every boundary.  Per tile segment the exit crossing is found as
// tile_runs finds it (binary64 estimates, exact values where their error
// intervals overlap), and the segment's exact state follows from repeated
// binary32 additions done by repeat_add: the exit axis's crossing before its
// last (Ep) and its last (Ee = Ep + vl), and for each other axis the number
// n of its crossings that precede Ee in the walk's (value, axis) order --
// from the binary64 quotient (Ee - V) / vl where that is decided beyond its
// error, by counting the crossings one by one where it is not -- with the
// last of them (P) and the next (P + vl).  V, the axes' next boundaries,
// are so exact at every tile entry, and a run there is (V, here = the entry
// crossing, the tile's local index, its step count) when its first step
// (min V - here) and its last (Ee - max P) have positive length; a segment
// with a zero-length first or last step (ties at its ends) is stepped
// through for its first and last pushed steps.  Walks with a non-positive
// boundary or voxel length (degenerate LORs) are stepped through by bin_lor.
template <class Emit>
__host__ __device__ __forceinline__ uint32_t bin_lor_jump(const Lor& l, uint32_t i, const DevFov& f,
                                                          const TileGeom& g, const RunClasses& rc, Emit&& emit,
                                                          uint32_t* n_steps = nullptr) {
  WalkC c;
  WalkS w;
  if (!walk_setup(l, f, c, w) || !bounded(w)) return 0;
  if (!(w.nb_x > 0.0f && w.nb_y > 0.0f && w.nb_z > 0.0f && c.vl_x > 0.0f && c.vl_y > 0.0f && c.vl_z > 0.0f))
    return bin_lor<false>(l, i, f, no_tof(), g, [&](const Run& ru, uint32_t tl) {
      emit(ru, tl, rc.of_steps((double)run_len(ru.code)));
    }, n_steps, nullptr);
  const double ivs = inv_vl_sum(c);
  const float ivl[3] = {1.0f / c.vl_x, 1.0f / c.vl_y, 1.0f / c.vl_z};
  double h_est = 0.0;  // estimate of the crossing that entered the tile, as tile_runs keeps it
  const float nb[3] = {w.nb_x, w.nb_y, w.nb_z}, vl[3] = {c.vl_x, c.vl_y, c.vl_z};
  int32_t cc[3], sgn[3], tile;
  {
    const int32_t nxy = f.nx * f.ny;
    const int32_t z = w.index / nxy;
    const int32_t rem = w.index - z * nxy;
    const int32_t y = rem / f.nx;
    const int32_t x = rem - y * f.nx;
    sgn[0] = c.d_x < 0 ? -1 : 1;
    sgn[1] = c.d_y < 0 ? -1 : 1;
    sgn[2] = c.d_z < 0 ? -1 : 1;
    cc[0] = sgn[0] > 0 ? g.tx - x % g.tx : x % g.tx + 1;
    cc[1] = sgn[1] > 0 ? g.ty - y % g.ty : y % g.ty + 1;
    cc[2] = sgn[2] > 0 ? g.tz - z % g.tz : z % g.tz + 1;
    tile = x / g.tx + g.ntx * (y / g.ty + g.nty * (z / g.tz));
  }
  const int32_t tt[3] = {g.tx, g.ty, g.tz};
  const int32_t dtile[3] = {sgn[0], sgn[1] * g.ntx, sgn[2] * g.ntx * g.nty};
  const uint32_t flips = (sgn[0] < 0 ? 1u : 0u) | (sgn[1] < 0 ? 2u : 0u) | (sgn[2] < 0 ? 4u : 0u);
  float V[3] = {w.nb_x, w.nb_y, w.nb_z};         // exact next boundary per axis
  uint32_t cnt[3] = {0u, 0u, 0u};                // crossings of each axis so far
  int32_t r[3] = {w.r_x, w.r_y, w.r_z};          // crossings left per axis (the last one leaves the FOV)
  float here = w.here;
  uint32_t count = 0, total = 0;
  uint32_t cls = 0;  // the segment's length class
  auto make_run = [&](const float s_nb[3], float h, const int32_t cq[3], uint32_t len) {
    Coord q;
    q.cx = cq[0];
    q.cy = cq[1];
    q.cz = cq[2];
    q.sx = sgn[0];
    q.sy = sgn[1];
    q.sz = sgn[2];
    Run ru;
    ru.nb_x = s_nb[0];
    ru.nb_y = s_nb[1];
    ru.nb_z = s_nb[2];
    ru.here = h;
    ru.vl_x = c.vl_x;
    ru.vl_y = c.vl_y;
    ru.vl_z = c.vl_z;
    ru.tof_peak = c.tof_peak;
    ru.code = (local_of(g, q) << 15) | (flips << 12) | len;
    ru.lor = i;
    emit(ru, (uint32_t)tile, cls);
    ++count;
  };
  for (;;) {
    // the segment-ending crossing of each axis: m[a]-th from now
    int32_t m[3];
    double b[3], err[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      m[a] = cc[a] < r[a] ? cc[a] : r[a];
      b[a] = crossing_estimate(nb[a], vl[a], cnt[a] + (uint32_t)m[a] - 1u, err[a]);
    }
    int e = 0;
    if (b[1] < b[e]) e = 1;
    if (b[2] < b[e]) e = 2;
    bool sure = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (a != e) sure = sure && b[e] + err[e] < b[a] - err[a];
    if (!sure) {
      float E[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) E[a] = repeat_add(V[a], vl[a], (uint32_t)m[a] - 1u);
      e = 0;
      if (E[1] < E[e]) e = 1;
      if (E[2] < E[e]) e = 2;
    }
    float Ve = V[0], vle = vl[0];
    int32_t me = m[0];
    double be = b[0];
#pragma unroll
    for (int a = 1; a < 3; ++a)
      if (a == e) {
        Ve = V[a];
        vle = vl[a];
        me = m[a];
        be = b[a];
      }
    cls = rc.of_crossings(h_est, be, ivs);
    // the exit axis's last two crossings of the segment
    const float Ep = me >= 2 ? repeat_add(Ve, vle, (uint32_t)me - 2u) : -INFINITY;
    const float Ee = me >= 2 ? Ep + vle : Ve;
    float P = here > Ep ? here : Ep;  // start of the segment's last step
    int32_t n[3];
    float Vn[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a == e) {
        n[a] = me;
        Vn[a] = V[a];
        continue;
      }
      const int32_t lim = m[a] - 1;
      const float v0 = V[a];
      // crossings of axis a before Ee in the walk's (value, axis) order, at most lim
      auto before = [&](float x) { return x < Ee || (x == Ee && a < e); };
      int32_t k = 0;
      float v = v0, pa = -INFINITY;
      if (lim > 0 && before(v0)) {
        // crossing j lies near v0 + j vl: about ceil((Ee - v0) / vl) of them
        // precede Ee.  The estimate is checked on the exact crossings it
        // names: lowered while crossing ke - 1 does not precede Ee, raised
        // while crossing ke does.
        const float xe = ceilf((Ee - v0) * ivl[a]);
        int32_t ke = xe >= 1.0f ? (xe < (float)lim ? (int32_t)xe : lim) : 1;
        for (;;) {
          pa = repeat_add(v0, vl[a], (uint32_t)ke - 1u);
          if (ke == 1 || before(pa)) break;
          --ke;
        }
        v = pa + vl[a];
        k = ke;
        while (k < lim && before(v)) {
          pa = v;
          v = v + vl[a];
          ++k;
        }
      }
      P = P > pa ? P : pa;
      n[a] = k;
      Vn[a] = v;
    }
    const uint32_t steps = (uint32_t)(n[0] + n[1] + n[2]);
    total += steps;
    const float F = fminf(fminf(V[0], V[1]), V[2]);  // end of the segment's first step
    if (F > here && Ee > P) {
      make_run(V, here, cc, steps);
    } else {
      // step through the segment (ties at its ends): first and last pushed step
      WalkS s = w;
      s.nb_x = V[0];
      s.nb_y = V[1];
      s.nb_z = V[2];
      s.here = here;
      WalkC cs = c;
      int32_t q3[3] = {cc[0], cc[1], cc[2]};
      float nb0[3] = {0.0f, 0.0f, 0.0f};
      float h0 = 0.0f;
      int32_t c0[3] = {0, 0, 0};
      int64_t first = -1, last = -1;
      for (uint32_t k = 0; k < steps; ++k) {
        const Step st = next_boundary(s);
        if (st.bnd - s.here > 0.0f) {
          if (first < 0) {
            first = k;
            nb0[0] = s.nb_x;
            nb0[1] = s.nb_y;
            nb0[2] = s.nb_z;
            h0 = s.here;
            c0[0] = q3[0];
            c0[1] = q3[1];
            c0[2] = q3[2];
          }
          last = k;
        }
        q3[0] -= st.ax ? 1 : 0;
        q3[1] -= st.ay ? 1 : 0;
        q3[2] -= st.az ? 1 : 0;
        move(s, cs, st);
      }
      if (first >= 0) make_run(nb0, h0, c0, (uint32_t)(last - first + 1));
    }
    bool out = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      r[a] -= n[a];
      cnt[a] += (uint32_t)n[a];
      if (a == e) {
        V[a] = Ee + vl[a];
        out = r[a] == 0;
        cc[a] = tt[a];
        tile += dtile[a];
      } else {
        V[a] = Vn[a];
        cc[a] -= n[a];
      }
    }
    here = Ee;
    h_est = be;
    if (out) break;
  }
  if (n_steps) *n_steps += total;
  return count;
}

// Length field of a run's sort key (bits 0-11): longer runs first, in
// buckets of 4 steps.  Lanes of a wavefront get runs of about equal length;
// inside a bucket the (stable) sort keeps LOR order, so neighbouring LORs'
// partials land side by side and the reduction's gathers share cache lines
// (exact lengths cost the reduction 15% more time; LOR order alone costs the
// replays 25-60%).
constexpr uint32_t kLenBucketShift = 2;
__host__ __device__ __forceinline__ uint64_t len_key(uint32_t code) {
  return (uint64_t)(((kRunMaxSteps - run_len(code)) >> kLenBucketShift) << kLenBucketShift);
}

// Binning walk: FILL = false counts runs per LOR into n_runs[i]; FILL = true
// writes them at runs[offset[i] .. offset[i + 1]) (LOR-major; a run past the
// end of that range is not written, a LOR left short is padded with empty
// runs, and a LOR whose run count differs from the range is counted in
// counters[CNT_RUN_MISMATCH]) with their sort keys:
// tile (bits 32 and up), then block of 2^lor_shift consecutive LORs (bits
// 12-31; LORs are ordered spatially, so a block's 1/(p*ac) values stay
// cache-resident while its runs in a tile are replayed), then len_key
// (bits 0-11).  With lor_shift >= 32 the key is (tile << 12) | len_key only.
// Each lane bins kBinLors LORs, 256 apart, one after the other: walks differ
// in length, and a lane's total over several LORs varies less than one
// walk's, so the lanes of a wavefront finish closer together.
constexpr uint32_t kBinLors = 8;
// Runs of a tile are sorted by blocks of 2^kLorBlockShift consecutive LORs
// before length.
constexpr uint32_t kLorBlockShift = 14;
// The partial projections are stored by blocks of 2^kWinShift LORs (the
// reduction's LDS windows).
constexpr uint32_t kWinShift = 12;
inline unsigned bin_grid(size_t n) { return (unsigned)((n + 256 * kBinLors - 1) / (256 * kBinLors)); }

template <bool TOF, bool FILL>
__global__ __launch_bounds__(256) void k_bin_runs(const Lor* __restrict__ lors, uint32_t n, DevFov f, DevTof t,
                                                  TileGeom g, uint32_t* __restrict__ n_runs,
                                                  const uint32_t* __restrict__ offset, Run* __restrict__ runs,
                                                  uint64_t* __restrict__ keys, uint32_t lor_shift,
                                                  unsigned long long* __restrict__ counters = nullptr) {
  uint32_t steps = 0, nz = 0, missed = 0, shifted = 0, mismatch = 0;
  const uint32_t first = blockIdx.x * blockDim.x * kBinLors + threadIdx.x;
  for (uint32_t u = 0; u < kBinLors; ++u) {
    const uint32_t i = first + u * blockDim.x;
    if (i >= n) break;
    const Lor l = load_lor(lors, i);
    const uint32_t base = FILL ? offset[i] : 0;
    const uint32_t room = FILL ? offset[i + 1] - base : 0;
    uint32_t k = 0;
    auto emit = [&](const Run& ru, uint32_t tile) {
      if (FILL && k < room) {
        runs[base + k] = ru;
        // lor_shift >= 32: no LOR-block field, key = (tile << 12) | ...
        const uint64_t blk = lor_shift >= 32 ? 0u : (uint64_t)(i >> lor_shift);
        const uint32_t tshift = lor_shift >= 32 ? 12u : 32u;
        keys[base + k] = ((uint64_t)tile << tshift) | (blk << 12) | len_key(ru.code);
      }
      ++k;
    };
    // filling without TOF: by tile jumps (the same runs; counting, the step
    // walk also gives the pushed-step count)
    const uint32_t count = (!TOF && FILL) ? bin_lor_jump(l, i, f, g, RunClasses{},
                                                         [&](const Run& ru, uint32_t tile, uint32_t) { emit(ru, tile); })
                                          : bin_lor<TOF>(l, i, f, t, g, emit, FILL ? nullptr : &steps,
                                                         FILL ? nullptr : &nz);
    if (FILL) {
      mismatch += count != room ? 1u : 0u;
      // a LOR left short holds empty runs (length 0, LOR 0, tile 0), not stale records
      for (; k < room; ++k) {
        runs[base + k] = Run{};
        keys[base + k] = len_key(0u);
      }
    }
    if (!FILL) {
      // unbounded walks are projected by the one-lane kernels instead
      WalkC c;
      WalkS w;
      bool sh = false;
      const bool hit = walk_setup(l, f, c, w, &sh);
      missed += hit ? 0u : 1u;
      shifted += (hit && sh) ? 1u : 0u;
      n_runs[i] = (hit && !bounded(w)) ? 0xffffffffu : count;
    }
  }
  if (!FILL && counters) {
    // exact work counters of the bounded walks (unbounded ones count where
    // they are projected)
    wave_add_u64(counters + CNT_STEPS, steps);
    wave_add_u64(counters + CNT_NZ, nz);
    wave_add_u64(counters + CNT_MISSED, missed);
    wave_add_u64(counters + CNT_SHIFTED, shifted);
  }
  if (FILL && counters) wave_add_u64(counters + CNT_RUN_MISMATCH, mismatch);
}

// Binning without TOF straight into tile order, for the sensitivity image.
// LOR i is lors[order[i]] (lors[i] without order).  Workgroup w of a grid of
// G takes the same LORs in both passes (as k_bin_runs), and its runs in tile
// t of length class c (RunClasses rc, C = rc.n classes) form segment
// ((t * C + C - 1 - c) * G + w) of a tile-major layout: in each tile the
// longest class first, and lanes that replay neighbouring runs of a tile get
// runs of about one length.  Segment sizes go to seg, their exclusive prefix
// sums are soff.  Count (FILL false): n_runs[i] as k_bin_runs<false, false>,
// and seg from a per-workgroup LDS histogram of the tile jumps (tile_runs);
// STATS also walks every step (bin_lor) for the step and pushed-step
// counters.  Fill (FILL true): each run (bin_lor_jump, the same tiles and
// classes) goes to the next free position of its segment (positions within
// a segment follow no order) and, with run_pos, LOR i's k-th run position to
// run_pos[offset[i] + k].  A run past its segment's end is not written (its
// run_pos entry is 0), a segment left short is padded with empty runs, and
// every segment whose fill differs from its count adds to
// counters[CNT_RUN_MISMATCH].  Dynamic LDS: 4 * ntiles * C bytes (count) or
// 8 * ntiles * C bytes (fill).
template <bool FILL, bool STATS>
__global__ __launch_bounds__(256) void k_bin_runs_seg(const Lor* __restrict__ lors, const uint32_t* __restrict__ order,
                                                      uint32_t n, DevFov f, TileGeom g, RunClasses rc,
                                                      uint32_t* __restrict__ n_runs, uint32_t* __restrict__ seg,
                                                      const uint32_t* __restrict__ soff,
                                                      const uint32_t* __restrict__ offset, Run* __restrict__ runs,
                                                      uint32_t* __restrict__ run_pos,
                                                      unsigned long long* __restrict__ counters) {
  // count: [0, ns) runs per segment; fill: [0, ns) cursor, [ns, 2 ns) segment end
  extern __shared__ uint32_t cur[];
  const uint32_t ncls = rc.n, ns = (uint32_t)g.ntiles * ncls, ng = gridDim.x, wg = blockIdx.x;
  for (uint32_t t = threadIdx.x; t < ns; t += blockDim.x) {
    if (FILL) {
      const uint32_t a = soff[(size_t)t * ng + wg];
      cur[t] = a;
      cur[ns + t] = a + seg[(size_t)t * ng + wg];
    } else {
      cur[t] = 0;
    }
  }
  __syncthreads();
  const DevTof notof = no_tof();
  uint32_t steps = 0, nz = 0, missed = 0, shifted = 0, mismatch = 0;
  const uint32_t first = wg * blockDim.x * kBinLors + threadIdx.x;
  for (uint32_t u = 0; u < kBinLors; ++u) {
    const uint32_t i = first + u * blockDim.x;
    if (i >= n) break;
    const Lor l = load_lor(lors, order ? order[i] : i);
    const uint32_t base = (FILL && run_pos) ? offset[i] : 0;
    const uint32_t room = (FILL && run_pos) ? offset[i + 1] - base : 0;
    uint32_t count;
    if (FILL) {
      uint32_t k = 0;
      count = bin_lor_jump(l, i, f, g, rc, [&](const Run& ru, uint32_t tile, uint32_t cls) {
        const uint32_t sg = tile * ncls + (ncls - 1u - cls);
        const uint32_t at = atomicAdd(&cur[sg], 1u);
        const bool fits = at < cur[ns + sg];
        if (fits) runs[at] = ru;
        if (run_pos && k < room) run_pos[base + k] = fits ? at : 0u;
        ++k;
      });
      if (run_pos) mismatch += count != room ? 1u : 0u;
    } else {
      count = tile_runs(l, i, f, g, rc,
                        [&](uint32_t tile, uint32_t cls) { atomicAdd(&cur[tile * ncls + (ncls - 1u - cls)], 1u); });
      if (STATS) bin_lor<false>(l, i, f, notof, g, [](const Run&, uint32_t) {}, &steps, &nz);
      WalkC c;
      WalkS w;
      bool sh = false;
      const bool hit = walk_setup(l, f, c, w, &sh);
      missed += hit ? 0u : 1u;
      shifted += (hit && sh) ? 1u : 0u;
      n_runs[i] = (hit && !bounded(w)) ? 0xffffffffu : count;
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < ns; t += blockDim.x) {
    if (FILL) {
      // the cursor ends at the segment's end exactly when the segment holds its count
      mismatch += cur[t] != cur[ns + t] ? 1u : 0u;
      // a segment left short holds empty runs (length 0, LOR 0), not stale records
      for (uint32_t x = cur[t]; x < cur[ns + t]; ++x) runs[x] = Run{};
    } else {
      seg[(size_t)t * ng + wg] = cur[t];
    }
  }
  if (!counters) return;
  if (FILL) {
    wave_add_u64(counters + CNT_RUN_MISMATCH, mismatch);
  } else {
    if (STATS) {
      wave_add_u64(counters + CNT_STEPS, steps);
      wave_add_u64(counters + CNT_NZ, nz);
    }
    wave_add_u64(counters + CNT_MISSED, missed);
    wave_add_u64(counters + CNT_SHIFTED, shifted);
  }
}

// Walk state of a run for a replay against an LDS tile of elements of
// 2^kShift bytes: the voxel index and its per-axis strides are kept in bytes,
// so a step's LDS address needs no scaling.
template <int kShift>
__host__ __device__ __forceinline__ void run_state(const Run& ru, const TileGeom& g, WalkS& w, WalkC& c,
                                                   uint32_t base = 0) {
  const int32_t dlx = 1 << kShift, dly = g.px << kShift, dlz = (g.px * g.ty) << kShift;
  w.nb_x = ru.nb_x;
  w.nb_y = ru.nb_y;
  w.nb_z = ru.nb_z;
  w.here = ru.here;
  c.vl_x = ru.vl_x;
  c.vl_y = ru.vl_y;
  c.vl_z = ru.vl_z;
  c.tof_peak = ru.tof_peak;
  const uint32_t flips = run_flips(ru.code);
  c.d_x = (flips & 1u) ? -dlx : dlx;
  c.d_y = (flips & 2u) ? -dly : dly;
  c.d_z = (flips & 4u) ? -dlz : dlz;
  w.index = (int32_t)(base + (run_local(ru.code) << kShift));
}

// LDS address (byte offset in the workgroup's LDS) of a pointer to shared
// memory, and LDS pointers made from such addresses: a replay keeps its
// voxel's LDS address itself, so that each step's access needs no add of the
// tile's base.
typedef __attribute__((address_space(3))) float lds_f32;
typedef __attribute__((address_space(3))) double lds_f64;
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// Replays a run from its recorded state, calling add(lds_address, w * rr)
// for every pushed step (addresses of elements of 8 bytes), where rr is the
// run's factor: r times the Gaussian's 666 * peak with TOF, times the
// accumulator's power-of-two scale (TileAcc); the tile's entries are
// 2^kShift bytes.  The voxel sequence and step
// lengths are walk #1's bit for bit; the TOF weight is
// length * tof_shape * (666 * peak), with the constant folded into rr once
// per run (a few ulp from walk #1's product).  Every step of a run lies
// inside the TOF window, so the window test of step_weight is skipped: the
// run spans the walk's first to last pushed step in the tile, pushed steps
// lie inside the window, `here` does not decrease along the walk and the
// window is an interval of here - peak.  The Gaussian is at least exp(-4.5)
// there, so wl > 0 exactly when the step length is; a step of length 0 is
// not pushed.  GUARD false adds it anyway, which adds +0 when r is finite
// (the caller's condition), saving the test.
template <bool TOF, bool GUARD, int kShift = 3, class Add>
__host__ __device__ __forceinline__ void replay_run(const Run& ru, float rr, const DevTof& t, const TileGeom& g,
                                                    uint32_t base, Add&& add) {
  WalkS w;
  WalkC c;
  run_state<kShift>(ru, g, w, c, base);
  const uint32_t m = run_len(ru.code);  // the run length bounds the replay
  auto step = [&] {
    const float bnd = fminf(fminf(w.nb_x, w.nb_y), w.nb_z);
    float wl = bnd - w.here;
    if (TOF) wl *= tof_shape(t, w.here, c.tof_peak);
    if (!GUARD || wl > 0.0f) add((uint32_t)w.index, wl * rr);
    cross(w, c, bnd);
  };
  // four steps per trip, written out: the compiler does not unroll a loop of
  // a lane-dependent trip count around the masked crossing
  uint32_t k = 0;
  for (; k + 4 <= m; k += 4) {
    step();
    step();
    step();
    step();
  }
  for (; k < m; ++k) step();
}

// Forward replay of a run against the LDS tile: p = sum of w * x over the
// run's pushed steps in walk order, as (666 * peak) * sum of
// (length * tof_shape) * x with fused multiply-adds.  Each step's LDS read is
// consumed one step later, so its latency overlaps the next step's
// arithmetic; every step of a run lies inside the tile, so the read is always
// in bounds.  A step that is not pushed adds nothing: with GUARD the select
// keeps x out of it, as the reference never reads it; without GUARD (the
// caller's tile holds only finite values) it adds 0 * x = +0.
template <bool TOF, bool GUARD>
__device__ __forceinline__ float replay_project(const Run& ru, const DevTof& t, const TileGeom& g,
                                                const float* __restrict__ timg) {
  WalkS w;
  WalkC c;
  run_state<2>(ru, g, w, c, lds_addr(timg));
  const uint32_t m = run_len(ru.code);
  float p = 0.0f, w_prev = 0.0f, x_prev = 0.0f;
  auto step = [&] {
    const float bnd = fminf(fminf(w.nb_x, w.nb_y), w.nb_z);
    float wl = bnd - w.here;  // in-window: see replay_run
    if (TOF) wl *= tof_shape(t, w.here, c.tof_peak);
    const float x = *(const lds_f32*)(uintptr_t)(uint32_t)w.index;
    p = fmaf(w_prev, (!GUARD || w_prev > 0.0f) ? x_prev : 0.0f, p);
    w_prev = wl;
    x_prev = x;
    cross(w, c, bnd);
  };
  uint32_t k = 0;  // unrolled by hand, as in replay_run
  for (; k + 4 <= m; k += 4) {
    step();
    step();
    step();
    step();
  }
  for (; k < m; ++k) step();
  p = fmaf(w_prev, (!GUARD || w_prev > 0.0f) ? x_prev : 0.0f, p);
  return TOF ? t.amp * p : p;
}

// Image voxels a lane of the forward kernel loads before writing them to LDS.
constexpr int32_t kTileLoads = 8;

// s_waitcnt vmcnt(0) (gfx9 encoding: vmcnt bits 0-3 and 14-15 zero, expcnt
// and lgkmcnt at their maxima, so not waited for).
constexpr int kWaitVmcnt0 = 0x0f70;

// A lane's runs j, j + stride, ... <= last of a forward work item: each
// run's partial is stored at lpos[j] (at j with PERM).  The position and the
// next run's record are loaded before the replay, so neither load's latency
// stalls it.  PERM: the runs are read through perm.
template <bool TOF, bool PERM>
__device__ __forceinline__ void fwd_runs(const Run* __restrict__ runs, const uint32_t* __restrict__ perm,
                                         const uint32_t* __restrict__ lpos, int32_t j, const int32_t last,
                                         const bool guard, const DevTof& t, const TileGeom& g,
                                         const float* __restrict__ timg, float* __restrict__ partial) {
  const int32_t stride = (int32_t)blockDim.x;
  Run cur = runs[PERM ? perm[j] : j];
  // Wait for the first record here: the loop's later records are register
  // copies made after their loads completed, and without this wait the
  // replay loop would wait on the vector-memory counter at every run (that
  // counter also holds the previous run's partial store).
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  for (;;) {
    const uint32_t at = (PERM || !lpos) ? (uint32_t)j : lpos[j];
    const int32_t jn = j + stride;
    const int32_t jc = jn < last ? jn : last;  // in bounds; unused past the end
    const Run nxt = runs[PERM ? perm[jc] : jc];
    const float p = guard ? replay_project<TOF, true>(cur, t, g, timg) : replay_project<TOF, false>(cur, t, g, timg);
    partial[at] = p;
    if (jn > last) break;
    cur = nxt;
    j = jn;
  }
}

// Forward, tiled: one workgroup per work item copies its tile of the image
// (dynamic LDS: tvox + 1 words) into LDS and replays its runs, writing each
// run's partial projection (sum of w*x over the run's pushed steps, in walk
// order) at lpos[j], its place in the (window block, tile) layout, or, with
// perm or without lpos, at its sorted position j (stores nearly contiguous;
// writing them LOR-major instead scatters 4-byte stores over the whole
// array and costs more than it saves in k_lor_reduce).  Each lane loads its next run record before
// replaying the current one, so the record's latency hides behind the
// replay (the replay itself touches only LDS).
template <bool TOF>
__device__ __forceinline__ void fwd_tile_body(const Run* __restrict__ runs, const int4 wi,
                                              const float* __restrict__ img, const DevFov& f, const DevTof& t,
                                              const TileGeom& g, float* __restrict__ partial,
                                              const uint32_t* __restrict__ perm,
                                              const uint32_t* __restrict__ lpos) {
  extern __shared__ float timg[];
  const int32_t tl = wi.x;
  const int32_t tzi = tl / (g.ntx * g.nty);
  const int32_t tyi = (tl / g.ntx) % g.nty;
  const int32_t txi = tl % g.ntx;
  const int32_t x0 = txi * g.tx, y0 = tyi * g.ty, z0 = tzi * g.tz;
  // one word after the tile: set when the tile holds a non-finite voxel
  volatile int32_t* odd_tile = reinterpret_cast<volatile int32_t*>(timg + g.tvox);
  if (threadIdx.x == 0) *odd_tile = 0;
  __syncthreads();
  // the tile is read kTileLoads entries per lane at a time, all loads in
  // flight before the first LDS write
  bool odd = false;
  const int32_t stride = (int32_t)blockDim.x;
  TileCursor cu(g, (int32_t)threadIdx.x, stride);
  while (cu.k < g.tvox) {
    const int32_t k0 = cu.k;
    float v[kTileLoads];
#pragma unroll
    for (int32_t u = 0; u < kTileLoads; ++u) {
      const int32_t x = x0 + cu.lx, y = y0 + cu.ly, z = z0 + cu.lz;
      const bool in = cu.k < g.tvox && cu.lx < g.tx && x < f.nx && y < f.ny && z < f.nz;
      v[u] = in ? img[x + f.nx * (y + f.ny * z)] : 0.0f;
      cu.next(g, stride);
    }
#pragma unroll
    for (int32_t u = 0; u < kTileLoads; ++u) {
      const int32_t k = k0 + u * stride;
      if (k < g.tvox) timg[k] = v[u];
      odd |= !(fabsf(v[u]) <= 3.402823466e38f);  // inf or NaN
    }
  }
  if (odd) *odd_tile = 1;
  __syncthreads();
  // the zero-length-step select is needed only where a voxel is not finite
  const bool guard = *odd_tile != 0;
  const int32_t j = wi.y + (int32_t)threadIdx.x;
  if (j < wi.z) {
    if (perm)
      fwd_runs<TOF, true>(runs, perm, lpos, j, wi.z - 1, guard, t, g, timg, partial);
    else
      fwd_runs<TOF, false>(runs, perm, lpos, j, wi.z - 1, guard, t, g, timg, partial);
  }
}

template <bool TOF>
__global__ __launch_bounds__(1024) void k_fwd_tile(const Run* __restrict__ runs, const int4* __restrict__ work,
                                                   const float* __restrict__ img, DevFov f, DevTof t, TileGeom g,
                                                   float* __restrict__ partial,
                                                   const uint32_t* __restrict__ perm,
                                                   const uint32_t* __restrict__ lpos) {
  const int4 wi = work[blockIdx.x];
  if (wi.z <= wi.y) return;  // an empty item (the grid may be an upper bound)
  fwd_tile_body<TOF>(runs, wi, img, f, t, g, partial, perm, lpos);
}

// One reconstruction of a batch, as the batched tile kernels see it: the runs
// and buffers of its current subset.
struct ReconDesc {
  const Run* runs;
  const float* image;
  float* partial;
  const float* r;
  double* bp;
  const uint32_t* lpos;  // partial position of each sorted run
  unsigned long long* rstats;  // r statistics of the update (kRStatWords)
  DevFov f;
  DevTof t;
  TileGeom g;
};

// Batched forward: work item w = (tile, first run, end run, reconstruction).
template <bool TOF>
__global__ __launch_bounds__(1024) void k_fwd_tile_batch(const ReconDesc* __restrict__ rd,
                                                         const int4* __restrict__ work) {
  const int4 wi = work[blockIdx.x];
  const ReconDesc& r = rd[wi.w];
  fwd_tile_body<TOF>(r.runs, wi, r.image, r.f, r.t, r.g, r.partial, nullptr, r.lpos);
}

// r = 1/(p*ac) per LOR (MLEM; 1/exp(-p) for the sensitivity image) with p the
// sum of its runs' partials in walk order;
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
// Positions of a LOR's first four runs' partials (kNoRun where it has
// fewer; bit 31 of the fourth set when it has more, which run_off/run_pos
// then list), so that the reduction waits on one 16-byte load before its
// gathers instead of two dependent ones.
constexpr uint32_t kNoRun = 0xffffffffu;
constexpr uint32_t kMoreRuns = 0x80000000u;

// Reduces LORs base + first + 64*u (u < kReduceLors, those below base + end).
template <bool SENS>
__device__ __forceinline__ void reduce_lors(const float* __restrict__ ac, uint32_t base, uint32_t first,
                                            uint32_t end, const uint32_t* __restrict__ run_off,
                                            const uint32_t* __restrict__ run_pos, const uint4* __restrict__ pos4,
                                            const float* __restrict__ partial, float* __restrict__ r_out,
                                            unsigned long long* rstats) {
  uint32_t idx[kReduceLors];
  bool on[kReduceLors];
  uint32_t q[kReduceLors][4];
#pragma unroll
  for (uint32_t u = 0; u < kReduceLors; ++u) {
    const uint32_t local = first + 64u * u;
    idx[u] = base + local;
    on[u] = local < end;
    const uint4 p4 = on[u] ? pos4[idx[u]] : make_uint4(kNoRun, kNoRun, kNoRun, kNoRun);
    q[u][0] = p4.x;
    q[u][1] = p4.y;
    q[u][2] = p4.z;
    q[u][3] = p4.w;
  }
  float v[kReduceLors][4];
#pragma unroll
  for (uint32_t u = 0; u < kReduceLors; ++u)
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) v[u][k] = q[u][k] != kNoRun ? partial[q[u][k] & ~kMoreRuns] : 0.0f;
  RStats rs;
#pragma unroll
  for (uint32_t u = 0; u < kReduceLors; ++u) {
    if (!on[u]) continue;
    float p = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (q[u][k] != kNoRun) p += v[u][k];
    if (q[u][3] != kNoRun && (q[u][3] & kMoreRuns)) {
      const uint32_t e = run_off[idx[u] + 1];
      for (uint32_t k = run_off[idx[u]] + 4; k < e; ++k) p += partial[run_pos[k]];
    }
    // MLEM: r = 1/(p*ac); sensitivity: p adapted to exp(-p), r = 1/exp(-p)
    // (projector.rs:67-74,144-149)
    const float rv = SENS ? 1.0f / expf(-p) : 1.0f / (p * ac[idx[u]]);
    r_out[idx[u]] = rv;
    if (q[u][0] != kNoRun) rs.note(rv);  // only LORs with runs backproject their r
  }
  if (rstats) block_r_stats(rstats, rs);
}

// pos4 of LORs 0..n-1 from run_off and run_pos.
__global__ void k_pack_pos4(const uint32_t* __restrict__ run_off, const uint32_t* __restrict__ run_pos, uint32_t n,
                            uint4* __restrict__ pos4) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = run_off[i], e = run_off[i + 1];
  uint32_t q[4];
  for (uint32_t k = 0; k < 4; ++k) q[k] = b + k < e ? run_pos[b + k] : kNoRun;
  if (e - b > 4) q[3] |= kMoreRuns;
  pos4[i] = make_uint4(q[0], q[1], q[2], q[3]);
}

// pos2 of LORs 0..n-1: the first four positions as 16-bit offsets into the
// LOR's window block (kNoRun16 where it has fewer; bit 15 of the fourth set
// when it has more).  Valid while every window holds under 2^15 - 1 entries.
constexpr uint32_t kNoRun16 = 0xffffu;
constexpr uint32_t kMoreRuns16 = 0x8000u;
constexpr uint32_t kMaxCompactWindow = 0x7fffu;
__global__ void k_pack_pos2(const uint32_t* __restrict__ run_off, const uint32_t* __restrict__ run_pos, uint32_t n,
                            const uint32_t* __restrict__ win, uint2* __restrict__ pos2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t ws = win[i >> kWinShift];
  const uint32_t b = run_off[i], e = run_off[i + 1];
  uint32_t q[4];
  for (uint32_t k = 0; k < 4; ++k) q[k] = b + k < e ? run_pos[b + k] - ws : kNoRun16;
  if (e - b > 4) q[3] |= kMoreRuns16;
  pos2[i] = make_uint2(q[0] | (q[1] << 16), q[2] | (q[3] << 16));
}

template <bool SENS>
__global__ __launch_bounds__(256) void k_lor_reduce(const float* __restrict__ ac, uint32_t n,
                                                    const uint32_t* __restrict__ run_off,
                                                    const uint32_t* __restrict__ run_pos,
                                                    const uint4* __restrict__ pos4,
                                                    const float* __restrict__ partial, float* __restrict__ r_out,
                                                    unsigned long long* rstats = nullptr) {
  const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3;
  const uint32_t share = (n + 7u) / 8u;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t first = slot * blockDim.x * kReduceLors + wave * 64u * kReduceLors + lane;
  const uint32_t base = xcd * share;
  const uint32_t end = base + share < n ? share : (base < n ? n - base : 0u);
  reduce_lors<SENS>(ac, base, first, end, run_off, run_pos, pos4, partial, r_out, rstats);
}

// The same reduction for the (window block, tile) partial layout: all
// partials of block b's LORs lie in [win[b], win[b + 1]).  One 1024-lane
// workgroup per block of 2^kWinShift LORs (kWinLors per lane: b * block + lane +
// 1024 * u) first issues its LORs' position and correction loads, then
// copies the block's window into LDS with coalesced loads (kTileLoads in
// flight per lane) and gathers from there: the partials cross the memory
// system once, in whole cache lines, instead of as one L2 request per
// partial.  A window over lds_entries entries is gathered from memory.
constexpr uint32_t kWinThreads = 1024;  // lanes of a k_lor_reduce_win workgroup
constexpr uint32_t kWinLors = (1u << kWinShift) / kWinThreads;
static_assert(kWinLors * kWinThreads == (1u << kWinShift), "LOR blocks of whole lanes");

// COMPACT: the first four positions come from pos2 as 16-bit offsets into
// the LOR's window (k_pack_pos2), else from pos4.
template <bool SENS, bool COMPACT>
__device__ __forceinline__ void reduce_window(const uint32_t b, const float* __restrict__ ac, uint32_t n,
                                              const uint32_t* __restrict__ win, uint32_t lds_entries,
                                              const uint32_t* __restrict__ run_off,
                                              const uint32_t* __restrict__ run_pos, const uint4* __restrict__ pos4,
                                              const uint2* __restrict__ pos2, const float* __restrict__ partial,
                                              float* __restrict__ r_out, unsigned long long* rstats) {
  extern __shared__ float wpart[];
  const uint32_t ws = win[b], size = win[b + 1] - ws;
  const uint32_t stride = blockDim.x;
  uint32_t q[kWinLors][4];
  float a[kWinLors];
#pragma unroll
  for (uint32_t u = 0; u < kWinLors; ++u) {
    const uint32_t i = (b << kWinShift) + threadIdx.x + u * stride;
    if (COMPACT) {
      const uint2 p2 = i < n ? pos2[i] : make_uint2(0xffffffffu, 0xffffffffu);
      const uint32_t h[4] = {p2.x & 0xffffu, p2.x >> 16, p2.y & 0xffffu, p2.y >> 16};
#pragma unroll
      for (uint32_t k = 0; k < 3; ++k) q[u][k] = h[k] == kNoRun16 ? kNoRun : ws + h[k];
      q[u][3] = h[3] == kNoRun16 ? kNoRun : ((ws + (h[3] & ~kMoreRuns16)) | ((h[3] & kMoreRuns16) ? kMoreRuns : 0u));
    } else {
      const uint4 p4 = i < n ? pos4[i] : make_uint4(kNoRun, kNoRun, kNoRun, kNoRun);
      q[u][0] = p4.x;
      q[u][1] = p4.y;
      q[u][2] = p4.z;
      q[u][3] = p4.w;
    }
    a[u] = (!SENS && i < n) ? ac[i] : 1.0f;
  }
  const bool staged = size <= lds_entries;  // the same for the whole workgroup
  if (staged) {
    for (uint32_t k0 = threadIdx.x; k0 < size; k0 += kTileLoads * stride) {
      float v[kTileLoads];
#pragma unroll
      for (uint32_t u = 0; u < (uint32_t)kTileLoads; ++u) {
        const uint32_t k = k0 + u * stride;
        v[u] = k < size ? partial[ws + k] : 0.0f;
      }
#pragma unroll
      for (uint32_t u = 0; u < (uint32_t)kTileLoads; ++u) {
        const uint32_t k = k0 + u * stride;
        if (k < size) wpart[k] = v[u];
      }
    }
    __syncthreads();
  }
  RStats rs;
#pragma unroll
  for (uint32_t u = 0; u < kWinLors; ++u) {
    const uint32_t i = (b << kWinShift) + threadIdx.x + u * stride;
    if (i >= n) continue;
    float p = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (q[u][k] != kNoRun) {
        const uint32_t at = q[u][k] & ~kMoreRuns;
        p += staged ? wpart[at - ws] : partial[at];
      }
    if (q[u][3] != kNoRun && (q[u][3] & kMoreRuns)) {
      const uint32_t e = run_off[i + 1];
      for (uint32_t k = run_off[i] + 4; k < e; ++k) p += staged ? wpart[run_pos[k] - ws] : partial[run_pos[k]];
    }
    // as reduce_lors (projector.rs:67-74,144-149)
    const float rv = SENS ? 1.0f / expf(-p) : 1.0f / (p * a[u]);
    r_out[i] = rv;
    if (q[u][0] != kNoRun) rs.note(rv);  // only LORs with runs backproject their r
  }
  if (rstats) block_r_stats(rstats, rs);
}

template <bool SENS, bool COMPACT>
__global__ __launch_bounds__(1024) void k_lor_reduce_win(const float* __restrict__ ac, uint32_t n,
                                                         const uint32_t* __restrict__ win, uint32_t lds_entries,
                                                         const uint32_t* __restrict__ run_off,
                                                         const uint32_t* __restrict__ run_pos,
                                                         const uint4* __restrict__ pos4, const uint2* __restrict__ pos2,
                                                         const float* __restrict__ partial, float* __restrict__ r_out,
                                                         unsigned long long* rstats) {
  reduce_window<SENS, COMPACT>(blockIdx.x, ac, n, win, lds_entries, run_off, run_pos, pos4, pos2, partial, r_out,
                               rstats);
}

// Per-reconstruction tables of the batched small kernels; blocks of one
// reconstruction are consecutive from block0.
struct ReduceDesc {
  const float* ac;
  const uint32_t* run_off;
  const uint32_t* run_pos;
  const uint4* pos4;
  const float* partial;
  float* r;
  const uint32_t* win;  // window blocks (k_lor_reduce_win_batch)
  const uint2* pos2;    // compact positions, or null for pos4
  unsigned long long* rstats;  // the reconstruction's r statistics (kRStatWords)
  uint32_t n;
  uint32_t block0;
};

struct UpdateDesc {
  float* x;
  double* bp;
  const float* s;
  unsigned long long* counters;
  unsigned long long* rstat;  // r statistics (kRStatWords)
  uint32_t nv;
  uint32_t block0;
};

// Index of the table entry whose block range holds block b (entries sorted
// by block0, the first at 0).
template <class D>
__device__ __forceinline__ uint32_t find_entry(const D* __restrict__ d, uint32_t n, uint32_t b) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (d[mid].block0 <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// Batched window reduction: one workgroup per window block of every
// reconstruction (blocks of one reconstruction consecutive from block0).
__global__ __launch_bounds__(1024) void k_lor_reduce_win_batch(const ReduceDesc* __restrict__ d, uint32_t n_rec,
                                                               uint32_t lds_entries) {
  const uint32_t r = find_entry(d, n_rec, blockIdx.x);
  const ReduceDesc q = d[r];
  if (q.pos2)
    reduce_window<false, true>(blockIdx.x - q.block0, q.ac, q.n, q.win, lds_entries, q.run_off, q.run_pos, q.pos4,
                               q.pos2, q.partial, q.r, q.rstats);
  else
    reduce_window<false, false>(blockIdx.x - q.block0, q.ac, q.n, q.win, lds_entries, q.run_off, q.run_pos, q.pos4,
                                q.pos2, q.partial, q.r, q.rstats);
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
// r of every run when none is given (never written; a global, so that
// loads through either table stay global loads).
__device__ float kOneR[1] = {1.0f};

typedef __attribute__((address_space(3))) unsigned long long lds_u64;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

// The backprojection tile's accumulator, the same for all lanes of a work
// item.  Fixed point (unit > 0): every contribution v = w * r is scaled by a
// power of two S and rounded to an integer -- in binary64, as the low 52 bits
// of the encoding of v * S + 2^52 (one fused multiply-add; the sum's
// exponent is fixed while v * S < 2^51) -- and added to a 64-bit integer with
// ds_add_u64, which the LDS executes at about twice the rate of ds_add_f64
// under a replay's scattered addresses (measured on gfx950; ds_add_f32 is an
// order of magnitude slower still).  S is the largest power of two that keeps
// the largest possible contribution -- the update's largest r, times the
// Gaussian's peak factor, times 1.25 voxel diagonals (no step inside a voxel
// is longer than its diagonal; the margin covers the boundaries' rounding) --
// below 2^t with t = min(51, 64 - ceil(log2(runs of the item))): a run adds
// at most once to a voxel, so no tile sum reaches 2^64.  Integer sums are
// exact and independent of order; a contribution's rounding to the unit is
// at most 2^-25 of a typical contribution (the guard below), less than its
// own binary32 rounding.  Binary64 sums (unit == 0) where the scale does not
// exist or could be too coarse: some r negative or not finite (a LOR with
// p = 0, projector.rs:145), no positive r, or a largest r more than 2^(t-26)
// times the geometric mean r (a few LORs of tiny projection would set a unit
// too coarse for the others; two bits of the margin cover the step lengths
// and Gaussian factors below their bound).  MLEM's LORs through voxels that
// have decayed towards zero reach about 2^20 times the geometric mean r on the
// 120^3 phantom workload; work items of up to 2^15 runs allow 2^23.
static_assert(kRSlots == 64, "one statistics slot per lane of a wavefront");

struct TileAcc {
  double s;     // the scale S (fixed point)
  double unit;  // 1 / S; 0 for binary64 sums
};

// The accumulator of a work item of n_item runs, from the statistics slots
// at rstats (kRStatWords words; ignored with unit_r, where every r is 1).
// Called by the first wavefront of the workgroup (all 64 lanes).
__device__ __forceinline__ TileAcc tile_acc(const unsigned long long* rstats, bool unit_r, uint32_t n_item,
                                            const DevFov& f, float amp) {
  const TileAcc f64{0.0, 0.0};
  float rmax = 1.0f;
  double lmean = 0.0;  // mean log2(r)
  if (!unit_r) {
    if (!rstats) return f64;
    const uint32_t lane = threadIdx.x & 63u;
    RStats rs;
    const unsigned long long* slot = rstats + 4 * lane;  // kRSlots == 64 slots, one per lane
    rs.rmax = __uint_as_float((uint32_t)slot[0]);
    rs.bad = slot[1] != 0;
    rs.lsum = __longlong_as_double((long long)slot[2]);
    rs.n = (uint32_t)slot[3];
    rs.wave_combine();
    if (rs.bad || rs.n == 0) return f64;
    rmax = rs.rmax;
    lmean = rs.lsum / (double)rs.n;
  }
  int lg = 0;
  while (lg < 32 && (1u << lg) < n_item) ++lg;  // n_item <= 2^lg
  const int t = min(51, 64 - lg);
  const float bound = rmax * amp * (1.25f * sqrtf(f.vs_x * f.vs_x + f.vs_y * f.vs_y + f.vs_z * f.vs_z));
  if (!(bound > 0.0f) || !(bound <= 3.402823466e38f) || !((double)__log2f(rmax) <= lmean + (double)(t - 26)))
    return f64;
  int k;
  frexpf(bound, &k);  // bound < 2^k
  const int e = t - k;
  if (e > 1000 || e < -1000) return f64;
  return TileAcc{ldexp(1.0, e), ldexp(1.0, -e)};
}

// round(v * s) for 0 <= v * s < 2^51, as a 64-bit integer
__device__ __forceinline__ unsigned long long fixed_units(float v, double s) {
  const double d = fma((double)v, s, 4503599627370496.0);  // 2^52 + v * s, rounded to an integer
  return (unsigned long long)__double_as_longlong(d) & 0x000fffffffffffffull;
}

// A lane's runs j, j + stride, ... <= last of a backprojection work item,
// added to the LDS tile with r = r_src[lor & r_mask].  Records are loaded
// two runs ahead and r one run ahead, so no load's latency stalls a replay;
// the first records are waited for before the loop, as in fwd_runs.
// PERM: the runs are read through perm.  ACC: the tile holds binary64 sums
// (kAccF64), 64-bit fixed point (kAccU64, TileAcc), where every r is finite,
// so zero-length steps add 0, or 32-bit fixed point (kAccU32, unit r: the
// scale sfix keeps a work item's sums below 2^32).
constexpr int kAccF64 = 0, kAccU64 = 1, kAccU32 = 2;
template <bool TOF, bool PERM, int ACC>
__device__ __forceinline__ void scatter_runs(const Run* __restrict__ runs, const uint32_t* __restrict__ perm,
                                             int32_t j, const int32_t last, const float* __restrict__ r_src,
                                             const uint32_t r_mask, const double sfix, const DevTof& t,
                                             const TileGeom& g, void* tile) {
  constexpr bool FIXED = ACC != kAccF64;
  const float sf32 = (float)sfix;
  const int32_t stride = (int32_t)blockDim.x;
  const int32_t j1 = j + stride;
  const int32_t j1c = j1 < last ? j1 : last;
  Run cur = runs[PERM ? perm[j] : j];
  Run nxt = runs[PERM ? perm[j1c] : j1c];
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  float r = r_src[cur.lor & r_mask];
  const uint32_t base = lds_addr(tile);
  auto add = [&](uint32_t at, float v) {
    if (ACC == kAccU32)  // v * sf32 is exact (a power of two), below 2^32
      __hip_atomic_fetch_add((lds_u32*)(uintptr_t)at, __float2uint_rn(v * sf32), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    else if (ACC == kAccU64)
      __hip_atomic_fetch_add((lds_u64*)(uintptr_t)at, fixed_units(v, sfix), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      __hip_atomic_fetch_add((lds_f64*)(uintptr_t)at, (double)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  for (;;) {
    const int32_t jn = j + stride;
    const int32_t jnn = jn + stride;
    const float rn = r_src[nxt.lor & r_mask];
    const int32_t jnc = jnn < last ? jnn : last;
    const Run nn = runs[PERM ? perm[jnc] : jnc];
    // without a non-finite r in the wavefront, zero-length steps may add +0
    const float rr = TOF ? r * t.amp : r;
    constexpr int kShift = ACC == kAccU32 ? 2 : 3;
    if (FIXED || __all(fabsf(r) <= 3.402823466e38f))
      replay_run<TOF, false, kShift>(cur, rr, t, g, base, add);
    else
      replay_run<TOF, true, kShift>(cur, rr, t, g, base, add);
    if (jn > last) break;
    cur = nxt;
    nxt = nn;
    r = rn;
    j = jn;
  }
}

// r_in == nullptr backprojects the weights themselves (r = 1).  With perm,
// sorted run j is runs[perm[j]] (runs left in LOR-major order).  rstats (the
// update's r statistics, kRStatWords words) chooses the accumulator
// (tile_acc); the first workgroup records its unit in word kRUnitWord.
template <bool TOF>
__device__ __forceinline__ void scatter_body(const Run* __restrict__ runs, const int4 wi,
                                             const float* __restrict__ r_in, const DevFov& f, const DevTof& t,
                                             const TileGeom& g, double* __restrict__ bp,
                                             const uint32_t* __restrict__ perm, unsigned long long* rstats) {
  extern __shared__ double tile[];
  __shared__ TileAcc s_acc;
  const int32_t tl = wi.x;
  if (threadIdx.x < 64) {
    const TileAcc a = tile_acc(rstats, r_in == nullptr, (uint32_t)(wi.z - wi.y), f, TOF ? t.amp : 1.0f);
    if (threadIdx.x == 0) {
      s_acc = a;
      if (rstats && blockIdx.x == 0) rstats[kRUnitWord] = (unsigned long long)__double_as_longlong(a.unit);
    }
  }
  // +0.0 and the integer 0 have the same bits
  for (int32_t k = threadIdx.x; k < g.tvox; k += blockDim.x) tile[k] = 0.0;
  __syncthreads();
  const TileAcc acc = s_acc;
  const int32_t j = wi.y + (int32_t)threadIdx.x;
  if (j < wi.z) {
    // without r every run reads the one entry of kOneR
    const float* r_src = r_in ? r_in : kOneR;
    const uint32_t r_mask = r_in ? 0xffffffffu : 0u;
    if (acc.unit > 0.0) {
      if (perm)
        scatter_runs<TOF, true, kAccU64>(runs, perm, j, wi.z - 1, r_src, r_mask, acc.s, t, g, tile);
      else
        scatter_runs<TOF, false, kAccU64>(runs, perm, j, wi.z - 1, r_src, r_mask, acc.s, t, g, tile);
    } else {
      if (perm)
        scatter_runs<TOF, true, kAccF64>(runs, perm, j, wi.z - 1, r_src, r_mask, 0.0, t, g, tile);
      else
        scatter_runs<TOF, false, kAccF64>(runs, perm, j, wi.z - 1, r_src, r_mask, 0.0, t, g, tile);
    }
  }
  __syncthreads();
  // flush the tile: consecutive lanes take consecutive x, so each wavefront
  // atomic touches few cache lines
  const int32_t tzi = tl / (g.ntx * g.nty);
  const int32_t tyi = (tl / g.ntx) % g.nty;
  const int32_t txi = tl % g.ntx;
  const int32_t x0 = txi * g.tx, y0 = tyi * g.ty, z0 = tzi * g.tz;
  const unsigned long long* tile_u = reinterpret_cast<const unsigned long long*>(tile);
  for (TileCursor cu(g, (int32_t)threadIdx.x, (int32_t)blockDim.x); cu.k < g.tvox; cu.next(g, (int32_t)blockDim.x)) {
    // a fixed-point sum (below 2^62) converts to within 2^-53
    const double v = acc.unit > 0.0 ? (double)tile_u[cu.k] * acc.unit : tile[cu.k];
    const int32_t x = x0 + cu.lx, y = y0 + cu.ly, z = z0 + cu.lz;
    if (v != 0.0 && cu.lx < g.tx && x < f.nx && y < f.ny && z < f.nz) unsafeAtomicAdd(bp + x + f.nx * (y + f.ny * z), v);
  }
}

template <bool TOF>
__global__ __launch_bounds__(1024) void k_scatter(const Run* __restrict__ runs, const int4* __restrict__ work,
                                                  const float* __restrict__ r_in, DevFov f, DevTof t, TileGeom g,
                                                  double* __restrict__ bp, const uint32_t* __restrict__ perm,
                                                  unsigned long long* rstats) {
  const int4 wi = work[blockIdx.x];
  if (wi.z <= wi.y) return;  // an empty item (the grid may be an upper bound)
  scatter_body<TOF>(runs, wi, r_in, f, t, g, bp, perm, rstats);
}

// Backprojection of the weights themselves (r = 1, no TOF: the sensitivity
// image without attenuation) with the tile in 32-bit fixed point: half the
// LDS of the 64-bit tile, so two workgroups share a CU at tiles of up to
// 20416 entries.  Each step adds round(w * s) units; s = 2^e comes from the
// host with every item of at most n runs and every step of at most the
// voxel diagonal d: n * (1.25 d) * s < 2^32, so no sum overflows.  A step's
// rounding is at most half a unit (unit = 1 / s, about 2^-16 of d for
// 16384-run items), which the flush's binary64 sum then carries exactly.
__global__ __launch_bounds__(1024) void k_scatter_unit32(const Run* __restrict__ runs, const int4* __restrict__ work,
                                                         DevFov f, TileGeom g, double* __restrict__ bp,
                                                         const uint32_t* __restrict__ perm, float s, double unit) {
  extern __shared__ uint32_t tile32[];
  const int4 wi = work[blockIdx.x];
  if (wi.z <= wi.y) return;  // an empty item (the grid may be an upper bound)
  for (int32_t k = threadIdx.x; k < g.tvox; k += blockDim.x) tile32[k] = 0u;
  __syncthreads();
  const int32_t j = wi.y + (int32_t)threadIdx.x;
  const DevTof notof = no_tof();
  if (j < wi.z) {
    if (perm)
      scatter_runs<false, true, kAccU32>(runs, perm, j, wi.z - 1, kOneR, 0u, (double)s, notof, g, tile32);
    else
      scatter_runs<false, false, kAccU32>(runs, perm, j, wi.z - 1, kOneR, 0u, (double)s, notof, g, tile32);
  }
  __syncthreads();
  const int32_t tl = wi.x;
  const int32_t tzi = tl / (g.ntx * g.nty);
  const int32_t tyi = (tl / g.ntx) % g.nty;
  const int32_t txi = tl % g.ntx;
  const int32_t x0 = txi * g.tx, y0 = tyi * g.ty, z0 = tzi * g.tz;
  for (TileCursor cu(g, (int32_t)threadIdx.x, (int32_t)blockDim.x); cu.k < g.tvox; cu.next(g, (int32_t)blockDim.x)) {
    const uint32_t u = tile32[cu.k];
    const int32_t x = x0 + cu.lx, y = y0 + cu.ly, z = z0 + cu.lz;
    if (u != 0u && cu.lx < g.tx && x < f.nx && y < f.ny && z < f.nz)
      unsafeAtomicAdd(bp + x + f.nx * (y + f.ny * z), (double)u * unit);
  }
}

// Batched backprojection: work item w = (tile, first run, end run, reconstruction).
template <bool TOF>
__global__ __launch_bounds__(1024) void k_scatter_batch(const ReconDesc* __restrict__ rd,
                                                        const int4* __restrict__ work) {
  const int4 wi = work[blockIdx.x];
  const ReconDesc& r = rd[wi.w];
  scatter_body<TOF>(r.runs, wi, r.r, r.f, r.t, r.g, r.bp, nullptr, r.rstats);
}

// Batched MLEM update (as k_update_f64) over every reconstruction's image.
__global__ __launch_bounds__(256) void k_update_batch(const UpdateDesc* __restrict__ d, uint32_t n_rec) {
  const uint32_t r = find_entry(d, n_rec, blockIdx.x);
  const UpdateDesc q = d[r];
  const uint32_t i = (blockIdx.x - q.block0) * blockDim.x + threadIdx.x;
  if (i >= q.nv) return;
  const float b = (float)q.bp[i];
  q.bp[i] = 0.0;
  const float sj = q.s ? q.s[i] : 1.0f;
  q.x[i] = sj > 0.0f ? q.x[i] * (b * sj) : 0.0f;
}

// Zeroes the work counters and r statistics of every reconstruction of a batch.
__global__ void k_zero_counters_batch(const UpdateDesc* __restrict__ d, uint32_t n_rec, uint32_t n_cnt) {
  for (uint32_t k = threadIdx.x; k < n_rec * n_cnt; k += blockDim.x) d[k / n_cnt].counters[k % n_cnt] = 0ull;
  for (uint32_t k = threadIdx.x; k < n_rec * kRStatWords; k += blockDim.x)
    d[k / kRStatWords].rstat[k % kRStatWords] = 0ull;
}

// ---------------------------------------------------------------------------
// Persistent schedule of a batch's subset update: one launch whose
// workgroups take tasks from an atomic queue in order -- every
// reconstruction's forward items, then its reduction blocks, its
// backprojection items and its update chunks -- each task of a later phase
// waiting until all tasks of the previous phase of its reconstruction have
// completed.  A task only waits for tasks earlier in the queue, which
// workgroups that are running have taken, so the schedule cannot deadlock
// whatever the number of resident workgroups.  Completion is published by
// every wavefront waiting for its own memory writes, a workgroup barrier,
// then one agent-scope release fence and atomic increment; a waiting task
// reads the counter with agent-scope loads and passes an agent-scope fence
// before its workgroup reads the data.  Every wait is bounded in time: past
// kSpinLimitNs the error word is set (sticky for the host) and waits stop,
// so a launch always ends.
//
// Measured on the 32-reconstruction workload it takes 28.8 ms per full
// iteration against 15.0 ms for one launch per phase: every task's
// agent-scope release and acquire write back and invalidate its XCD's L2,
// which the other workgroups there then miss (without the fences, which
// makes it incorrect, 16.4 ms; taking 4, 8 or 16 tasks per queue access is
// slower still, 31.8-50.6 ms, by load imbalance).  It is therefore an
// option (petalo_batch_set_schedule), not the default.
//
// Eight wavefronts per SIMD (__launch_bounds__(1024, 8): at most 64 VGPRs),
// so that two workgroups share a CU as the phase kernels do.
//
// Task encoding (int4): x = kind (bits 0-1: 0 forward, 1 reduction,
// 2 backprojection, 3 update) | TOF (bit 2) | reconstruction << 3;
// forward and backprojection: y, z, w = tile, first run, end run;
// reduction: y = window block; update: y, z = voxel range.
constexpr uint32_t kTaskForward = 0, kTaskReduce = 1, kTaskScatter = 2, kTaskUpdate = 3;
// sync words: [0] queue head, [1] error (this update), [2] error (sticky),
// [3 + 3 r + k] completed tasks of phase k of reconstruction r
constexpr uint32_t kSyncHead = 0, kSyncError = 1, kSyncSticky = 2, kSyncDone = 3;
constexpr uint64_t kSpinLimitNs = 200000000ull;  // 0.2 s; an update takes about a millisecond

__device__ __forceinline__ uint64_t wall_ns() {
  return (uint64_t)__builtin_amdgcn_s_memrealtime() * 10ull;  // 100 MHz constant clock
}

// Thread 0: waits until *ctr >= need (or an error), then fences (acquire).
// Bounded twice: by the constant clock and by a count of sleeps (about a
// second at 64 cycles each), should that clock not advance.
constexpr uint32_t kSpinLimitIters = 1u << 25;
__device__ __forceinline__ void task_wait(const uint32_t* ctr, uint32_t need, uint32_t* sync) {
  const uint64_t t0 = wall_ns();
  for (uint32_t it = 0;; ++it) {
    if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need) break;
    if (__hip_atomic_load(&sync[kSyncError], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
    if (wall_ns() - t0 > kSpinLimitNs || it >= kSpinLimitIters) {
      atomicOr(&sync[kSyncError], 1u);
      atomicOr(&sync[kSyncSticky], 1u);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  __threadfence();
}

__global__ __launch_bounds__(1024, 8) void k_batch_persistent(const int4* __restrict__ tasks, uint32_t n_tasks,
                                                              const ReconDesc* __restrict__ rd,
                                                              const ReduceDesc* __restrict__ red,
                                                              const UpdateDesc* __restrict__ ud,
                                                              const uint32_t* __restrict__ need, uint32_t* sync,
                                                              uint32_t lds_entries) {
  __shared__ uint32_t s_task;
  for (;;) {
    if (threadIdx.x == 0) s_task = atomicAdd(&sync[kSyncHead], 1u);
    __syncthreads();
    const uint32_t t = s_task;
    __syncthreads();  // every lane has read s_task before it is taken again
    if (t >= n_tasks) break;
    const int4 tk = tasks[t];
    const uint32_t kind = (uint32_t)tk.x & 3u, r = (uint32_t)tk.x >> 3;
    const bool tof = ((uint32_t)tk.x & 4u) != 0u;
    if (kind != kTaskForward) {
      if (threadIdx.x == 0) task_wait(&sync[kSyncDone + 3 * r + (kind - 1)], need[3 * r + (kind - 1)], sync);
      __syncthreads();
    }
    if (kind == kTaskForward || kind == kTaskScatter) {
      const ReconDesc& d = rd[r];
      const int4 wi = make_int4(tk.y, tk.z, tk.w, (int)r);
      if (kind == kTaskForward) {
        if (tof)
          fwd_tile_body<true>(d.runs, wi, d.image, d.f, d.t, d.g, d.partial, nullptr, d.lpos);
        else
          fwd_tile_body<false>(d.runs, wi, d.image, d.f, d.t, d.g, d.partial, nullptr, d.lpos);
      } else {
        if (tof)
          scatter_body<true>(d.runs, wi, d.r, d.f, d.t, d.g, d.bp, nullptr, d.rstats);
        else
          scatter_body<false>(d.runs, wi, d.r, d.f, d.t, d.g, d.bp, nullptr, d.rstats);
      }
    } else if (kind == kTaskReduce) {
      const ReduceDesc q = red[r];
      if (q.pos2)
        reduce_window<false, true>((uint32_t)tk.y, q.ac, q.n, q.win, lds_entries, q.run_off, q.run_pos, q.pos4,
                                   q.pos2, q.partial, q.r, q.rstats);
      else
        reduce_window<false, false>((uint32_t)tk.y, q.ac, q.n, q.win, lds_entries, q.run_off, q.run_pos, q.pos4,
                                    q.pos2, q.partial, q.r, q.rstats);
    } else {
      const UpdateDesc q = ud[r];
      for (uint32_t i = (uint32_t)tk.y + threadIdx.x; i < (uint32_t)tk.z; i += blockDim.x) {
        const float b = (float)q.bp[i];
        q.bp[i] = 0.0;
        const float sj = q.s ? q.s[i] : 1.0f;
        q.x[i] = sj > 0.0f ? q.x[i] * (b * sj) : 0.0f;
      }
    }
    // publish: every wavefront's writes done, then one release and count
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();
    if (threadIdx.x == 0 && kind != kTaskUpdate) {
      __threadfence();
      atomicAdd(&sync[kSyncDone + 3 * r + kind], 1u);
    }
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
