// Host-side check that the fill pass's tile-jump binning (bin_lor_jump) emits
// exactly the runs and step counts of the step-by-step binning walk
// (bin_lor<false>), bit for bit, and that the count pass's tile jump
// (tile_runs) finds the same tiles and length classes, over random LORs and LORs built to tie: axis-parallel, on
// voxel planes, through voxel edges and corners, diagonal, degenerate
// (p1 == p2), grazing a face, and the irregular LORs of the sensitivity
// tests.  Usage: binning_axes_check [lors per case]; exit status 1 on any
// mismatch.
#include "tiles.hpp"
#include <cstdio>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>
using namespace petalo;
struct R { Run ru; uint32_t tile; };
static bool same(const R& a, const R& b) { return std::memcmp(&a.ru, &b.ru, sizeof(Run)) == 0 && a.tile == b.tile; }
int main(int argc, char** argv) {
  const int n_lors = argc > 1 ? atoi(argv[1]) : 200000;
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<float> U(-1.0f, 1.0f);
  long bad = 0, bad_tiles = 0, runs = 0, lors_hit = 0, fallback_like = 0;
  // repeat_add against m binary32 additions one by one: values and
  // increments across binades, increments with a half-ulp remainder (ties),
  // increments above the value, and long runs that cross several binades
  long bad_add = 0, adds = 0;
  {
    std::uniform_real_distribution<float> L(-30.0f, 12.0f), M(0.0f, 1.0f);
    for (int k = 0; k < 400000; ++k) {
      float v = std::ldexp(1.0f + M(rng), (int)L(rng));
      float d = std::ldexp(1.0f + M(rng), (int)L(rng));
      const int kind = k % 5;
      if (kind == 1) d = v * (0.001f + M(rng));                 // a fraction of v
      if (kind == 2) {                                           // d / ulp(v) ends in exactly 1/2
        int e;
        std::frexp(v, &e);
        const float u = std::ldexp(1.0f, e - 24);
        d = ((float)(1 + (int)(M(rng) * 5000)) + 0.5f) * u;
      }
      if (kind == 3) v = std::nextafter(std::ldexp(1.0f, (int)L(rng)), 0.0f);  // just below a power of two
      if (kind == 4) { v = 1e-7f * (1.0f + M(rng)); d = 0.5f + M(rng); }        // a walk's first boundary
      const uint32_t m = (uint32_t)(M(rng) * (k % 7 == 0 ? 5000.0f : 60.0f));
      float ref = v;
      for (uint32_t j = 0; j < m; ++j) ref = ref + d;
      const float got = repeat_add(v, d, m);
      ++adds;
      if (std::memcmp(&ref, &got, 4) != 0) {
        if (bad_add < 5) printf("REPEAT_ADD MISMATCH v %.9g d %.9g m %u: %.9g vs %.9g\n", v, d, m, ref, got);
        ++bad_add;
      }
    }
    const float inf = INFINITY;
    if (!(repeat_add(inf, 1.0f, 3) == inf && repeat_add(1.0f, inf, 2) == inf && repeat_add(2.5f, inf, 0) == 2.5f)) ++bad_add;
  }
  printf("repeat_add: %ld cases, repeat_add mismatches %ld\n", adds, bad_add);
  uint64_t steps_a = 0, steps_b = 0;
  long cls_hist[9] = {0}, bad_cls = 0;
  struct Case { float full[3]; uint32_t n[3]; int32_t tile[3]; };
  std::vector<Case> cases = {
    {{30, 30, 30}, {6, 6, 6}, {6, 6, 6}}, {{30, 30, 30}, {6, 6, 6}, {2, 3, 4}},
    {{180, 180, 180}, {120, 120, 120}, {40, 40, 6}}, {{180, 180, 180}, {60, 60, 60}, {4, 4, 4}},
    {{100, 130, 170}, {37, 51, 29}, {5, 7, 3}},       {{30, 30, 30}, {3, 3, 3}, {2, 2, 2}},
    {{307.2f, 307.2f, 153.6f}, {256, 256, 128}, {40, 40, 6}}, {{60, 60, 48}, {20, 20, 16}, {8, 16, 4}}};
  for (const Case& cs : cases) {
    const DevFov f = make_fov(cs.full, cs.n);
    TileGeom g{};
    g.tx = cs.tile[0]; g.ty = cs.tile[1]; g.tz = cs.tile[2]; g.px = g.tx | 1;
    g.ntx = (int32_t)((cs.n[0] + g.tx - 1) / g.tx); g.nty = (int32_t)((cs.n[1] + g.ty - 1) / g.ty);
    g.ntz = (int32_t)((cs.n[2] + g.tz - 1) / g.tz); g.ntiles = g.ntx * g.nty * g.ntz; g.tvox = g.px * g.ty * g.tz;
    const DevTof t = no_tof();
    const float vsx = cs.full[0] / cs.n[0], vsy = cs.full[1] / cs.n[1], vsz = cs.full[2] / cs.n[2];
    for (int i = 0; i < n_lors; ++i) {
      Lor l;
      const int kind = i % 8;
      float R0 = 1.2f * cs.full[0];
      l.x1 = U(rng) * R0; l.y1 = U(rng) * R0; l.z1 = U(rng) * cs.full[2];
      l.x2 = U(rng) * R0; l.y2 = U(rng) * R0; l.z2 = U(rng) * cs.full[2];
      if (kind == 1) l.y2 = l.y1;                      // parallel to a plane
      if (kind == 2) { l.z2 = l.z1; l.y2 = l.y1; }     // parallel to x
      if (kind == 3) {                                  // through voxel corners (ties)
        const int a = (int)((U(rng) + 1) * 5), b = (int)((U(rng) + 1) * 5), h = (int)((U(rng) + 1) * 3);
        l.x1 = -cs.full[0]; l.y1 = -cs.full[1] / 2 + a * vsy; l.z1 = -cs.full[2] / 2 + h * vsz;
        l.x2 = cs.full[0]; l.y2 = l.y1 + 2 * cs.full[0] / vsx * vsy * (b % 3 == 0 ? 1.0f : 0.5f); l.z2 = l.z1 + (b % 2) * 2 * cs.full[0];
      }
      if (kind == 4) {                                  // exact diagonals
        const float d = (U(rng) + 1) * 4;
        l.x1 = -cs.full[0]; l.y1 = -cs.full[0] + d * vsy; l.z1 = -cs.full[0] * vsz / vsx;
        l.x2 = cs.full[0]; l.y2 = cs.full[0] + d * vsy; l.z2 = cs.full[0] * vsz / vsx;
      }
      if (kind == 5) { l.x1 = (int)(U(rng) * 10) * vsx; l.x2 = l.x1; }  // on a voxel plane
      if (kind == 6 && (i % 64) == 6) { l.x2 = l.x1; l.y2 = l.y1; l.z2 = l.z1; }  // degenerate
      if (kind == 7 && (i % 64) == 7) { l.x1 = -cs.full[0]; l.x2 = cs.full[0]; l.y1 = l.y2 = cs.full[1] / 2; l.z1 = l.z2 = 0; }  // grazing a face
      if (i < 10) {  // the irregular LORs of the sensitivity test
        const float P1[10][3] = {{-100, 5, 0}, {5, 5, 5}, {-100, 0, 0}, {15.0f, -100, 1.0f}, {-100, -100, -100}, {-100, 2, 3},
                                 {15.0f, 0, 0}, {-15.0f, -100.0f, -15.0f}, {0, 0, -100.0f}, {-100, 50, 0}};
        const float P2[10][3] = {{100, 5, 0}, {5, 5, 5}, {100, 0, 0}, {15.0f, 100, 1.0f}, {100, 100, 100}, {100, -2, -3},
                                 {100.0f, 0, 0}, {-15.0f, 100.0f, -15.0f}, {0, 0, 100.0f}, {100, 50, 0}};
        l.x1 = P1[i][0]; l.y1 = P1[i][1]; l.z1 = P1[i][2]; l.x2 = P2[i][0]; l.y2 = P2[i][1]; l.z2 = P2[i][2];
      }
      l.dt = 0.0f; l.ac = 1.0f;
      std::vector<R> a, b;
      std::vector<uint32_t> cls_fill;
      uint32_t sa = 0, sb = 0;
      RunClasses rc;
      rc.n = 8;
      rc.scale = 8.0 / (double)(g.tx + g.ty + g.tz);
      bin_lor<false>(l, (uint32_t)i, f, t, g, [&](const Run& ru, uint32_t tl) { a.push_back({ru, tl}); }, &sa, nullptr);
      bin_lor_jump(l, (uint32_t)i, f, g, rc, [&](const Run& ru, uint32_t tl, uint32_t cl) {
        b.push_back({ru, tl});
        cls_fill.push_back(cl);
      }, &sb);
      WalkC c; WalkS w;
      const bool bnd = walk_setup(l, f, c, w) && bounded(w);
      if (!bnd) { sa = 0; }  // bin_lor counts nothing for unbounded walks either
      lors_hit += a.empty() ? 0 : 1;
      runs += (long)a.size();
      steps_a += sa; steps_b += sb;
      bool ok = a.size() == b.size() && sa == sb;
      for (size_t k = 0; ok && k < a.size(); ++k) ok = same(a[k], b[k]);
      // the tile jump: the same tiles, in walk order
      // the tile jump: the same tiles, in walk order, in the fill's length classes
      std::vector<uint32_t> tj, cls_count;
      const uint32_t nj = tile_runs(l, (uint32_t)i, f, g, rc, [&](uint32_t tl, uint32_t cl) {
        tj.push_back(tl);
        cls_count.push_back(cl);
      });
      bool tok = nj == tj.size() && tj.size() == a.size() && cls_count.size() == cls_fill.size();
      for (size_t k = 0; tok && k < a.size(); ++k) tok = tj[k] == a[k].tile && cls_count[k] == cls_fill[k];
      for (size_t k = 0; k < cls_count.size(); ++k) {
        ++cls_hist[cls_count[k] < 8 ? cls_count[k] : 8];
        // for LORs in general position (no tied crossings, which the class
        // counts as steps) the class is the run's length scaled, give or
        // take the estimate's error: the crossings of one axis over a
        // distance D number D / vl within one, so the run's length lies
        // within 3 steps of the estimate
        const double len = (double)run_len(a[k].ru.code), c = (double)cls_count[k];
        if (kind == 0 && i >= 10 && k < a.size() &&
            (c > (len + 3.0) * rc.scale + 1e-9 || (cls_count[k] < 7 && c + 1.0 < (len - 3.0) * rc.scale - 1e-9)))
          ++bad_cls;
      }
      if (!tok) {
        if (bad_tiles < 5) {
          printf("TILE MISMATCH case %ux%ux%u tile %dx%dx%d lor %d kind %d: runs %zu vs %zu\n", cs.n[0], cs.n[1], cs.n[2],
                 g.tx, g.ty, g.tz, i, kind, a.size(), tj.size());
          for (size_t k = 0; k < a.size() || k < tj.size(); ++k)
            printf("  step walk %d  jump %d\n", k < a.size() ? (int)a[k].tile : -1, k < tj.size() ? (int)tj[k] : -1);
        }
        ++bad_tiles;
        if (bad_tiles <= 5 && tj.size() == a.size())
          for (size_t k = 0; k < cls_count.size() && k < cls_fill.size(); ++k)
            printf("  class count %u fill %u (len %u)\n", cls_count[k], cls_fill[k], run_len(a[k].ru.code));
      }
      if (!ok) {
        if (bad < 5) {
          printf("MISMATCH case %ux%ux%u tile %dx%dx%d lor %d kind %d: runs %zu vs %zu steps %u vs %u\n", cs.n[0], cs.n[1], cs.n[2],
                 g.tx, g.ty, g.tz, i, kind, a.size(), b.size(), sa, sb);
          for (size_t k = 0; k < a.size() || k < b.size(); ++k) {
            if (k < a.size()) printf("  A %zu: tile %u len %u local %u nb %.9g %.9g %.9g here %.9g\n", k, a[k].tile, run_len(a[k].ru.code), run_local(a[k].ru.code), a[k].ru.nb_x, a[k].ru.nb_y, a[k].ru.nb_z, a[k].ru.here);
            if (k < b.size()) printf("  J %zu: tile %u len %u local %u nb %.9g %.9g %.9g here %.9g\n", k, b[k].tile, run_len(b[k].ru.code), run_local(b[k].ru.code), b[k].ru.nb_x, b[k].ru.nb_y, b[k].ru.nb_z, b[k].ru.here);
          }
        }
        ++bad;
      }
    }
  }
  printf("lors with runs %ld, runs %ld, steps %llu vs %llu, mismatches %ld, tile mismatches %ld\n", lors_hit, runs,
         (unsigned long long)steps_a, (unsigned long long)steps_b, bad, bad_tiles);
  printf("length classes:");
  for (int k = 0; k < 9; ++k) printf(" %ld", cls_hist[k]);
  printf(", class far from the run length %ld\n", bad_cls);
  return (bad || bad_tiles || bad_add || bad_cls) ? 1 : 0;
}
