// Seeded scene and request-stream generation (host side).
//
// The reference has no RNG and no scenes (nothing in src/ draws random numbers);
// these generators define the workloads the tick is measured on.  Sequential
// draws use SplitMix64 (state += golden gamma; output = SplitMix64 finaliser);
// per-cell draws are counter based (the i-th SplitMix64 output of a stream
// keyed by the seed), so device and host produce identical grids.
#include <cmath>
#include <cstring>

#include "fs3d_internal.h"

namespace fs3d {

u64 mix64(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

u64 sm64_next(u64 *state) {
  *state += kGamma;
  return mix64(*state);
}

u64 mulhi64(u64 a, u64 b) { return (u64)(((unsigned __int128)a * b) >> 64); }

u64 uniform_incl(u64 r, u64 lo, u64 hi) { return lo + mulhi64(r, hi - lo + 1); }

u64 prob_threshold(u64 num, u64 den) { return (u64)(((unsigned __int128)num << 64) / den); }

// Discrete log-uniform edge length: P(e) proportional to floor(2^32 / e).
int loguniform_int(u64 r, int lo, int hi) {
  u64 total = 0;
  for (int e = lo; e <= hi; e++) total += (1ull << 32) / (u64)e;
  const u64 u = mulhi64(r, total);
  u64 cum = 0;
  for (int e = lo; e <= hi; e++) {
    cum += (1ull << 32) / (u64)e;
    if (u < cum) return e;
  }
  return hi;
}

// Boxes over a STATIC floor, painted in generation order until SAND, FLUID and
// STATIC each cover floor(N * pct / 100) cells; the material with the largest
// remaining deficit is painted next (ties in the order SAND, FLUID, STATIC).
long long paint_boxes(u8 *g, int nx, int ny, int nz, u64 seed, const fs3d_box_params &p) {
  if (nz < 2 || p.side_min < 1 || p.side_min > nx || p.side_min > ny || p.side_min > nz - 1 ||
      p.side_max < p.side_min || p.pct_sand < 0 || p.pct_fluid < 0 || p.pct_static < 0)
    return -1;
  const size_t layer = (size_t)nx * ny, N = layer * (size_t)nz;
  std::memset(g, FS3D_EMPTY, N);
  std::memset(g, FS3D_STATIC, layer);
  long long count[4] = {(long long)(N - layer), (long long)layer, 0, 0};
  const u8 order[3] = {FS3D_SAND, FS3D_FLUID, FS3D_STATIC};
  const long long goal[3] = {(long long)(N * (u64)p.pct_sand / 100u),
                             (long long)(N * (u64)p.pct_fluid / 100u),
                             (long long)(N * (u64)p.pct_static / 100u)};
  const int cap_x = std::min(p.side_max, nx), cap_y = std::min(p.side_max, ny),
            cap_z = std::min(p.side_max, nz - 1);
  u64 st = seed;
  long long boxes = 0;
  for (; boxes < p.max_boxes; boxes++) {
    int pick = -1;
    long long deficit = 0;
    for (int m = 0; m < 3; m++) {
      const long long d = goal[m] - count[order[m]];
      if (d > deficit) deficit = d, pick = m;
    }
    if (pick < 0) break;
    const u8 mat = order[pick];
    const int sx = (int)uniform_incl(sm64_next(&st), p.side_min, cap_x);
    const int sy = (int)uniform_incl(sm64_next(&st), p.side_min, cap_y);
    const int sz = (int)uniform_incl(sm64_next(&st), p.side_min, cap_z);
    const int x0 = (int)uniform_incl(sm64_next(&st), 0, nx - sx);
    const int y0 = (int)uniform_incl(sm64_next(&st), 0, ny - sy);
    const int z0 = (int)uniform_incl(sm64_next(&st), 1, nz - sz);
    for (int z = z0; z < z0 + sz; z++)
      for (int y = y0; y < y0 + sy; y++) {
        u8 *row = g + ((size_t)z * ny + y) * nx + x0;
        for (int x = 0; x < sx; x++) count[row[x]]--;
        std::memset(row, mat, (size_t)sx);
      }
    count[mat] += (long long)sx * sy * sz;
  }
  return boxes;
}

// Pool box then pillars, drawn from one SplitMix64 stream (draw order: pool W, D,
// H, x0, y0; per pillar cx, cy, r, h).
int pillar_list(int nx, int ny, int nz, u64 seed, const fs3d_pillar_params &p, int pool[5],
                std::vector<Pillar> *out) {
  if (nz < 2 || p.pool_w_min < 1 || p.pool_d_min < 1 || p.pool_w_max > nx || p.pool_d_max > ny ||
      p.pool_w_max < p.pool_w_min || p.pool_d_max < p.pool_d_min || p.pool_h_max < p.pool_h_min ||
      p.r_min < 0 || p.r_max < p.r_min || p.h_max < p.h_min || p.n_pillars < 0)
    return -1;
  u64 st = seed;
  const int W = (int)uniform_incl(sm64_next(&st), p.pool_w_min, p.pool_w_max);
  const int D = (int)uniform_incl(sm64_next(&st), p.pool_d_min, p.pool_d_max);
  const int H = (int)uniform_incl(sm64_next(&st), p.pool_h_min, p.pool_h_max);
  pool[0] = (int)uniform_incl(sm64_next(&st), 0, nx - W);
  pool[1] = (int)uniform_incl(sm64_next(&st), 0, ny - D);
  pool[2] = W;
  pool[3] = D;
  pool[4] = std::min(H, nz - 1);
  out->clear();
  for (int k = 0; k < p.n_pillars; k++) {
    Pillar q;
    q.cx = (int)uniform_incl(sm64_next(&st), 0, nx - 1);
    q.cy = (int)uniform_incl(sm64_next(&st), 0, ny - 1);
    q.r = (int)uniform_incl(sm64_next(&st), p.r_min, p.r_max);
    const int h = (int)uniform_incl(sm64_next(&st), p.h_min, p.h_max);
    q.zt = std::min(h, nz - 1);
    out->push_back(q);
  }
  return 0;
}

}  // namespace fs3d

using namespace fs3d;

extern "C" int fs3d_threshold(uint64_t num, uint64_t den, uint64_t *thr) {
  if (!thr || den == 0 || num > den) return set_error(FS3D_ERR_INVALID, "threshold needs 0 <= num <= den, den > 0");
  *thr = num == den ? ~0ull : prob_threshold(num, den);
  return clear_error();
}

// Per world: three edges log-uniform in [lo, hi], then the sand threshold in
// [2^64/20, 2^64/2) and the world seed.
extern "C" int fs3d_batch_specs(int32_t n, uint64_t seed, int32_t lo, int32_t hi, int32_t *dims,
                                uint64_t *thr, uint64_t *seeds) {
  if (n < 0 || lo < 1 || hi < lo || !dims || !thr || !seeds)
    return set_error(FS3D_ERR_INVALID, "batch_specs: bad arguments");
  const u64 tlo = prob_threshold(1, 20), thi = prob_threshold(1, 2);
  u64 st = seed;
  for (int w = 0; w < n; w++) {
    for (int a = 0; a < 3; a++) dims[3 * w + a] = loguniform_int(sm64_next(&st), lo, hi);
    thr[w] = tlo + mulhi64(sm64_next(&st), thi - tlo);
    seeds[w] = sm64_next(&st);
  }
  return clear_error();
}

// Per request: exponential inter-arrival gap at rate lam (seconds), three edges,
// the tick count uniform in [tick_lo, tick_hi], the threshold and the seed.
extern "C" int fs3d_serve_specs(int32_t n, uint64_t seed, int32_t lo, int32_t hi, int32_t tick_lo,
                                int32_t tick_hi, double lam, int32_t *dims, uint64_t *thr,
                                uint64_t *seeds, int32_t *ticks, double *arrival) {
  if (n < 0 || lo < 1 || hi < lo || tick_lo < 0 || tick_hi < tick_lo || !(lam > 0) || !dims ||
      !thr || !seeds || !ticks || !arrival)
    return set_error(FS3D_ERR_INVALID, "serve_specs: bad arguments");
  const u64 tlo = prob_threshold(1, 20), thi = prob_threshold(1, 2);
  u64 st = seed;
  double now = 0.0;
  for (int w = 0; w < n; w++) {
    const double u = (double)(sm64_next(&st) >> 11) * (1.0 / 9007199254740992.0);
    now += -std::log1p(-u) / lam;
    arrival[w] = now;
    for (int a = 0; a < 3; a++) dims[3 * w + a] = loguniform_int(sm64_next(&st), lo, hi);
    ticks[w] = (int32_t)uniform_incl(sm64_next(&st), (u64)tick_lo, (u64)tick_hi);
    thr[w] = tlo + mulhi64(sm64_next(&st), thi - tlo);
    seeds[w] = sm64_next(&st);
  }
  return clear_error();
}
