/*
 * fs3d oracle (TEST INFRASTRUCTURE ONLY) -- CPU restatement of the fs3d tick.
 *
 * THE SPEC (project-authored; the reference snapshot has no simulation code --
 * its engine loop src/engine/engine.cpp:54-74 calls only Window::handleEvents and
 * Renderer::draw, and the tick would sit between them):
 *
 * Grid.  uint8 v[z][y][x], x fastest, linear index i = (z*ny + y)*nx + x, z up
 * (gravity toward -z).  Material bytes: 0 EMPTY, 1 STATIC, 2 SAND, 3 FLUID;
 * 4..255 are reserved, behave as STATIC and keep their byte value.  Cells
 * outside the grid behave as STATIC.
 *
 * Tick t (t = the world's tick counter before the step, parity p = t & 1).
 * The grid is partitioned into 2x2x2 blocks whose origin is shifted by
 * (-p,-p,-p): block (bx,by,bz) covers x in {2bx-p, 2bx-p+1} (same for y, z).
 * Block counts per axis: (n + p + 1) >> 1.  Inside a block, cell b = ox + 2*oy
 * + 4*oz (oz = 1 is the top layer); a block state packs the 2-bit categories as
 * s = sum cat_b << 2b.  Every block is updated independently from the previous
 * grid into a second grid (double buffering), so the tick has no serial order.
 *
 * Variant.  w = fmix32(g*0x9E3779B1 + by*0x85EBCA77 + bz*0xC2B2AE3D +
 * t*0x27D4EB2F) over uint32 with g = bx >> 4; variant = (w >> 2*(bx & 15)) & 3.
 * fmix32 is the MurmurHash3 32-bit finaliser.
 *
 * Block rule R(s, variant).  The four columns of a block (col = ox + 2*oy) are
 * walked on the cycle CYC = [0, 1, 3, 2] (cycle position -> column).  Variant v
 * gives direction d = +1 (v even) or -1 (v odd) and start position v.  Scan
 * order: positions v, v+d, v+2d, v+3d (mod 4).  A cell at cycle position k
 * prefers targets at positions k+d, k-d, k+2 (mod 4), in that order.  Each cell
 * takes part in at most one swap per tick ("moved" cells are skipped):
 *   1 fall   : in every column, top SAND/FLUID over bottom EMPTY swap, and top
 *              SAND over bottom FLUID swap;
 *   2 slide  : in scan order, each unmoved top SAND swaps with the first unmoved
 *              EMPTY or FLUID bottom cell of another column, in its preference
 *              order;
 *   3 topple : in scan order, each unmoved top FLUID swaps with the first unmoved
 *              EMPTY bottom cell of another column, in preference order;
 *   4 spread : bottom layer, then top layer: in scan order, each unmoved FLUID
 *              swaps with the first unmoved EMPTY cell of another column in the
 *              same layer, in preference order.
 * STATIC (and reserved) cells never move and nothing moves into them.
 *
 * Hash.  H = sum over cells with v_i != 0 of mix64((i << 8) | v_i) mod 2^64,
 * mix64 being the SplitMix64 finaliser.  Order-independent, so it can be summed
 * in any order.
 *
 * RNG.  Sequential streams are SplitMix64 (state += GAMMA; out = mix64(state)).
 * Per-cell draws are counter based: cell_rand(key, i) = mix64(key + (i+1)*GAMMA)
 * (the i-th SplitMix64 output of a stream seeded with key).  A probability
 * num/den is the integer threshold floor(num * 2^64 / den); a draw r hits when
 * r < threshold.  uniform(r, lo, hi) = lo + mulhi64(r, hi - lo + 1).
 */
#include "fs3d_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------- */
/* hashing / RNG                                                           */
/* ---------------------------------------------------------------------- */

/* SplitMix64 output function (Steele, Lea & Flood, OOPSLA 2014; Vigna's
 * public-domain splitmix64.c). */
uint64_t fso_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t fso_sm64_next(uint64_t *state) {
  *state += FSO_GAMMA;
  return fso_mix64(*state);
}

uint64_t fso_cell_rand(uint64_t key, uint64_t i) { return fso_mix64(key + (i + 1) * FSO_GAMMA); }

uint64_t fso_mulhi64(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * (unsigned __int128)b) >> 64);
}

uint64_t fso_uniform(uint64_t r, uint64_t lo, uint64_t hi) {
  return lo + fso_mulhi64(r, hi - lo + 1);
}

uint64_t fso_threshold(uint64_t num, uint64_t den) {
  return (uint64_t)((((unsigned __int128)num) << 64) / den);
}

/* MurmurHash3 fmix32 (Appleby, public domain). */
uint32_t fso_fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

uint32_t fso_group_word(uint32_t g, uint32_t by, uint32_t bz, uint32_t t) {
  return fso_fmix32(g * 0x9E3779B1u + by * 0x85EBCA77u + bz * 0xC2B2AE3Du + t * 0x27D4EB2Fu);
}

int fso_variant(uint32_t bx, uint32_t by, uint32_t bz, uint32_t t) {
  uint32_t w = fso_group_word(bx >> 4, by, bz, t);
  return (int)((w >> (2u * (bx & 15u))) & 3u);
}

/* Discrete log-uniform integer in [lo, hi]: P(e) proportional to
 * floor(2^32 / e), sampled by mulhi64(r, total) against the cumulative sums. */
int fso_loguniform(uint64_t r, int lo, int hi) {
  uint64_t total = 0;
  for (int e = lo; e <= hi; e++) total += (uint64_t)(4294967296ull / (uint64_t)e);
  uint64_t u = fso_mulhi64(r, total);
  uint64_t cum = 0;
  for (int e = lo; e <= hi; e++) {
    cum += (uint64_t)(4294967296ull / (uint64_t)e);
    if (u < cum) return e;
  }
  return hi;
}

/* ---------------------------------------------------------------------- */
/* block rule                                                              */
/* ---------------------------------------------------------------------- */

static const int CYC[4] = {0, 1, 3, 2};

static void swap_cells(uint8_t *c, int *moved, int a, int b) {
  uint8_t t = c[a];
  c[a] = c[b];
  c[b] = t;
  moved[a] = 1;
  moved[b] = 1;
}

void fso_block_rule(const uint8_t in[8], int v, uint8_t c[8]) {
  int moved[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < 8; b++) c[b] = in[b];
  const int d = (v & 1) ? 3 : 1; /* +1 or -1 on the 4-cycle */
  const int pref[3] = {d, 4 - d, 2};

  /* 1 fall */
  for (int col = 0; col < 4; col++) {
    int T = col + 4, B = col;
    int movable = (c[T] == FSO_SAND || c[T] == FSO_FLUID);
    if ((movable && c[B] == FSO_EMPTY) || (c[T] == FSO_SAND && c[B] == FSO_FLUID))
      swap_cells(c, moved, T, B);
  }
  /* 2 slide */
  for (int n = 0; n < 4; n++) {
    int k = (v + n * d) & 3;
    int T = CYC[k] + 4;
    if (c[T] != FSO_SAND || moved[T]) continue;
    for (int m = 0; m < 3; m++) {
      int B = CYC[(k + pref[m]) & 3];
      if (!moved[B] && (c[B] == FSO_EMPTY || c[B] == FSO_FLUID)) {
        swap_cells(c, moved, T, B);
        break;
      }
    }
  }
  /* 3 topple */
  for (int n = 0; n < 4; n++) {
    int k = (v + n * d) & 3;
    int T = CYC[k] + 4;
    if (c[T] != FSO_FLUID || moved[T]) continue;
    for (int m = 0; m < 3; m++) {
      int B = CYC[(k + pref[m]) & 3];
      if (!moved[B] && c[B] == FSO_EMPTY) {
        swap_cells(c, moved, T, B);
        break;
      }
    }
  }
  /* 4 spread: bottom layer, then top layer */
  for (int L = 0; L < 2; L++) {
    for (int n = 0; n < 4; n++) {
      int k = (v + n * d) & 3;
      int C = CYC[k] + 4 * L;
      if (c[C] != FSO_FLUID || moved[C]) continue;
      for (int m = 0; m < 3; m++) {
        int D = CYC[(k + pref[m]) & 3] + 4 * L;
        if (!moved[D] && c[D] == FSO_EMPTY) {
          swap_cells(c, moved, C, D);
          break;
        }
      }
    }
  }
}

uint16_t fso_rule_state(uint16_t s, int v) {
  uint8_t in[8], out[8];
  for (int b = 0; b < 8; b++) in[b] = (uint8_t)((s >> (2 * b)) & 3);
  fso_block_rule(in, v, out);
  uint16_t r = 0;
  for (int b = 0; b < 8; b++) r |= (uint16_t)(out[b] << (2 * b));
  return r;
}

void fso_build_lut(uint16_t *lut) {
  for (int v = 0; v < 4; v++)
    for (uint32_t s = 0; s < 65536u; s++) lut[v * 65536 + s] = fso_rule_state((uint16_t)s, v);
}

/* ---------------------------------------------------------------------- */
/* tick                                                                    */
/* ---------------------------------------------------------------------- */

static int g_threads = 0;

void fso_set_threads(int n) { g_threads = n; }

int fso_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

typedef struct {
  int nx, ny, nz, p;
  uint32_t t;
} tick_geom;

/* Gathers one block: raw bytes, categories (OOB -> STATIC) and in-bounds flags. */
static void gather_block(const uint8_t *src, const tick_geom *G, int bx, int by, int bz,
                         uint8_t raw[8], uint8_t cat[8], size_t idx[8], int inb[8]) {
  for (int b = 0; b < 8; b++) {
    int x = 2 * bx - G->p + (b & 1);
    int y = 2 * by - G->p + ((b >> 1) & 1);
    int z = 2 * bz - G->p + (b >> 2);
    inb[b] = (x >= 0 && x < G->nx && y >= 0 && y < G->ny && z >= 0 && z < G->nz);
    if (inb[b]) {
      idx[b] = ((size_t)z * (size_t)G->ny + (size_t)y) * (size_t)G->nx + (size_t)x;
      raw[b] = src[idx[b]];
      cat[b] = raw[b] <= 3 ? raw[b] : FSO_STATIC;
    } else {
      idx[b] = 0;
      raw[b] = FSO_STATIC;
      cat[b] = FSO_STATIC;
    }
  }
}

static void scatter_block(uint8_t *dst, const uint8_t raw[8], const uint8_t cat[8],
                          const uint8_t out[8], const size_t idx[8], const int inb[8]) {
  for (int b = 0; b < 8; b++)
    if (inb[b]) dst[idx[b]] = (cat[b] == FSO_STATIC) ? raw[b] : out[b];
}

void fso_tick(const uint8_t *src, uint8_t *dst, int nx, int ny, int nz, uint64_t t) {
  tick_geom G = {nx, ny, nz, (int)(t & 1), (uint32_t)t};
  const int nbx = (nx + G.p + 1) >> 1, nby = (ny + G.p + 1) >> 1, nbz = (nz + G.p + 1) >> 1;
  const long long nrows = (long long)nbz * nby;
#ifdef _OPENMP
  int nt = fso_get_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (long long row = 0; row < nrows; row++) {
    int bz = (int)(row / nby), by = (int)(row % nby);
    for (int bx = 0; bx < nbx; bx++) {
      uint8_t raw[8], cat[8], out[8];
      size_t idx[8];
      int inb[8];
      gather_block(src, &G, bx, by, bz, raw, cat, idx, inb);
      fso_block_rule(cat, fso_variant((uint32_t)bx, (uint32_t)by, (uint32_t)bz, G.t), out);
      scatter_block(dst, raw, cat, out, idx, inb);
    }
  }
}

void fso_tick_lut(const uint8_t *src, uint8_t *dst, int nx, int ny, int nz, uint64_t t,
                  const uint16_t *lut) {
  tick_geom G = {nx, ny, nz, (int)(t & 1), (uint32_t)t};
  const int nbx = (nx + G.p + 1) >> 1, nby = (ny + G.p + 1) >> 1, nbz = (nz + G.p + 1) >> 1;
  const long long nrows = (long long)nbz * nby;
#ifdef _OPENMP
  int nt = fso_get_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (long long row = 0; row < nrows; row++) {
    int bz = (int)(row / nby), by = (int)(row % nby);
      for (int bx = 0; bx < nbx; bx++) {
        uint8_t raw[8], cat[8], out[8];
        size_t idx[8];
        int inb[8];
        gather_block(src, &G, bx, by, bz, raw, cat, idx, inb);
        uint16_t s = 0;
        for (int b = 0; b < 8; b++) s |= (uint16_t)(cat[b] << (2 * b));
        int v = fso_variant((uint32_t)bx, (uint32_t)by, (uint32_t)bz, G.t);
        uint16_t o = lut[v * 65536 + s];
        for (int b = 0; b < 8; b++) out[b] = (uint8_t)((o >> (2 * b)) & 3);
        scatter_block(dst, raw, cat, out, idx, inb);
      }
  }
}

uint64_t fso_hash(const uint8_t *g, uint64_t n) {
  uint64_t h = 0;
#ifdef _OPENMP
  int nt = fso_get_threads();
#pragma omp parallel for schedule(static) reduction(+ : h) num_threads(nt)
#endif
  for (long long i = 0; i < (long long)n; i++)
    if (g[i]) h += fso_mix64(((uint64_t)i << 8) | g[i]);
  return h;
}

void fso_counts(const uint8_t *g, uint64_t n, uint64_t counts[256]) {
  memset(counts, 0, 256 * sizeof(uint64_t));
  for (uint64_t i = 0; i < n; i++) counts[g[i]]++;
}

int fso_run(uint8_t *grid, int nx, int ny, int nz, uint64_t t0, int nticks, int hash_every,
            uint64_t *hashes) {
  return fso_run_lut(grid, nx, ny, nz, t0, nticks, hash_every, hashes, NULL);
}

int fso_run_lut(uint8_t *grid, int nx, int ny, int nz, uint64_t t0, int nticks, int hash_every,
                uint64_t *hashes, const uint16_t *lut) {
  size_t n = (size_t)nx * ny * nz;
  uint8_t *tmp = (uint8_t *)malloc(n ? n : 1);
  if (!tmp) return -1;
  uint8_t *a = grid, *b = tmp;
  for (int k = 0; k < nticks; k++) {
    if (lut)
      fso_tick_lut(a, b, nx, ny, nz, t0 + (uint64_t)k, lut);
    else
      fso_tick(a, b, nx, ny, nz, t0 + (uint64_t)k);
    uint8_t *s = a;
    a = b;
    b = s;
    if (hash_every > 0 && (k + 1) % hash_every == 0) hashes[(k + 1) / hash_every - 1] = fso_hash(a, n);
  }
  if (a != grid) memcpy(grid, a, n);
  free(tmp);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* scenes                                                                  */
/* ---------------------------------------------------------------------- */

/* Rain fill: z = 0 is STATIC floor; every cell with z >= max(1, nz/2) is SAND
 * when cell_rand(seed, i) < thr, else EMPTY; all other cells EMPTY. */
void fso_scene_rain(uint8_t *g, int nx, int ny, int nz, uint64_t seed, uint64_t thr) {
  size_t layer = (size_t)nx * ny;
  int zr = nz >> 1;
  if (zr < 1) zr = 1;
  for (int z = 0; z < nz; z++) {
    uint8_t *L = g + (size_t)z * layer;
    if (z == 0) {
      memset(L, FSO_STATIC, layer);
    } else if (z < zr) {
      memset(L, FSO_EMPTY, layer);
    } else {
      size_t base = (size_t)z * layer;
      for (size_t j = 0; j < layer; j++) L[j] = fso_cell_rand(seed, base + j) < thr ? FSO_SAND : FSO_EMPTY;
    }
  }
}

/* Box fill: STATIC floor at z = 0, the rest EMPTY; then boxes are drawn from the
 * SplitMix64 stream seeded with `seed` and painted in generation order until the
 * cell counts of SAND, FLUID and STATIC all reach floor(N * pct / 100).  Before
 * each box the material with the largest remaining deficit is chosen (ties:
 * SAND, FLUID, STATIC).  Draw order per box: sx, sy, sz (uniform in
 * [side_min, min(side_max, n_axis)], z side capped at nz - 1), then x0 in
 * [0, nx - sx], y0 in [0, ny - sy], z0 in [1, nz - sz]. */
int64_t fso_scene_boxes(uint8_t *g, int nx, int ny, int nz, uint64_t seed, const fso_box_params *P) {
  if (nz < 2 || P->side_min < 1 || P->side_min > nx || P->side_min > ny || P->side_min > nz - 1)
    return -1;
  const size_t layer = (size_t)nx * ny, N = layer * (size_t)nz;
  memset(g, FSO_EMPTY, N);
  memset(g, FSO_STATIC, layer);
  int64_t cnt[4] = {(int64_t)(N - layer), (int64_t)layer, 0, 0};
  const int mats[3] = {FSO_SAND, FSO_FLUID, FSO_STATIC};
  const int pct[3] = {P->pct_sand, P->pct_fluid, P->pct_static};
  int64_t target[3];
  for (int m = 0; m < 3; m++) target[m] = (int64_t)((N * (uint64_t)pct[m]) / 100u);
  uint64_t st = seed;
  int64_t boxes = 0;
  const int smax_x = P->side_max < nx ? P->side_max : nx;
  const int smax_y = P->side_max < ny ? P->side_max : ny;
  const int smax_z = P->side_max < nz - 1 ? P->side_max : nz - 1;
  while (boxes < P->max_boxes) {
    int best = -1;
    int64_t bestd = 0;
    for (int m = 0; m < 3; m++) {
      int64_t d = target[m] - cnt[mats[m]];
      if (d > bestd) {
        bestd = d;
        best = m;
      }
    }
    if (best < 0) break;
    const uint8_t mat = (uint8_t)mats[best];
    int sx = (int)fso_uniform(fso_sm64_next(&st), P->side_min, smax_x);
    int sy = (int)fso_uniform(fso_sm64_next(&st), P->side_min, smax_y);
    int sz = (int)fso_uniform(fso_sm64_next(&st), P->side_min, smax_z);
    int x0 = (int)fso_uniform(fso_sm64_next(&st), 0, nx - sx);
    int y0 = (int)fso_uniform(fso_sm64_next(&st), 0, ny - sy);
    int z0 = (int)fso_uniform(fso_sm64_next(&st), 1, nz - sz);
    for (int z = z0; z < z0 + sz; z++)
      for (int y = y0; y < y0 + sy; y++) {
        uint8_t *row = g + ((size_t)z * ny + y) * nx;
        for (int x = x0; x < x0 + sx; x++) {
          cnt[row[x]]--;
          row[x] = mat;
        }
      }
    cnt[mat] += (int64_t)sx * sy * sz;
    boxes++;
  }
  return boxes;
}

/* Pillar fill: STATIC floor at z = 0; then one FLUID pool box (draw order W, D,
 * H, x0 in [0, nx-W], y0 in [0, ny-D]; cells x0<=x<x0+W, y0<=y<y0+D,
 * 1<=z<=min(H, nz-1)); then n_pillars SAND pillars (draw order cx in [0,nx-1],
 * cy in [0,ny-1], r, h; cells with (x-cx)^2 + (y-cy)^2 <= r^2 and
 * 1 <= z <= min(h, nz-1)), painted over the pool. */
int fso_scene_pillars(uint8_t *g, int nx, int ny, int nz, uint64_t seed, const fso_pillar_params *P) {
  if (nz < 2 || P->pool_w_max > nx || P->pool_d_max > ny || P->pool_w_min < 1 || P->pool_d_min < 1)
    return -1;
  const size_t layer = (size_t)nx * ny, N = layer * (size_t)nz;
  memset(g, FSO_EMPTY, N);
  memset(g, FSO_STATIC, layer);
  uint64_t st = seed;
  int W = (int)fso_uniform(fso_sm64_next(&st), P->pool_w_min, P->pool_w_max);
  int D = (int)fso_uniform(fso_sm64_next(&st), P->pool_d_min, P->pool_d_max);
  int H = (int)fso_uniform(fso_sm64_next(&st), P->pool_h_min, P->pool_h_max);
  int x0 = (int)fso_uniform(fso_sm64_next(&st), 0, nx - W);
  int y0 = (int)fso_uniform(fso_sm64_next(&st), 0, ny - D);
  int ztop = H < nz - 1 ? H : nz - 1;
  for (int z = 1; z <= ztop; z++)
    for (int y = y0; y < y0 + D; y++) memset(g + ((size_t)z * ny + y) * nx + x0, FSO_FLUID, (size_t)W);
  for (int k = 0; k < P->n_pillars; k++) {
    int cx = (int)fso_uniform(fso_sm64_next(&st), 0, nx - 1);
    int cy = (int)fso_uniform(fso_sm64_next(&st), 0, ny - 1);
    int r = (int)fso_uniform(fso_sm64_next(&st), P->r_min, P->r_max);
    int h = (int)fso_uniform(fso_sm64_next(&st), P->h_min, P->h_max);
    int zt = h < nz - 1 ? h : nz - 1;
    for (int y = cy - r; y <= cy + r; y++) {
      if (y < 0 || y >= ny) continue;
      for (int x = cx - r; x <= cx + r; x++) {
        if (x < 0 || x >= nx) continue;
        long long dx = x - cx, dy = y - cy;
        if (dx * dx + dy * dy > (long long)r * r) continue;
        for (int z = 1; z <= zt; z++) g[((size_t)z * ny + y) * nx + x] = FSO_SAND;
      }
    }
  }
  return 0;
}

void fso_batch_specs(int n, uint64_t seed, int lo, int hi, int32_t *dims, uint64_t *thr,
                     uint64_t *seeds) {
  const uint64_t tlo = fso_threshold(1, 20), thi = fso_threshold(1, 2);
  uint64_t st = seed;
  for (int w = 0; w < n; w++) {
    for (int a = 0; a < 3; a++) dims[3 * w + a] = fso_loguniform(fso_sm64_next(&st), lo, hi);
    thr[w] = tlo + fso_mulhi64(fso_sm64_next(&st), thi - tlo);
    seeds[w] = fso_sm64_next(&st);
  }
}

void fso_serve_specs(int n, uint64_t seed, int lo, int hi, int tick_lo, int tick_hi, double lam,
                     int32_t *dims, uint64_t *thr, uint64_t *seeds, int32_t *ticks,
                     double *arrival) {
  const uint64_t tlo = fso_threshold(1, 20), thi = fso_threshold(1, 2);
  uint64_t st = seed;
  double now = 0.0;
  for (int w = 0; w < n; w++) {
    double u = (double)(fso_sm64_next(&st) >> 11) * (1.0 / 9007199254740992.0);
    now += -log1p(-u) / lam;
    arrival[w] = now;
    for (int a = 0; a < 3; a++) dims[3 * w + a] = fso_loguniform(fso_sm64_next(&st), lo, hi);
    ticks[w] = (int32_t)fso_uniform(fso_sm64_next(&st), (uint64_t)tick_lo, (uint64_t)tick_hi);
    thr[w] = tlo + fso_mulhi64(fso_sm64_next(&st), thi - tlo);
    seeds[w] = fso_sm64_next(&st);
  }
}
