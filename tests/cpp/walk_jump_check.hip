// Host-side check of the voxel ray caster's box jumps: for random rays over
// random grid sizes, Walk::leave_box must leave every 8^3 brick and every 32^3
// region in exactly the state (cell, tnext, t, face, end of walk) that
// Walk::advance() reaches one step at a time.  The rays include axis-aligned and diagonal directions and
// integer origins, which make tnext values tie across axes.
//
// Usage: walk_jump_check [trials] [seed]; prints one summary line, exit 1 on a
// mismatch.  Built for the host from the library's own render.hip.
#include "render.hip"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

using fs3d::Walk;

namespace {

bool same(const Walk &a, const Walk &b) {
  return a.i0 == b.i0 && a.i1 == b.i1 && a.i2 == b.i2 && !std::memcmp(&a.tn0, &b.tn0, 4) &&
         !std::memcmp(&a.tn1, &b.tn1, 4) && !std::memcmp(&a.tn2, &b.tn2, 4) && !std::memcmp(&a.t, &b.t, 4) &&
         a.f == b.f;
}

// The ray set-up of dda_kernel: slab test, start cell, step, tdelta, tnext.
bool start_walk(const float o[3], const float d[3], int nx, int ny, int nz, Walk *w) {
  const float nd[3] = {(float)nx, (float)ny, (float)nz};
  const int ni[3] = {nx, ny, nz};
  float tmin = 0.0f, tmax = __builtin_inff();
  int enter = -1;
  for (int a = 0; a < 3; a++) {
    if (d[a] == 0.0f) {
      if (o[a] < 0.0f || o[a] > nd[a]) return false;
      continue;
    }
    const float t1 = (0.0f - o[a]) / d[a];
    const float t2 = (nd[a] - o[a]) / d[a];
    const float lo = std::fmin(t1, t2), hi = std::fmax(t1, t2);
    if (lo > tmin) {
      tmin = lo;
      enter = a;
    }
    if (hi < tmax) tmax = hi;
  }
  if (tmin > tmax) return false;
  int idx[3], step[3];
  float tnext[3], tdelta[3];
  for (int a = 0; a < 3; a++) {
    const float p = o[a] + tmin * d[a];
    int i = (int)std::floor(p);
    i = i < 0 ? 0 : (i > ni[a] - 1 ? ni[a] - 1 : i);
    idx[a] = i;
    if (d[a] > 0.0f) {
      step[a] = 1;
      tdelta[a] = std::fabs(1.0f / d[a]);
      tnext[a] = ((float)(i + 1) - o[a]) / d[a];
    } else if (d[a] < 0.0f) {
      step[a] = -1;
      tdelta[a] = std::fabs(1.0f / d[a]);
      tnext[a] = ((float)i - o[a]) / d[a];
    } else {
      step[a] = 0;
      tdelta[a] = __builtin_inff();
      tnext[a] = __builtin_inff();
    }
  }
  const int f0 = enter >= 0 ? 2 * enter + (d[enter] > 0.0f ? 0 : 1) : 6;
  *w = Walk::start(idx, tnext, tmin, f0, tdelta, step, nx, ny, nz);
  return true;
}

// Walks the whole ray box by box (side lm + 1: 8 for bricks, 32 for regions):
// each box left once by Walk::leave_box<L>(lm) and once by single steps, then
// compared.  L is the longest chain the jump is compiled for.
template <int L>
void walk_boxes(Walk w, int lm, const int n[3], const float o[3], const float d[3], long long *boxes,
                long long *ties, long long *bad) {
  const int sh = lm == 7 ? 3 : 5;
  for (;;) {
    Walk a = w, b = w;
    bool crossed = false, go;
    do {
      go = a.advance(crossed);
    } while (go && (a.i0 >> sh) == (w.i0 >> sh) && (a.i1 >> sh) == (w.i1 >> sh) && (a.i2 >> sh) == (w.i2 >> sh));
    const bool gb = b.leave_box<L>(lm);
    (*boxes)++;
    *ties += go && (a.tn0 == a.tn1 || a.tn1 == a.tn2 || a.tn0 == a.tn2);
    if (go != gb || (go && !same(a, b))) {
      if ((*bad)++ < 5)
        std::printf("mismatch (L=%d, side %d): dims %d %d %d o %.9g %.9g %.9g d %.9g %.9g %.9g\n", L, lm + 1, n[0],
                    n[1], n[2], o[0], o[1], o[2], d[0], d[1], d[2]);
      return;
    }
    if (!go) return;
    w = a;
  }
}

}  // namespace

int main(int argc, char **argv) {
  const long trials = argc > 1 ? std::atol(argv[1]) : 100000;
  std::mt19937_64 rng(argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1);
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  long long rays = 0, bricks = 0, regions = 0, ties = 0, bad = 0;
  for (long trial = 0; trial < trials; trial++) {
    const int n[3] = {1 + (int)(rng() % 150), 1 + (int)(rng() % 150), 1 + (int)(rng() % 150)};
    const int mode = (int)(rng() % 8);
    float o[3], dd[3], d[3];
    for (int a = 0; a < 3; a++)  // integer and half-integer origins, or anywhere around the grid
      o[a] = mode < 2 ? (float)(rng() % (n[a] + 1)) + (mode == 1 ? 0.5f : 0.0f) : (U(rng) * 1.6f - 0.3f) * n[a];
    for (int a = 0; a < 3; a++)  // axis-aligned and diagonal, small integer ratios, or any direction
      dd[a] = mode == 3 ? (float)((int)(rng() % 3) - 1)
                        : (mode == 4 ? (float)((int)(rng() % 7) - 3) : U(rng) * 2.0f - 1.0f);
    if (mode >= 6) {  // near-ties: two or three components equal up to a few ulps
      const float base = (U(rng) * 0.8f + 0.2f) * (rng() & 1 ? 1.0f : -1.0f);
      for (int a = 0; a < (mode == 6 ? 2 : 3); a++)
        dd[a] = std::nextafter(base, (rng() & 1) ? 2.0f : -2.0f) * ((rng() & 3) ? 1.0f : -1.0f);
      for (int a = 0; a < 3; a++)
        if (rng() & 1) dd[a] = std::nextafter(dd[a], 0.0f);
      for (int a = 0; a < 3; a++)
        if (mode == 7 || a < 2) o[a] = (float)(rng() % (n[a] + 1));
    }
    const float l = std::sqrt((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
    if (!(l > 0.0f)) continue;
    for (int a = 0; a < 3; a++) d[a] = dd[a] / l;
    Walk w;
    if (!start_walk(o, d, n[0], n[1], n[2], &w) || !w.no_nan()) continue;
    rays++;
    walk_boxes<8>(w, 7, n, o, d, &bricks, &ties, &bad);
    walk_boxes<32>(w, 7, n, o, d, &bricks, &ties, &bad);  // bricks in the kernel's shared jump pass
    walk_boxes<32>(w, 31, n, o, d, &regions, &ties, &bad);
  }
  std::printf("rays %lld bricks %lld regions %lld tied %lld mismatches %lld\n", rays, bricks, regions, ties, bad);
  return bad != 0;
}
