// Renderers: the HIP port of the reference's full-screen sphere tracer and the
// voxel-grid ray caster that hands a world's grid to the screen.
//
// Pixel -> ray follows the reference exactly (shaders/fs_raymarch.frag:67-75,
// shaders/fs_raymarch.vert:30-36, quad UVs src/engine/rendering/renderer.cpp:1253-1267):
// inUV = ((W - px - 0.5)/W, (py + 0.5)/H), py = 0 at the top; uv = inUV*2 - 1;
// uv.y *= 1/aspect; dir = normalize(uv, 1) (camera +x = screen-left, +y =
// screen-down, +z = forward).  normalize(v) = v / sqrt((x*x + y*y) + z*z) and
// dot(a,b) = (ax*bx + ay*by) + az*bz, each float op a single correctly rounded
// IEEE-754 operation (the library is built with -ffp-contract=off; HIP's fp32
// division and sqrt are correctly rounded by default).
//
// sdf_kernel: shaders/fs_raymarch.frag:10-81 -- sphere of radius 0.5 at the
// origin, <= 64 steps, hit below 0.001, stop once t > 1000 (checked after the
// hit test), central-difference normal with h = 0.001, colour (1,0,0) *
// max(0.05, dot(n, normalize(p - (2,5,3)))), miss = black, alpha 1.
//
// dda_kernel: one thread per pixel; world dir = dx*b0 + dy*b1 + dz*b2 from the
// camera basis; slab test against [0,nx]x[0,ny]x[0,nz]; Amanatides & Woo voxel
// traversal (ties x, then y, then z) to the first non-EMPTY cell; face = the one
// entered through (0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z, 6 = started inside);
// colour = palette entry (STATIC 127, SAND 162, FLUID 137, other bytes their own
// index) * max(0.05, dot(n, normalize(p - light))) -- fs_raymarch's lighting with
// the face normal (-dir when inside).  Miss: black, hit = UINT64_MAX, face 255.
#include "fs3d_internal.h"
#include "fs3d_render.h"

namespace fs3d {
namespace {

__constant__ u32 kPaletteBits[1024] = {
#include "palette.inc"
};

typedef u32 u32x2 __attribute__((ext_vector_type(2)));
typedef u32x2 __attribute__((aligned(1))) u32x2_u;  // rows start at any byte

struct F3 {
  float x, y, z;
};

__device__ __forceinline__ F3 nrm3(float x, float y, float z) {
  const float l = sqrtf((x * x + y * y) + z * z);
  return F3{x / l, y / l, z / l};
}

__device__ __forceinline__ float dot3(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

__device__ __forceinline__ F3 pixel_dir(int px, int py, int W, int H, float aspect) {
  const float u = ((float)(W - px) - 0.5f) / (float)W;
  const float v = ((float)py + 0.5f) / (float)H;
  const float ux = u * 2.0f - 1.0f;
  float uy = v * 2.0f - 1.0f;
  const float inv = 1.0f / aspect;
  uy = uy * inv;
  return nrm3(ux, uy, 1.0f);
}

__device__ __forceinline__ float sdf_map(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z) - 0.5f; }

__global__ __launch_bounds__(256) void sdf_kernel(float ox, float oy, float oz, float aspect, int W, int H,
                                                  float *rgba) {
  const int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (px >= W || py >= H) return;
  const F3 d = pixel_dir(px, py, W, H, aspect);
  float r = 0.0f, gg = 0.0f, b = 0.0f;
  float t = 0.0f;
  for (int i = 0; i < 64; i++) {
    const float p0 = ox + t * d.x, p1 = oy + t * d.y, p2 = oz + t * d.z;
    const float dist = sdf_map(p0, p1, p2);
    if (dist < 0.001f) {
      const float h = 0.001f;
      const float gx = sdf_map(p0 + h, p1 + 0.0f, p2 + 0.0f) - sdf_map(p0 - h, p1 - 0.0f, p2 - 0.0f);
      const float gy = sdf_map(p0 + 0.0f, p1 + h, p2 + 0.0f) - sdf_map(p0 - 0.0f, p1 - h, p2 - 0.0f);
      const float gz = sdf_map(p0 + 0.0f, p1 + 0.0f, p2 + h) - sdf_map(p0 - 0.0f, p1 - 0.0f, p2 - h);
      const F3 n = nrm3(gx, gy, gz);
      const F3 dl = nrm3(p0 - 2.0f, p1 - 5.0f, p2 - 3.0f);
      const float diff = fmaxf(0.05f, dot3(n, dl));
      r = 1.0f * diff;
      gg = 0.0f * diff;
      b = 0.0f * diff;
      break;
    } else if (t > 1000.0f) {
      break;
    }
    t = t + dist;
  }
  float *out = rgba + ((size_t)py * W + px) * 4;
  out[0] = r;
  out[1] = gg;
  out[2] = b;
  out[3] = 1.0f;
}

// shaders/raymarch_sphere.frag:9-53 -- the scene's active marcher: per fragment,
// dir = normalize(local - origin), at most 64 box-SDF steps (half-size 0.5), hit
// below 0.001 -> (max(0.05, density), 0, 0) with density 1 - 0.05 per step taken,
// stop once t > 1000 (tested after the hit test), miss = black, alpha 1.
__device__ __forceinline__ float box_sdf(float x, float y, float z) {
  const float qx = fabsf(x) - 0.5f, qy = fabsf(y) - 0.5f, qz = fabsf(z) - 0.5f;
  const float mx = fmaxf(qx, 0.0f), my = fmaxf(qy, 0.0f), mz = fmaxf(qz, 0.0f);
  const float len = sqrtf((mx * mx + my * my) + mz * mz);
  return len + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
}

__global__ __launch_bounds__(256) void box_march_kernel(const float *__restrict__ origin,
                                                        const float *__restrict__ local, int n, float *rgba) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const float o0 = origin[3 * k], o1 = origin[3 * k + 1], o2 = origin[3 * k + 2];
  const F3 d = nrm3(local[3 * k] - o0, local[3 * k + 1] - o1, local[3 * k + 2] - o2);
  float r = 0.0f, t = 0.0f, density = 1.0f;
  for (int i = 0; i < 64; i++) {
    const float dist = box_sdf(o0 + d.x * t, o1 + d.y * t, o2 + d.z * t);
    if (dist < 0.001f) {
      r = fmaxf(0.05f, density);
      break;
    } else if (t > 1000.0f) {
      break;
    }
    t = t + dist;
    density = density - 0.05f;
  }
  float *out = rgba + (size_t)k * 4;
  out[0] = r;
  out[1] = 0.0f;
  out[2] = 0.0f;
  out[3] = 1.0f;
}

__device__ __forceinline__ int palette_index(u32 m) {
  return m == FS3D_STATIC ? 127 : m == FS3D_SAND ? 162 : m == FS3D_FLUID ? 137 : (int)m;
}

// Occupancy maps of a grid, two levels:
//   coarse: one bit per 8x8x8 brick (set when any cell is not EMPTY),
//           coarse[(bz * bny + by) * bwords + bx / 64] bit bx % 64;
//   fine:   one u64 per brick, fine[(bz * bny + by) * bnx + bx], with bit
//           ((x & 7) >> 1) | (((y & 7) >> 1) << 2) | (((z & 7) >> 1) << 4) set
//           when the 2x2x2 sub-brick holding (x, y, z) has a non-EMPTY cell.
// A coarse bit is set exactly when the brick's fine word is non-zero.  Cells
// outside the grid count as EMPTY.  One wave covers 64 bricks along x of one
// brick row: each lane ORs the 64 rows of 8 bytes of its brick (loads coalesced
// across the wave, eight in flight per lane), bytes past nx masked off.
//
// With update set, only the bricks whose bit is set in dirty (same layout as the
// coarse map) are read again; the other bricks keep their fine words.  Either
// way the kernel clears the dirty words it covers.
__global__ __launch_bounds__(256) void occ_kernel(const u8 *__restrict__ g, int nx, int ny, int nz, int bnx, int bny,
                                                  int bnz, int bwords, u64 *coarse, u64 *fine, u64 *dirty,
                                                  int update) {
  const int lane = threadIdx.x & 63;
  const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long nwords = (long long)bwords * bny * bnz;
  if (wave >= nwords) return;
  const u64 dw = update ? dirty[wave] : ~0ull;  // one word per wave
  if (dw == 0) return;
  const int wx = (int)(wave % bwords);
  const long long rest = wave / bwords;
  const int by = (int)(rest % bny), bz = (int)(rest / bny);
  const int bx = wx * 64 + lane;
  const long long fi = ((long long)bz * bny + by) * bnx + bx;
  const bool redo = bx < bnx && ((dw >> lane) & 1ull);
  u64 acc = 0;
  if (bx < bnx && !redo) acc = fine[fi];
  if (redo) {
    const int x0 = bx * 8;
    const int keep = min(8, nx - x0);
    const u32 mlo = keep >= 4 ? ~0u : ((1u << (8 * keep)) - 1u);
    const u32 mhi = keep >= 8 ? ~0u : (keep <= 4 ? 0u : ((1u << (8 * (keep - 4))) - 1u));
    const int ylim = min(8, ny - by * 8), zlim = min(8, nz - bz * 8);  // rows of the brick inside the grid
    u32 al[4] = {0, 0, 0, 0}, ah[4] = {0, 0, 0, 0};  // ORed rows of each 2-row pair of the current 2-layer pair
    for (int dz = 0; dz < zlim; dz++) {
      u32x2 v[8];
#pragma unroll
      for (int dy = 0; dy < 8; dy++) {  // eight loads in flight; rows past ny are read clamped, then dropped
        const int y = min(by * 8 + dy, ny - 1);
        // unaligned 8-byte load; at most 7 bytes past the row end, inside the
        // next row or the buffer's guard bytes, and masked off
        v[dy] = *reinterpret_cast<const u32x2_u *>(g + ((size_t)(bz * 8 + dz) * ny + y) * nx + x0);
      }
#pragma unroll
      for (int sy = 0; sy < 4; sy++) {
        if (2 * sy < ylim) {
          al[sy] |= v[2 * sy].x;
          ah[sy] |= v[2 * sy].y;
        }
        if (2 * sy + 1 < ylim) {
          al[sy] |= v[2 * sy + 1].x;
          ah[sy] |= v[2 * sy + 1].y;
        }
      }
      if ((dz & 1) || dz + 1 == zlim) {  // a 2-layer pair is complete: its 16 sub-brick bits
#pragma unroll
        for (int sy = 0; sy < 4; sy++) {
          const u32 lo = al[sy] & mlo, hi = ah[sy] & mhi;
          const u32 b4 = ((lo & 0xFFFFu) != 0 ? 1u : 0u) | ((lo >> 16) != 0 ? 2u : 0u) |
                         ((hi & 0xFFFFu) != 0 ? 4u : 0u) | ((hi >> 16) != 0 ? 8u : 0u);
          acc |= (u64)b4 << (4 * sy + 16 * (dz >> 1));
          al[sy] = ah[sy] = 0;
        }
      }
    }
    fine[fi] = acc;
  }
  const u64 bits = __ballot(acc != 0);
  if (lane == 0) {
    coarse[wave] = bits;
    dirty[wave] = 0;
  }
}

// After a skipping tick: marks in dirty (coarse-map layout) every brick that
// holds a lane segment whose change bit the tick set.  The tick's masks are
// those of its parity p: row quad (yq, zq) covers rows 2yq - p, 2yq - p + 1 and
// layers 2zq - p, 2zq - p + 1, bit s covers x in [s * seg, (s + 1) * seg), and
// every cell the tick changed lies in a segment whose bit it set.  One thread
// per brick row (by, bz): it ORs the masks of the (up to 5 x 5) row quads that
// share rows with it, then spreads each segment bit over the seg / 8 bricks of
// that segment in the row's dirty words.
__global__ __launch_bounds__(256) void occ_dirty_kernel(const u64 *__restrict__ m0, const u64 *__restrict__ m1,
                                                        u64 t_host, const u64 *tick_dev, u32 k_off, int ny, int nz,
                                                        int seg, int bnx, int bny, int bnz, int bwords,
                                                        u64 *dirty) {
  const long long row = (long long)blockIdx.x * 256 + threadIdx.x;
  if (row >= (long long)bny * bnz) return;
  const int by = (int)(row % bny), bz = (int)(row / bny);
  const u64 t = tick_dev ? *tick_dev + k_off : t_host;
  const int p = (int)(t & 1u);
  const u64 *m = p ? m1 : m0;
  const int nyq = (ny + p + 1) >> 1, nzq = (nz + p + 1) >> 1;
  const int yq0 = (by * 8 + p) >> 1, yq1 = min((min(by * 8 + 7, ny - 1) + p) >> 1, nyq - 1);
  const int zq0 = (bz * 8 + p) >> 1, zq1 = min((min(bz * 8 + 7, nz - 1) + p) >> 1, nzq - 1);
  // all 25 loads issued at once (indices clamped into the rows' range, the
  // ones outside it dropped afterwards)
  u64 v[5][5];
#pragma unroll
  for (int j = 0; j < 5; j++)
#pragma unroll
    for (int i = 0; i < 5; i++) v[j][i] = m[(long long)min(zq0 + j, zq1) * nyq + min(yq0 + i, yq1)];
  u64 segs = 0;
#pragma unroll
  for (int j = 0; j < 5; j++)
#pragma unroll
    for (int i = 0; i < 5; i++) segs |= (zq0 + j <= zq1 && yq0 + i <= yq1) ? v[j][i] : 0ull;
  if (segs == 0) return;
  // segment s covers bricks [s << sh, (s + 1) << sh), sh = log2(seg / 8)
  const int sh = __builtin_ctz((unsigned)seg) - 3;
  const int per = 1 << sh, spw = 64 >> sh;  // bricks per segment, segments per dirty word
  const u64 run = per == 64 ? ~0ull : ((1ull << per) - 1ull);
  for (int wx = 0; wx < bwords; wx++) {
    const u64 sw = spw == 64 ? segs : (segs >> (wx * spw)) & ((1ull << spw) - 1ull);
    if (sw == 0) continue;
    u64 bits = 0;
    for (int j = 0; j < spw; j++)
      if ((sw >> j) & 1ull) bits |= run << (j * per);
    const int left = bnx - wx * 64;  // bricks of the row in this word
    if (left < 64) bits &= (1ull << left) - 1ull;
    if (bits) dirty[row * bwords + wx] |= bits;
  }
}

// The region map: one bit per 32^3 region (4 x 4 x 4 bricks), set when any of
// its bricks' coarse bits is, region[(rz * rny + ry) * rwords + rx / 64] bit
// rx % 64.  One wave per 64 regions along x; each lane ORs the 4-bit groups of
// its 16 brick rows.
__global__ __launch_bounds__(256) void region_kernel(const u64 *__restrict__ coarse, int bny, int bnz, int bwords,
                                                     int rnx, int rny, int rnz, int rwords, u64 *region) {
  const int lane = threadIdx.x & 63;
  const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= (long long)rwords * rny * rnz) return;
  const int wx = (int)(wave % rwords);
  const long long rest = wave / rwords;
  const int ry = (int)(rest % rny), rz = (int)(rest / rny);
  const int rx = wx * 64 + lane;
  u64 any = 0;
  if (rx < rnx) {
    const int bx = rx * 4;  // the group's 4 bits never straddle a word: bx % 64 is a multiple of 4
    for (int dz = 0; dz < 4; dz++) {
      const int bz = rz * 4 + dz;
      if (bz >= bnz) break;
#pragma unroll
      for (int dy = 0; dy < 4; dy++) {
        const int by = ry * 4 + dy;
        const u64 w = by < bny ? coarse[((long long)bz * bny + by) * bwords + (bx >> 6)] : 0ull;
        any |= (w >> (bx & 63)) & 0xFull;
      }
    }
  }
  const u64 bits = __ballot(any != 0);
  if (lane == 0) region[wave] = bits;
}

template <typename T>
__host__ __device__ __forceinline__ T pick(bool cx, bool cy, T x, T y, T z) {  // values, not lvalues: no address selects
  return cx ? x : (cy ? y : z);
}

// State of one ray's cell walk.  advance() takes one step: the axis with the
// smallest tnext (ties x, then y, then z); false when the walk ends (a zero
// step or leaving the grid), else t = tnext[a], idx[a] += step[a],
// tnext[a] += tdelta[a], f = the face entered through, and crossed says whether
// the new cell lies in another 8x8x8 brick.
struct Walk {
  int i0, i1, i2;
  float tn0, tn1, tn2, t;
  int f;
  float td0, td1, td2;
  int s0, s1, s2;
  int nx, ny, nz;
  int fc0, fc1, fc2;  // face entered by a step along each axis
  int up0, up1, up2;  // -1 on an axis stepping up, else 0: (i ^ (up & (L - 1))) & (L - 1) = steps left in an L-box
  int hi0, hi1, hi2;  // last index on an axis stepping up, else a bound no index reaches
  float rtd0, rtd1, rtd2;  // 1 / tdelta on a moving axis (estimates only), else 0

  // The walk from its first cell, as dda_kernel sets it up.
  static __host__ __device__ __forceinline__ Walk start(const int idx[3], const float tnext[3], float t, int f,
                                                        const float tdelta[3], const int step[3], int nx, int ny,
                                                        int nz) {
    const int big = 1 << 30;
    return Walk{idx[0], idx[1], idx[2], tnext[0], tnext[1], tnext[2], t, f, tdelta[0], tdelta[1], tdelta[2],
                step[0], step[1], step[2], nx, ny, nz,
                step[0] > 0 ? 0 : 1, step[1] > 0 ? 2 : 3, step[2] > 0 ? 4 : 5,
                step[0] > 0 ? -1 : 0, step[1] > 0 ? -1 : 0, step[2] > 0 ? -1 : 0,
                step[0] > 0 ? nx - 1 : big, step[1] > 0 ? ny - 1 : big, step[2] > 0 ? nz - 1 : big,
                step[0] ? 1.0f / tdelta[0] : 0.0f, step[1] ? 1.0f / tdelta[1] : 0.0f,
                step[2] ? 1.0f / tdelta[2] : 0.0f};
  }

  // Branch-free: every field is written on every step by a select, so the
  // lanes of a wave stay together and the state stays in registers.
  __host__ __device__ __forceinline__ bool advance(bool &crossed) {
    const bool cx = (tn0 <= tn1) & (tn0 <= tn2);
    const bool cy = !cx & (tn1 <= tn2);
    const int sa = pick(cx, cy, s0, s1, s2);
    const int ia = pick(cx, cy, i0, i1, i2);
    const int na = ia + sa;
    const bool ok = (sa != 0) & ((unsigned)na < (unsigned)pick(cx, cy, nx, ny, nz));
    const bool mx = ok & cx, my = ok & cy, mz = ok & !cx & !cy;
    t = ok ? pick(cx, cy, tn0, tn1, tn2) : t;
    f = ok ? pick(cx, cy, fc0, fc1, fc2) : f;
    i0 = mx ? na : i0;
    i1 = my ? na : i1;
    i2 = mz ? na : i2;
    tn0 = mx ? tn0 + td0 : tn0;
    tn1 = my ? tn1 + td1 : tn1;
    tn2 = mz ? tn2 + td2 : tn2;
    crossed = ((na ^ ia) & ~7) != 0;
    return ok;
  }

  // True when no chain tnext, tnext + tdelta, ... can reach NaN: tnext is not
  // NaN, and tdelta is finite on every moving axis (on a still axis both are
  // +inf, and +inf + +inf stays +inf).
  __host__ __device__ __forceinline__ bool no_nan() const {
    return tn0 == tn0 && tn1 == tn1 && tn2 == tn2 && (s0 == 0 || td0 < __builtin_inff()) &&
           (s1 == 0 || td1 < __builtin_inff()) && (s2 == 0 || td2 < __builtin_inff());
  }

  // advance() repeated until the step that leaves the current L x L x L box
  // (an 8^3 brick or a 32^3 region, aligned to multiples of L), for a box whose
  // cells are all EMPTY, on a walk where no_nan() holds: false when the walk
  // ends inside the box or on that step.
  // The steps along one axis happen at the values of that axis's own addition
  // chain tnext, tnext + tdelta, ..., which never decreases, and advance()
  // takes them in increasing (value, axis) order.  With k the axis's number of
  // steps to leave the box's part of the grid, its leaving step comes at
  // T = tnext after k - 1 additions; the walk's first leaving step is the
  // least (T, axis), and by then each other axis has taken the steps of its
  // chain that come before it.  An axis with a zero step has T = tnext = +inf:
  // when it is least, advance() would pick it and end the walk.  The additions
  // are the same ones advance() makes, so t, tnext and the indices come out
  // bit for bit as after the single steps.
  //
  // The side is lm + 1 (8 or 32, up to L, at run time: lanes leaving bricks and
  // regions share one pass); the chains run up to L - 1 additions.
  template <int L>
  __host__ __device__ __forceinline__ bool leave_box(int lm = L - 1) {
    // steps to leave the box's part of the grid along each axis (on a still
    // axis any value in 1..lm+1: its chain is +inf throughout)
    const int k0 = min((i0 ^ (up0 & lm)) & lm, hi0 - i0) + 1;
    const int k1 = min((i1 ^ (up1 & lm)) & lm, hi1 - i1) + 1;
    const int k2 = min((i2 ^ (up2 & lm)) & lm, hi2 - i2) + 1;
    // The leaving axis e and its leaving time Te.  First from estimates: on a
    // moving axis E = tnext + (k - 1) * tdelta is within (L + 2) * 2^-24 *
    // max(|tnext|, |E|) of the chain value T (k - 1 <= L - 1 rounded additions
    // against one rounded product and sum), so when the least E is below each
    // other one by more than `margin` (2^-20 for L = 8, 2^-18 for L = 32) times
    // the two axes' scales summed, the chain values order the same way with no
    // tie, and only e's chain is needed.  A still axis (T = E = +inf) is apart
    // from any finite E.  Near-ties take all three chains.
    constexpr float margin = L <= 8 ? 0x1p-20f : 0x1p-18f;
    static_assert(L <= 32, "the margin covers at most 31 additions");
    const float E0 = s0 ? tn0 + (float)(k0 - 1) * td0 : tn0;
    const float E1 = s1 ? tn1 + (float)(k1 - 1) * td1 : tn1;
    const float E2 = s2 ? tn2 + (float)(k2 - 1) * td2 : tn2;
    bool ex = (E0 <= E1) & (E0 <= E2);
    bool ey = !ex & (E1 <= E2);
    bool ez = !ex & !ey;
    const float Ee = pick(ex, ey, E0, E1, E2);
    const float me = fmaxf(fabsf(pick(ex, ey, tn0, tn1, tn2)), fabsf(Ee));
    const bool fin = Ee < 1e38f;
    const bool a0 = ex | (s0 == 0 ? fin : (E0 - Ee > margin * (fmaxf(fabsf(tn0), fabsf(E0)) + me)));
    const bool a1 = ey | (s1 == 0 ? fin : (E1 - Ee > margin * (fmaxf(fabsf(tn1), fabsf(E1)) + me)));
    const bool a2 = ez | (s2 == 0 ? fin : (E2 - Ee > margin * (fmaxf(fabsf(tn2), fabsf(E2)) + me)));
    float Te =
        chain_at<L>(pick(ex, ey, tn0, tn1, tn2), pick(ex, ey, td0, td1, td2), pick(ex, ey, k0, k1, k2) - 1);
    if (!(a0 & a1 & a2)) {
      const float T0 = chain_at<L>(tn0, td0, k0 - 1), T1 = chain_at<L>(tn1, td1, k1 - 1),
                  T2 = chain_at<L>(tn2, td2, k2 - 1);
      ex = (T0 <= T1) & (T0 <= T2);
      ey = !ex & (T1 <= T2);
      ez = !ex & !ey;
      Te = pick(ex, ey, T0, T1, T2);
    }
    const int se = pick(ex, ey, s0, s1, s2);
    const int ne = pick(ex, ey, i0, i1, i2) + pick(ex, ey, k0, k1, k2) * se;
    // (the state is not used after the walk ends)
    const bool ok = (se != 0) & ((unsigned)ne < (unsigned)pick(ex, ey, nx, ny, nz));
    // The other two axes, a < b (e = 0: 1, 2; e = 1: 0, 2; e = 2: 0, 1): steps
    // at chain values before (Te, e); an axis below e also takes the steps that
    // tie with Te (a unless e = 0, b when e = 2).
    const float ta = ex ? tn1 : tn0, da = ex ? td1 : td0, ra = ex ? rtd1 : rtd0;
    const float tb = ez ? tn1 : tn2, db = ez ? td1 : td2, rb = ez ? rtd1 : rtd2;
    const int ka = ex ? k1 : k0, kb = ez ? k1 : k2;
    const bool mova = (ex ? s1 : s0) != 0, movb = (ez ? s1 : s2) != 0;
    int na, nb;
    float ca, cb;
    const bool fa = count_near<L>(ta, da, ra, ka, Te, !ex, margin, mova, &na, &ca);
    const bool fb = count_near<L>(tb, db, rb, kb, Te, ez, margin, movb, &nb, &cb);
    if (!(fa & fb)) {  // tdelta tiny against the ray's scale: value by value
      ca = chain_before<L>(ta, da, ka, Te, !ex, &na);
      cb = chain_before<L>(tb, db, kb, Te, ez, &nb);
    }
    const float Tn = Te + pick(ex, ey, td0, td1, td2);
    i0 = ex ? ne : i0 + na * s0;
    i1 = ey ? ne : i1 + (ex ? na : nb) * s1;
    i2 = ez ? ne : i2 + nb * s2;
    tn0 = ex ? Tn : ca;
    tn1 = ey ? Tn : (ex ? ca : cb);
    tn2 = ez ? Tn : cb;
    t = Te;
    f = pick(ex, ey, fc0, fc1, fc2);
    return ok;
  }

  // The number of values of the chain c, c + d, ... (its first k) that come
  // before T (below it, or equal when tie), and the value after that many
  // additions.  Chain value j is within (L + 3) * 2^-24 * s of c + j d, and
  // x = (T - c) / d computed here within 3 * 2^-24 * s / d of the exact
  // quotient, s = |c| + |T| + L d; so with d above twice `margin` * s, value
  // n - 2 is below T and value n + 1 above it for n = ceil(x) (clamped to
  // 0..k-1), and the count is n - 1, n or n + 1, settled by comparing values
  // n - 1 and n with T.  One chain of additions yields both.  False (nothing
  // set) when d is not that large.  A still axis (d = c = +inf) takes no step.
  template <int L>
  static __host__ __device__ __forceinline__ bool count_near(float c, float d, float rd, int k, float T, bool tie,
                                                             float margin, bool moving, int *count, float *value) {
    const float x = (T - c) * rd;
    const int n = moving ? (int)fminf(fmaxf(ceilf(x), 0.0f), (float)(k - 1)) : 0;
    // value n - 1 by the chain; value n is one more of its additions
    const float prev = chain_at<L>(c, d, n - 1);
    const float cur = n == 0 ? c : prev + d;
    const float next = cur + d;
    const bool pp = (n == 0) | (prev < T) | (tie & (prev == T));  // value n - 1 comes before T
    const bool pc = (cur < T) | (tie & (cur == T));               // value n does
    *count = moving ? n - 1 + (int)pp + (int)pc : 0;
    *value = moving ? (!pp ? prev : (pc ? next : cur)) : c;
    return !moving | (d > 2.0f * margin * (fabsf(c) + fabsf(T) + (float)L * d));
  }

  // c after m (at most L - 1; none when m <= 0) additions of d
  template <int L>
  static __host__ __device__ __forceinline__ float chain_at(float c, float d, int m) {
#pragma unroll
    for (int j = 1; j < L; j++) c = j <= m ? c + d : c;
    return c;
  }

  // The chain c, c + d, ... over its first k (<= L) values: *n = how many come
  // before T (or equal it when tie), and the value after that many additions.
  template <int L>
  static __host__ __device__ __forceinline__ float chain_before(float c, float d, int k, float T, bool tie, int *n) {
    float r = c;
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < L; j++) {
      const bool before = (j < k) & ((c < T) | (tie & (c == T)));  // & and |: no branches
      const float c1 = c + d;
      r = before ? c1 : r;
      cnt = before ? j + 1 : cnt;
      c = c1;
    }
    *n = cnt;
    return r;
  }
};

// One thread per pixel.  The walk is the cell-by-cell Amanatides & Woo
// traversal: at each cell, test it, then take the axis with the smallest
// tnext (ties x, then y, then z), t = tnext[a], idx[a] += step[a] (leaving the
// grid or a zero step ends the walk), tnext[a] += tdelta[a].  It is organised
// per brick: on entering a brick the region bit, then the coarse bit and, when
// set, the fine word are read once.  An all-EMPTY 32^3 region or 8^3 brick is
// left in one jump (Walk::leave_box); in an occupied brick a cell is loaded
// only when its sub-brick bit is set, every other cell is EMPTY by
// construction of the maps.  The jumps reproduce the single steps' float
// operations exactly, so t, the hit cell and the face are those of the plain
// walk.
//
// Bounds: every step moves one axis by one cell in that axis's fixed direction,
// so a walk ends after at most nx + ny + nz steps, and a brick is left after at
// most 7 in-brick steps per axis.  The plain walk's cap of nx + ny + nz + 3
// tested cells is therefore never reached, and the outer loop, each iteration
// of which leaves at least one brick, is capped by the number of bricks a
// walk can cross.
__global__ __launch_bounds__(256) void dda_kernel(const u8 *__restrict__ g, int nx, int ny, int nz, RenderCam cam,
                                                  int W, int H, float *rgb, u64 *hit, u8 *face,
                                                  const u64 *__restrict__ coarse, const u64 *__restrict__ fine,
                                                  const u64 *__restrict__ region, int bnx, int bny, int bwords, int rny,
                                                  int rwords) {
  const int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (px >= W || py >= H) return;
  const size_t pix = (size_t)py * W + px;
  const F3 dc = pixel_dir(px, py, W, H, cam.aspect);
  const float *bx = cam.basis, *by = cam.basis + 3, *bz = cam.basis + 6;
  const F3 dd = nrm3((dc.x * bx[0] + dc.y * by[0]) + dc.z * bz[0], (dc.x * bx[1] + dc.y * by[1]) + dc.z * bz[1],
                     (dc.x * bx[2] + dc.y * by[2]) + dc.z * bz[2]);
  const float d[3] = {dd.x, dd.y, dd.z};
  const float o[3] = {cam.origin[0], cam.origin[1], cam.origin[2]};
  const float nd[3] = {(float)nx, (float)ny, (float)nz};
  const int ni[3] = {nx, ny, nz};
  float outc[3] = {0.0f, 0.0f, 0.0f};
  u64 hout = ~0ull;
  u8 fout = 255;

  float tmin = 0.0f, tmax = __builtin_inff();
  int enter = -1;
  bool miss = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (d[a] == 0.0f) {
      if (o[a] < 0.0f || o[a] > nd[a]) miss = true;
      continue;
    }
    const float t1 = (0.0f - o[a]) / d[a];
    const float t2 = (nd[a] - o[a]) / d[a];
    const float lo = fminf(t1, t2), hi = fmaxf(t1, t2);
    if (lo > tmin) {
      tmin = lo;
      enter = a;
    }
    if (hi < tmax) tmax = hi;
  }
  if (!miss && !(tmin > tmax)) {
    int idx[3], step[3];
    float tnext[3], tdelta[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const float p = o[a] + tmin * d[a];
      int i = (int)floorf(p);
      i = i < 0 ? 0 : (i > ni[a] - 1 ? ni[a] - 1 : i);
      idx[a] = i;
      if (d[a] > 0.0f) {
        step[a] = 1;
        tdelta[a] = fabsf(1.0f / d[a]);
        tnext[a] = ((float)(i + 1) - o[a]) / d[a];
      } else if (d[a] < 0.0f) {
        step[a] = -1;
        tdelta[a] = fabsf(1.0f / d[a]);
        tnext[a] = ((float)i - o[a]) / d[a];
      } else {
        step[a] = 0;
        tdelta[a] = __builtin_inff();
        tnext[a] = __builtin_inff();
      }
    }
    const int f0 = enter >= 0 ? 2 * enter + (d[enter] > 0.0f ? 0 : 1) : 6;
    Walk wk = Walk::start(idx, tnext, tmin, f0, tdelta, step, nx, ny, nz);
    const bool jump = wk.no_nan();
    const int maxb = (nx >> 3) + (ny >> 3) + (nz >> 3) + 6;  // bricks per axis + 1 each, + 3
    long long cw = -1;
    u64 cword = 0;
    bool found = false;
    size_t hli = 0;
    u32 hm = 0;
    long long crw = -1;
    u64 crword = 0;
    for (int nb = 0; nb < maxb; nb++) {
      bool empty_region = false;
      if (jump) {
        const int rx = wk.i0 >> 5, ry = wk.i1 >> 5, rz = wk.i2 >> 5;
        const long long rw = ((long long)rz * rny + ry) * rwords + (rx >> 6);
        if (rw != crw) {
          crw = rw;
          crword = region[rw];
        }
        empty_region = !((crword >> (rx & 63)) & 1ull);
      }
      u64 fw = 0;
      if (!empty_region) {
        const int ux = wk.i0 >> 3, uy = wk.i1 >> 3, uz = wk.i2 >> 3;
        const long long w = ((long long)uz * bny + uy) * bwords + (ux >> 6);
        if (w != cw) {  // rays along x reuse one coarse word for 64 bricks
          cw = w;
          cword = coarse[w];
        }
        fw = ((cword >> (ux & 63)) & 1ull) ? fine[((long long)uz * bny + uy) * bnx + ux] : 0ull;
      }
      bool crossed = false, go = true;
      if (fw == 0 && jump) {  // an all-EMPTY 32^3 region or 8^3 brick, left in one jump
        go = wk.leave_box<32>(empty_region ? 31 : 7);
      } else if (fw == 0) {  // rays whose chains could reach NaN: the single steps
        do {
          go = wk.advance(crossed);
        } while (go && !crossed);
      } else {
        do {
          const int sb = ((wk.i0 >> 1) & 3) | (((wk.i1 >> 1) & 3) << 2) | (((wk.i2 >> 1) & 3) << 4);
          if ((fw >> sb) & 1ull) {
            const size_t li = ((size_t)wk.i2 * ny + wk.i1) * nx + wk.i0;
            const u32 m = g[li];
            if (m != 0) {
              hm = m;
              hli = li;
              found = true;
              break;
            }
          }
          go = wk.advance(crossed);
        } while (go && !crossed);
      }
      if (found || !go) break;
    }
    if (found) {
      const int ci = palette_index(hm);
      const int f = wk.f;
      const float t = wk.t;
      F3 n{0.0f, 0.0f, 0.0f};
      if (f == 6) {
        n = F3{-d[0], -d[1], -d[2]};
      } else {
        const float s = (f & 1) ? 1.0f : -1.0f;
        n.x = (f >> 1) == 0 ? s : 0.0f;
        n.y = (f >> 1) == 1 ? s : 0.0f;
        n.z = (f >> 1) == 2 ? s : 0.0f;
      }
      const float h0 = o[0] + t * d[0], h1 = o[1] + t * d[1], h2 = o[2] + t * d[2];
      const F3 dl = nrm3(h0 - cam.light[0], h1 - cam.light[1], h2 - cam.light[2]);
      const float diff = fmaxf(0.05f, dot3(n, dl));
      outc[0] = __uint_as_float(kPaletteBits[4 * ci + 0]) * diff;
      outc[1] = __uint_as_float(kPaletteBits[4 * ci + 1]) * diff;
      outc[2] = __uint_as_float(kPaletteBits[4 * ci + 2]) * diff;
      hout = (u64)hli;
      fout = (u8)f;
    }
  }
  rgb[pix * 3 + 0] = outc[0];
  rgb[pix * 3 + 1] = outc[1];
  rgb[pix * 3 + 2] = outc[2];
  hit[pix] = hout;
  face[pix] = fout;
}

}  // namespace

hipError_t launch_sdf(const float cam_pos[3], float aspect, int W, int H, float *d_rgba, hipStream_t s) {
  const dim3 grid((W + 15) / 16, (H + 15) / 16);
  hipLaunchKernelGGL(sdf_kernel, grid, dim3(256), 0, s, cam_pos[0], cam_pos[1], cam_pos[2], aspect, W, H, d_rgba);
  return hipGetLastError();
}

hipError_t launch_box_march(const float *d_origin, const float *d_local, int n, float *d_rgba, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(box_march_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_origin, d_local, n, d_rgba);
  return hipGetLastError();
}

static size_t bricks(int n) { return ((size_t)n + 7) / 8; }  // 8-cell bricks along an axis of n cells

static size_t coarse_words(int nx, int ny, int nz) {
  const size_t bwords = (bricks(nx) + 63) / 64;
  return bwords * bricks(ny) * bricks(nz);
}

static size_t regions(int n) { return (bricks(n) + 3) / 4; }  // 32-cell regions along an axis of n cells

static size_t region_words(int nx, int ny, int nz) { return (regions(nx) + 63) / 64 * regions(ny) * regions(nz); }

size_t occ_words(int nx, int ny, int nz) {
  return coarse_words(nx, ny, nz) + bricks(nx) * bricks(ny) * bricks(nz) + region_words(nx, ny, nz);
}

size_t dirty_words(int nx, int ny, int nz) { return coarse_words(nx, ny, nz); }

hipError_t launch_occ_dirty(u64 *const masks[2], u64 t, const u64 *tick_dev, u32 k_off, int nx, int ny, int nz,
                            int seg, u64 *d_dirty, hipStream_t s) {
  const int bnx = (int)bricks(nx), bny = (int)bricks(ny), bnz = (int)bricks(nz), bwords = (bnx + 63) / 64;
  const long long rows = (long long)bny * bnz;
  hipLaunchKernelGGL(occ_dirty_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, masks[0], masks[1], t,
                     tick_dev, k_off, ny, nz, seg, bnx, bny, bnz, bwords, d_dirty);
  return hipGetLastError();
}

hipError_t launch_dda(const u8 *g, int nx, int ny, int nz, const RenderCam &cam, int W, int H, float *d_rgb,
                      u64 *d_hit, u8 *d_face, u64 *d_occ, u64 *d_dirty, bool update, hipStream_t s) {
  const int bnx = (int)bricks(nx), bny = (int)bricks(ny), bnz = (int)bricks(nz), bwords = (bnx + 63) / 64;
  const long long waves = (long long)bwords * bny * bnz;
  u64 *coarse = d_occ, *fine = d_occ + coarse_words(nx, ny, nz);
  u64 *region = fine + (size_t)bnx * bny * bnz;
  const int rnx = (int)regions(nx), rny = (int)regions(ny), rnz = (int)regions(nz), rwords = (rnx + 63) / 64;
  hipLaunchKernelGGL(occ_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, g, nx, ny, nz, bnx, bny, bnz,
                     bwords, coarse, fine, d_dirty, update ? 1 : 0);
  const long long rwaves = (long long)rwords * rny * rnz;
  hipLaunchKernelGGL(region_kernel, dim3((unsigned)((rwaves + 3) / 4)), dim3(256), 0, s, coarse, bny, bnz, bwords,
                     rnx, rny, rnz, rwords, region);
  const dim3 grid((W + 15) / 16, (H + 15) / 16);
  hipLaunchKernelGGL(dda_kernel, grid, dim3(256), 0, s, g, nx, ny, nz, cam, W, H, d_rgb, d_hit, d_face, coarse, fine,
                     region, bnx, bny, bwords, rny, rwords);
  return hipGetLastError();
}

void palette_host(float *rgba) {
  static const u32 bits[1024] = {
#include "palette.inc"
  };
  for (int i = 0; i < 1024; i++) {
    u32 b = bits[i];
    float f;
    __builtin_memcpy(&f, &b, 4);
    rgba[i] = f;
  }
}

}  // namespace fs3d
