// Two ticks per HBM pass (temporal blocking) for the dense tick.
//
// Tick A (parity pA) and tick B (parity pB = 1 - pA) run in one launch.  A
// workgroup owns TY row quads of tick B along y (yq' in [Y0, Y0 + TY)) over a
// range of z row-quads, and sweeps z.  Its TY + 1 waves compute tick A for the
// row quads those need (tick B's row quad yq' takes rows from tick A's yq' - 1 +
// pA and yq' + pA; the extra row quad is a halo that the neighbouring tile also
// computes, from the same inputs), keeping A's output rows in LDS only; after a
// barrier, TY waves gather tick B's row quads from LDS and write them to HBM.
// Tick B's z-row b reads the top layer of tick A's z-row b - 1 + pA and the
// bottom layer of A's z-row b + pA, so the sweep computes A's z-row s and then
// B's z-row s - pA; three LDS layer buffers rotate (the top layer of the
// previous A row, the bottom and the top of the current one), each holding the
// 2 * TY rows that tick B reads (the outermost A rows are computed but not kept).
// Every wave of the workgroup runs every step of the sweep (a uniform loop with
// two barriers per step) and fetches its next tick-A row quad from HBM before
// updating the current one, so the loads overlap a step of work.
//
// HBM bytes per pair of ticks: one read of the grid (plus 1/TY of y-halo, which
// the neighbouring tile on the same XCD has usually just read, and one A z-row
// per z-range) and one write: ~1.2 bytes per voxel-update instead of 2.  Rows
// are whole (one wave per row quad: lanes_per_row > 32); hashing and
// settled-segment skipping use the single-tick kernels.
#include "tick_core.h"

namespace fs3d {
namespace {

using namespace tickcore;

// The grids these passes serve do not fit in cache: row quads are read from and
// written to HBM with streaming (nontemporal) accesses; the LDS side ignores the
// hint (2048^3 pillar scene, ms per fused tick: 2.36 streaming loads and stores,
// 2.40 streaming stores only, 2.39 neither; 1024^3 boxes: 0.313 / 0.315 / 0.323).
template <int S>
using StreamRQ = RowQuad<S, size_t, kStreamLoads | kStreamStores>;

constexpr int kTY = 3;              // tick-B row quads per workgroup along y
constexpr int kWaves = kTY + 1;     // waves per workgroup (one tick-A row quad each)
constexpr int kLayerRows = 2 * kTY; // y rows per LDS layer: the rows tick B reads
constexpr u32 kXcds = 8;            // workgroup i is dispatched to XCD i % 8

struct Tick2Args {
  const u8 *src;
  u8 *dst;
  int nx, ny, nz;
  int lanes_per_row;
  int stride;   // LDS bytes per row (lanes_per_row * S rounded up to 16: the last lane's
                // full segment fits; its bytes past nx are never used)
  int ytiles;   // tiles along y
  int zlen;     // tick-B z row-quads per workgroup
  u64 t;        // tick A's counter (when tick_dev is NULL)
  u32 k_off;
  const u64 *tick_dev;
  const u64 *rule;
};

template <int S>
__global__ __launch_bounds__(kWaves * 64) void tick2_kernel(Tick2Args a) {
  extern __shared__ __attribute__((aligned(16))) u8 lds[];
  constexpr int ND = S / 4;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const u64 t64 = a.tick_dev ? *a.tick_dev + a.k_off : a.t;
  const u32 tA = (u32)t64, tB = tA + 1u;
  const int pA = (int)(tA & 1u), pB = pA ^ 1;
  const int nyqA = (a.ny + pA + 1) >> 1, nzqA = (a.nz + pA + 1) >> 1;
  const int nyqB = (a.ny + pB + 1) >> 1, nzqB = (a.nz + pB + 1) >> 1;
  // Each XCD takes a contiguous range of tiles, so y-neighbouring tiles (which
  // read the same halo row quad) run together behind one L2.
  const u32 G = gridDim.x, bid = blockIdx.x;
  const u32 xcd = bid % kXcds, per = G / kXcds, extra = G % kXcds;
  const int tile = (int)(xcd * per + min(xcd, extra) + bid / kXcds);
  const int ytile = tile % a.ytiles, zseg = tile / a.ytiles;
  const int Y0 = ytile * kTY;             // first tick-B row quad along y
  const int YA0 = Y0 - 1 + pA;            // first tick-A row quad along y
  const int y_lo = 2 * YA0 - pA + 1;      // first y row held in LDS
  const int Z0 = zseg * a.zlen;
  const int Z1 = min(Z0 + a.zlen, nzqB);
  const int LR = a.lanes_per_row;
  const size_t layer_bytes = (size_t)kLayerRows * a.stride;
  const int s_first = Z0 + pA - 1, s_last = Z1 - 1 + pA;
  const int yqA = YA0 + wv;
  // rows of this wave's tick-A row quads that no tick-B row of the tile reads
  const int unused = (wv == 0 ? 0x5 : 0) | (wv == kWaves - 1 ? 0xA : 0);

  // tick A's row quad (yqA, s), fetched one step ahead of its update
  auto fetch_a = [&](StreamRQ<S> &R, int seg, int s) {
    R.seg = seg;
    R.yq = yqA;
    R.zq = s;
    R.active = seg < LR && yqA >= 0 && yqA < nyqA && s >= 0 && s < nzqA && s <= s_last;
    R.q = 0;
    R.skip = false;
    rq_fetch<S>(R, a.src, a.nx, a.ny, a.nz, pA, true);
  };
  StreamRQ<S> R;
  fetch_a(R, lane, s_first);
  for (int s = s_first; s <= s_last; s++) {
    const int i = s - s_first;
    // An opaque copy of the lane's segment index: otherwise the per-lane
    // predicates derived from it are hoisted out of the loop into SGPR pairs.
    int seg = lane;
    asm volatile("" : "+v"(seg));
    u8 *const prev_top = lds + (size_t)((2 * i + 2) % 3) * layer_bytes;  // top layer of A's z-row s - 1
    u8 *const cur_bot = lds + (size_t)((2 * i) % 3) * layer_bytes;       // bottom layer of A's z-row s
    u8 *const cur_top = lds + (size_t)((2 * i + 1) % 3) * layer_bytes;   // top layer of A's z-row s
    StreamRQ<S> Rn;
    fetch_a(Rn, seg, s + 1);
    // ---- tick A: row quad (yqA, s) into LDS ----
    R.seg = seg;
    R.valid &= ~unused;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = max(2 * wv + (r & 1) - 1, 0);
      u8 *layer = (r >> 1) ? cur_top : cur_bot;
      R.rowoff[r] = (size_t)(layer - lds) + (size_t)row * a.stride + (size_t)seg * S;
    }
    rq_update<S, false>(R, lds, LR, tA, a.rule);
    __syncthreads();
    // ---- tick B: row quad (Y0 + wv, s - pA) from LDS into HBM ----
    const int b = s - pA;
    if (wv < kTY && b >= Z0 && b < Z1) {  // uniform per wave; no barrier inside
      StreamRQ<S> Q;
      Q.seg = seg;
      Q.yq = Y0 + wv;
      Q.zq = b;
      Q.active = seg < LR && Q.yq < nyqB;
      Q.q = 0;
      Q.skip = false;
      Q.loaded = true;
      const int y0 = 2 * Q.yq - pB, z0 = 2 * b - pB;
      const int xs = Q.seg * S;
      Q.m = min(S, a.nx - xs);
      Q.valid = 0;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int y = y0 + (r & 1), z = z0 + (r >> 1);
        const bool ok = Q.active && y >= 0 && y < a.ny && z >= 0 && z < a.nz;
        Q.valid |= ok ? (1 << r) : 0;
        Q.rowoff[r] = ((size_t)(ok ? z : 0) * a.ny + (ok ? y : 0)) * a.nx + xs;
#pragma unroll
        for (int j = 0; j < ND; j++) Q.own[r][j] = kStatic4;
        if (ok) {
          // z0 + 0 is the top layer of A's z-row s - 1, z0 + 1 the bottom of A's z-row s;
          // y - y_lo = 2 * wv + (r & 1)
          const u8 *layer = (r >> 1) ? cur_bot : prev_top;
          load_seg<S>(layer + (size_t)(y - y_lo) * a.stride + xs, Q.own[r]);
        }
        if (Q.m < S) {
#pragma unroll
          for (int j = 0; j < ND; j++) Q.own[r][j] = keep_bytes(Q.own[r][j], Q.m - 4 * j);
        }
      }
      rq_update<S, false>(Q, a.dst, LR, tB, a.rule);
    }
    __syncthreads();
    R = Rn;
  }
}

// The fused pass for rows longer than one wave holds (nx > 1024): each tick-A
// and tick-B row quad is split into x-chunks of 64 * S bytes, one wave each, so
// a workgroup has (TY + 1) * chunks waves over the same LDS ring, and a wave's
// registers hold S bytes of a row instead of a whole row's share.  Tick A's
// chunk ends read the neighbouring chunk's edge bytes from HBM (odd ticks), tick
// B's from the LDS rows the neighbouring chunk's tick-A wave wrote before the
// barrier; in both, the straddling block is evaluated by both end lanes from
// the same inputs and each stores its own bytes.  Wave w takes row quad w /
// chunks and chunk w % chunks.
template <int S, int TY>
__global__ __launch_bounds__(1024) void tick2c_kernel(Tick2Args a, int chunks) {
  constexpr int kTYc = TY, kWavesC = TY + 1, kLayerRowsC = 2 * TY;
  extern __shared__ __attribute__((aligned(16))) u8 lds[];
  constexpr int ND = S / 4;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int wv = w / chunks, c = w - wv * chunks;
  const u64 t64 = a.tick_dev ? *a.tick_dev + a.k_off : a.t;
  const u32 tA = (u32)t64, tB = tA + 1u;
  const int pA = (int)(tA & 1u), pB = pA ^ 1;
  const int nyqA = (a.ny + pA + 1) >> 1, nzqA = (a.nz + pA + 1) >> 1;
  const int nyqB = (a.ny + pB + 1) >> 1, nzqB = (a.nz + pB + 1) >> 1;
  const u32 G = gridDim.x, bid = blockIdx.x;
  const u32 xcd = bid % kXcds, per = G / kXcds, extra = G % kXcds;
  const int tile = (int)(xcd * per + min(xcd, extra) + bid / kXcds);
  const int ytile = tile % a.ytiles, zseg = tile / a.ytiles;
  const int Y0 = ytile * kTYc;
  const int YA0 = Y0 - 1 + pA;
  const int y_lo = 2 * YA0 - pA + 1;
  const int Z0 = zseg * a.zlen;
  const int Z1 = min(Z0 + a.zlen, nzqB);
  const int LR = a.lanes_per_row;  // S-byte segments per row
  const size_t layer_bytes = (size_t)kLayerRowsC * a.stride;
  const int s_first = Z0 + pA - 1, s_last = Z1 - 1 + pA;
  const int yqA = YA0 + wv;
  const int unused = (wv == 0 ? 0x5 : 0) | (wv == kWavesC - 1 ? 0xA : 0);

  // tick A's row quad (yqA, s), chunk c, fetched one step ahead of its update;
  // the chunk ends' edge bytes come from HBM on odd ticks
  auto fetch_a = [&](StreamRQ<S> &R, int seg, int s) {
    R.seg = seg;
    R.yq = yqA;
    R.zq = s;
    R.active = seg < LR && yqA >= 0 && yqA < nyqA && s >= 0 && s < nzqA && s <= s_last;
    R.q = 0;
    R.skip = false;
    R.nl = R.nr = R.sl = R.sr = false;
    R.ml = R.active && pA && lane == 0 && seg > 0;
    R.mr = R.active && pA && lane == 63 && seg < LR - 1;
    rq_fetch<S>(R, a.src, a.nx, a.ny, a.nz, pA, true);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      R.edge[r] = kStatic4;
      if ((R.ml || R.mr) && (R.valid >> r & 1)) {
        const u8 *row = a.src + R.rowoff[r];
        R.edge[r] = ((R.ml ? (u32)row[-1] : 0u) << 24) | (R.mr ? (u32)row[S] : 0u);
      }
    }
  };
  StreamRQ<S> R;
  fetch_a(R, c * 64 + lane, s_first);
  for (int s = s_first; s <= s_last; s++) {
    const int i = s - s_first;
    // An opaque copy of the lane's segment index: otherwise the per-lane
    // predicates derived from it are hoisted out of the loop into SGPR pairs.
    int seg = c * 64 + lane;
    asm volatile("" : "+v"(seg));
    u8 *const prev_top = lds + (size_t)((2 * i + 2) % 3) * layer_bytes;  // top layer of A's z-row s - 1
    u8 *const cur_bot = lds + (size_t)((2 * i) % 3) * layer_bytes;       // bottom layer of A's z-row s
    u8 *const cur_top = lds + (size_t)((2 * i + 1) % 3) * layer_bytes;   // top layer of A's z-row s
    StreamRQ<S> Rn;
    fetch_a(Rn, seg, s + 1);
    // ---- tick A: row quad (yqA, s), chunk c, into LDS ----
    R.valid &= ~unused;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = max(2 * wv + (r & 1) - 1, 0);
      u8 *layer = (r >> 1) ? cur_top : cur_bot;
      R.rowoff[r] = (size_t)(layer - lds) + (size_t)row * a.stride + (size_t)seg * S;
    }
    rq_update<S, false, false, false, true>(R, lds, LR, tA, a.rule);
    __syncthreads();
    // ---- tick B: row quad (Y0 + wv, s - pA), chunk c, from LDS into HBM ----
    const int b = s - pA;
    if (wv < kTYc && b >= Z0 && b < Z1) {  // uniform per wave; no barrier inside
      StreamRQ<S> Q;
      Q.seg = seg;
      Q.yq = Y0 + wv;
      Q.zq = b;
      Q.active = seg < LR && Q.yq < nyqB;
      Q.q = 0;
      Q.skip = false;
      Q.loaded = true;
      Q.nl = Q.nr = Q.sl = Q.sr = false;
      Q.ml = Q.active && pB && lane == 0 && seg > 0;
      Q.mr = Q.active && pB && lane == 63 && seg < LR - 1;
      const int y0 = 2 * Q.yq - pB, z0 = 2 * b - pB;
      const int xs = Q.seg * S;
      Q.m = min(S, a.nx - xs);
      Q.valid = 0;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int y = y0 + (r & 1), z = z0 + (r >> 1);
        const bool ok = Q.active && y >= 0 && y < a.ny && z >= 0 && z < a.nz;
        Q.valid |= ok ? (1 << r) : 0;
        Q.rowoff[r] = ((size_t)(ok ? z : 0) * a.ny + (ok ? y : 0)) * a.nx + xs;
#pragma unroll
        for (int j = 0; j < ND; j++) Q.own[r][j] = kStatic4;
        Q.edge[r] = kStatic4;
        if (ok) {
          // z0 + 0 is the top layer of A's z-row s - 1, z0 + 1 the bottom of A's z-row s;
          // y - y_lo = 2 * wv + (r & 1)
          const u8 *lrow = ((r >> 1) ? cur_bot : prev_top) + (size_t)(y - y_lo) * a.stride;
          load_seg<S>(lrow + xs, Q.own[r]);
          if (Q.ml || Q.mr)
            Q.edge[r] = ((Q.ml ? (u32)lrow[xs - 1] : 0u) << 24) | (Q.mr ? (u32)lrow[xs + S] : 0u);
        }
        if (Q.m < S) {
#pragma unroll
          for (int j = 0; j < ND; j++) Q.own[r][j] = keep_bytes(Q.own[r][j], Q.m - 4 * j);
        }
      }
      rq_update<S, false, false, false, true>(Q, a.dst, LR, tB, a.rule);
    }
    __syncthreads();
    R = Rn;
  }
}

}  // namespace

size_t tick2_stride(const TickShape &sh) { return ((size_t)sh.lanes_per_row * sh.seg + 15) & ~(size_t)15; }

// Rows longer than this run the chunked fused pass (16-byte segments).
constexpr int kTick2ChunkMinNx = 1024;
constexpr int kTick2ChunkSeg = 16;

size_t tick2c_stride(int nx) { return ((size_t)nx + 15) & ~(size_t)15; }

// Tick-B row quads per workgroup along y of the chunked pass: 5 where the LDS
// ring (3 layers x 10 rows) fits 64 KiB and a workgroup of 6 x chunks waves fits
// 1024 threads, else 3.  Deeper tiles recompute fewer tick-A halo rows.
int tick2c_ty(int nx) {
  const int chunks = (int)((tick2c_stride(nx) / kTick2ChunkSeg + 63) / 64);
  return 3 * 10 * tick2c_stride(nx) <= 64 * 1024 && 6 * chunks <= 16 ? 5 : 3;
}

bool tick2_supported(const TickShape &sh, int nx, size_t *lds_bytes) {
  size_t bytes = 0;
  if (nx > kTick2ChunkMinNx) {
    bytes = 3 * 2 * (size_t)tick2c_ty(nx) * tick2c_stride(nx);
  } else {
    if (sh.rq_per_wave != 1) return false;
    bytes = 3 * (size_t)kLayerRows * tick2_stride(sh);
  }
  if (lds_bytes) *lds_bytes = bytes;
  return bytes <= 64 * 1024;
}

hipError_t launch_tick2(const TickShape &sh, const u8 *src, u8 *dst, int nx, int ny, int nz, u64 t,
                        const u64 *tick_dev, u32 k_off, const u64 *rule, int zlen, hipStream_t s) {
  size_t lds = 0;
  if (!tick2_supported(sh, nx, &lds)) return hipErrorInvalidValue;
  Tick2Args a;
  a.src = src;
  a.dst = dst;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.lanes_per_row = sh.lanes_per_row;
  a.stride = (int)tick2_stride(sh);
  // graph replays read the tick from the device: size the tiles for the larger
  // parity's row-quad counts (the kernel bounds itself by tick B's)
  const int nyq_max = (ny + 2) >> 1, nzq_max = (nz + 2) >> 1;
  const int ty = nx > kTick2ChunkMinNx ? tick2c_ty(nx) : kTY;
  a.ytiles = (nyq_max + ty - 1) / ty;
  a.zlen = zlen;
  a.t = t;
  a.k_off = k_off;
  a.tick_dev = tick_dev;
  a.rule = rule;
  const long long blocks = (long long)a.ytiles * ((nzq_max + zlen - 1) / zlen);
  if (nx > kTick2ChunkMinNx) {
    constexpr int S = kTick2ChunkSeg;
    a.lanes_per_row = (nx + S - 1) / S;
    a.stride = (int)tick2c_stride(nx);
    const int chunks = (a.lanes_per_row + 63) / 64;  // 2 at nx = 2048, at most 4 within the LDS limit
    if (ty == 5)
      hipLaunchKernelGGL((tick2c_kernel<S, 5>), dim3((unsigned)blocks), dim3(6 * chunks * 64), lds, s, a, chunks);
    else
      hipLaunchKernelGGL((tick2c_kernel<S, 3>), dim3((unsigned)blocks), dim3(4 * chunks * 64), lds, s, a, chunks);
    return hipGetLastError();
  }
  const dim3 grid((unsigned)blocks), block(kWaves * 64);
  switch (sh.seg) {  // nx <= 1024 in whole rows: 8 bytes up to 512, else 16
    case 8: hipLaunchKernelGGL((tick2_kernel<8>), grid, block, lds, s, a); break;
    case 16: hipLaunchKernelGGL((tick2_kernel<16>), grid, block, lds, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace fs3d
