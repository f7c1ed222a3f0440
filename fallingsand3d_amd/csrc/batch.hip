// Batched worlds: many independent grids of uneven sizes stepped by one
// persistent kernel.
//
// Each workgroup takes the next world from an atomic queue (worlds ordered
// largest first, so the long ones start early) and runs all requested ticks of
// that world before taking another: the waves of the workgroup sweep the
// world's row quads, a workgroup barrier separates ticks, and the update is in
// place (a row quad is loaded whole by one wave before any of it is stored).  A
// world of up to a few MiB stays hot in L2 / Infinity Cache across its ticks, so
// the batch streams far less than one HBM pass per tick.  The state hash of the
// final tick is fused into that tick and reduced in the workgroup.
//
// The reference has nothing of the kind (one world, one process, no simulation);
// this is the batching mode the tick is required to support.
#include <algorithm>

#include "fs3d_batch.h"
#include "tick_core.h"

namespace fs3d {
namespace {

using namespace tickcore;

template <int S, int NT, int NR>
__global__ __launch_bounds__(NT) void batch_kernel(BWorld *worlds, const int *order, int n, int nticks, u32 *queue,
                                                   const u64 *__restrict__ rule, u64 *hashes, int flags) {
  constexpr int NW = NT / 64;
  const bool want_hash = flags & kBatchHash, settle = flags & kBatchSettle;
  __shared__ int s_idx;
  __shared__ u64 s_sum[NW];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (;;) {
    if (threadIdx.x == 0) s_idx = (int)atomicAdd(queue, 1u);
    __syncthreads();
    // The queue index is the same for the whole workgroup; readfirstlane makes
    // that visible to the compiler, so the loop exit below is a uniform (scalar)
    // branch.  Left as a per-lane value, the exit looks divergent and the
    // compiler may route thread 0 and the other lanes of wave 0 through the
    // loop's barriers on different paths.
    const int idx = __builtin_amdgcn_readfirstlane(s_idx);
    __syncthreads();
    if (idx >= n) break;
    const int wi = order[idx];
    const BWorld w = worlds[wi];
    u8 *const g = grid_of(w);
    u64 acc = 0;
    int quiet = 0;  // consecutive ticks in which no cell of the world changed
    for (int k = 0; k < nticks; k++) {
      // Two quiet ticks in a row (one of each parity) make the state a fixed
      // point of both parities, whatever the variants: the remaining ticks
      // change nothing, so only the last (hashed) one is run.
      if (settle && quiet >= 2) {
        if (!want_hash) break;
        k = nticks - 1;
      }
      const u32 t = (u32)(w.tick + (u64)k);
      const int p = (int)(t & 1u);
      const long long nrq = (long long)((w.ny + p + 1) >> 1) * ((w.nz + p + 1) >> 1);
      const long long nwav = (nrq + w.rqpw - 1) / w.rqpw;
      bool ch = false;
      // each wave takes NR row-quad groups per step (wv, wv + NW, ...)
      if (want_hash && k == nticks - 1) {
        for (long long wv = wid; wv < nwav; wv += (long long)NW * NR)
          acc += tick_waves<S, true, NR>(g, g, w.nx, w.ny, w.nz, w.lr, w.rqpw, t, wv, NW, lane, rule, &ch);
      } else {
        for (long long wv = wid; wv < nwav; wv += (long long)NW * NR)
          tick_waves<S, false, NR>(g, g, w.nx, w.ny, w.nz, w.lr, w.rqpw, t, wv, NW, lane, rule, &ch);
      }
      quiet = __syncthreads_or(ch) ? 0 : quiet + 1;
    }
    if (want_hash) {
      acc = wave_sum_u64(acc);
      if (lane == 0) s_sum[wid] = acc;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (want_hash) {
        u64 tot = 0;
#pragma unroll
        for (int i = 0; i < NW; i++) tot += s_sum[i];
        hashes[wi] = tot;
      }
      worlds[wi].tick = w.tick + (u64)nticks;
    }
    __syncthreads();
  }
}

// One workgroup per world: hash of the current grid.
__global__ __launch_bounds__(256) void batch_hash_kernel(const BWorld *worlds, int n, u64 *hashes) {
  __shared__ u64 s_sum[4];
  for (int wi = blockIdx.x; wi < n; wi += gridDim.x) {
    const BWorld w = worlds[wi];
    const size_t cnt = (size_t)w.nx * w.ny * w.nz;
    u64 acc = 0;
    for (size_t i = threadIdx.x; i < cnt; i += 256) {
      const u32 v = grid_of(w)[i];
      if (v) acc += dmix64(((u64)i << 8) | v);
    }
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) hashes[wi] = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
    __syncthreads();
  }
}

// Rain fill of every world (same cell rule as the single-world fill).
__global__ __launch_bounds__(256) void batch_fill_rain_kernel(const BWorld *worlds, int n, const u64 *seeds,
                                                              const u64 *thr) {
  for (int wi = blockIdx.y; wi < n; wi += gridDim.y) {
    const BWorld w = worlds[wi];
    const size_t layer = (size_t)w.nx * w.ny, cnt = layer * w.nz;
    const int zr = max(1, w.nz >> 1);
    const u64 seed = seeds[wi], th = thr[wi];
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < cnt; i += (size_t)gridDim.x * 256) {
      const size_t z = i / layer;
      u8 v = FS3D_EMPTY;
      if (z == 0)
        v = FS3D_STATIC;
      else if (z >= (size_t)zr && dmix64(seed + (i + 1) * kGamma) < th)
        v = FS3D_SAND;
      grid_of(w)[i] = v;
    }
  }
}

__global__ void reset_u32_kernel(u32 *p) {
  if (threadIdx.x == 0) *p = 0u;
}

template <int S, int NT, int NR>
hipError_t launch_batch_t(BWorld *worlds, const int *order, int n, int nticks, u32 *queue, const u64 *rule,
                          u64 *hashes, int flags, int grid, hipStream_t s) {
  hipLaunchKernelGGL((batch_kernel<S, NT, NR>), dim3(grid), dim3(NT), 0, s, worlds, order, n, nticks, queue, rule,
                     hashes, flags);
  return hipGetLastError();
}

template <int S, int NT, int NR>
int occupancy_t() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<S, NT, NR>, NT, 0) != hipSuccess) return 1;
  return nb > 0 ? nb : 1;
}

}  // namespace

// Supported persistent-kernel shapes: segment bytes x workgroup threads x
// row-quad groups per wave step.
bool batch_config_ok(int seg, int threads, int groups) {
  const bool t_ok = threads == 256 || threads == 512 || threads == 1024;
  const bool g_ok = groups == 1 || groups == 2 || groups == 4;
  if (seg == 8 || seg == 16) return t_ok && g_ok;
  return seg == 32 && threads == 256 && groups == 1;
}

#define FS3D_BATCH_DISPATCH(SEG, NT, NR, CALL)     \
  switch ((SEG) * 100000 + (NT) * 10 + (NR)) {      \
    case 802561: CALL(8, 256, 1);                   \
    case 802562: CALL(8, 256, 2);                   \
    case 802564: CALL(8, 256, 4);                   \
    case 805121: CALL(8, 512, 1);                   \
    case 805122: CALL(8, 512, 2);                   \
    case 805124: CALL(8, 512, 4);                   \
    case 810241: CALL(8, 1024, 1);                  \
    case 810242: CALL(8, 1024, 2);                  \
    case 810244: CALL(8, 1024, 4);                  \
    case 1602561: CALL(16, 256, 1);                 \
    case 1602562: CALL(16, 256, 2);                 \
    case 1602564: CALL(16, 256, 4);                 \
    case 1605121: CALL(16, 512, 1);                 \
    case 1605122: CALL(16, 512, 2);                 \
    case 1605124: CALL(16, 512, 4);                 \
    case 1610241: CALL(16, 1024, 1);                \
    case 1610242: CALL(16, 1024, 2);                \
    case 1610244: CALL(16, 1024, 4);                \
    default: CALL(32, 256, 1);                      \
  }

int batch_occupancy(int seg, int threads, int groups) {
#define OCC(S_, N_, R_) return occupancy_t<S_, N_, R_>()
  FS3D_BATCH_DISPATCH(seg, threads, groups, OCC)
#undef OCC
}

hipError_t launch_batch(int seg, int threads, int groups, BWorld *worlds, const int *order, int n, int nticks,
                        u32 *queue, const u64 *rule, u64 *hashes, int flags, int grid, hipStream_t s) {
  hipLaunchKernelGGL(reset_u32_kernel, dim3(1), dim3(64), 0, s, queue);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
#define LB(S_, N_, R_) \
  return launch_batch_t<S_, N_, R_>(worlds, order, n, nticks, queue, rule, hashes, flags, grid, s)
  FS3D_BATCH_DISPATCH(seg, threads, groups, LB)
#undef LB
}

hipError_t launch_batch_hash(const BWorld *worlds, int n, u64 *hashes, hipStream_t s) {
  const int grid = n < 8192 ? n : 8192;
  hipLaunchKernelGGL(batch_hash_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, worlds, n, hashes);
  return hipGetLastError();
}

hipError_t launch_batch_fill_rain(const BWorld *worlds, int n, const u64 *seeds, const u64 *thr, size_t max_cells,
                                  hipStream_t s) {
  unsigned gx = (unsigned)std::min<size_t>((max_cells + 255) / 256, 64);
  unsigned gy = (unsigned)std::min(n, 65535);
  hipLaunchKernelGGL(batch_fill_rain_kernel, dim3(gx ? gx : 1, gy ? gy : 1), dim3(256), 0, s, worlds, n, seeds, thr);
  return hipGetLastError();
}

}  // namespace fs3d
