// Wide multi-world tick: one launch advances a set of worlds by one tick each,
// spreading all their row quads over the whole GPU.
//
// A launch is described by a table of world entries, a prefix of their
// workgroup counts for this launch (bpre[i] = first workgroup of entry i, each
// entry padded to whole 4-wave workgroups; an entry that sits the launch out
// has none) and blk2ent, the entry of every workgroup, expanded on the device
// from the prefix.  Each workgroup reads its entry with one scalar load and its
// waves run the ordinary row-quad tick in place.  An entry whose final tick
// this is (k + 1 == entry.nt with entry.fin set) hashes its new state; every
// wave adds its share to hashes[entry.slot] with one atomic.
//
// Used by the serving loop (continuous batching at tick granularity) and as the
// batch's third mode.
#include "fs3d_batch.h"
#include "tick_core.h"

namespace fs3d {
namespace {

using namespace tickcore;

// chg (optional): per hash slot, four change flags indexed by tick mod 4.  A
// launch at tick t reads the flags of ticks t-1 and t-2; when both are clear the
// world is a fixed point of both parities and the tick is not run, unless it is
// the world's final, hashed one.  Otherwise the tick runs and sets flag t when a
// cell changed.  The first workgroup of each entry clears flag t+1, which no wave
// of this launch reads or writes.  Admission sets all four flags.
template <int S>
__global__ __launch_bounds__(256) void wide_kernel(const BWorld *__restrict__ ents, const long long *__restrict__ bpre,
                                                   const int *__restrict__ blk2ent, int k, const u64 *__restrict__ rule,
                                                   u64 *hashes, u32 *chg) {
  const int e = __builtin_amdgcn_readfirstlane(blk2ent[blockIdx.x]);
  const BWorld w = ents[e];
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long long wave = ((long long)blockIdx.x - bpre[e]) * 4 + wid;
  const u32 t = (u32)(w.tick + (u64)k);
  const bool final_hash = k + 1 == w.nt && w.fin && hashes;
  u32 *flags = chg ? chg + (size_t)w.slot * 4 : nullptr;
  if (flags) {
    if (wave == 0 && threadIdx.x == 0) flags[(t + 1u) & 3u] = 0;
    if (!final_hash && flags[(t - 1u) & 3u] == 0 && flags[(t - 2u) & 3u] == 0) return;  // settled world
  }
  if (wave >= world_waves_dev(w.ny, w.nz, w.rqpw, (int)(t & 1u))) return;  // wave-uniform; no barriers below
  const int lane = threadIdx.x & 63;
  bool ch = false;
  if (final_hash) {
    u64 acc = tick_wave<S, true, true>(grid_of(w), grid_of(w), w.nx, w.ny, w.nz, w.lr, w.rqpw, t, wave, lane, rule, &ch);
    acc = wave_sum_u64(acc);
    if (lane == 0 && acc) atomicAdd((unsigned long long *)(hashes + w.slot), (unsigned long long)acc);
  } else {
    tick_wave<S, false, true>(grid_of(w), grid_of(w), w.nx, w.ny, w.nz, w.lr, w.rqpw, t, wave, lane, rule, &ch);
  }
  if (flags && __ballot(ch) && lane == 0) flags[t & 3u] = 1;
}

// blk2ent[b] = i for every workgroup b of entry i.
__global__ __launch_bounds__(256) void expand_kernel(const long long *__restrict__ bpre, int n, int *blk2ent) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (long long b = bpre[i]; b < bpre[i + 1]; b++) blk2ent[b] = i;
}

// The workgroup -> entry tables of K launches at once: launch k's workgroups are
// [koff[k], koff[k+1]) of blk2ent and its prefix is bpre[k*(n+1) .. k*(n+1)+n];
// one thread per workgroup finds its entry by binary search (the last i with
// bpre[i] <= b: an entry with blocks covering b has bpre[i+1] > b).
__global__ __launch_bounds__(256) void expand_round_kernel(const long long *__restrict__ bpre, int n, int K,
                                                           const long long *__restrict__ koff, long long total,
                                                           int *blk2ent) {
  for (long long gb = (long long)blockIdx.x * blockDim.x + threadIdx.x; gb < total;
       gb += (long long)gridDim.x * blockDim.x) {
    int k = 0;
    while (k + 1 < K && koff[k + 1] <= gb) k++;
    const long long b = gb - koff[k];
    const long long *pre = bpre + (size_t)k * (n + 1);
    int lo = 0, hi = n;  // pre[lo] <= b < pre[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (pre[mid] <= b) lo = mid; else hi = mid;
    }
    blk2ent[gb] = lo;
  }
}

// hashes[ents[i].slot] = 0 for i < n (before a final tick accumulates into them).
__global__ void zero_slots_kernel(u64 *hashes, const BWorld *__restrict__ ents, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) hashes[ents[i].slot] = 0;
}

}  // namespace

hipError_t launch_wide(int seg, const BWorld *ents, const long long *bpre, const int *blk2ent, long long blocks, int k,
                       const u64 *rule, u64 *hashes, hipStream_t s, u32 *chg) {
  if (blocks <= 0) return hipSuccess;
  const dim3 grid((unsigned)blocks);
  switch (seg) {
    case 8: hipLaunchKernelGGL(wide_kernel<8>, grid, dim3(256), 0, s, ents, bpre, blk2ent, k, rule, hashes, chg); break;
    case 16:
      hipLaunchKernelGGL(wide_kernel<16>, grid, dim3(256), 0, s, ents, bpre, blk2ent, k, rule, hashes, chg);
      break;
    default:
      hipLaunchKernelGGL(wide_kernel<32>, grid, dim3(256), 0, s, ents, bpre, blk2ent, k, rule, hashes, chg);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_expand(const long long *bpre, int n, int *blk2ent, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
  hipLaunchKernelGGL(expand_kernel, dim3(grid), dim3(256), 0, s, bpre, n, blk2ent);
  return hipGetLastError();
}

hipError_t launch_expand_round(const long long *bpre, int n, int K, const long long *koff, long long total,
                               int *blk2ent, hipStream_t s) {
  if (total <= 0 || n <= 0) return hipSuccess;
  const long long want = (total + 255) / 256;
  const unsigned grid = (unsigned)(want < 4096 ? want : 4096);
  hipLaunchKernelGGL(expand_round_kernel, dim3(grid), dim3(256), 0, s, bpre, n, K, koff, total, blk2ent);
  return hipGetLastError();
}

hipError_t launch_zero_slots(u64 *hashes, const BWorld *ents, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_slots_kernel, dim3((n + 255) / 256), dim3(256), 0, s, hashes, ents, n);
  return hipGetLastError();
}

long long world_waves(const BWorld &w, int p) { return world_waves_dev(w.ny, w.nz, w.rqpw, p); }

long long world_blocks(const BWorld &w, int p) { return (world_waves(w, p) + 3) / 4; }

}  // namespace fs3d
