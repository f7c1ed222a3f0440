// fs3d tick kernels for gfx950 (CDNA4, wave64).
//
// One tick updates every 2x2x2 block of the grid (blocks shifted by (-p,-p,-p),
// p = tick parity) from the source buffer into the destination buffer.  The
// reference has no tick; its engine loop (src/engine/engine.cpp:59-69) is the
// call site this replaces.
//
// The per-wave work (row quads, DPP exchange, table lookup) lives in
// tick_core.h; this file holds the single-world launches, the hash passes and
// small helper kernels.
#include "tick_core.h"

namespace fs3d {

namespace {

using namespace tickcore;

struct TickArgs {
  const u8 *src;
  u8 *dst;
  int nx, ny, nz;
  int lanes_per_row, rq_per_wave;
  u32 t;
  u32 k_off;
  const u64 *tick_dev;  // when set, tick = *tick_dev + k_off (graph replays)
  const u64 *rule;      // 4 x 65536 entries
  u64 *parts;           // hash partials (HASH only)
};

template <int S, bool HASH>
__global__ __launch_bounds__(256) void tick_kernel(TickArgs a) {
  const int lane = threadIdx.x & 63;
  const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const u32 t = a.tick_dev ? (u32)(*a.tick_dev + a.k_off) : a.t;
  const int p = (int)(t & 1u);
  const long long nrq = (long long)((a.ny + p + 1) >> 1) * ((a.nz + p + 1) >> 1);
  // Whole waves past the end leave early; the hashing variant keeps them until
  // its workgroup barrier (they load and store nothing).
  if (!HASH && wave * a.rq_per_wave >= nrq) return;
  u64 acc = tick_wave<S, HASH>(a.src, a.dst, a.nx, a.ny, a.nz, a.lanes_per_row, a.rq_per_wave, t, wave, lane,
                               a.rule);
  if constexpr (HASH) {
    __shared__ u64 wsum[4];
    acc = wave_sum_u64(acc);
    if (lane == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      const u64 tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
      if (tot) atomicAdd((unsigned long long *)(a.parts + (size_t)(blockIdx.x % kHashParts) * kHashStride),
                         (unsigned long long)tot);
    }
  }
}

__global__ __launch_bounds__(256) void hash_kernel(const u8 *g, size_t n, u64 *parts) {
  u64 acc = 0;
  const size_t nchunk = (n + 15) / 16;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nchunk;
       c += (size_t)gridDim.x * blockDim.x) {
    const size_t base = c * 16;
    const u32x4 v = *reinterpret_cast<const u32x4 *>(g + base);  // buffers are 256-byte aligned
    const u32 d[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const u32 b = (d[k >> 2] >> (8 * (k & 3))) & 0xFFu;
      if (b && base + k < n) acc += dmix64(((u64)(base + k) << 8) | b);
    }
  }
  __shared__ u64 wsum[4];
  acc = wave_sum_u64(acc);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const u64 tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (tot) atomicAdd((unsigned long long *)(parts + (size_t)(blockIdx.x % kHashParts) * kHashStride),
                       (unsigned long long)tot);
  }
}

// Sums the partial slots into *out and clears them for the next hash.
__global__ __launch_bounds__(256) void hash_finalize_kernel(u64 *parts, u64 *out) {
  u64 acc = 0;
  for (int i = threadIdx.x; i < kHashParts; i += 256) {
    acc += parts[(size_t)i * kHashStride];
    parts[(size_t)i * kHashStride] = 0;
  }
  __shared__ u64 wsum[4];
  acc = wave_sum_u64(acc);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ void set_u64_kernel(u64 *p, u64 v) {
  if (threadIdx.x == 0) *p = v;
}

__global__ __launch_bounds__(256) void counts_kernel(const u8 *g, size_t n, u64 *counts) {
  __shared__ u32 h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    atomicAdd(&h[g[i]], 1u);
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd((unsigned long long *)&counts[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

template <int S>
hipError_t launch_s(const TickShape &sh, const TickArgs &a, bool hash, hipStream_t s) {
  if (hash)
    hipLaunchKernelGGL((tick_kernel<S, true>), dim3(sh.blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((tick_kernel<S, false>), dim3(sh.blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace

int plan_tick(int nx, int ny, int nz, int seg_req, TickShape *out) {
  int seg = seg_req;
  if (seg == 0) seg = nx <= 1024 ? 16 : (nx <= 2048 ? 32 : 64);
  if (seg != 8 && seg != 16 && seg != 32 && seg != 64)
    return set_error(FS3D_ERR_INVALID, "segment must be 8, 16, 32 or 64 bytes (got %d)", seg);
  const int lpr = (nx + seg - 1) / seg;
  if (lpr > 64)
    return set_error(FS3D_ERR_INVALID, "nx=%d needs %d lanes per row with %d-byte segments (max 64)", nx,
                     lpr, seg);
  TickShape sh;
  sh.seg = seg;
  sh.lanes_per_row = lpr;
  sh.rq_per_wave = 64 / lpr;
  const long long nrq = (long long)((ny + 2) >> 1) * ((nz + 2) >> 1);
  sh.waves = (nrq + sh.rq_per_wave - 1) / sh.rq_per_wave;
  const long long blocks = (sh.waves + 3) / 4;
  if (blocks > 0x7FFFFFFFLL) return set_error(FS3D_ERR_INVALID, "grid too large");
  sh.blocks = (unsigned)blocks;
  *out = sh;
  return FS3D_OK;
}

hipError_t launch_tick(const TickShape &sh, const u8 *src, u8 *dst, int nx, int ny, int nz, u32 t,
                       const u64 *tick_dev, u32 k_off, const u64 *rule, u64 *hash_parts,
                       hipStream_t s) {
  TickArgs a;
  a.src = src;
  a.dst = dst;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.lanes_per_row = sh.lanes_per_row;
  a.rq_per_wave = sh.rq_per_wave;
  a.t = t;
  a.k_off = k_off;
  a.tick_dev = tick_dev;
  a.rule = rule;
  a.parts = hash_parts;
  const bool hash = hash_parts != nullptr;
  switch (sh.seg) {
    case 8: return launch_s<8>(sh, a, hash, s);
    case 16: return launch_s<16>(sh, a, hash, s);
    case 32: return launch_s<32>(sh, a, hash, s);
    default: return launch_s<64>(sh, a, hash, s);
  }
}

hipError_t launch_hash(const u8 *g, size_t n, u64 *parts, hipStream_t s) {
  const size_t nchunk = (n + 15) / 16;
  size_t blocks = (nchunk + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(hash_kernel, dim3((unsigned)blocks), dim3(256), 0, s, g, n, parts);
  return hipGetLastError();
}

hipError_t launch_hash_finalize(u64 *parts, u64 *out, hipStream_t s) {
  hipLaunchKernelGGL(hash_finalize_kernel, dim3(1), dim3(256), 0, s, parts, out);
  return hipGetLastError();
}

hipError_t launch_set_u64(u64 *p, u64 v, hipStream_t s) {
  hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

hipError_t launch_counts(const u8 *g, size_t n, u64 *counts, hipStream_t s) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(counts_kernel, dim3((unsigned)blocks), dim3(256), 0, s, g, n, counts);
  return hipGetLastError();
}

}  // namespace fs3d
