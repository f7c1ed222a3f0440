// Bandwidth calibration: a plain streaming copy, the practical HBM ceiling the
// tick kernel is compared with and the known-size kernel that scales rocprof's
// FETCH_SIZE / WRITE_SIZE.  One 16-byte load and store per lane and one lane
// per 16 bytes (no grid-stride loop), with the same (temporal) load and store
// instructions as the tick kernels, so that the counters' scale carries over.
// Shapes measured on MI355X for an 8 GiB copy: this one 6.25 TB/s; the same with
// nontemporal hints 6.42; a grid-stride loop over 256K workgroups 5.66;
// persistent grids of 4-32 workgroups per CU 4.6-5.4; 2-8 loads in flight per
// lane 5.4-5.9.
#include <algorithm>

#include "fs3d_internal.h"

namespace fs3d {
namespace {

typedef u32 u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__global__ __launch_bounds__(256) void copy_kernel(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                   size_t n16) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16) {
    if constexpr (NT)
      __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
    else
      dst[i] = src[i];
  }
}

__device__ __forceinline__ u32 probe_mix(u32 h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

// The access shape of the packed skipping tick without its arithmetic: item i
// is the k-th of `per_rq` (k = i % per_rq) 32-byte segments of row quad i /
// per_rq of an n^3 grid (even parity: rows (2yq, 2yq + 1) x (2zq, 2zq + 1)), at
// segment k * span + h, span = floor(64 / per_rq), h a hash in [0, span), so the
// segments are distinct, ascending and scattered.  One lane per item copies
// the four rows' 32 bytes (SECTOR false) or their whole 64-byte sectors
// (SECTOR true) from src to dst.  n = 2048: 64 segments per row.
template <bool SECTOR>
__global__ __launch_bounds__(256) void scatter_kernel(const u8 *__restrict__ src, u8 *__restrict__ dst, int n,
                                                      size_t nitems, int per_rq) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nitems) return;
  const size_t rq = i / (size_t)per_rq;
  const int k = (int)(i % (size_t)per_rq);
  const int span = 64 / per_rq;
  const int seg = k * span + (int)(probe_mix((u32)rq * 131u + (u32)k) % (u32)span);
  const int nrq_y = n / 2;
  const size_t yq = rq % (size_t)nrq_y, zq = rq / (size_t)nrq_y;
  const size_t base = ((2 * zq) * (size_t)n + 2 * yq) * (size_t)n;
  const size_t off[4] = {base, base + n, base + (size_t)n * n, base + (size_t)n * n + n};
  constexpr int kV = SECTOR ? 4 : 2;  // 16-byte vectors per row
  const size_t x = SECTOR ? (size_t)(seg >> 1) * 64 : (size_t)seg * 32;
  u32x4 v[4][kV];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < kV; c++) v[r][c] = *(const u32x4 *)(src + off[r] + x + 16 * c);
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < kV; c++) *(u32x4 *)(dst + off[r] + x + 16 * c) = v[r][c];
}

// The dense chunked tick's access shape without its arithmetic, on an n^3 grid
// pair (even parity): wave w copies x-chunk w % 2 (1 KiB: 16 bytes per lane) of
// the four rows of row quad w / 2, with streaming (nontemporal) loads (NTL)
// and stores (NTS) or temporal ones.  n = 2048: two chunks per row.
template <bool NTL, bool NTS>
__global__ __launch_bounds__(256) void rowquad_chunk_kernel(const u8 *__restrict__ src, u8 *__restrict__ dst,
                                                            int n) {
  const int lane = threadIdx.x & 63;
  const size_t w = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const size_t chunk = w & 1, rq = w >> 1;
  const size_t nrq_y = (size_t)n / 2;
  const size_t yq = rq % nrq_y, zq = rq / nrq_y;
  const size_t base = ((2 * zq) * (size_t)n + 2 * yq) * (size_t)n + chunk * 1024 + (size_t)lane * 16;
  const size_t off[4] = {base, base + n, base + (size_t)n * n, base + (size_t)n * n + n};
  u32x4 v[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const u32x4 *p = (const u32x4 *)(src + off[r]);
    v[r] = NTL ? __builtin_nontemporal_load(p) : *p;
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    u32x4 *p = (u32x4 *)(dst + off[r]);
    if constexpr (NTS)
      __builtin_nontemporal_store(v[r], p);
    else
      *p = v[r];
  }
}

}  // namespace
}  // namespace fs3d

using namespace fs3d;

namespace {

int copy_probe(size_t nbytes, int32_t iters, double *ms_per_copy, bool streaming) {
  if (!ms_per_copy || iters < 1 || nbytes < 16 || (nbytes & 15))
    return set_error(FS3D_ERR_INVALID, "copy probe needs nbytes >= 16, a multiple of 16, and iters >= 1");
  const size_t n16 = nbytes / 16;
  const size_t nblocks = (n16 + 255) / 256;  // one lane per 16 bytes, one launch
  if (nblocks > 0xFFFFFFFFull) return set_error(FS3D_ERR_INVALID, "copy probe: %zu bytes exceed one launch", nbytes);
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (int rc = ensure_device(dev)) return rc;
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, nbytes);
  if (e == hipSuccess) e = hipMalloc(&b, nbytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(a, 1, nbytes, s);
  const unsigned blocks = (unsigned)nblocks;
  auto launch = [&]() {
    if (streaming)
      hipLaunchKernelGGL(copy_kernel<true>, dim3(blocks), dim3(256), 0, s, (const u32x4 *)a, (u32x4 *)b, n16);
    else
      hipLaunchKernelGGL(copy_kernel<false>, dim3(blocks), dim3(256), 0, s, (const u32x4 *)a, (u32x4 *)b, n16);
    return hipGetLastError();
  };
  if (e == hipSuccess) e = launch();  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int i = 0; i < iters && e == hipSuccess; i++) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  const hipError_t first = e;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (first != hipSuccess) return set_error(FS3D_ERR_HIP, "copy probe failed: %s", hipGetErrorString(first));
  *ms_per_copy = (double)ms / iters;
  return clear_error();
}

}  // namespace

extern "C" int fs3d_copy_probe(size_t nbytes, int32_t iters, double *ms_per_copy) {
  return copy_probe(nbytes, iters, ms_per_copy, false);
}

extern "C" int fs3d_copy_probe_streaming(size_t nbytes, int32_t iters, double *ms_per_copy) {
  return copy_probe(nbytes, iters, ms_per_copy, true);
}

namespace {

// Mean ms per launch of `launch(src, dst, stream)` over a 2048^3 grid pair
// (16 GiB, allocated and freed here), after one warm-up launch.
template <class F>
int time_on_pair(int32_t iters, const char *what, F launch, double *ms_per_pass) {
  constexpr size_t nbytes = (size_t)2048 * 2048 * 2048;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (int rc = ensure_device(dev)) return rc;
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, nbytes);
  if (e == hipSuccess) e = hipMalloc(&b, nbytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(a, 1, nbytes, s);
  if (e == hipSuccess) e = hipMemsetAsync(b, 0, nbytes, s);
  for (int i = 0; i <= iters && e == hipSuccess; i++) {  // launch 0 warms up
    if (i == 1) e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = launch((const u8 *)a, (u8 *)b, s);
  }
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  const hipError_t first = e;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (first != hipSuccess) return set_error(FS3D_ERR_HIP, "%s failed: %s", what, hipGetErrorString(first));
  *ms_per_pass = (double)ms / iters;
  return clear_error();
}

}  // namespace

extern "C" int fs3d_scatter_probe(int32_t per_rq, int32_t sector, int32_t iters, double *ms_per_pass,
                                  double *item_bytes) {
  constexpr int n = 2048;
  if (!ms_per_pass || !item_bytes || iters < 1 || per_rq < 1 || per_rq > 64)
    return set_error(FS3D_ERR_INVALID, "scatter probe needs 1 <= per_rq <= 64 and iters >= 1");
  const size_t nitems = (size_t)(n / 2) * (n / 2) * (size_t)per_rq;
  const unsigned blocks = (unsigned)((nitems + 255) / 256);
  auto launch = [&](const u8 *a, u8 *b, hipStream_t s) {
    if (sector)
      hipLaunchKernelGGL(scatter_kernel<true>, dim3(blocks), dim3(256), 0, s, a, b, n, nitems, per_rq);
    else
      hipLaunchKernelGGL(scatter_kernel<false>, dim3(blocks), dim3(256), 0, s, a, b, n, nitems, per_rq);
    return hipGetLastError();
  };
  if (int rc = time_on_pair(iters, "scatter probe", launch, ms_per_pass)) return rc;
  *item_bytes = (double)nitems * 4 * (sector ? 64 : 32) * 2;
  return clear_error();
}

extern "C" int fs3d_rowquad_probe(int32_t streaming_loads, int32_t streaming_stores, int32_t iters,
                                  double *ms_per_pass) {
  constexpr int n = 2048;
  if (!ms_per_pass || iters < 1) return set_error(FS3D_ERR_INVALID, "row-quad probe needs iters >= 1");
  const unsigned blocks = (unsigned)((size_t)(n / 2) * (n / 2) * 2 / 4);  // two chunk waves per row quad
  auto launch = [&](const u8 *a, u8 *b, hipStream_t s) {
    if (streaming_loads && streaming_stores)
      hipLaunchKernelGGL((rowquad_chunk_kernel<true, true>), dim3(blocks), dim3(256), 0, s, a, b, n);
    else if (streaming_loads)
      hipLaunchKernelGGL((rowquad_chunk_kernel<true, false>), dim3(blocks), dim3(256), 0, s, a, b, n);
    else if (streaming_stores)
      hipLaunchKernelGGL((rowquad_chunk_kernel<false, true>), dim3(blocks), dim3(256), 0, s, a, b, n);
    else
      hipLaunchKernelGGL((rowquad_chunk_kernel<false, false>), dim3(blocks), dim3(256), 0, s, a, b, n);
    return hipGetLastError();
  };
  return time_on_pair(iters, "row-quad probe", launch, ms_per_pass);
}
