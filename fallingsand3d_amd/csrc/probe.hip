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

__global__ __launch_bounds__(256) void copy_kernel(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                   size_t n16) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}

}  // namespace
}  // namespace fs3d

using namespace fs3d;

extern "C" int fs3d_copy_probe(size_t nbytes, int32_t iters, double *ms_per_copy) {
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
  if (e == hipSuccess) {
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, s, (const u32x4 *)a, (u32x4 *)b, n16);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int i = 0; i < iters && e == hipSuccess; i++) {
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, s, (const u32x4 *)a, (u32x4 *)b, n16);
    e = hipGetLastError();
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
  if (first != hipSuccess) return set_error(FS3D_ERR_HIP, "copy probe failed: %s", hipGetErrorString(first));
  *ms_per_copy = (double)ms / iters;
  return clear_error();
}
