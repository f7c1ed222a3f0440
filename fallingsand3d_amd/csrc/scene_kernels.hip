// Device-side scene fills (rain, boxes, pillars) for the seeded workloads.
#include "fs3d_internal.h"

namespace fs3d {

namespace {

__device__ __forceinline__ u64 dmix64(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Rain fill: STATIC floor at z = 0; cells with z >= max(1, nz/2) are SAND when the
// counter-based draw mix64(seed + (i+1)*gamma) is below thr; all others EMPTY.
__global__ __launch_bounds__(256) void fill_rain_kernel(u8 *g, int nx, int ny, int nz, u64 seed, u64 thr) {
  const size_t layer = (size_t)nx * ny, n = layer * nz;
  const int zr = max(1, nz >> 1);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t z = i / layer;
    u8 v = FS3D_EMPTY;
    if (z == 0)
      v = FS3D_STATIC;
    else if (z >= (size_t)zr && dmix64(seed + (i + 1) * kGamma) < thr)
      v = FS3D_SAND;
    g[i] = v;
  }
}

__global__ __launch_bounds__(256) void fill_box_kernel(u8 *g, int nx, int ny, int x0, int y0, int z0, int sx,
                                                       int sy, int sz, u8 v) {
  const size_t cnt = (size_t)sx * sy * sz;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % sx);
    const size_t r = i / sx;
    const int y = (int)(r % sy), z = (int)(r / sy);
    g[((size_t)(z0 + z) * ny + (y0 + y)) * nx + (x0 + x)] = v;
  }
}

// One grid column per blockIdx.y = pillar; threads sweep the pillar's bounding
// prism and paint SAND inside the disc (x-cx)^2 + (y-cy)^2 <= r^2, 1 <= z <= zt.
// All pillars paint the same value, so their order does not matter.
__global__ __launch_bounds__(256) void fill_pillars_kernel(u8 *g, int nx, int ny, const Pillar *pillars) {
  const Pillar q = pillars[blockIdx.y];
  const int side = 2 * q.r + 1;
  const size_t cnt = (size_t)side * side * q.zt;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x) {
    const int dx = (int)(i % side) - q.r;
    const size_t rest = i / side;
    const int dy = (int)(rest % side) - q.r;
    const int z = 1 + (int)(rest / side);
    const int x = q.cx + dx, y = q.cy + dy;
    if (x < 0 || x >= nx || y < 0 || y >= ny) continue;
    if ((long long)dx * dx + (long long)dy * dy > (long long)q.r * q.r) continue;
    g[((size_t)z * ny + y) * nx + x] = FS3D_SAND;
  }
}

unsigned grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return b ? (unsigned)b : 1u;
}

}  // namespace

hipError_t launch_fill_rain(u8 *g, int nx, int ny, int nz, u64 seed, u64 thr, hipStream_t s) {
  hipLaunchKernelGGL(fill_rain_kernel, dim3(grid_for((size_t)nx * ny * nz)), dim3(256), 0, s, g, nx, ny, nz, seed, thr);
  return hipGetLastError();
}

hipError_t launch_fill_box(u8 *g, int nx, int ny, int x0, int y0, int z0, int sx, int sy, int sz, u8 v,
                           hipStream_t s) {
  hipLaunchKernelGGL(fill_box_kernel, dim3(grid_for((size_t)sx * sy * sz)), dim3(256), 0, s, g, nx, ny, x0, y0, z0,
                     sx, sy, sz, v);
  return hipGetLastError();
}

hipError_t launch_fill_pillars(u8 *g, int nx, int ny, const Pillar *dev_pillars, int n, int max_r, int max_zt,
                               hipStream_t s) {
  const size_t per = (size_t)(2 * max_r + 1) * (2 * max_r + 1) * (size_t)max_zt;
  unsigned gx = grid_for(per);
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(fill_pillars_kernel, dim3(gx, (unsigned)n), dim3(256), 0, s, g, nx, ny, dev_pillars);
  return hipGetLastError();
}

}  // namespace fs3d
