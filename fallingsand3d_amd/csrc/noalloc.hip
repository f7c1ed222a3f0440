// fs3d_grid_run_noalloc: dense ticks of a grid held in caller-owned device
// memory, for callers that cannot allocate device memory.  The device memory the
// kernels need beyond the caller's buffers -- the fused hash's partial sums and
// one slot per requested hash -- lives in one file-scope __device__ array of
// fixed size, so the entry point calls no hipMalloc/hipFree.
//
// Every GPU holds its own copy of the array at its own address.  ensure_device
// (the library's per-device set-up, run by fs3d_init and the first world or
// batch creation) calls noalloc_setup with the device current; it looks the
// copy's address up once and keeps it per device.  Host code never passes the
// array's name to a kernel or copy: in host code that name is a stand-in, not a
// device address.
//
// All calls on one GPU share its copy, which keeps the last call's contents.
// Each call clears the partial sums before its first tick.  Calls that overlap
// on one GPU (other threads, other streams) race on the array.
#include <hip/hip_runtime.h>

#include "fs3d_internal.h"

namespace fs3d {

namespace {

constexpr size_t kPartWords = (size_t)kHashParts * kHashStride;
constexpr size_t kScratchWords = kPartWords + (size_t)FS3D_NOALLOC_MAX_HASHES;

// [0, kPartWords): hash partial sums (kHashParts slots, kHashStride words
// apart); then FS3D_NOALLOC_MAX_HASHES hash outputs.  96 KiB per device.
__device__ __align__(256) u64 scratch[kScratchWords];

u64 *g_scratch[kMaxDevices] = {};  // per device: the address of its copy

}  // namespace

int noalloc_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return set_error(FS3D_ERR_INVALID, "device %d out of range", dev);
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_error(FS3D_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                     hipGetErrorString(e));
  // partial sums start at zero (a fill kernel, not a memset node)
  FS3D_HIP(launch_fill_u64(static_cast<u64 *>(p), kScratchWords, 0, nullptr));
  FS3D_HIP(hipStreamSynchronize(nullptr));
  g_scratch[dev] = static_cast<u64 *>(p);
  return FS3D_OK;
}

}  // namespace fs3d

using namespace fs3d;

extern "C" int fs3d_grid_run_noalloc(uint8_t *a, uint8_t *b, const int32_t dims[3], uint64_t tick, int32_t nticks,
                                     int32_t hash_every, uint64_t *hashes, int32_t *final_index, void *stream) {
  if (!a || !b || !dims || !final_index) return set_error(FS3D_ERR_INVALID, "a, b, dims and final_index are required");
  if (a == b) return set_error(FS3D_ERR_INVALID, "a and b must be two distinct buffers");
  if (nticks < 0 || hash_every < 0) return set_error(FS3D_ERR_INVALID, "nticks/hash_every < 0");
  for (int i = 0; i < 3; i++)
    if (dims[i] < 1) return set_error(FS3D_ERR_INVALID, "dims must be >= 1 (got %d,%d,%d)", dims[0], dims[1], dims[2]);
  if (dims[0] > kMaxNx) return set_error(FS3D_ERR_INVALID, "nx=%d exceeds %d", dims[0], kMaxNx);
  if ((u64)dims[0] * (u64)dims[1] * (u64)dims[2] > (1ull << 40)) return set_error(FS3D_ERR_INVALID, "grid too large");
  const int nh = hash_every > 0 ? nticks / hash_every : 0;
  if (nh > FS3D_NOALLOC_MAX_HASHES)
    return set_error(FS3D_ERR_INVALID,
                     "noalloc run: %d hashes requested; the static scratch holds at most %d (FS3D_NOALLOC_MAX_HASHES)",
                     nh, FS3D_NOALLOC_MAX_HASHES);
  if (nh > 0 && !hashes) return set_error(FS3D_ERR_INVALID, "hashes is NULL");
  int dev = -1;
  FS3D_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices || !g_scratch[dev] || !device_rule_table(dev))
    return set_error(FS3D_ERR_STATE, "device %d is not set up: call fs3d_init(%d) once before this entry point", dev,
                     dev);
  TickShape sh;
  if (int rc = plan_tick(dims[0], dims[1], dims[2], 0, &sh)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  u64 *parts = g_scratch[dev];
  u64 *slots = parts + kPartWords;
  const u64 *rule = device_rule_table(dev);
  // a call that stopped between a hash tick and its finaliser leaves partial sums
  FS3D_HIP(launch_fill_u64(parts, kPartWords, 0, s));
  u8 *buf[2] = {a, b};
  int src = 0;
  for (int k = 0; k < nticks; k++) {
    const bool hashed = hash_every > 0 && (k + 1) % hash_every == 0;
    FS3D_HIP(launch_tick(sh, buf[src], buf[src ^ 1], dims[0], dims[1], dims[2], tick + (u64)k, nullptr, 0, rule,
                         hashed ? parts : nullptr, s));
    if (hashed) FS3D_HIP(launch_hash_finalize(parts, slots + ((k + 1) / hash_every - 1), s));
    src ^= 1;
  }
  if (nh > 0) FS3D_HIP(hipMemcpyAsync(hashes, slots, (size_t)nh * sizeof(u64), hipMemcpyDeviceToHost, s));
  FS3D_HIP(hipStreamSynchronize(s));
  *final_index = src;
  return clear_error();
}
