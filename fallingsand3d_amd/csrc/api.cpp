// C ABI of the fs3d library: errors, device context, world grid, tick, hash,
// captured multi-tick runs and scene fills.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fs3d_internal.h"
#include "fs3d_render.h"

namespace fs3d {

namespace {
thread_local std::string g_error;

struct DeviceCtx {
  bool ready = false;
  u64 *rule = nullptr;  // 4 x 65536 x 8 bytes
};
DeviceCtx g_dev[kMaxDevices];
std::mutex g_dev_mutex;
}  // namespace

int set_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

int clear_error() {
  g_error.clear();
  return FS3D_OK;
}

int ensure_device(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return set_error(FS3D_ERR_INVALID, "device %d out of range", dev);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return set_error(FS3D_ERR_NODEVICE, "no HIP device is visible");
  if (dev >= count) return set_error(FS3D_ERR_NODEVICE, "device %d not present (%d visible)", dev, count);
  FS3D_HIP(hipSetDevice(dev));
  std::lock_guard<std::mutex> lock(g_dev_mutex);
  DeviceCtx &c = g_dev[dev];
  if (c.ready) return FS3D_OK;
  hipDeviceProp_t prop;
  FS3D_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(FS3D_ERR_NODEVICE, "device %d is %s; this library is built for gfx950 only", dev,
                     prop.gcnArchName);
  std::vector<u64> host((size_t)4 * 65536);
  build_rule_entries(host.data());
  FS3D_HIP(hipMalloc(&c.rule, host.size() * sizeof(u64)));
  FS3D_HIP(hipMemcpy(c.rule, host.data(), host.size() * sizeof(u64), hipMemcpyHostToDevice));
  if (int rc = noalloc_setup(dev)) return rc;
  c.ready = true;
  return FS3D_OK;
}

const u64 *device_rule_table(int dev) { return g_dev[dev].rule; }

}  // namespace fs3d

using namespace fs3d;

namespace {

int check_world(const fs3d_world *w) {
  if (!w) return set_error(FS3D_ERR_INVALID, "world handle is NULL");
  return FS3D_OK;
}

int ensure_small(fs3d_world *w, int slots) {
  // d_small[0] = device tick counter, d_small[1..slots] = hash outputs
  if (slots <= w->small_cap) return FS3D_OK;
  int cap = std::max(slots, 64);
  u64 *d = nullptr;
  FS3D_HIP(hipStreamSynchronize(w->stream));
  FS3D_HIP(hipMalloc(&d, (size_t)(cap + 1) * sizeof(u64)));
  FS3D_HIP(hipMemset(d, 0, (size_t)(cap + 1) * sizeof(u64)));
  if (w->d_small) FS3D_HIP(hipFree(w->d_small));
  w->d_small = d;
  w->small_cap = cap;
  // cached graphs captured the old pointers
  for (auto &kv : w->graphs) (void)hipGraphExecDestroy(kv.second);
  w->graphs.clear();
  if (w->h_pinned_cap < cap + 1) {
    if (w->h_pinned) FS3D_HIP(hipHostFree(w->h_pinned));
    w->h_pinned = nullptr;
    FS3D_HIP(hipHostMalloc(&w->h_pinned, (size_t)(cap + 1) * sizeof(u64), hipHostMallocDefault));
    w->h_pinned_cap = cap + 1;
  }
  return FS3D_OK;
}

void release(fs3d_world *w) {
  (void)hipSetDevice(w->dev);
  if (w->stream) (void)hipStreamSynchronize(w->stream);
  for (auto &kv : w->graphs) (void)hipGraphExecDestroy(kv.second);
  w->graphs.clear();
  for (int i = 0; i < 2; i++)
    if (w->buf[i]) (void)hipFree(w->buf[i]);
  if (w->d_parts) (void)hipFree(w->d_parts);
  if (w->d_rgb) (void)hipFree(w->d_rgb);
  if (w->d_hit) (void)hipFree(w->d_hit);
  if (w->d_face) (void)hipFree(w->d_face);
  if (w->d_occ) (void)hipFree(w->d_occ);
  if (w->d_dirty) (void)hipFree(w->d_dirty);
  if (w->d_small) (void)hipFree(w->d_small);
  for (int i = 0; i < 2; i++)
    if (w->d_masks[i]) (void)hipFree(w->d_masks[i]);
  if (w->d_stats) (void)hipFree(w->d_stats);
  if (w->d_list) (void)hipFree(w->d_list);
  if (w->d_ctl) (void)hipFree(w->d_ctl);
  if (w->h_pinned) (void)hipHostFree(w->h_pinned);
  if (w->ev0) (void)hipEventDestroy(w->ev0);
  if (w->ev1) (void)hipEventDestroy(w->ev1);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  delete w;
}

constexpr size_t kStatWords = (size_t)kHashParts * kHashStride + 1;

// Every change bit set: the next two ticks evaluate every segment, after which
// the masks describe real changes again.
int mark_all(fs3d_world *w) {
  w->occ_valid = false;  // a host-side change: the renderer rebuilds its maps
  for (int i = 0; i < 2; i++) FS3D_HIP(launch_fill_u64(w->d_masks[i], 2 * w->mask_words, ~0ull, w->stream));
  // work-list counts start from zero after any host-side change (tick, shape)
  FS3D_HIP(hipMemsetAsync(w->d_ctl, 0, 4 * sizeof(u32), w->stream));
  return FS3D_OK;
}

// Sets the change bits of every lane segment whose blocks contain cell (x,y,z),
// at both parities (segments x/S - 1 .. x/S + 1 of the row quads holding it).
int mark_cell(fs3d_world *w, int x, int y, int z) {
  w->occ_valid = false;
  long long idx[2];
  for (int q = 0; q < 2; q++) {
    const long long nyq = (w->ny + q + 1) >> 1;
    idx[q] = (long long)((z + q) >> 1) * nyq + ((y + q) >> 1);
  }
  const int s = x / w->shape.seg, lr = w->shape.lanes_per_row;
  u64 bits = 1ull << s;  // s <= 63
  if (s > 0) bits |= 1ull << (s - 1);
  if (s + 1 < lr) bits |= 1ull << (s + 1);
  FS3D_HIP(launch_mark_cells(w->d_masks[0], w->d_masks[1], idx[0], idx[1], bits, w->mask_words, w->stream));
  return FS3D_OK;
}

// Persistent waves of the listed tick: as many as stay resident at its register
// use, no more than one pass over every row quad needs.
// listed row quads per wave and pass of the packed tick: small chunks keep every
// resident wave busy on small lists (chunks of 8, 16, 32, 64 at 512^3: 53, 56,
// 74, 120 us per tick) at the cost of partly filled last groups
constexpr int kPackChunk = 8;

int size_list(fs3d_world *w) {
  int cus = 0;
  FS3D_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, w->dev));
  const long long need = ((long long)w->mask_words + w->shape.rq_per_wave - 1) / w->shape.rq_per_wave;
  const long long resident = (long long)std::max(cus, 1) * list_blocks_per_cu(w->shape.seg) * 4;
  long long lw = std::min<long long>(need, resident);
  lw = std::max<long long>((lw + 3) & ~3LL, 4);
  w->list_waves = (u32)lw;
  // the packed tick: a wave per chunk of listed row quads
  const long long pk_need = ((long long)w->mask_words + kPackChunk - 1) / kPackChunk;
  const long long pk_res = (long long)std::max(cus, 1) * pack_blocks_per_cu(w->shape.seg) * 4;
  long long pw = std::min<long long>(pk_need, pk_res);
  pw = std::max<long long>((pw + 3) & ~3LL, 4);
  w->pack_waves = (u32)pw;
  return FS3D_OK;
}

void drop_graphs(fs3d_world *w) { world_drop_graphs(w); }

// The schedule of the non-hash ticks of a skipping world: one wave per row-quad
// group for small grids (no plan pass), packed segments for large ones (2048^3
// pillar scene: 1.01 ms per tick against 1.37 ms listed; 512^3 box scene: waves
// 39 us, listed 44 us, packed 52 us).
int skip_schedule(const fs3d_world *w) {
  if (w->skip != FS3D_SKIP_AUTO) return w->skip;
  return w->shape.waves > (1ll << 17) ? FS3D_SKIP_SEGS : FS3D_SKIP_WAVES;
}

bool fuse_ok(const fs3d_world *w) { return !w->skip && w->fuse && tick2_supported(w->shape2, w->nx, nullptr); }

// Grids up to this many voxels keep both buffers within about twice the 256 MB
// MALL, where the single-tick kernel is not HBM-bound and fusion's redundant
// halo work costs more than the traffic it saves (512^3: 48.4 vs 47.7 us per
// tick; 1024^3: 324 vs 390 us; 2048^3: 2.40 vs 3.22 ms).
constexpr u64 kFuseMinVoxels = 1ull << 28;

// The fused pass's launch shape for rows up to 1024 bytes: the smallest segment
// that puts a whole row in one wave (it shares no state with the single-tick
// kernel's shape).  Longer rows run in 16-byte x-chunks and need no shape.  The
// default for pairing ticks is on for large grids whose rows the pass supports.
void plan_fused(fs3d_world *w) {
  const int seg = w->nx <= 512 ? 8 : 16;
  (void)plan_tick(w->nx, w->ny, w->nz, seg, &w->shape2);
  w->fuse = (u64)w->nx * (u64)w->ny * (u64)w->nz > kFuseMinVoxels && tick2_supported(w->shape2, w->nx, nullptr);
}

// z row-quads per workgroup of the fused two-tick pass.  Eight keep enough
// workgroups in flight on smaller grids (us per fused tick, z = 4 / 8 / 16 /
// 32: 512^3 47.3 / 48.4 / 49.2 / 69.5; 1024^3 336 / 324 / 326 / 341); the z
// halo costs 1/8 more tick-A rows.  Above 2^32 voxels there are workgroups
// enough for sixteen, whose halo costs half as much (2048^3 with streaming
// accesses: 2.30-2.31 against 2.36 ms per tick at eight and 2.33 at 32; 1024^3
// at sixteen: 0.319 against 0.312).
int tick2_zlen(const fs3d_world *w) { return (u64)w->nx * (u64)w->ny * (u64)w->nz > (1ull << 32) ? 16 : 8; }

// Dense single ticks of rows longer than 1024 bytes (with the automatic
// segment) run one wave per 16-byte x-chunk of 1 KiB instead of one wave per
// whole row quad (launch_tick_chunked): each wave moves a quarter of the bytes
// at nx = 2048, and the tick takes 3.21 instead of 3.30 ms at 2048^3 (four
// alternating runs each on one MI355X).  Shorter rows keep the row-quad kernel,
// which packs several row quads into a wave.  0: the row-quad kernel.
int dense_chunk_seg(const fs3d_world *w) { return w->seg_req == 0 && w->nx > 1024 ? 16 : 0; }

// Whole-row-quad ticks of grids up to kInPlaceMaxBytes update in place: every
// wave loads its row quads before it stores any of them, and the row quads of
// one tick share no rows.  One buffer instead of two keeps a 512^3 world within
// the 256 MB last-level cache (box scene, us per tick in place / from one buffer
// into the other: ticks 5-25 with the last hashed 47.4 / 50.4; ticks 100-1100
// 37.3 / 37.1; dense 42.9 / 47.3).  Larger grids live in HBM either way, where
// reading and writing the same lines is slower (1024^3 dense: 405.7 / 388.9).
// The packed schedule reads neighbouring segments' edge bytes from memory, and
// the chunked and fused dense kernels neighbouring chunks' and tiles' rows, so
// they always run from one buffer into the other.  A skipping world keeps one
// mode for all its ticks (in place unless packed): a skipped segment relies on
// the destination already holding its bytes.
constexpr size_t kInPlaceMaxBytes = size_t(1) << 27;

// Enqueues nticks ticks starting from buffer w->cur; *final receives the index of
// the buffer holding the result (a fused pair of dense ticks moves the state to
// the other buffer once, a single tick from one buffer into the other once, a
// tick in place not at all).  With launch = false it only computes *final.
int enqueue_ticks(fs3d_world *w, int nticks, int hash_every, const u64 *tick_dev, int *final, bool launch = true) {
  const u64 *rule = device_rule_table(w->dev);
  u64 *const *masks = w->skip ? w->d_masks : nullptr;
  const int sched = w->skip ? skip_schedule(w) : FS3D_SKIP_OFF;
  const bool fuse = fuse_ok(w);
  const bool inplace = w->n <= kInPlaceMaxBytes && !(masks && sched == FS3D_SKIP_SEGS);
  auto hashed_after = [&](int k) { return hash_every > 0 && (k + 1) % hash_every == 0; };
  int src = w->cur;
  for (int k = 0; k < nticks;) {
    if (fuse && k + 1 < nticks && !hashed_after(k) && !hashed_after(k + 1)) {
      if (launch)
        FS3D_HIP(launch_tick2(w->shape2, w->buf[src], w->buf[src ^ 1], w->nx, w->ny, w->nz, w->tick + (u64)k,
                              tick_dev, (u32)k, rule, tick2_zlen(w), w->stream));
      src ^= 1;
      k += 2;
      continue;
    }
    const bool hashed = hashed_after(k);
    const bool chunked = !masks && dense_chunk_seg(w);
    const bool packed = masks && sched == FS3D_SKIP_SEGS && !hashed;
    const int dst = inplace && !chunked && !packed ? src : src ^ 1;
    if (launch) {
      if (masks && sched == FS3D_SKIP_LISTED && !hashed) {
        // settled row quads are listed out; hash ticks read every segment anyway
        FS3D_HIP(launch_tick_listed(w->shape, w->buf[src], w->buf[dst], w->nx, w->ny, w->nz,
                                    w->tick + (u64)k, tick_dev, (u32)k, rule, w->stream, masks,
                                    w->d_stats, w->d_list, w->d_ctl, w->mask_words, w->list_waves));
      } else if (packed) {
        FS3D_HIP(launch_tick_packed(w->shape, w->buf[src], w->buf[dst], w->nx, w->ny, w->nz, w->tick + (u64)k,
                                    tick_dev, (u32)k, rule, w->stream, masks, w->d_stats, w->d_list, w->d_ctl,
                                    w->mask_words, w->pack_waves, kPackChunk));
      } else {
        if (chunked)
          FS3D_HIP(launch_tick_chunked(dense_chunk_seg(w), w->buf[src], w->buf[dst], w->nx, w->ny, w->nz,
                                       w->tick + (u64)k, tick_dev, (u32)k, rule, hashed ? w->d_parts : nullptr,
                                       w->stream));
        else
          FS3D_HIP(launch_tick(w->shape, w->buf[src], w->buf[dst], w->nx, w->ny, w->nz,
                               w->tick + (u64)k, tick_dev, (u32)k, rule,
                               hashed ? w->d_parts : nullptr, w->stream, masks, w->d_stats, w->d_ctl));
        if (hashed) FS3D_HIP(launch_hash_finalize(w->d_parts, w->d_small + (k + 1) / hash_every, w->stream));
      }
      // once the world has been rendered: the bricks this tick changed
      if (masks && w->occ_track)
        FS3D_HIP(launch_occ_dirty(masks, w->tick + (u64)k, tick_dev, (u32)k, w->nx, w->ny, w->nz, w->shape.seg,
                                  w->d_dirty, w->stream));
    }
    src = dst;
    k++;
  }
  if (final) *final = src;
  return FS3D_OK;
}

}  // namespace

namespace {

// The cached graph of nticks ticks with a hash every hash_every from the current
// buffer, captured and instantiated on first use.
int world_graph(fs3d_world *w, int nticks, int hash_every, hipGraphExec_t *out) {
  const auto key = std::make_tuple(nticks, hash_every, w->cur);
  auto it = w->graphs.find(key);
  if (it == w->graphs.end()) {
    hipGraph_t graph = nullptr;
    FS3D_HIP(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_ticks(w, nticks, hash_every, w->d_small, nullptr);
    hipError_t e = hipStreamEndCapture(w->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return set_error(FS3D_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return set_error(FS3D_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    it = w->graphs.emplace(key, exec).first;
  }
  *out = it->second;
  return FS3D_OK;
}

}  // namespace

void world_drop_graphs(fs3d_world *w) {
  for (auto &kv : w->graphs) (void)hipGraphExecDestroy(kv.second);
  w->graphs.clear();
}

extern "C" {

const char *fs3d_last_error(void) { return g_error.c_str(); }

int fs3d_abi_version(void) { return FS3D_ABI_VERSION; }

int fs3d_device_count(int *count) {
  if (!count) return set_error(FS3D_ERR_INVALID, "count is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return clear_error();
}

int fs3d_init(int device) {
  int rc = ensure_device(device);
  return rc ? rc : clear_error();
}

int fs3d_rule_lut(uint16_t *lut) {
  if (!lut) return set_error(FS3D_ERR_INVALID, "lut is NULL");
  for (int v = 0; v < 4; v++)
    for (u32 s = 0; s < 65536u; s++) lut[(size_t)v * 65536u + s] = rule_state_packed((u16)s, v);
  return clear_error();
}

int fs3d_world_create(const int32_t dims[3], uint32_t flags, fs3d_world **out) {
  if (!dims || !out) return set_error(FS3D_ERR_INVALID, "dims/out is NULL");
  if (flags & ~(uint32_t)FS3D_WORLD_DENSE) return set_error(FS3D_ERR_INVALID, "unknown flags 0x%x", flags);
  *out = nullptr;
  for (int a = 0; a < 3; a++)
    if (dims[a] < 1) return set_error(FS3D_ERR_INVALID, "dims must be >= 1 (got %d,%d,%d)", dims[0], dims[1], dims[2]);
  if (dims[0] > kMaxNx) return set_error(FS3D_ERR_INVALID, "nx=%d exceeds %d", dims[0], kMaxNx);
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  if (n > (1ull << 40)) return set_error(FS3D_ERR_INVALID, "grid too large");
  int dev = 0;
  (void)hipGetDevice(&dev);
  int rc = ensure_device(dev);
  if (rc) return rc;
  fs3d_world *w = new fs3d_world();
  w->dev = dev;
  w->nx = dims[0];
  w->ny = dims[1];
  w->nz = dims[2];
  w->n = n;
  w->alloc = (n + kGuardBytes + 255) & ~(size_t)255;
  w->mask_words = mask_words_of(w->ny, w->nz);
  // work-list counts are 32-bit: beyond 2^31 row quads a world ticks densely
  w->skip = (flags & FS3D_WORLD_DENSE) || w->mask_words >= (1ull << 31) ? FS3D_SKIP_OFF : FS3D_SKIP_AUTO;
  rc = plan_tick(w->nx, w->ny, w->nz, 0, &w->shape);
  if (rc) {
    delete w;
    return rc;
  }
  hipError_t e = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking);
  for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipMalloc(&w->buf[i], w->alloc);
  for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipMalloc(&w->d_masks[i], 2 * w->mask_words * sizeof(u64));
  if (e == hipSuccess) e = hipMalloc(&w->d_list, w->mask_words * sizeof(TickEntry));
  if (e == hipSuccess) e = hipMalloc(&w->d_ctl, 4 * sizeof(u32));
  if (e == hipSuccess) e = hipMalloc(&w->d_parts, (size_t)kHashParts * kHashStride * sizeof(u64));
  if (e == hipSuccess) e = hipEventCreate(&w->ev0);
  if (e == hipSuccess) e = hipEventCreate(&w->ev1);
  if (e != hipSuccess) {
    const bool oom = e == hipErrorOutOfMemory;
    release(w);
    (void)hipGetLastError();
    return set_error(oom ? FS3D_ERR_NOMEM : FS3D_ERR_HIP, "world allocation (%zu bytes x2) failed: %s",
                     (size_t)n, hipGetErrorString(e));
  }
  e = hipMemsetAsync(w->d_parts, 0, (size_t)kHashParts * kHashStride * sizeof(u64), w->stream);
  for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipMemsetAsync(w->buf[i], 0, w->alloc, w->stream);
  for (int i = 0; i < 2 && e == hipSuccess; i++)
    e = launch_fill_u64(w->d_masks[i], 2 * w->mask_words, ~0ull, w->stream);
  if (e == hipSuccess) e = hipMemsetAsync(w->d_ctl, 0, 4 * sizeof(u32), w->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
  if (e != hipSuccess) {
    release(w);
    return set_error(FS3D_ERR_HIP, "world init failed: %s", hipGetErrorString(e));
  }
  plan_fused(w);
  rc = ensure_small(w, 64);
  if (!rc) rc = size_list(w);
  if (rc) {
    release(w);
    return rc;
  }
  *out = w;
  return clear_error();
}

int fs3d_world_destroy(fs3d_world *w) {
  if (w) release(w);
  return clear_error();
}

int fs3d_world_dims(const fs3d_world *w, int32_t dims[3]) {
  if (int rc = check_world(w)) return rc;
  if (!dims) return set_error(FS3D_ERR_INVALID, "dims is NULL");
  dims[0] = w->nx;
  dims[1] = w->ny;
  dims[2] = w->nz;
  return clear_error();
}

static int voxel_offset(const fs3d_world *w, int32_t x, int32_t y, int32_t z, size_t *off) {
  if (x < 0 || y < 0 || z < 0 || x >= w->nx || y >= w->ny || z >= w->nz)
    return set_error(FS3D_ERR_INVALID, "voxel (%d,%d,%d) outside %dx%dx%d", x, y, z, w->nx, w->ny, w->nz);
  *off = ((size_t)z * w->ny + y) * w->nx + x;
  return FS3D_OK;
}

int fs3d_world_set_voxel(fs3d_world *w, int32_t x, int32_t y, int32_t z, uint8_t v) {
  if (int rc = check_world(w)) return rc;
  size_t off;
  if (int rc = voxel_offset(w, x, y, z, &off)) return rc;
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(hipMemcpyAsync(w->buf[w->cur] + off, &v, 1, hipMemcpyHostToDevice, w->stream));
  if (int rc = mark_cell(w, x, y, z)) return rc;
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return clear_error();
}

int fs3d_world_get_voxel(fs3d_world *w, int32_t x, int32_t y, int32_t z, uint8_t *v) {
  if (int rc = check_world(w)) return rc;
  if (!v) return set_error(FS3D_ERR_INVALID, "v is NULL");
  size_t off;
  if (int rc = voxel_offset(w, x, y, z, &off)) return rc;
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(hipMemcpyAsync(v, w->buf[w->cur] + off, 1, hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return clear_error();
}

int fs3d_world_upload(fs3d_world *w, const uint8_t *host, size_t nbytes) {
  if (int rc = check_world(w)) return rc;
  if (!host || nbytes != w->n) return set_error(FS3D_ERR_INVALID, "upload needs %zu bytes (got %zu)", w->n, nbytes);
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(hipMemcpyAsync(w->buf[w->cur], host, w->n, hipMemcpyHostToDevice, w->stream));
  if (int rc = mark_all(w)) return rc;
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return clear_error();
}

int fs3d_world_download(fs3d_world *w, uint8_t *host, size_t nbytes) {
  if (int rc = check_world(w)) return rc;
  if (!host || nbytes != w->n) return set_error(FS3D_ERR_INVALID, "download needs %zu bytes (got %zu)", w->n, nbytes);
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(hipMemcpyAsync(host, w->buf[w->cur], w->n, hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return clear_error();
}

int fs3d_world_get_tick(const fs3d_world *w, uint64_t *tick) {
  if (int rc = check_world(w)) return rc;
  if (!tick) return set_error(FS3D_ERR_INVALID, "tick is NULL");
  *tick = w->tick;
  return clear_error();
}

int fs3d_world_set_tick(fs3d_world *w, uint64_t tick) {
  if (int rc = check_world(w)) return rc;
  FS3D_HIP(hipSetDevice(w->dev));
  if (int rc = mark_all(w)) return rc;
  w->tick = tick;
  return clear_error();
}

int fs3d_world_set_segment(fs3d_world *w, int32_t seg) {
  if (int rc = check_world(w)) return rc;
  TickShape sh;
  if (int rc = plan_tick(w->nx, w->ny, w->nz, seg, &sh)) return rc;
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  w->shape = sh;
  w->seg_req = seg;
  drop_graphs(w);
  if (int rc = size_list(w)) return rc;
  if (int rc = mark_all(w)) return rc;  // the mask bits index lane segments
  return clear_error();
}

int fs3d_world_set_skip(fs3d_world *w, int32_t mode) {
  if (int rc = check_world(w)) return rc;
  if (mode < FS3D_SKIP_OFF || mode > FS3D_SKIP_SEGS) return set_error(FS3D_ERR_INVALID, "skip mode %d", mode);
  if (mode && w->mask_words >= (1ull << 31))
    return set_error(FS3D_ERR_INVALID, "skipping needs fewer than 2^31 row quads (ny*nz/4)");
  FS3D_HIP(hipSetDevice(w->dev));
  if (mode != w->skip) {
    FS3D_HIP(hipStreamSynchronize(w->stream));
    drop_graphs(w);
    w->skip = mode;
  }
  if (int rc = mark_all(w)) return rc;
  return clear_error();
}

int fs3d_world_set_fuse(fs3d_world *w, int32_t on) {
  if (int rc = check_world(w)) return rc;
  FS3D_HIP(hipSetDevice(w->dev));
  if ((on != 0) != w->fuse) {
    FS3D_HIP(hipStreamSynchronize(w->stream));
    drop_graphs(w);
    w->fuse = on != 0;
  }
  return clear_error();
}

int fs3d_world_get_fuse(const fs3d_world *w, int32_t *on) {
  if (int rc = check_world(w)) return rc;
  if (!on) return set_error(FS3D_ERR_INVALID, "on is NULL");
  *on = fuse_ok(w) ? 1 : 0;
  return clear_error();
}

int fs3d_world_get_skip(const fs3d_world *w, int32_t *mode) {
  if (int rc = check_world(w)) return rc;
  if (!mode) return set_error(FS3D_ERR_INVALID, "mode is NULL");
  *mode = w->skip;
  return clear_error();
}

int fs3d_world_sweep_stats(fs3d_world *w, int32_t enable, uint64_t out[3]) {
  if (int rc = check_world(w)) return rc;
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  u64 tot[3] = {0, 0, 0};
  if (w->d_stats) {
    std::vector<u64> h(kStatWords);
    FS3D_HIP(hipMemcpy(h.data(), w->d_stats, kStatWords * sizeof(u64), hipMemcpyDeviceToHost));
    for (int i = 0; i < kHashParts; i++) {
      tot[0] += h[(size_t)i * kHashStride];
      tot[1] += h[(size_t)i * kHashStride + 1];
    }
    tot[2] = h[kStatWords - 1];
  }
  if (enable && !w->d_stats) {
    FS3D_HIP(hipMalloc(&w->d_stats, kStatWords * sizeof(u64)));
    drop_graphs(w);
  } else if (!enable && w->d_stats) {
    FS3D_HIP(hipFree(w->d_stats));
    w->d_stats = nullptr;
    drop_graphs(w);
  }
  if (w->d_stats) {
    FS3D_HIP(hipMemset(w->d_stats, 0, kStatWords * sizeof(u64)));
  }
  if (out)
    for (int i = 0; i < 3; i++) out[i] = tot[i];
  return clear_error();
}

int fs3d_world_get_segment(const fs3d_world *w, int32_t *seg) {
  if (int rc = check_world(w)) return rc;
  if (!seg) return set_error(FS3D_ERR_INVALID, "seg is NULL");
  *seg = w->shape.seg;
  return clear_error();
}

int fs3d_world_step(fs3d_world *w, int32_t nticks) {
  if (int rc = check_world(w)) return rc;
  if (nticks < 0) return set_error(FS3D_ERR_INVALID, "nticks < 0");
  FS3D_HIP(hipSetDevice(w->dev));
  int fin = w->cur;
  if (int rc = enqueue_ticks(w, nticks, 0, nullptr, &fin)) return rc;
  note_ticks(w);
  w->cur = fin;
  w->tick += (u64)nticks;
  return clear_error();
}

int fs3d_world_tick(fs3d_world *w, uint64_t *hash) {
  if (int rc = check_world(w)) return rc;
  if (!hash) return set_error(FS3D_ERR_INVALID, "hash is NULL");
  FS3D_HIP(hipSetDevice(w->dev));
  int fin = w->cur;
  if (int rc = enqueue_ticks(w, 1, 1, nullptr, &fin)) return rc;
  note_ticks(w);
  FS3D_HIP(hipMemcpyAsync(w->h_pinned, w->d_small + 1, sizeof(u64), hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  w->cur = fin;
  w->tick += 1;
  *hash = w->h_pinned[0];
  return clear_error();
}

int fs3d_world_hash(fs3d_world *w, uint64_t *hash) {
  if (int rc = check_world(w)) return rc;
  if (!hash) return set_error(FS3D_ERR_INVALID, "hash is NULL");
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(launch_hash(w->buf[w->cur], w->n, w->d_parts, w->stream));
  FS3D_HIP(launch_hash_finalize(w->d_parts, w->d_small + 1, w->stream));
  FS3D_HIP(hipMemcpyAsync(w->h_pinned, w->d_small + 1, sizeof(u64), hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  *hash = w->h_pinned[0];
  return clear_error();
}

int fs3d_world_prepare_run(fs3d_world *w, int32_t nticks, int32_t hash_every) {
  if (int rc = check_world(w)) return rc;
  if (nticks < 0 || hash_every < 0) return set_error(FS3D_ERR_INVALID, "nticks/hash_every < 0");
  if (nticks == 0) return clear_error();
  FS3D_HIP(hipSetDevice(w->dev));
  const int nh = hash_every > 0 ? nticks / hash_every : 0;
  if (int rc = ensure_small(w, nh + 1)) return rc;
  hipGraphExec_t exec = nullptr;
  if (int rc = world_graph(w, nticks, hash_every, &exec)) return rc;
  return clear_error();
}

int fs3d_world_run(fs3d_world *w, int32_t nticks, int32_t hash_every, uint64_t *hashes) {
  if (int rc = check_world(w)) return rc;
  if (nticks < 0 || hash_every < 0) return set_error(FS3D_ERR_INVALID, "nticks/hash_every < 0");
  const int nh = hash_every > 0 ? nticks / hash_every : 0;
  if (nh > 0 && !hashes) return set_error(FS3D_ERR_INVALID, "hashes is NULL");
  if (nticks == 0) return clear_error();
  FS3D_HIP(hipSetDevice(w->dev));
  if (int rc = ensure_small(w, nh + 1)) return rc;
  hipGraphExec_t exec = nullptr;
  if (int rc = world_graph(w, nticks, hash_every, &exec)) return rc;
  FS3D_HIP(launch_set_u64(w->d_small, w->tick, w->stream));
  FS3D_HIP(hipEventRecord(w->ev0, w->stream));
  FS3D_HIP(hipGraphLaunch(exec, w->stream));
  note_ticks(w);
  FS3D_HIP(hipEventRecord(w->ev1, w->stream));
  if (nh > 0)
    FS3D_HIP(hipMemcpyAsync(w->h_pinned, w->d_small + 1, (size_t)nh * sizeof(u64), hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  float ms = 0.f;
  FS3D_HIP(hipEventElapsedTime(&ms, w->ev0, w->ev1));
  w->last_run_ms = (double)ms / nticks;
  for (int i = 0; i < nh; i++) hashes[i] = w->h_pinned[i];
  int fin = w->cur;
  if (int rc = enqueue_ticks(w, nticks, hash_every, nullptr, &fin, false)) return rc;
  w->cur = fin;
  w->tick += (u64)nticks;
  return clear_error();
}

int fs3d_world_last_run_ms(const fs3d_world *w, double *ms) {
  if (int rc = check_world(w)) return rc;
  if (!ms) return set_error(FS3D_ERR_INVALID, "ms is NULL");
  *ms = w->last_run_ms;
  return clear_error();
}

int fs3d_world_counts(fs3d_world *w, uint64_t *counts) {
  if (int rc = check_world(w)) return rc;
  if (!counts) return set_error(FS3D_ERR_INVALID, "counts is NULL");
  FS3D_HIP(hipSetDevice(w->dev));
  u64 *d = nullptr;
  FS3D_HIP(hipMallocAsync((void **)&d, 256 * sizeof(u64), w->stream));
  FS3D_HIP(hipMemsetAsync(d, 0, 256 * sizeof(u64), w->stream));
  FS3D_HIP(launch_counts(w->buf[w->cur], w->n, d, w->stream));
  FS3D_HIP(hipMemcpyAsync(counts, d, 256 * sizeof(u64), hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipFreeAsync(d, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return clear_error();
}

int fs3d_world_sync(fs3d_world *w) {
  if (int rc = check_world(w)) return rc;
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return clear_error();
}

int fs3d_world_device_ptr(fs3d_world *w, const void **dptr) {
  if (int rc = check_world(w)) return rc;
  if (!dptr) return set_error(FS3D_ERR_INVALID, "dptr is NULL");
  *dptr = w->buf[w->cur];
  return clear_error();
}

static int save_chunk(void *ctx, size_t off, u8 *buf, size_t n) {
  fs3d_world *w = static_cast<fs3d_world *>(ctx);
  FS3D_HIP(hipMemcpyAsync(buf, w->buf[w->cur] + off, n, hipMemcpyDeviceToHost, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return FS3D_OK;
}

// A load streams the payload into the world's idle buffer.
static int load_chunk(void *ctx, size_t off, const u8 *buf, size_t n) {
  fs3d_world *w = static_cast<fs3d_world *>(ctx);
  FS3D_HIP(hipMemcpyAsync(w->buf[w->cur ^ 1] + off, buf, n, hipMemcpyHostToDevice, w->stream));
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return FS3D_OK;
}

int fs3d_world_save(fs3d_world *w, const char *path) {
  if (int rc = check_world(w)) return rc;
  u64 h = 0;
  if (int rc = fs3d_world_hash(w, &h)) return rc;
  const int32_t dims[3] = {w->nx, w->ny, w->nz};
  if (int rc = file_write(path, dims, w->tick, h, nullptr, w->n, save_chunk, w)) return rc;
  return clear_error();
}

int fs3d_world_load(fs3d_world *w, const char *path) {
  if (int rc = check_world(w)) return rc;
  int32_t dims[3];
  u64 tick = 0, want = 0;
  if (int rc = file_info(path, dims, &tick, &want, nullptr)) return rc;
  if (dims[0] != w->nx || dims[1] != w->ny || dims[2] != w->nz)
    return set_error(FS3D_ERR_INVALID, "%s holds a %dx%dx%d world, this one is %dx%dx%d", path, dims[0], dims[1],
                     dims[2], w->nx, w->ny, w->nz);
  FS3D_HIP(hipSetDevice(w->dev));
  // The payload goes to the idle buffer and becomes the world's state only once
  // its hash matches the stored one; on any failure the grid and tick stay as
  // they were.  Skipped segments rely on the destination buffer holding their
  // bytes, so every change bit is set either way (mark_all): the next two ticks
  // evaluate every segment and rewrite whichever buffers they write.
  int rc = file_read(path, nullptr, w->n, load_chunk, w);
  u64 h = 0;
  if (!rc) {
    u8 *idle = w->buf[w->cur ^ 1];
    hipError_t e = launch_hash(idle, w->n, w->d_parts, w->stream);
    if (e == hipSuccess) e = launch_hash_finalize(w->d_parts, w->d_small + 1, w->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(w->h_pinned, w->d_small + 1, sizeof(u64), hipMemcpyDeviceToHost, w->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) rc = set_error(FS3D_ERR_HIP, "world load hash failed: %s", hipGetErrorString(e));
    h = w->h_pinned[0];
    if (!rc && h != want)
      rc = set_error(FS3D_ERR_STATE, "%s: grid hash %016llx does not match the stored %016llx", path,
                     (unsigned long long)h, (unsigned long long)want);
  }
  if (!rc) {
    w->cur ^= 1;
    w->tick = tick;
  }
  if (int rc2 = mark_all(w)) return rc ? rc : rc2;
  if (hipStreamSynchronize(w->stream) != hipSuccess && !rc) rc = set_error(FS3D_ERR_HIP, "world load sync failed");
  return rc ? rc : clear_error();
}

int fs3d_world_fill_rain(fs3d_world *w, uint64_t seed, uint64_t thr) {
  if (int rc = check_world(w)) return rc;
  FS3D_HIP(hipSetDevice(w->dev));
  FS3D_HIP(launch_fill_rain(w->buf[w->cur], w->nx, w->ny, w->nz, seed, thr, w->stream));
  if (int rc = mark_all(w)) return rc;
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return clear_error();
}

int fs3d_world_fill_boxes(fs3d_world *w, uint64_t seed, const fs3d_box_params *p, int64_t *nboxes) {
  if (int rc = check_world(w)) return rc;
  if (!p) return set_error(FS3D_ERR_INVALID, "params is NULL");
  FS3D_HIP(hipSetDevice(w->dev));
  u8 *host = nullptr;
  FS3D_HIP(hipHostMalloc(&host, w->n, hipHostMallocDefault));
  const long long nb = paint_boxes(host, w->nx, w->ny, w->nz, seed, *p);
  if (nb < 0) {
    (void)hipHostFree(host);
    return set_error(FS3D_ERR_INVALID, "box parameters do not fit the grid");
  }
  w->occ_valid = false;
  hipError_t e = hipMemcpyAsync(w->buf[w->cur], host, w->n, hipMemcpyHostToDevice, w->stream);
  for (int i = 0; i < 2 && e == hipSuccess; i++) e = launch_fill_u64(w->d_masks[i], 2 * w->mask_words, ~0ull, w->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
  (void)hipHostFree(host);
  if (e != hipSuccess) return set_error(FS3D_ERR_HIP, "box upload failed: %s", hipGetErrorString(e));
  if (nboxes) *nboxes = nb;
  return clear_error();
}

int fs3d_world_fill_pillars(fs3d_world *w, uint64_t seed, const fs3d_pillar_params *p) {
  if (int rc = check_world(w)) return rc;
  if (!p) return set_error(FS3D_ERR_INVALID, "params is NULL");
  int pool[5];
  std::vector<Pillar> pillars;
  if (pillar_list(w->nx, w->ny, w->nz, seed, *p, pool, &pillars) != 0)
    return set_error(FS3D_ERR_INVALID, "pillar parameters do not fit the grid");
  FS3D_HIP(hipSetDevice(w->dev));
  u8 *g = w->buf[w->cur];
  const size_t layer = (size_t)w->nx * w->ny;
  if (int rc = mark_all(w)) return rc;
  FS3D_HIP(hipMemsetAsync(g, FS3D_EMPTY, w->n, w->stream));
  FS3D_HIP(hipMemsetAsync(g, FS3D_STATIC, layer, w->stream));
  if (pool[4] >= 1)
    FS3D_HIP(launch_fill_box(g, w->nx, w->ny, pool[0], pool[1], 1, pool[2], pool[3], pool[4], FS3D_FLUID, w->stream));
  if (!pillars.empty()) {
    Pillar *d = nullptr;
    int max_r = 0, max_zt = 0;
    for (const Pillar &q : pillars) max_r = std::max(max_r, q.r), max_zt = std::max(max_zt, q.zt);
    FS3D_HIP(hipMalloc(&d, pillars.size() * sizeof(Pillar)));
    hipError_t e = hipMemcpyAsync(d, pillars.data(), pillars.size() * sizeof(Pillar), hipMemcpyHostToDevice, w->stream);
    if (e == hipSuccess) e = launch_fill_pillars(g, w->nx, w->ny, d, (int)pillars.size(), max_r, max_zt, w->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return set_error(FS3D_ERR_HIP, "pillar fill failed: %s", hipGetErrorString(e));
  }
  FS3D_HIP(hipStreamSynchronize(w->stream));
  return clear_error();
}

}  // extern "C"
