// C ABI of batched worlds: one device arena holding many independent grids,
// stepped either by the persistent batch kernel (one launch for all worlds and
// all ticks, captured in a HIP graph) or by the plain loop of one single-world
// tick launch per world per tick (the baseline the batch is measured against).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

#include "fs3d_batch.h"

using namespace fs3d;

struct fs3d_batch {
  int dev = 0;
  int n = 0;
  std::vector<int> nx, ny, nz;
  std::vector<size_t> off, cells;
  std::vector<TickShape> plain_shape;  // single-world launch shape per world (plain loop)
  u8 *arena = nullptr;
  size_t arena_bytes = 0;
  BWorld *d_worlds = nullptr;
  int *d_order = nullptr;
  u32 *d_queue = nullptr;
  u64 *d_hashes = nullptr;
  long long *d_prefix = nullptr;  // wide mode: workgroup prefix per parity, 2 x (n + 1)
  int *d_blk2ent[2] = {nullptr, nullptr};  // wide mode: workgroup -> world, per parity
  long long total_blocks[2] = {0, 0};
  int table_nt = -1;              // nt stored in the device table (wide mode's final tick)
  int tuned_mode = FS3D_BATCH_PERSISTENT;
  std::vector<BWorld> h_worlds;
  u64 tick = 0;  // all worlds of a batch share the tick counter
  int seg = 16, threads = 256, groups = 1, grid = 0;
  bool settle = true;  // persistent mode: stop ticking worlds that have settled
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  // (mode, nticks, hash, tick parity, seg, threads, groups)
  std::map<std::tuple<int, int, int, int, int, int, int>, hipGraphExec_t> graphs;
};

namespace {

int max_nx(const fs3d_batch *b) { return b->n ? *std::max_element(b->nx.begin(), b->nx.end()) : 1; }

int check_batch(const fs3d_batch *b) {
  if (!b) return set_error(FS3D_ERR_INVALID, "batch handle is NULL");
  return FS3D_OK;
}

void drop_graphs(fs3d_batch *b) {
  for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
}

// Rebuilds the device world table (and the wide mode's wave prefixes) for the
// current segment width and tick; nt is the tick count the wide mode's final
// hash is keyed to.
int sync_table(fs3d_batch *b, int nt = 0) {
  for (int i = 0; i < b->n; i++) {
    BWorld &w = b->h_worlds[i];
    w.g = b->arena + b->off[i];
    w.nx = b->nx[i];
    w.ny = b->ny[i];
    w.nz = b->nz[i];
    w.lr = (w.nx + b->seg - 1) / b->seg;
    w.rqpw = 64 / w.lr;
    w.nt = nt;
    w.slot = i;
    w.fin = 1;  // a batch run's last tick is the one hashed
    w.tick = b->tick;
  }
  FS3D_HIP(hipMemcpyAsync(b->d_worlds, b->h_worlds.data(), sizeof(BWorld) * b->n, hipMemcpyHostToDevice, b->stream));
  std::vector<long long> pre(2 * (size_t)(b->n + 1));
  for (int p = 0; p < 2; p++) {
    long long acc = 0;
    for (int i = 0; i < b->n; i++) {
      pre[(size_t)p * (b->n + 1) + i] = acc;
      acc += world_blocks(b->h_worlds[i], p);
    }
    pre[(size_t)p * (b->n + 1) + b->n] = acc;
    if (acc > b->total_blocks[p] || !b->d_blk2ent[p]) {
      drop_graphs(b);  // captured launches hold the old table pointer
      if (b->d_blk2ent[p]) FS3D_HIP(hipFree(b->d_blk2ent[p]));
      b->d_blk2ent[p] = nullptr;
      FS3D_HIP(hipMalloc(&b->d_blk2ent[p], sizeof(int) * (size_t)std::max(acc, 1LL)));
    }
    b->total_blocks[p] = acc;
  }
  FS3D_HIP(hipMemcpyAsync(b->d_prefix, pre.data(), sizeof(long long) * pre.size(), hipMemcpyHostToDevice, b->stream));
  for (int p = 0; p < 2; p++)
    FS3D_HIP(launch_expand(b->d_prefix + (size_t)p * (b->n + 1), b->n, b->d_blk2ent[p], b->stream));
  FS3D_HIP(hipStreamSynchronize(b->stream));  // pre is a local buffer
  b->table_nt = nt;
  return FS3D_OK;
}

int apply_config(fs3d_batch *b, int seg, int threads, int groups) {
  if (!batch_config_ok(seg, threads, groups))
    return set_error(FS3D_ERR_INVALID, "batch config (seg %d, threads %d, groups %d) unsupported", seg, threads,
                     groups);
  if ((max_nx(b) + seg - 1) / seg > 64)
    return set_error(FS3D_ERR_INVALID, "segment %d too narrow for nx=%d", seg, max_nx(b));
  hipDeviceProp_t prop;
  FS3D_HIP(hipGetDeviceProperties(&prop, b->dev));
  const int occ = batch_occupancy(seg, threads, groups);
  b->seg = seg;
  b->threads = threads;
  b->groups = groups;
  b->grid = std::max(1, std::min(b->n, prop.multiProcessorCount * occ));
  drop_graphs(b);
  return sync_table(b);
}

void release(fs3d_batch *b) {
  (void)hipSetDevice(b->dev);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  if (b->arena) (void)hipFree(b->arena);
  if (b->d_worlds) (void)hipFree(b->d_worlds);
  if (b->d_order) (void)hipFree(b->d_order);
  if (b->d_queue) (void)hipFree(b->d_queue);
  if (b->d_hashes) (void)hipFree(b->d_hashes);
  if (b->d_prefix) (void)hipFree(b->d_prefix);
  for (int p = 0; p < 2; p++)
    if (b->d_blk2ent[p]) (void)hipFree(b->d_blk2ent[p]);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int run_persistent(fs3d_batch *b, int nticks, int want_hash) {
  const int flags = (want_hash ? kBatchHash : 0) | (b->settle ? kBatchSettle : 0);
  const auto key = std::make_tuple((int)FS3D_BATCH_PERSISTENT, nticks, flags, 0, b->seg, b->threads, b->groups);
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    hipGraph_t graph = nullptr;
    FS3D_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch_batch(b->seg, b->threads, b->groups, b->d_worlds, b->d_order, b->n, nticks, b->d_queue,
                                device_rule_table(b->dev), b->d_hashes, flags, b->grid, b->stream);
    hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return set_error(FS3D_ERR_HIP, "batch graph capture failed: %s",
                       hipGetErrorString(e != hipSuccess ? e : e2));
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return set_error(FS3D_ERR_HIP, "batch graph instantiate failed: %s", hipGetErrorString(e));
    it = b->graphs.emplace(key, exec).first;
  }
  FS3D_HIP(hipEventRecord(b->ev0, b->stream));
  FS3D_HIP(hipGraphLaunch(it->second, b->stream));
  FS3D_HIP(hipEventRecord(b->ev1, b->stream));
  return FS3D_OK;
}

// One wide launch per tick, every world in each launch; captured in a graph.
int run_wide(fs3d_batch *b, int nticks, int want_hash) {
  if (b->table_nt != nticks)
    if (int rc = sync_table(b, nticks)) return rc;
  const int par = (int)(b->tick & 1u);
  const auto key = std::make_tuple((int)FS3D_BATCH_WIDE, nticks, want_hash, par, b->seg, 0, 0);
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    hipGraph_t graph = nullptr;
    FS3D_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    if (want_hash) e = launch_zero_slots(b->d_hashes, b->d_worlds, b->n, b->stream);
    const u64 *rule = device_rule_table(b->dev);
    for (int k = 0; k < nticks && e == hipSuccess; k++) {
      const int p = (par + k) & 1;
      e = launch_wide(b->seg, b->d_worlds, b->d_prefix + (size_t)p * (b->n + 1), b->d_blk2ent[p], b->total_blocks[p],
                      k, rule, want_hash ? b->d_hashes : nullptr, b->stream);
    }
    hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return set_error(FS3D_ERR_HIP, "wide graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return set_error(FS3D_ERR_HIP, "wide graph instantiate failed: %s", hipGetErrorString(e));
    it = b->graphs.emplace(key, exec).first;
  }
  FS3D_HIP(hipEventRecord(b->ev0, b->stream));
  FS3D_HIP(hipGraphLaunch(it->second, b->stream));
  FS3D_HIP(hipEventRecord(b->ev1, b->stream));
  return FS3D_OK;
}

int run_plain(fs3d_batch *b, int nticks, int want_hash) {
  const u64 *rule = device_rule_table(b->dev);
  FS3D_HIP(hipEventRecord(b->ev0, b->stream));
  for (int k = 0; k < nticks; k++)
    for (int i = 0; i < b->n; i++) {
      u8 *g = b->arena + b->off[i];
      FS3D_HIP(launch_tick(b->plain_shape[i], g, g, b->nx[i], b->ny[i], b->nz[i], (u32)(b->tick + (u64)k), nullptr,
                           0, rule, nullptr, b->stream));
    }
  if (want_hash) FS3D_HIP(launch_batch_hash(b->d_worlds, b->n, b->d_hashes, b->stream));
  FS3D_HIP(hipEventRecord(b->ev1, b->stream));
  return FS3D_OK;
}

}  // namespace

extern "C" {

int fs3d_batch_create(int32_t n, const int32_t *dims, fs3d_batch **out) {
  if (!out || (n > 0 && !dims) || n < 1) return set_error(FS3D_ERR_INVALID, "batch_create needs n >= 1, dims, out");
  *out = nullptr;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (int rc = ensure_device(dev)) return rc;
  fs3d_batch *b = new fs3d_batch();
  b->dev = dev;
  b->n = n;
  size_t total = 0;
  for (int i = 0; i < n; i++) {
    const int x = dims[3 * i], y = dims[3 * i + 1], z = dims[3 * i + 2];
    if (x < 1 || y < 1 || z < 1 || x > 2048) {
      delete b;
      return set_error(FS3D_ERR_INVALID, "world %d: dims (%d,%d,%d) out of range (1..2048 in x)", i, x, y, z);
    }
    b->nx.push_back(x);
    b->ny.push_back(y);
    b->nz.push_back(z);
    const size_t c = (size_t)x * y * z;
    b->cells.push_back(c);
    b->off.push_back(total);
    total += (c + kGuardBytes + 255) & ~(size_t)255;
    TickShape sh;
    if (int rc = plan_tick(x, y, z, 0, &sh)) {
      delete b;
      return rc;
    }
    b->plain_shape.push_back(sh);
  }
  b->arena_bytes = total;
  b->h_worlds.resize(n);
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&b->arena, total);
  if (e == hipSuccess) e = hipMalloc(&b->d_worlds, sizeof(BWorld) * n);
  if (e == hipSuccess) e = hipMalloc(&b->d_order, sizeof(int) * n);
  if (e == hipSuccess) e = hipMalloc(&b->d_queue, 256);
  if (e == hipSuccess) e = hipMalloc(&b->d_hashes, sizeof(u64) * n);
  if (e == hipSuccess) e = hipMalloc(&b->d_prefix, sizeof(long long) * 2 * (size_t)(n + 1));
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipMemsetAsync(b->arena, 0, total, b->stream);
  if (e != hipSuccess) {
    const bool oom = e == hipErrorOutOfMemory;
    release(b);
    (void)hipGetLastError();
    return set_error(oom ? FS3D_ERR_NOMEM : FS3D_ERR_HIP, "batch allocation (%zu bytes) failed: %s", total,
                     hipGetErrorString(e));
  }
  // largest worlds first: they start early and the short ones fill the tail
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->cells[a] > b->cells[c]; });
  e = hipMemcpyAsync(b->d_order, order.data(), sizeof(int) * n, hipMemcpyHostToDevice, b->stream);
  if (e != hipSuccess) {
    release(b);
    return set_error(FS3D_ERR_HIP, "batch init failed: %s", hipGetErrorString(e));
  }
  const int seg0 = max_nx(b) <= 512 ? 16 : 32;
  if (int rc = apply_config(b, seg0, 256, 1)) {
    release(b);
    return rc;
  }
  if (hipStreamSynchronize(b->stream) != hipSuccess) {
    release(b);
    return set_error(FS3D_ERR_HIP, "batch init sync failed");
  }
  *out = b;
  return clear_error();
}

int fs3d_batch_destroy(fs3d_batch *b) {
  if (b) release(b);
  return clear_error();
}

int fs3d_batch_size(const fs3d_batch *b, int32_t *n) {
  if (int rc = check_batch(b)) return rc;
  if (!n) return set_error(FS3D_ERR_INVALID, "n is NULL");
  *n = b->n;
  return clear_error();
}

int fs3d_batch_upload(fs3d_batch *b, int32_t i, const uint8_t *host, size_t nbytes) {
  if (int rc = check_batch(b)) return rc;
  if (i < 0 || i >= b->n || !host || nbytes != b->cells[i])
    return set_error(FS3D_ERR_INVALID, "batch upload: bad world index or size");
  FS3D_HIP(hipSetDevice(b->dev));
  FS3D_HIP(hipMemcpyAsync(b->arena + b->off[i], host, nbytes, hipMemcpyHostToDevice, b->stream));
  FS3D_HIP(hipStreamSynchronize(b->stream));
  return clear_error();
}

int fs3d_batch_download(fs3d_batch *b, int32_t i, uint8_t *host, size_t nbytes) {
  if (int rc = check_batch(b)) return rc;
  if (i < 0 || i >= b->n || !host || nbytes != b->cells[i])
    return set_error(FS3D_ERR_INVALID, "batch download: bad world index or size");
  FS3D_HIP(hipSetDevice(b->dev));
  FS3D_HIP(hipMemcpyAsync(host, b->arena + b->off[i], nbytes, hipMemcpyDeviceToHost, b->stream));
  FS3D_HIP(hipStreamSynchronize(b->stream));
  return clear_error();
}

int fs3d_batch_fill_rain(fs3d_batch *b, const uint64_t *seeds, const uint64_t *thr) {
  if (int rc = check_batch(b)) return rc;
  if (!seeds || !thr) return set_error(FS3D_ERR_INVALID, "seeds/thr is NULL");
  FS3D_HIP(hipSetDevice(b->dev));
  u64 *d = nullptr;
  FS3D_HIP(hipMalloc(&d, sizeof(u64) * 2 * b->n));
  hipError_t e = hipMemcpyAsync(d, seeds, sizeof(u64) * b->n, hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d + b->n, thr, sizeof(u64) * b->n, hipMemcpyHostToDevice, b->stream);
  const size_t maxc = *std::max_element(b->cells.begin(), b->cells.end());
  if (e == hipSuccess) e = launch_batch_fill_rain(b->d_worlds, b->n, d, d + b->n, maxc, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return set_error(FS3D_ERR_HIP, "batch fill failed: %s", hipGetErrorString(e));
  return clear_error();
}

int fs3d_batch_get_tick(const fs3d_batch *b, uint64_t *tick) {
  if (int rc = check_batch(b)) return rc;
  if (!tick) return set_error(FS3D_ERR_INVALID, "tick is NULL");
  *tick = b->tick;
  return clear_error();
}

int fs3d_batch_set_tick(fs3d_batch *b, uint64_t tick) {
  if (int rc = check_batch(b)) return rc;
  FS3D_HIP(hipSetDevice(b->dev));
  b->tick = tick;
  if (int rc = sync_table(b)) return rc;
  FS3D_HIP(hipStreamSynchronize(b->stream));
  return clear_error();
}

int fs3d_batch_set_config(fs3d_batch *b, int32_t seg, int32_t threads, int32_t groups) {
  if (int rc = check_batch(b)) return rc;
  FS3D_HIP(hipSetDevice(b->dev));
  FS3D_HIP(hipStreamSynchronize(b->stream));
  if (int rc = apply_config(b, seg, threads, groups)) return rc;
  FS3D_HIP(hipStreamSynchronize(b->stream));
  return clear_error();
}

int fs3d_batch_get_config(const fs3d_batch *b, int32_t *seg, int32_t *threads, int32_t *groups, int32_t *grid) {
  if (int rc = check_batch(b)) return rc;
  if (seg) *seg = b->seg;
  if (threads) *threads = b->threads;
  if (groups) *groups = b->groups;
  if (grid) *grid = b->grid;
  return clear_error();
}

int fs3d_batch_set_skip(fs3d_batch *b, int32_t on) {
  if (int rc = check_batch(b)) return rc;
  b->settle = on != 0;
  return clear_error();
}

int fs3d_batch_get_skip(const fs3d_batch *b, int32_t *on) {
  if (int rc = check_batch(b)) return rc;
  if (!on) return set_error(FS3D_ERR_INVALID, "on is NULL");
  *on = b->settle ? 1 : 0;
  return clear_error();
}

int fs3d_batch_run(fs3d_batch *b, int32_t nticks, int32_t mode, uint64_t *hashes) {
  if (int rc = check_batch(b)) return rc;
  if (mode == FS3D_BATCH_TUNED) mode = b ? b->tuned_mode : mode;
  if (nticks < 0 || (mode != FS3D_BATCH_PERSISTENT && mode != FS3D_BATCH_PLAIN && mode != FS3D_BATCH_WIDE))
    return set_error(FS3D_ERR_INVALID, "batch_run: nticks >= 0 and a known mode required");
  FS3D_HIP(hipSetDevice(b->dev));
  const int want_hash = hashes != nullptr;
  if (nticks == 0) {
    if (want_hash) {
      FS3D_HIP(launch_batch_hash(b->d_worlds, b->n, b->d_hashes, b->stream));
      FS3D_HIP(hipMemcpyAsync(hashes, b->d_hashes, sizeof(u64) * b->n, hipMemcpyDeviceToHost, b->stream));
    }
    FS3D_HIP(hipStreamSynchronize(b->stream));
    b->last_ms = 0.0;
    return clear_error();
  }
  int rc = mode == FS3D_BATCH_PERSISTENT ? run_persistent(b, nticks, want_hash)
           : mode == FS3D_BATCH_WIDE     ? run_wide(b, nticks, want_hash)
                                         : run_plain(b, nticks, want_hash);
  if (rc) return rc;
  if (want_hash)
    FS3D_HIP(hipMemcpyAsync(hashes, b->d_hashes, sizeof(u64) * b->n, hipMemcpyDeviceToHost, b->stream));
  FS3D_HIP(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  FS3D_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  b->tick += (u64)nticks;
  if (mode != FS3D_BATCH_PERSISTENT) {
    // the persistent kernel advances the device table itself; the others do not
    if (int rc2 = sync_table(b, b->table_nt < 0 ? 0 : b->table_nt)) return rc2;
    FS3D_HIP(hipStreamSynchronize(b->stream));
  }
  return clear_error();
}

int fs3d_batch_last_run_ms(const fs3d_batch *b, double *ms) {
  if (int rc = check_batch(b)) return rc;
  if (!ms) return set_error(FS3D_ERR_INVALID, "ms is NULL");
  *ms = b->last_ms;
  return clear_error();
}

// Times every supported persistent-kernel shape on a scratch copy of the arena
// (the batch's own state is not advanced) and keeps the fastest.
int fs3d_batch_autotune(fs3d_batch *b, int32_t trial_ticks, int32_t *seg_out, int32_t *threads_out,
                        int32_t *groups_out, int32_t *mode_out, double *best_ms) {
  if (int rc = check_batch(b)) return rc;
  if (trial_ticks < 1) return set_error(FS3D_ERR_INVALID, "trial_ticks < 1");
  FS3D_HIP(hipSetDevice(b->dev));
  FS3D_HIP(hipStreamSynchronize(b->stream));
  u8 *scratch = nullptr;
  FS3D_HIP(hipMalloc(&scratch, b->arena_bytes));
  u8 *const real = b->arena;
  const int seg0 = b->seg, thr0 = b->threads, grp0 = b->groups;
  double best = 1e30;
  int bseg = seg0, bthr = thr0, bgrp = grp0, bmode = FS3D_BATCH_PERSISTENT;
  int rc = FS3D_OK;
  const int segs[3] = {8, 16, 32}, thrs[3] = {256, 512, 1024}, grps[3] = {1, 2, 4};
  // trial index 27 + si: the wide mode at segment segs[si]
  for (int si = 0; si < 3 && !rc; si++)
    for (int ti = 0; ti < 4 && !rc; ti++)
      for (int gi = 0; gi < 3 && !rc; gi++) {
        const bool wide = ti == 3;
        if (wide && gi > 0) continue;
        if (!wide && !batch_config_ok(segs[si], thrs[ti], grps[gi])) continue;
        if ((max_nx(b) + segs[si] - 1) / segs[si] > 64) continue;
        hipError_t e = hipMemcpyAsync(scratch, real, b->arena_bytes, hipMemcpyDeviceToDevice, b->stream);
        if (e != hipSuccess) {
          rc = set_error(FS3D_ERR_HIP, "autotune copy failed: %s", hipGetErrorString(e));
          break;
        }
        b->arena = scratch;
        rc = wide ? apply_config(b, segs[si], 256, 1) : apply_config(b, segs[si], thrs[ti], grps[gi]);
        if (!rc) rc = wide ? run_wide(b, trial_ticks, 0) : run_persistent(b, trial_ticks, 0);
        if (!rc && hipStreamSynchronize(b->stream) != hipSuccess) rc = set_error(FS3D_ERR_HIP, "autotune sync failed");
        float ms = 0.f;
        if (!rc) (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
        b->arena = real;
        if (!rc && ms < best) {
          best = ms;
          bseg = segs[si];
          bthr = wide ? 256 : thrs[ti];  // persistent shape unused by the wide mode
          bgrp = wide ? 1 : grps[gi];
          bmode = wide ? FS3D_BATCH_WIDE : FS3D_BATCH_PERSISTENT;
        }
      }
  b->arena = real;
  drop_graphs(b);
  (void)hipStreamSynchronize(b->stream);
  (void)hipFree(scratch);
  if (rc) {
    (void)apply_config(b, seg0, thr0, grp0);
    return rc;
  }
  if (int rc2 = apply_config(b, bseg, bthr, bgrp)) return rc2;
  FS3D_HIP(hipStreamSynchronize(b->stream));
  if (seg_out) *seg_out = bseg;
  if (threads_out) *threads_out = bthr;
  if (groups_out) *groups_out = bgrp;
  b->tuned_mode = bmode;
  if (mode_out) *mode_out = bmode;
  if (best_ms) *best_ms = best;
  return clear_error();
}

}  // extern "C"
