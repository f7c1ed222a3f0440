// Serving loop: continuous batching of simulation requests on one GPU.
//
// Requests (dims, seed, density threshold, tick count) are admitted in arrival
// order while their grids fit the HBM budget, filled on a side stream, and kept
// resident until their last tick.  Work advances in rounds of up to K ticks:
// each round is K wide launches (every active world advances one tick per
// launch, all worlds' row quads spread over the whole GPU), so a request that
// arrives while others run joins at the next round boundary and its ticks
// overlap with theirs on the device.  Rounds are planned on the host ahead of
// the GPU (tick counts are known), two at a time, with double-buffered tables.
// A request's final tick hashes its state (fused), and its memory returns to the
// grid arena once that round has completed.  The HBM budget bounds the device
// memory the arena has actually allocated (GridArena, arena.h); a request that
// does not fit waits at the head of the queue until finished grids come back.
//
// Latency accounting: every launch is bracketed by HIP events, so each launch
// duration is one tick-latency sample for the worlds in it; request latency is
// host time from submission to the poll that observes its final round done.
//
// The reference has no serving of any kind (one interactive window,
// src/engine/engine.cpp:54-74); this is the serving mode the tick supports.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <vector>

#include "arena.h"
#include "fs3d_batch.h"

using namespace fs3d;

namespace {

constexpr int kStatusQueued = FS3D_REQ_QUEUED, kStatusActive = FS3D_REQ_ACTIVE, kStatusDone = FS3D_REQ_DONE,
              kStatusRejected = FS3D_REQ_REJECTED;

struct Request {
  int nx = 0, ny = 0, nz = 0;
  u64 seed = 0, thr = 0;
  int ticks = 0;
  int remaining = 0;
  u64 tick = 0;
  int status = kStatusQueued;
  u64 hash = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  u8 *g = nullptr;
  size_t bytes = 0;
  int slot = -1;
};

struct Round {
  u8 *h_buf = nullptr;  // pinned: entries | bpre (K x (n+1)) | launch offsets (K+1)
  u8 *d_buf = nullptr;  // device copy of the same
  size_t cap = 0;
  int *d_blk = nullptr;  // device: blk2ent for all launches
  size_t blk_cap = 0;
  u64 *h_hash = nullptr;  // pinned: hashes of finishing requests
  int hash_cap = 0;
  std::vector<hipEvent_t> ev;  // K + 1 launch boundaries
  hipEvent_t done = nullptr;
  std::vector<long long> finishing;  // request ids
  int nlaunch = 0;
  bool inflight = false;
};

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// A request's grid bytes in the arena: the voxels plus the tick's guard bytes.
size_t grid_bytes(const Request &r) { return GridArena::round_up((size_t)r.nx * r.ny * r.nz + kGuardBytes); }

// Device chunks of the grid arena.  A failed hipMalloc returns nullptr and
// leaves no sticky error behind.
void *chunk_alloc(size_t bytes, void *) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}
void chunk_free(void *p, void *) { (void)hipFree(p); }

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

struct fs3d_server {
  int dev = 0;
  fs3d_server_config cfg{};
  int seg = 8;
  hipStream_t s_tick = nullptr, s_fill = nullptr;
  hipEvent_t fill_ev = nullptr;
  GridArena arena;  // resident grids, budget on allocated bytes
  // slots and hashes
  u64 *d_hashes = nullptr;
  u32 *d_chg = nullptr;  // settled-world flags, four per slot (tick mod 4)
  std::vector<int> free_slots;
  // requests
  std::vector<Request> reqs;
  std::deque<long long> queue;
  std::vector<long long> pending, active;
  Round rounds[2];
  int next_round = 0;
  std::deque<int> inflight;  // round indices in launch order
  // stats
  std::vector<double> tick_ms;
  double busy_ms = 0;
  long long n_rounds = 0, n_launches = 0, n_done = 0, n_rejected = 0;
  double voxel_updates = 0;
  int peak_active = 0;
  double t0 = 0;
};

namespace {

int check_server(const fs3d_server *s) {
  if (!s) return set_error(FS3D_ERR_INVALID, "server handle is NULL");
  return FS3D_OK;
}

void free_grid(fs3d_server *s, Request &r) {
  if (!r.g) return;
  s->arena.release(r.g);
  r.g = nullptr;
  s->free_slots.push_back(r.slot);
  r.slot = -1;
}

int ensure_round_cap(Round &rd, size_t bytes, size_t blks, int nhash, int nev) {
  if (bytes > rd.cap) {
    if (rd.h_buf) FS3D_HIP(hipHostFree(rd.h_buf));
    if (rd.d_buf) FS3D_HIP(hipFree(rd.d_buf));
    rd.cap = std::max(bytes, (size_t)1 << 16) * 2;
    FS3D_HIP(hipHostMalloc(&rd.h_buf, rd.cap, hipHostMallocDefault));
    FS3D_HIP(hipMalloc(&rd.d_buf, rd.cap));
  }
  if (blks > rd.blk_cap) {
    if (rd.d_blk) FS3D_HIP(hipFree(rd.d_blk));
    rd.blk_cap = std::max(blks, (size_t)1 << 16) * 2;
    FS3D_HIP(hipMalloc(&rd.d_blk, rd.blk_cap * sizeof(int)));
  }
  if (nhash > rd.hash_cap) {
    if (rd.h_hash) FS3D_HIP(hipHostFree(rd.h_hash));
    rd.hash_cap = std::max(nhash, 256) * 2;
    FS3D_HIP(hipHostMalloc(&rd.h_hash, rd.hash_cap * sizeof(u64), hipHostMallocDefault));
  }
  while ((int)rd.ev.size() < nev) {
    hipEvent_t e;
    FS3D_HIP(hipEventCreate(&e));
    rd.ev.push_back(e);
  }
  if (!rd.done) FS3D_HIP(hipEventCreate(&rd.done));
  return FS3D_OK;
}

// Observes completed rounds: records tick latencies, hands out final hashes and
// frees finished grids.
int reap(fs3d_server *s) {
  while (!s->inflight.empty()) {
    Round &rd = s->rounds[s->inflight.front()];
    const hipError_t q = hipEventQuery(rd.done);
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) return set_error(FS3D_ERR_HIP, "round failed: %s", hipGetErrorString(q));
    const double t = now_s() - s->t0;
    for (int k = 0; k < rd.nlaunch; k++) {
      float ms = 0.f;
      FS3D_HIP(hipEventElapsedTime(&ms, rd.ev[k], rd.ev[k + 1]));
      s->tick_ms.push_back(ms);
      s->busy_ms += ms;
    }
    for (size_t i = 0; i < rd.finishing.size(); i++) {
      Request &r = s->reqs[rd.finishing[i]];
      r.hash = rd.h_hash[i];
      r.t_done = t;
      r.status = kStatusDone;
      free_grid(s, r);
      s->n_done++;
    }
    rd.finishing.clear();
    rd.inflight = false;
    s->inflight.pop_front();
  }
  return FS3D_OK;
}

// Admits queued requests in arrival order while a hash slot is free and the
// arena can place their grids within the budget; stops at the first that does
// not fit (it waits for finished grids to come back).  A request larger than
// the whole budget, or whose chunk allocation fails while no other request
// holds memory that could come back, is rejected.
int admit(fs3d_server *s) {
  bool any = false;
  while (!s->queue.empty()) {
    Request &r = s->reqs[s->queue.front()];
    const size_t need = grid_bytes(r);
    if (need > (size_t)s->cfg.budget_bytes) {
      r.status = kStatusRejected;
      s->n_rejected++;
      s->queue.pop_front();
      continue;
    }
    if (s->free_slots.empty()) break;
    u8 *g = nullptr;
    const GridArena::Result res = s->arena.alloc(need, &g);
    if (res != GridArena::kOk) {
      const bool others = s->arena.live_blocks() > 0;
      if (res == GridArena::kNoMemory && !others) {
        r.status = kStatusRejected;  // nothing will come back: the device cannot hold it
        s->n_rejected++;
        s->queue.pop_front();
        continue;
      }
      break;  // back-pressure until finished grids are released
    }
    r.g = g;
    r.bytes = need;
    r.slot = s->free_slots.back();
    s->free_slots.pop_back();
    FS3D_HIP(launch_fill_rain(r.g, r.nx, r.ny, r.nz, r.seed, r.thr, s->s_fill));
    FS3D_HIP(hipMemsetAsync(s->d_hashes + r.slot, 0, sizeof(u64), s->s_fill));
    // all four change flags set: the first ticks of a new world always run
    FS3D_HIP(hipMemsetAsync(s->d_chg + (size_t)r.slot * 4, 1, 4 * sizeof(u32), s->s_fill));
    r.status = kStatusActive;
    r.t_admit = now_s() - s->t0;
    r.remaining = r.ticks;
    r.tick = 0;
    s->pending.push_back(s->queue.front());
    s->queue.pop_front();
    any = true;
  }
  if (any) FS3D_HIP(hipEventRecord(s->fill_ev, s->s_fill));
  return FS3D_OK;
}

// Plans and enqueues one round of up to K ticks over the active worlds.
int launch_round(fs3d_server *s) {
  if (s->inflight.size() >= 2) return FS3D_OK;
  if (s->active.empty() && s->pending.empty()) return FS3D_OK;
  if (!s->pending.empty()) {
    FS3D_HIP(hipStreamWaitEvent(s->s_tick, s->fill_ev, 0));
    for (long long id : s->pending) s->active.push_back(id);
    s->pending.clear();
  }
  const int n = (int)s->active.size();
  int K = 0;
  for (long long id : s->active) K = std::max(K, std::min(s->cfg.round_ticks, s->reqs[id].remaining));
  const int ri = s->next_round;
  Round &rd = s->rounds[ri];
  const size_t ent_bytes = sizeof(BWorld) * (size_t)n, pre_bytes = sizeof(long long) * (size_t)(n + 1) * K;
  // workgroup counts per launch
  std::vector<long long> tot(K, 0);
  std::vector<BWorld> ents(n);
  for (int i = 0; i < n; i++) {
    const Request &r = s->reqs[s->active[i]];
    BWorld &w = ents[i];
    w.g = r.g;
    w.nx = r.nx;
    w.ny = r.ny;
    w.nz = r.nz;
    w.lr = (r.nx + s->seg - 1) / s->seg;
    w.rqpw = 64 / w.lr;
    w.nt = std::min(s->cfg.round_ticks, r.remaining);
    w.slot = r.slot;
    w.fin = r.remaining == w.nt ? 1 : 0;
    w.tick = r.tick;
  }
  std::vector<long long> pre((size_t)(n + 1) * K);
  for (int k = 0; k < K; k++) {
    long long acc = 0;
    for (int i = 0; i < n; i++) {
      pre[(size_t)k * (n + 1) + i] = acc;
      if (ents[i].nt > k) acc += world_blocks(ents[i], (int)((ents[i].tick + (u64)k) & 1u));
    }
    pre[(size_t)k * (n + 1) + n] = acc;
    tot[k] = acc;
  }
  size_t blks = 0;
  for (int k = 0; k < K; k++) blks += (size_t)tot[k];
  // requests whose last tick falls in this round
  std::vector<long long> fin;
  for (int i = 0; i < n; i++)
    if (s->reqs[s->active[i]].remaining == ents[i].nt) fin.push_back(s->active[i]);
  const size_t koff_bytes = sizeof(long long) * (size_t)(K + 1);
  if (int rc = ensure_round_cap(rd, ent_bytes + pre_bytes + koff_bytes, blks, (int)fin.size(), K + 1)) return rc;
  std::copy(ents.begin(), ents.end(), (BWorld *)rd.h_buf);
  std::copy(pre.begin(), pre.end(), (long long *)(rd.h_buf + ent_bytes));
  long long *koff = (long long *)(rd.h_buf + ent_bytes + pre_bytes);
  koff[0] = 0;
  for (int k = 0; k < K; k++) koff[k + 1] = koff[k] + tot[k];
  FS3D_HIP(hipMemcpyAsync(rd.d_buf, rd.h_buf, ent_bytes + pre_bytes + koff_bytes, hipMemcpyHostToDevice, s->s_tick));
  const BWorld *d_ents = (const BWorld *)rd.d_buf;
  const long long *d_pre = (const long long *)(rd.d_buf + ent_bytes);
  // one pass fills the workgroup -> entry tables of all K launches
  FS3D_HIP(launch_expand_round(d_pre, n, K, (const long long *)(rd.d_buf + ent_bytes + pre_bytes), (long long)blks,
                               rd.d_blk, s->s_tick));
  const u64 *rule = device_rule_table(s->dev);
  size_t boff = 0;
  FS3D_HIP(hipEventRecord(rd.ev[0], s->s_tick));
  for (int k = 0; k < K; k++) {
    FS3D_HIP(launch_wide(s->seg, d_ents, d_pre + (size_t)k * (n + 1), rd.d_blk + boff, tot[k], k, rule, s->d_hashes,
                         s->s_tick, s->d_chg));
    FS3D_HIP(hipEventRecord(rd.ev[k + 1], s->s_tick));
    boff += (size_t)tot[k];
  }
  for (size_t i = 0; i < fin.size(); i++)
    FS3D_HIP(hipMemcpyAsync(rd.h_hash + i, s->d_hashes + s->reqs[fin[i]].slot, sizeof(u64), hipMemcpyDeviceToHost,
                            s->s_tick));
  FS3D_HIP(hipEventRecord(rd.done, s->s_tick));
  rd.finishing = fin;
  rd.nlaunch = K;
  rd.inflight = true;
  s->inflight.push_back(ri);
  s->next_round ^= 1;
  s->n_rounds++;
  s->n_launches += K;
  s->peak_active = std::max(s->peak_active, n);
  // advance host state
  std::vector<long long> still;
  for (int i = 0; i < n; i++) {
    Request &r = s->reqs[s->active[i]];
    r.tick += (u64)ents[i].nt;
    r.remaining -= ents[i].nt;
    s->voxel_updates += (double)r.nx * r.ny * r.nz * ents[i].nt;
    if (r.remaining > 0) still.push_back(s->active[i]);
  }
  s->active.swap(still);
  return FS3D_OK;
}

int poll_once(fs3d_server *s) {
  if (int rc = reap(s)) return rc;
  if (int rc = admit(s)) return rc;
  if (int rc = launch_round(s)) return rc;
  return FS3D_OK;
}

bool idle(const fs3d_server *s) {
  return s->queue.empty() && s->pending.empty() && s->active.empty() && s->inflight.empty();
}

}  // namespace

extern "C" {

int fs3d_server_create(const fs3d_server_config *cfg, fs3d_server **out) {
  if (!cfg || !out) return set_error(FS3D_ERR_INVALID, "cfg/out is NULL");
  *out = nullptr;
  if (cfg->budget_bytes <= 0 || cfg->round_ticks < 1 || cfg->max_active < 1 ||
      (cfg->seg != 0 && cfg->seg != 8 && cfg->seg != 16 && cfg->seg != 32))
    return set_error(FS3D_ERR_INVALID, "server config: budget > 0, round_ticks >= 1, max_active >= 1, seg 0/8/16/32");
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (int rc = ensure_device(dev)) return rc;
  fs3d_server *s = new fs3d_server();
  s->dev = dev;
  s->cfg = *cfg;
  s->seg = cfg->seg ? cfg->seg : 8;
  s->arena.init((size_t)cfg->budget_bytes, std::min<size_t>((size_t)cfg->budget_bytes, (size_t)1 << 30),
                 chunk_alloc, chunk_free, nullptr);
  hipError_t e = hipStreamCreateWithFlags(&s->s_tick, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->s_fill, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->fill_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&s->d_hashes, sizeof(u64) * (size_t)cfg->max_active);
  if (e == hipSuccess) e = hipMalloc(&s->d_chg, 4 * sizeof(u32) * (size_t)cfg->max_active);
  if (e != hipSuccess) {
    fs3d_server_destroy(s);
    return set_error(FS3D_ERR_HIP, "server init failed: %s", hipGetErrorString(e));
  }
  for (int i = cfg->max_active - 1; i >= 0; i--) s->free_slots.push_back(i);
  s->t0 = now_s();
  *out = s;
  return clear_error();
}

int fs3d_server_destroy(fs3d_server *s) {
  if (!s) return clear_error();
  (void)hipSetDevice(s->dev);
  if (s->s_tick) (void)hipStreamSynchronize(s->s_tick);
  if (s->s_fill) (void)hipStreamSynchronize(s->s_fill);
  for (Round &rd : s->rounds) {
    if (rd.h_buf) (void)hipHostFree(rd.h_buf);
    if (rd.d_buf) (void)hipFree(rd.d_buf);
    if (rd.d_blk) (void)hipFree(rd.d_blk);
    if (rd.h_hash) (void)hipHostFree(rd.h_hash);
    for (hipEvent_t e : rd.ev) (void)hipEventDestroy(e);
    if (rd.done) (void)hipEventDestroy(rd.done);
  }
  s->arena.clear();
  if (s->d_hashes) (void)hipFree(s->d_hashes);
  if (s->d_chg) (void)hipFree(s->d_chg);
  if (s->fill_ev) (void)hipEventDestroy(s->fill_ev);
  if (s->s_tick) (void)hipStreamDestroy(s->s_tick);
  if (s->s_fill) (void)hipStreamDestroy(s->s_fill);
  delete s;
  return clear_error();
}

int fs3d_server_submit(fs3d_server *s, const int32_t dims[3], uint64_t seed, uint64_t thr, int32_t ticks,
                       int64_t *id) {
  if (int rc = check_server(s)) return rc;
  if (!dims || dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || ticks < 1)
    return set_error(FS3D_ERR_INVALID, "request: dims >= 1 and ticks >= 1 required");
  if ((dims[0] + s->seg - 1) / s->seg > 64)
    return set_error(FS3D_ERR_INVALID, "request nx=%d exceeds 64 x %d-byte segments", dims[0], s->seg);
  Request r;
  r.nx = dims[0];
  r.ny = dims[1];
  r.nz = dims[2];
  r.seed = seed;
  r.thr = thr;
  r.ticks = ticks;
  r.t_submit = now_s() - s->t0;
  s->reqs.push_back(r);
  const long long rid = (long long)s->reqs.size() - 1;
  s->queue.push_back(rid);
  if (id) *id = rid;
  return clear_error();
}

int fs3d_server_poll(fs3d_server *s, int32_t *n_active, int32_t *n_queued, int64_t *n_done) {
  if (int rc = check_server(s)) return rc;
  FS3D_HIP(hipSetDevice(s->dev));
  if (int rc = poll_once(s)) return rc;
  if (n_active) *n_active = (int32_t)(s->active.size() + s->pending.size());
  if (n_queued) *n_queued = (int32_t)s->queue.size();
  if (n_done) *n_done = s->n_done;
  return clear_error();
}

int fs3d_server_drain(fs3d_server *s) {
  if (int rc = check_server(s)) return rc;
  FS3D_HIP(hipSetDevice(s->dev));
  while (!idle(s)) {
    if (int rc = poll_once(s)) return rc;
  }
  return clear_error();
}

int fs3d_server_result(const fs3d_server *s, int64_t id, int32_t *status, uint64_t *hash, double *latency_ms) {
  if (int rc = check_server(s)) return rc;
  if (id < 0 || id >= (int64_t)s->reqs.size()) return set_error(FS3D_ERR_INVALID, "unknown request id %lld", (long long)id);
  const Request &r = s->reqs[id];
  if (status) *status = r.status;
  if (hash) *hash = r.hash;
  if (latency_ms) *latency_ms = r.status == kStatusDone ? (r.t_done - r.t_submit) * 1e3 : -1.0;
  return clear_error();
}

int fs3d_server_replay(fs3d_server *s, int32_t n, const int32_t *dims, const uint64_t *seeds, const uint64_t *thr,
                       const int32_t *ticks, const double *arrival, int64_t *ids) {
  if (int rc = check_server(s)) return rc;
  if (n < 0 || (n > 0 && (!dims || !seeds || !thr || !ticks || !arrival)))
    return set_error(FS3D_ERR_INVALID, "replay: bad arguments");
  FS3D_HIP(hipSetDevice(s->dev));
  const double start = now_s();
  int next = 0;
  while (next < n || !idle(s)) {
    const double t = now_s() - start;
    while (next < n && arrival[next] <= t) {
      int64_t id = -1;
      if (int rc = fs3d_server_submit(s, dims + 3 * next, seeds[next], thr[next], ticks[next], &id)) return rc;
      if (ids) ids[next] = id;
      next++;
    }
    if (int rc = poll_once(s)) return rc;
  }
  return clear_error();
}

int fs3d_server_get_stats(const fs3d_server *s, fs3d_server_stats *out) {
  if (int rc = check_server(s)) return rc;
  if (!out) return set_error(FS3D_ERR_INVALID, "out is NULL");
  fs3d_server_stats st{};
  st.submitted = (int64_t)s->reqs.size();
  st.completed = s->n_done;
  st.rejected = s->n_rejected;
  st.rounds = s->n_rounds;
  st.launches = s->n_launches;
  st.tick_p50_ms = percentile(s->tick_ms, 0.50);
  st.tick_p99_ms = percentile(s->tick_ms, 0.99);
  st.tick_max_ms = s->tick_ms.empty() ? 0.0 : *std::max_element(s->tick_ms.begin(), s->tick_ms.end());
  std::vector<double> lat;
  for (const Request &r : s->reqs)
    if (r.status == kStatusDone) lat.push_back((r.t_done - r.t_submit) * 1e3);
  st.req_p50_ms = percentile(lat, 0.50);
  st.req_p99_ms = percentile(lat, 0.99);
  st.voxel_updates = s->voxel_updates;
  st.busy_ms = s->busy_ms;
  st.peak_bytes = (int64_t)s->arena.peak_in_use();
  st.used_bytes = (int64_t)s->arena.in_use();
  st.alloc_bytes = (int64_t)s->arena.allocated();
  st.peak_alloc_bytes = (int64_t)s->arena.peak_allocated();
  st.peak_active = s->peak_active;
  *out = st;
  return clear_error();
}

}  // extern "C"
