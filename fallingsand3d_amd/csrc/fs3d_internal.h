// Internal declarations shared by the fs3d library sources (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/fs3d.h"

namespace fs3d {

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;

constexpr u64 kGamma = 0x9E3779B97F4A7C15ull;
// Hash partial sums: NPART slots, one 64-byte line apart, summed by a finaliser.
constexpr int kHashParts = 1024;
constexpr int kHashStride = 8;  // u64 per slot
// Bytes allocated past the end of every grid buffer: tick loads may read up to
// one lane segment (<= 64 bytes) beyond the last row.
constexpr size_t kGuardBytes = 256;
constexpr int kMaxNx = 4096;  // 64 lanes x 64-byte segments per row
constexpr int kMaxDevices = 16;

int set_error(int code, const char *fmt, ...);
int clear_error();

#define FS3D_HIP(expr)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::fs3d::set_error(FS3D_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---- host math shared by the runtime ------------------------------------------
u64 mix64(u64 z);
u64 sm64_next(u64 *state);
u64 mulhi64(u64 a, u64 b);
u64 uniform_incl(u64 r, u64 lo, u64 hi);
u64 prob_threshold(u64 num, u64 den);
int loguniform_int(u64 r, int lo, int hi);

// ---- rule table ------------------------------------------------------------------
// entries[v*65536 + s]: byte b of the u64 is the output category of cell b;
// bit 7 of byte 0 (kRuleChanged) is set when the block changes.
constexpr u64 kRuleChanged = 0x80ull;
void build_rule_entries(u64 *entries);
u16 rule_state_packed(u16 s, int v);

// ---- world files ------------------------------------------------------------------
u64 rule_hash();
// fill (when grid is NULL) provides the payload in chunks; sink (when grid is
// NULL) receives it.  Both return an fs3d status.
int file_write(const char *path, const int32_t dims[3], u64 tick, u64 state_hash, const u8 *grid, size_t nbytes,
               int (*fill)(void *ctx, size_t off, u8 *buf, size_t n), void *ctx);
int file_info(const char *path, int32_t dims[3], u64 *tick, u64 *state_hash, u64 *rule);
int file_read(const char *path, u8 *grid, size_t nbytes, int (*sink)(void *ctx, size_t off, const u8 *buf, size_t n),
              void *ctx);

// ---- device context -------------------------------------------------------------
int ensure_device(int dev);            // sets the device current, uploads the rule table
const u64 *device_rule_table(int dev);  // valid after ensure_device
// Looks up this device's copy of the no-allocation entry point's static scratch
// (with dev current) and clears it; called once per device by ensure_device.
int noalloc_setup(int dev);

// ---- tick kernels -----------------------------------------------------------------
struct TickShape {
  int seg;            // bytes per lane along x
  int lanes_per_row;  // ceil(nx / seg) (<= 64)
  int rq_per_wave;    // row quads (2 rows x 2 layers) per 64-lane wave
  long long waves;    // waves covering the larger parity's row-quad count
  unsigned blocks;    // workgroups of 4 waves
};
int plan_tick(int nx, int ny, int nz, int seg_req, TickShape *out);
// masks (optional): the two per-parity change-mask arrays, one u64 per row quad
// of the larger parity; when given, settled lane segments are skipped.  stats
// (optional, with masks): kHashParts slots of {swept voxels, read-only voxels}
// plus one violation counter.
// ctl (with masks): the listed tick's work-list counts, which a hash tick keeps
// consistent (see launch_tick_listed).
hipError_t launch_tick(const TickShape &sh, const u8 *src, u8 *dst, int nx, int ny, int nz,
                       u64 t, const u64 *tick_dev, u32 k_off, const u64 *rule,
                       u64 *hash_parts, hipStream_t s, u64 *const *masks = nullptr, u64 *stats = nullptr,
                       u32 *ctl = nullptr);
// The dense tick (no skipping) with one wave per x-chunk of 64 * seg bytes of a
// row quad (seg 8, 16 or 32; rows of any length up to kMaxNx).
hipError_t launch_tick_chunked(int seg, const u8 *src, u8 *dst, int nx, int ny, int nz, u64 t, const u64 *tick_dev,
                               u32 k_off, const u64 *rule, u64 *hash_parts, hipStream_t s);
hipError_t launch_fill_u64(u64 *p, size_t n, u64 v, hipStream_t s);
// Work-list entry of the listed skipping tick: active lanes of one row quad and
// its position (zq << 32 | yq).  On odd ticks of the packed schedule, edge bit s
// is set unless the block straddling x = s * seg is known not to change: it did
// not change two ticks before, and none of its cells changed one tick before.
struct TickEntry {
  u64 act;
  u64 pos;
  u64 edge;
};
// Change masks of one parity: mask_words change words (bit s: a block that lane
// segment s evaluated changed at the last tick of that parity), then mask_words
// edge words (bit s: a block holding cell x = s * seg - 1 or x = s * seg changed
// then; every writer may set more bits than that, never fewer).
inline size_t mask_words_of(int ny, int nz) { return (size_t)((ny + 2) >> 1) * (size_t)((nz + 2) >> 1); }
// Skipping tick in two launches: a plan pass over all row quads that lists
// those with active lanes (list capacity mask_words), then a persistent pass of
// list_waves waves (a multiple of 4) over the list.  No hash.  ctl holds three
// list counts, used by ticks t mod 3; the plan pass of tick t (or the hash tick
// at t) clears slot (t+1) mod 3, so all three must be zero before the first
// tick after any change of the tick counter.
hipError_t launch_tick_listed(const TickShape &sh, const u8 *src, u8 *dst, int nx, int ny, int nz, u64 t,
                              const u64 *tick_dev, u32 k_off, const u64 *rule, hipStream_t s,
                              u64 *const *masks, u64 *stats, TickEntry *list, u32 *ctl, size_t mask_words,
                              u32 list_waves);
// Resident 256-thread workgroups per compute unit of the listed tick's
// persistent kernel for a segment size.
int list_blocks_per_cu(int seg);
// Skipping tick that packs the active lane segments of the listed row quads
// into full waves: the listed tick's plan pass, then a persistent pass of
// pack_waves waves (a multiple of 4) over the list, `chunk` (<= 64) entries per
// wave and pass.  No hash.  Same list, capacity and count slots as the listed
// tick.
hipError_t launch_tick_packed(const TickShape &sh, const u8 *src, u8 *dst, int nx, int ny, int nz, u64 t,
                              const u64 *tick_dev, u32 k_off, const u64 *rule, hipStream_t s, u64 *const *masks,
                              u64 *stats, TickEntry *list, u32 *ctl, size_t mask_words, u32 pack_waves,
                              int chunk);
int pack_blocks_per_cu(int seg);
// Sets change bits `bits` of row quad i0 (parity 0) and i1 (parity 1), and the
// edge bits next to them (edge words at offset edge_off).
hipError_t launch_mark_cells(u64 *m0, u64 *m1, long long i0, long long i1, u64 bits, size_t edge_off, hipStream_t s);
// Two dense ticks (t and t + 1) in one pass: src -> dst holds the state after
// both.  Supported for one row quad per wave and LDS tiles within 64 KiB.
bool tick2_supported(const TickShape &sh, int nx, size_t *lds_bytes);
hipError_t launch_tick2(const TickShape &sh, const u8 *src, u8 *dst, int nx, int ny, int nz, u64 t,
                        const u64 *tick_dev, u32 k_off, const u64 *rule, int zlen, hipStream_t s);
hipError_t launch_hash(const u8 *g, size_t n, u64 *hash_parts, hipStream_t s);
hipError_t launch_hash_finalize(u64 *hash_parts, u64 *out, hipStream_t s);
hipError_t launch_set_u64(u64 *p, u64 v, hipStream_t s);
hipError_t launch_counts(const u8 *g, size_t n, u64 *counts, hipStream_t s);

// ---- scene kernels ----------------------------------------------------------------
hipError_t launch_fill_rain(u8 *g, int nx, int ny, int nz, u64 seed, u64 thr, hipStream_t s);
struct Pillar {
  int cx, cy, r, zt;
};
hipError_t launch_fill_box(u8 *g, int nx, int ny, int x0, int y0, int z0, int sx, int sy, int sz,
                           u8 v, hipStream_t s);
hipError_t launch_fill_pillars(u8 *g, int nx, int ny, const Pillar *dev_pillars, int n,
                               int max_r, int max_zt, hipStream_t s);
// Host box painter; returns the number of boxes or -1 on bad parameters.
long long paint_boxes(u8 *g, int nx, int ny, int nz, u64 seed, const fs3d_box_params &p);
int pillar_list(int nx, int ny, int nz, u64 seed, const fs3d_pillar_params &p,
                int pool[5] /* x0,y0,W,D,ztop */, std::vector<Pillar> *out);

}  // namespace fs3d

// ---- world object ----------------------------------------------------------------
struct fs3d_world {
  int dev = 0;
  int nx = 0, ny = 0, nz = 0;
  size_t n = 0;      // voxels
  size_t alloc = 0;  // bytes per buffer (voxels + guard, rounded)
  fs3d::u8 *buf[2] = {nullptr, nullptr};
  int cur = 0;  // index of the buffer holding the current state
  fs3d::u64 tick = 0;
  hipStream_t stream = nullptr;
  fs3d::u64 *d_parts = nullptr;  // hash partials
  fs3d::u64 *d_small = nullptr;  // [0] device tick, [1..] hash outputs
  int small_cap = 0;             // number of hash output slots
  fs3d::u64 *h_pinned = nullptr; // pinned staging for small read-backs
  int h_pinned_cap = 0;
  fs3d::TickShape shape{};
  fs3d::TickShape shape2{};  // the fused two-tick pass: smallest segment with whole rows per wave
  int seg_req = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_run_ms = 0.0;
  // render outputs on the device (grown on demand)
  float *d_rgb = nullptr;
  fs3d::u64 *d_hit = nullptr;
  fs3d::u8 *d_face = nullptr;
  size_t render_px = 0;
  fs3d::u64 *d_occ = nullptr;    // occupancy maps of the renderer (coarse, then fine)
  fs3d::u64 *d_dirty = nullptr;  // bricks changed since the maps were built (coarse-map layout)
  int track_mode = FS3D_TRACK_AUTO;
  bool occ_track = false;  // set by a render when tracking: skipping ticks record changed bricks in d_dirty
  bool occ_valid = false;  // d_occ, updated at the bricks of d_dirty, describes the current grid
  double last_render_ms = 0.0;
  // settled-segment skipping: mode (FS3D_SKIP_*) and per-parity change masks
  // (one change word and one edge word per row quad: 2 * mask_words u64)
  int skip = FS3D_SKIP_AUTO;
  bool fuse = false;  // dense ticks: pairs of ticks in one pass where supported
  fs3d::u64 *d_masks[2] = {nullptr, nullptr};
  size_t mask_words = 0;
  fs3d::u64 *d_stats = nullptr;  // sweep counters, allocated when enabled
  fs3d::TickEntry *d_list = nullptr;  // work list of the listed tick (mask_words entries)
  fs3d::u32 *d_ctl = nullptr;         // list counts, one per tick mod 3
  fs3d::u32 list_waves = 0;           // persistent waves of the listed tick
  fs3d::u32 pack_waves = 0;           // persistent waves of the packed tick
  // cached graphs keyed by (nticks, hash_every, cur)
  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;
};

// Destroys the world's cached tick graphs (they are captured again on next use).
void world_drop_graphs(fs3d_world *w);
// Ticks that record no changed bricks (skipping off, or tracking off) leave
// the renderer's occupancy maps stale.
inline void note_ticks(fs3d_world *w) {
  if (!w->skip || !w->occ_track) w->occ_valid = false;
}
