// Internal declarations of the batched-world, wide multi-world and serving kernels.
#pragma once

#include "fs3d_internal.h"

namespace fs3d {

// Device world-table entry.  The grid is updated in place.
struct BWorld {
  u8 *g;
  int nx, ny, nz;
  int lr, rqpw;  // lanes per row / row quads per wave for the launch's segment width
  int nt;        // wide launches: ticks this entry runs in the round
  int slot;      // wide launches: index of the entry's hash slot
  int fin;       // wide launches: 1 when the round's last tick is the entry's final one (hash it)
  u64 tick;      // tick counter before the next run (round)
};

// The grid of a world-table entry as a global-memory pointer.  A pointer read
// from a table is generic to the compiler, which then emits flat loads and
// stores (they also wait on the LDS counter); passed through an integer into
// an address_space(1) pointer, it lets the compiler emit global ones.
__device__ __forceinline__ u8 *grid_of(const BWorld &w) {
  return (u8 *)(__attribute__((address_space(1))) u8 *)(uintptr_t)w.g;
}

bool batch_config_ok(int seg, int threads, int groups);
int batch_occupancy(int seg, int threads, int groups);
// flags of the persistent batch launch: hash every world's final tick; stop
// ticking a world after two ticks in which nothing changed (its state is then a
// fixed point of both parities) and run only its final tick.
constexpr int kBatchHash = 1, kBatchSettle = 2;
hipError_t launch_batch(int seg, int threads, int groups, BWorld *worlds, const int *order, int n, int nticks,
                        u32 *queue, const u64 *rule, u64 *hashes, int flags, int grid, hipStream_t s);
hipError_t launch_batch_hash(const BWorld *worlds, int n, u64 *hashes, hipStream_t s);
hipError_t launch_batch_fill_rain(const BWorld *worlds, int n, const u64 *seeds, const u64 *thr, size_t max_cells,
                                  hipStream_t s);

// Waves of rqpw row quads each that a tick of parity p needs.
__host__ __device__ inline long long world_waves_dev(int ny, int nz, int rqpw, int p) {
  const long long nrq = (long long)((ny + p + 1) >> 1) * ((nz + p + 1) >> 1);
  return (nrq + rqpw - 1) / rqpw;
}

// Wide launch k of a round: entry i owns workgroups [bpre[i], bpre[i+1]) (4 waves
// each; none when it sits the launch out); blk2ent maps workgroup -> entry.
// chg (optional): settled-world flags, four u32 per hash slot (see wide.hip).
hipError_t launch_wide(int seg, const BWorld *ents, const long long *bpre, const int *blk2ent, long long blocks, int k,
                       const u64 *rule, u64 *hashes, hipStream_t s, u32 *chg = nullptr);
hipError_t launch_expand(const long long *bpre, int n, int *blk2ent, hipStream_t s);
// Tables of K launches in one pass (bpre: K prefixes of n+1; koff: K+1 launch
// offsets into blk2ent; total = koff[K]).
hipError_t launch_expand_round(const long long *bpre, int n, int K, const long long *koff, long long total,
                               int *blk2ent, hipStream_t s);
hipError_t launch_zero_slots(u64 *hashes, const BWorld *ents, int n, hipStream_t s);
long long world_waves(const BWorld &w, int p);
long long world_blocks(const BWorld &w, int p);

}  // namespace fs3d
