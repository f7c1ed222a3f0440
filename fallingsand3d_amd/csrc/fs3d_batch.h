// Internal declarations of the batched-world kernels.
#pragma once

#include "fs3d_internal.h"

namespace fs3d {

// Device world-table entry.  The grid is updated in place.
struct BWorld {
  u8 *g;
  int nx, ny, nz;
  int lr, rqpw;  // lanes per row / row quads per wave for the batch's segment width
  int pad;
  u64 tick;      // tick counter before the next run
};

bool batch_config_ok(int seg, int threads);
int batch_occupancy(int seg, int threads);
hipError_t launch_batch(int seg, int threads, BWorld *worlds, const int *order, int n, int nticks, u32 *queue,
                        const u64 *rule, u64 *hashes, int want_hash, int grid, hipStream_t s);
hipError_t launch_batch_hash(const BWorld *worlds, int n, u64 *hashes, hipStream_t s);
hipError_t launch_batch_fill_rain(const BWorld *worlds, int n, const u64 *seeds, const u64 *thr, size_t max_cells,
                                  hipStream_t s);

}  // namespace fs3d
