// Internal declarations of the render kernels.
#pragma once

#include "fs3d_internal.h"

namespace fs3d {

// Camera of the voxel ray caster: ray origin, basis (screen-left, screen-down,
// forward; world dir = dx*b0 + dy*b1 + dz*b2 before normalising), light position
// and the aspect ratio the pixel mapping divides uv.y by.
struct RenderCam {
  float origin[3];
  float basis[9];
  float light[3];
  float aspect;
};

hipError_t launch_sdf(const float cam_pos[3], float aspect, int W, int H, float *d_rgba, hipStream_t s);
// u64 words of a grid's occupancy maps: the coarse map (one bit per 8x8x8
// brick, 64 bricks per word along x) followed by the fine map (one word per
// brick, one bit per 2x2x2 sub-brick).
size_t occ_words(int nx, int ny, int nz);
// u64 words of the map of bricks changed since the occupancy maps were built
// (one bit per 8x8x8 brick, the coarse map's layout).
size_t dirty_words(int nx, int ny, int nz);
// ORs into d_dirty the bricks holding a lane segment that tick t (t = *tick_dev +
// k_off when tick_dev is set) marked changed in its parity's masks.
hipError_t launch_occ_dirty(u64 *const masks[2], u64 t, const u64 *tick_dev, u32 k_off, int nx, int ny, int nz,
                            int seg, u64 *d_dirty, hipStream_t s);
// Builds the occupancy maps in d_occ (occ_words u64) -- all bricks, or with
// update only those marked in d_dirty -- and clears d_dirty; then casts the rays.
hipError_t launch_dda(const u8 *g, int nx, int ny, int nz, const RenderCam &cam, int W, int H, float *d_rgb,
                      u64 *d_hit, u8 *d_face, u64 *d_occ, u64 *d_dirty, bool update, hipStream_t s);
void palette_host(float *rgba);
hipError_t launch_box_march(const float *d_origin, const float *d_local, int n, float *d_rgba, hipStream_t s);

}  // namespace fs3d
