/*
 * fs3d oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, CPU restatement of the fs3d voxel tick, its state hash, the seeded
 * scene generators and the render transcriptions.  It exists to check the HIP
 * product path; the product never links, loads or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only as
 * the checker (or, in bench.py, as the timed CPU baseline).
 *
 * Provenance.  The reference snapshot (amarde05/FallingSand3D) holds no voxel
 * grid, no tick, no hash and no RNG: it is a Vulkan renderer skeleton whose only
 * per-frame update is camera motion (src/engine/rendering/renderer.cpp:437-467)
 * and whose only GPU program near this path is the analytic sphere tracer
 * shaders/fs_raymarch.frag.  The tick rules, hash and scenes restated here are
 * therefore project-authored; they are written out in full in the comments of
 * fs3d_oracle.c.  What pins each function (no reference result can):
 *
 *   fso_mix64 / fso_sm64_next / fso_cell_rand   SplitMix64 known-answer vectors
 *   fso_fmix32                                   MurmurHash3_x86_32 known answers
 *   fso_block_rule / fso_build_lut               exhaustive invariants over all
 *                                                65,536 states x 4 variants
 *                                                (conservation, statics fixed,
 *                                                gravity, one move per cell,
 *                                                variant-independent fixedness)
 *                                                and hand-worked cases
 *   fso_tick                                     closed forms (falling grain,
 *                                                settled fixed points), brute
 *                                                force against fso_tick_lut and a
 *                                                committed 64^3 x 256-tick hash
 *                                                fixture
 *   fso_hash                                     closed form on 1-2 cell grids
 *   scenes                                       exact cell counts, closed forms
 *   render (fs3d_oracle_render.c)                analytic ray/sphere and ray/box
 *                                                intersection, axis-aligned rays
 */
#ifndef FS3D_ORACLE_H
#define FS3D_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FSO_EMPTY = 0, FSO_STATIC = 1, FSO_SAND = 2, FSO_FLUID = 3 };

#define FSO_GAMMA 0x9E3779B97F4A7C15ull

/* ---- hashing / RNG ---------------------------------------------------- */
uint64_t fso_mix64(uint64_t z);
uint64_t fso_sm64_next(uint64_t *state);
uint64_t fso_cell_rand(uint64_t key, uint64_t i);
uint64_t fso_mulhi64(uint64_t a, uint64_t b);
uint64_t fso_uniform(uint64_t r, uint64_t lo, uint64_t hi);
uint64_t fso_threshold(uint64_t num, uint64_t den);
uint32_t fso_fmix32(uint32_t h);
uint32_t fso_group_word(uint32_t g, uint32_t by, uint32_t bz, uint32_t t);
int fso_variant(uint32_t bx, uint32_t by, uint32_t bz, uint32_t t);
int fso_loguniform(uint64_t r, int lo, int hi);

/* ---- block rule -------------------------------------------------------- */
void fso_block_rule(const uint8_t cat_in[8], int variant, uint8_t cat_out[8]);
uint16_t fso_rule_state(uint16_t s, int variant);
void fso_build_lut(uint16_t *lut /* 4 * 65536 */);

/* ---- tick / hash --------------------------------------------------------- */
void fso_set_threads(int n);
int fso_get_threads(void);
void fso_tick(const uint8_t *src, uint8_t *dst, int nx, int ny, int nz, uint64_t t);
void fso_tick_lut(const uint8_t *src, uint8_t *dst, int nx, int ny, int nz, uint64_t t,
                  const uint16_t *lut);
uint64_t fso_hash(const uint8_t *g, uint64_t n);
void fso_counts(const uint8_t *g, uint64_t n, uint64_t counts[256]);
/* Runs nticks ticks in place starting at tick t0; when hash_every > 0 writes the
 * hash of the state after every hash_every-th tick to hashes[]. Returns 0, or -1
 * when the scratch buffer cannot be allocated. */
int fso_run(uint8_t *grid, int nx, int ny, int nz, uint64_t t0, int nticks, int hash_every,
            uint64_t *hashes);
/* Same, evaluating blocks through a table built by fso_build_lut (lut != NULL). */
int fso_run_lut(uint8_t *grid, int nx, int ny, int nz, uint64_t t0, int nticks, int hash_every,
                uint64_t *hashes, const uint16_t *lut);

/* ---- scenes -------------------------------------------------------------- */
void fso_scene_rain(uint8_t *g, int nx, int ny, int nz, uint64_t seed, uint64_t thr);
typedef struct {
  int32_t side_min, side_max;
  int32_t pct_sand, pct_fluid, pct_static; /* percent of all cells */
  int32_t max_boxes;
} fso_box_params;
/* returns number of boxes painted, or -1 on invalid parameters */
int64_t fso_scene_boxes(uint8_t *g, int nx, int ny, int nz, uint64_t seed, const fso_box_params *p);
typedef struct {
  int32_t n_pillars, r_min, r_max, h_min, h_max;
  int32_t pool_w_min, pool_w_max, pool_d_min, pool_d_max, pool_h_min, pool_h_max;
} fso_pillar_params;
int fso_scene_pillars(uint8_t *g, int nx, int ny, int nz, uint64_t seed, const fso_pillar_params *p);

/* Batched worlds: per-axis edges drawn log-uniform in [lo,hi]; per-world sand
 * threshold drawn in [2^64/20, 2^64/2); per-world seed. */
void fso_batch_specs(int n, uint64_t seed, int lo, int hi, int32_t *dims /* n*3 */,
                     uint64_t *thr, uint64_t *seeds);
/* Serving requests: as fso_batch_specs plus tick counts uniform in
 * [tick_lo, tick_hi] and Poisson arrival times (seconds) at rate lam. */
void fso_serve_specs(int n, uint64_t seed, int lo, int hi, int tick_lo, int tick_hi, double lam,
                     int32_t *dims, uint64_t *thr, uint64_t *seeds, int32_t *ticks,
                     double *arrival);

/* ---- render (fs3d_oracle_render.c) ----------------------------------------- */
typedef struct {
  float origin[3];
  float basis[9]; /* screen-left, screen-down, forward (world dir = dx*b0 + dy*b1 + dz*b2) */
  float light[3];
  float aspect;
} fso_camera;
void fso_pixel_dir(int px, int py, int W, int H, float aspect, float d[3]);
void fso_raymarch_sdf(const float cam_pos[3], float aspect, int W, int H, float *rgba);
/* shaders/raymarch_sphere.{vert,frag}: object-space ray origin of the vertex
 * stage, and the fragment stage over n (origin, local position) pairs. */
void fso_box_vertex(const float cam_pos[3], const float inv_model[16], float origin[3]);
void fso_box_march(const float *origin, const float *local, int n, float *rgba);
void fso_render_dda(const uint8_t *g, int nx, int ny, int nz, const fso_camera *cam, const float *palette, int W,
                    int H, float *rgb, uint64_t *hit, uint8_t *face);
void fso_orbit_camera(int nx, int ny, int nz, double angle, double dist_scale, double elev, float aspect,
                      fso_camera *out);

#ifdef __cplusplus
}
#endif
#endif
