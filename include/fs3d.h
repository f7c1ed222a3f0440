/*
 * fs3d.h -- C ABI of the MI355X-native (gfx950) falling-sand voxel library.
 *
 * Plain C types only (pointers, sizes, fixed-width integers).  Every call returns
 * an fs3d_status; on failure fs3d_last_error() returns a thread-local message.
 *
 * The reference (amarde05/FallingSand3D) is one monolithic Vulkan executable with
 * no plugin, operator or FFI API (src/CMakeLists.txt:1-8) and no voxel grid.  The
 * seams this ABI replaces are named per entry point:
 *   - the engine loop call site, where a world step belongs between event
 *     handling and drawing (src/engine/engine.cpp:59-69);
 *   - the renderer's hand-off of data to shaders through per-material set-2
 *     bindings (MaterialLayout::addDataBinding / Material::writeBuffer,
 *     src/engine/rendering/materials.cpp:26-59, :388-412) and the fs_raymarch
 *     full-screen draw (shaders/fs_raymarch.frag:67-81);
 *   - the GlobalData camera block written each frame
 *     (src/engine/rendering/materials.cpp:531-575, materials.h:19-28);
 *   - error reporting, which in the reference logs and throws
 *     (util::displayError, src/util/debug.cpp:23-27); here codes + message, and
 *     include/fs3d.hpp rethrows as std::runtime_error.
 *
 * Threading: calls on different handles may run concurrently; calls on one
 * handle must be serialised by the caller.  fs3d_world_step is asynchronous on
 * the world's HIP stream; hash, download, get_voxel, run and render synchronise.
 */
#ifndef FS3D_H
#define FS3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS3D_ABI_VERSION 1

typedef enum fs3d_status {
  FS3D_OK = 0,
  FS3D_ERR_INVALID = 1,   /* bad argument (dims, coordinates, sizes, NULL) */
  FS3D_ERR_NOMEM = 2,     /* device or host allocation failed */
  FS3D_ERR_BUDGET = 3,    /* request exceeds the configured HBM budget */
  FS3D_ERR_HIP = 4,       /* a HIP runtime call failed */
  FS3D_ERR_NODEVICE = 5,  /* no usable gfx950 device */
  FS3D_ERR_STATE = 6      /* call not valid in the object's current state */
} fs3d_status;

/* Material bytes. 4..255 are reserved: they behave as STATIC and keep their value. */
enum { FS3D_EMPTY = 0, FS3D_STATIC = 1, FS3D_SAND = 2, FS3D_FLUID = 3 };

typedef struct fs3d_world fs3d_world; /* opaque */

/* ---- library ------------------------------------------------------------- */
/* Replaces: util::displayError's message (src/util/debug.cpp:23-27). */
const char *fs3d_last_error(void);
int fs3d_abi_version(void);
/* Number of visible HIP devices (0 without a GPU; never an error for that). */
int fs3d_device_count(int *count);
/* Selects the device for this thread and uploads the rule table to it.  Called
 * implicitly with device 0 by the first world creation.  Replaces the device
 * pick in VulkanDevice (src/engine/rendering/device.cpp:172-201). */
int fs3d_init(int device);
/* The packed block rule, lut[v*65536 + s] = R(s, v) (2 bits per cell); host
 * only, no device needed.  Exposes the table the tick kernels read. */
int fs3d_rule_lut(uint16_t *lut /* 4*65536 */);

/* ---- world grid ------------------------------------------------------------ */
/* uint8 v[z][y][x], x fastest, z up.  No reference counterpart exists (the
 * reference Renderer holds meshes, materials and a camera only,
 * src/engine/rendering/renderer.h:100-154).  flags: FS3D_WORLD_DENSE turns
 * settled-segment skipping off (see fs3d_world_set_skip). */
enum { FS3D_WORLD_DENSE = 1 };
int fs3d_world_create(const int32_t dims[3], uint32_t flags, fs3d_world **out);
int fs3d_world_destroy(fs3d_world *w); /* NULL is a no-op */
int fs3d_world_dims(const fs3d_world *w, int32_t dims[3]);
int fs3d_world_set_voxel(fs3d_world *w, int32_t x, int32_t y, int32_t z, uint8_t v);
int fs3d_world_get_voxel(fs3d_world *w, int32_t x, int32_t y, int32_t z, uint8_t *v);
int fs3d_world_upload(fs3d_world *w, const uint8_t *host, size_t nbytes);
int fs3d_world_download(fs3d_world *w, uint8_t *host, size_t nbytes);
int fs3d_world_get_tick(const fs3d_world *w, uint64_t *tick);
int fs3d_world_set_tick(fs3d_world *w, uint64_t tick);

/* ---- tick ------------------------------------------------------------------ */
/* Attaches at the engine loop seam (src/engine/engine.cpp:59-69): advances the
 * world by nticks ticks; asynchronous. */
int fs3d_world_step(fs3d_world *w, int32_t nticks);
/* One tick with the state hash of the new state fused into the same kernel. */
int fs3d_world_tick(fs3d_world *w, uint64_t *hash);
/* State hash of the current grid (synchronises). */
int fs3d_world_hash(fs3d_world *w, uint64_t *hash);
/* nticks ticks as one captured HIP graph (cached per shape), writing the hash
 * after every hash_every-th tick to hashes[nticks / hash_every] (hash_every = 0:
 * none).  Synchronises. */
int fs3d_world_run(fs3d_world *w, int32_t nticks, int32_t hash_every, uint64_t *hashes);
/* Captures and instantiates the graph fs3d_world_run(w, nticks, hash_every)
 * would launch from the current state, without running it (one-time set-up
 * kept out of a timed run).  No counterpart in the reference. */
int fs3d_world_prepare_run(fs3d_world *w, int32_t nticks, int32_t hash_every);
/* Per-material cell counts of the current grid (counts[256]); synchronises. */
int fs3d_world_counts(fs3d_world *w, uint64_t *counts);
/* Blocks until all work queued on the world's stream has finished. */
int fs3d_world_sync(fs3d_world *w);
/* Device pointer of the current grid, for a renderer on the same device (the
 * set-2 data binding hand-off, src/engine/rendering/materials.cpp:388-412).
 * Valid until the next step/run/tick on this world. */
int fs3d_world_device_ptr(fs3d_world *w, const void **dptr);

/* Tick kernel launch shape: bytes per lane along x (8, 16, 32 or 64; 0 = auto). */
int fs3d_world_set_segment(fs3d_world *w, int32_t seg_bytes);
int fs3d_world_get_segment(const fs3d_world *w, int32_t *seg_bytes);

/* Settled-segment skipping (FS3D_SKIP_AUTO by default).  A tick leaves a lane
 * segment (S bytes of 4 rows) unread and unwritten when none of the blocks it
 * evaluates changed at the previous tick of the same parity and none of their
 * cells changed at the tick before; results are bit-identical to the dense tick
 * (whether a block is a fixed point does not depend on its variant).  Three
 * schedules: LISTED runs a plan pass that lists the row quads with work and a
 * persistent pass over the list (settled regions cost nothing); SEGS reads the
 * same list and packs the listed row quads' active 64-byte sectors into full
 * waves (the block update is paid per active sector, not per row quad with
 * one); WAVES launches one wave per row-quad group that skips on its own (lower
 * per-tick overhead; best for small grids).  AUTO picks SEGS above 2^17 waves
 * of work per tick, else WAVES.  Hash ticks read settled segments in all
 * three.  Worlds of up to 2^27 voxels tick in place with WAVES and LISTED (and
 * dense single ticks of rows up to 1024 bytes); the grid stays where
 * fs3d_world_device_ptr last pointed only in that case.
 * Every host-side edit (set_voxel, upload, fills, set_tick, set_segment) keeps
 * the bookkeeping exact; writing the grid through fs3d_world_device_ptr is not
 * supported. */
enum { FS3D_SKIP_OFF = 0, FS3D_SKIP_AUTO = 1, FS3D_SKIP_LISTED = 2, FS3D_SKIP_WAVES = 3, FS3D_SKIP_SEGS = 4 };
int fs3d_world_set_skip(fs3d_world *w, int32_t mode);
int fs3d_world_get_skip(const fs3d_world *w, int32_t *mode);
/* Two-tick fusion of the dense tick (used when skipping is off): a pair of
 * ticks, neither hashed, runs as one pass that keeps the first tick's output in
 * LDS and writes only the second's, so the grid crosses HBM once per two ticks.
 * On by default for grids of more than 2^28 voxels that it supports (smaller
 * grids stay in the last-level cache, where single ticks are faster).  Needs
 * 256 < nx <= 3632 (rows up to 1024 bytes whole in a wave, longer ones in
 * 16-byte x-chunks of 1 KiB; LDS tiles within 64 KiB); get_fuse reports
 * whether fusion is in effect. */
int fs3d_world_set_fuse(fs3d_world *w, int32_t on);
int fs3d_world_get_fuse(const fs3d_world *w, int32_t *on);
/* Sweep counters of the skipping tick: enable != 0 starts (or keeps) counting,
 * 0 stops.  Returns and clears what was counted so far: out[0] = voxels of the
 * segments the skip rule updates (the whole-sector loads and stores around them
 * are not counted), out[1] = voxels only read (settled segments on hash ticks),
 * out[2] = settled segments whose blocks changed where the full-grid schedule
 * evaluated them (hash ticks, and beside an active segment of their 64-byte
 * sector; always 0 unless the skipping bookkeeping is wrong).  Synchronises. */
int fs3d_world_sweep_stats(fs3d_world *w, int32_t enable, uint64_t out[3]);

/* ---- scenes (seeded, bit-exact, generated by this library) ------------------- */
/* Rain fill: STATIC floor at z = 0; cells with z >= max(1, nz/2) are SAND when
 * cell_rand(seed, i) < thr.  Generated on the device. */
int fs3d_world_fill_rain(fs3d_world *w, uint64_t seed, uint64_t thr);
typedef struct fs3d_box_params {
  int32_t side_min, side_max;
  int32_t pct_sand, pct_fluid, pct_static;
  int32_t max_boxes;
} fs3d_box_params;
/* Seeded axis-aligned boxes over a floor, painted until target fractions are
 * met.  Generated on the host and uploaded.  *nboxes receives the box count. */
int fs3d_world_fill_boxes(fs3d_world *w, uint64_t seed, const fs3d_box_params *p, int64_t *nboxes);
typedef struct fs3d_pillar_params {
  int32_t n_pillars, r_min, r_max, h_min, h_max;
  int32_t pool_w_min, pool_w_max, pool_d_min, pool_d_max, pool_h_min, pool_h_max;
} fs3d_pillar_params;
/* Seeded sand pillars standing in a fluid pool; generated on the device. */
int fs3d_world_fill_pillars(fs3d_world *w, uint64_t seed, const fs3d_pillar_params *p);
/* Host-side spec streams for batched worlds and serving requests. */
int fs3d_batch_specs(int32_t n, uint64_t seed, int32_t lo, int32_t hi, int32_t *dims,
                     uint64_t *thr, uint64_t *seeds);
int fs3d_serve_specs(int32_t n, uint64_t seed, int32_t lo, int32_t hi, int32_t tick_lo,
                     int32_t tick_hi, double lam, int32_t *dims, uint64_t *thr, uint64_t *seeds,
                     int32_t *ticks, double *arrival);
int fs3d_threshold(uint64_t num, uint64_t den, uint64_t *thr);

/* ---- batched worlds -------------------------------------------------------- */
/* Many independent worlds of uneven sizes in one device arena.  No reference
 * counterpart: the reference runs one scene in one process (src/main.cpp:6-16).
 * FS3D_BATCH_PERSISTENT steps every world for nticks ticks in one persistent
 * kernel launch (captured in a HIP graph): workgroups pull worlds from an atomic
 * queue, largest first, and run all ticks of a world in place.
 * FS3D_BATCH_PLAIN launches the single-world tick kernel once per world per
 * tick (the baseline).  Both give bit-identical grids.  Worlds share the tick
 * counter; nx <= 2048. */
typedef struct fs3d_batch fs3d_batch; /* opaque */
/* FS3D_BATCH_WIDE launches one wide multi-world tick per tick (all worlds'
 * row quads over the whole GPU, captured in a graph); FS3D_BATCH_TUNED runs
 * the mode the last autotune picked (persistent until then). */
enum { FS3D_BATCH_PERSISTENT = 0, FS3D_BATCH_PLAIN = 1, FS3D_BATCH_WIDE = 2, FS3D_BATCH_TUNED = 3 };
int fs3d_batch_create(int32_t n, const int32_t *dims /* n*3: nx,ny,nz */, fs3d_batch **out);
int fs3d_batch_destroy(fs3d_batch *b); /* NULL is a no-op */
int fs3d_batch_size(const fs3d_batch *b, int32_t *n);
int fs3d_batch_upload(fs3d_batch *b, int32_t i, const uint8_t *host, size_t nbytes);
int fs3d_batch_download(fs3d_batch *b, int32_t i, uint8_t *host, size_t nbytes);
/* Rain fill of every world i with seeds[i], thr[i] (as fs3d_world_fill_rain). */
int fs3d_batch_fill_rain(fs3d_batch *b, const uint64_t *seeds, const uint64_t *thr);
int fs3d_batch_get_tick(const fs3d_batch *b, uint64_t *tick);
int fs3d_batch_set_tick(fs3d_batch *b, uint64_t tick);
/* Persistent-kernel shape: bytes per lane (8 or 16; 32 with 256 threads and 1
 * group), workgroup size (256, 512, 1024) and row-quad groups each wave loads
 * before updating (1, 2, 4: memory-level parallelism).  *grid receives the
 * number of resident workgroups. */
int fs3d_batch_set_config(fs3d_batch *b, int32_t seg, int32_t threads, int32_t groups);
int fs3d_batch_get_config(const fs3d_batch *b, int32_t *seg, int32_t *threads, int32_t *groups, int32_t *grid);
/* Settled worlds (on by default): in FS3D_BATCH_PERSISTENT runs a world in which
 * no cell changed for two consecutive ticks is a fixed point of both parities,
 * so its remaining ticks are not run (only its final, hashed tick is); results
 * are bit-identical either way. */
int fs3d_batch_set_skip(fs3d_batch *b, int32_t on);
int fs3d_batch_get_skip(const fs3d_batch *b, int32_t *on);
/* Advances every world by nticks; when hashes != NULL writes each world's state
 * hash after the last tick (hashes[n]).  Synchronises. */
int fs3d_batch_run(fs3d_batch *b, int32_t nticks, int32_t mode, uint64_t *hashes);
/* Device time (ms, HIP events) of the last fs3d_batch_run. */
int fs3d_batch_last_run_ms(const fs3d_batch *b, double *ms);
/* Times every persistent-kernel shape and the wide mode at each segment width
 * on a scratch copy of the arena for trial_ticks ticks (the batch itself does
 * not advance), keeps the fastest and makes its mode the FS3D_BATCH_TUNED one. */
int fs3d_batch_autotune(fs3d_batch *b, int32_t trial_ticks, int32_t *seg, int32_t *threads, int32_t *groups,
                        int32_t *mode, double *best_ms);

/* ---- rendering ---------------------------------------------------------------- */
/* Pixel -> ray as the reference's full-screen ray marcher
 * (shaders/fs_raymarch.frag:67-75, shaders/fs_raymarch.vert:30-36, quad UVs
 * src/engine/rendering/renderer.cpp:1253-1267): py = 0 is the top row; camera
 * +x is screen-left, +y screen-down, +z forward; uv.y is divided by aspect. */
typedef struct fs3d_camera {
  float origin[3];
  float basis[9]; /* screen-left, screen-down, forward: world dir = dx*b0 + dy*b1 + dz*b2 */
  float light[3];
  float aspect;
} fs3d_camera;
/* HIP port of shaders/fs_raymarch.frag (sphere tracer; origin = cam_pos as in
 * shaders/fs_raymarch.vert:33): rgba[H][W][4] float, written to host memory.
 * Images are 1..65536 pixels per side here and in fs3d_world_render. */
int fs3d_render_sdf(const float cam_pos[3], float aspect, int32_t W, int32_t H, float *rgba);
/* The scene's active marcher, shaders/raymarch_sphere.{vert,frag}, at its stage
 * boundaries (the rasteriser between them stays the caller's).  Vertex stage:
 * the fragments' ray origin, invModel * vec4(-camPos, 1) (inv_model column-major
 * as in src/engine/rendering/materials.h:30-33; host only).  Fragment stage: n
 * fragments given their origin[3] and interpolated object-space position
 * local[3]; rgba[4] each. */
int fs3d_box_vertex(const float cam_pos[3], const float inv_model[16], float origin[3]);
int fs3d_render_box_march(const float *origin, const float *local, int32_t n, float *rgba);
/* Voxel ray caster over the world's current grid (the grid -> renderer hand-off
 * the reference's set-2 bindings would carry, src/engine/rendering/materials.cpp:26-59):
 * rgb[H][W][3] float, hit[H][W] = linear voxel index of the first non-EMPTY cell
 * (UINT64_MAX on a miss), face[H][W] = entry face (0 -x, 1 +x, 2 -y, 3 +y, 4 -z,
 * 5 +z, 6 started inside, 255 miss).  Colours from the reference palette
 * (src/engine/rendering/renderer.cpp:136-393): STATIC 127, SAND 162, FLUID 137,
 * other bytes their own index.  Synchronises. */
int fs3d_world_render(fs3d_world *w, const fs3d_camera *cam, int32_t W, int32_t H, float *rgb, uint64_t *hit,
                      uint8_t *face);
int fs3d_world_last_render_ms(const fs3d_world *w, double *ms);
/* Occupancy tracking between frames (no counterpart in the reference).  Each
 * render reads the grid into occupancy maps before casting rays.  With tracking
 * on, once a world has been rendered its skipping ticks record the 8^3 bricks
 * they change (about 10 us per tick at 2048^3), and the next render reads only
 * those bricks again; dense ticks and host-side edits make it read the whole
 * grid.  Frames are identical either way.  FS3D_TRACK_AUTO (the default) tracks
 * grids of at least 2^29 voxels, where a full read costs more than ten ticks of
 * recording (2048^3: 1.55 ms against 0.56 ms per frame after 10 ticks). */
enum { FS3D_TRACK_AUTO = 0, FS3D_TRACK_ON = 1, FS3D_TRACK_OFF = 2 };
int fs3d_world_set_render_track(fs3d_world *w, int32_t mode);
int fs3d_world_get_render_track(const fs3d_world *w, int32_t *mode);
/* Diagnostics (no counterpart in the reference): the occupancy maps of the last
 * render, which describe the grid as it was then.  With bnx = ceil(nx / 8) (and
 * likewise bny, bnz), bw = ceil(bnx / 64), rnx = ceil(bnx / 4) (and likewise
 * rny, rnz) and rw = ceil(rnx / 64):
 *   coarse[(bz * bny + by) * bw + bx / 64] bit bx % 64: the 8^3 brick (bx, by, bz)
 *     holds a non-EMPTY cell;
 *   fine[(bz * bny + by) * bnx + bx] bit ((x & 7) >> 1) | (((y & 7) >> 1) << 2) |
 *     (((z & 7) >> 1) << 4): the 2x2x2 sub-brick holding (x, y, z) does;
 *   region[(rz * rny + ry) * rw + rx / 64] bit rx % 64: the 32^3 region (rx, ry,
 *     rz) does.
 * n_coarse, n_fine and n_region are the capacities in words (at least
 * bw * bny * bnz, bnx * bny * bnz and rw * rny * rnz).  FS3D_ERR_STATE before the
 * first render.  Synchronises. */
int fs3d_world_render_maps(fs3d_world *w, uint64_t *coarse, size_t n_coarse, uint64_t *fine, size_t n_fine,
                           uint64_t *region, size_t n_region);
/* Seeded-orbit camera: eye at angle/elevation (radians) around the grid centre,
 * dist_scale * max(dims) away, looking at the centre with +z up on screen. */
int fs3d_orbit_camera(const int32_t dims[3], double angle, double dist_scale, double elev, float aspect,
                      fs3d_camera *out);
/* Voxel camera from the reference's key-driven camera state (cam_pos, cam_rot
 * as fs3d_camera_integrate advances them; the view matrix of
 * fs3d_camera_globaldata, src/engine/rendering/materials.cpp:532-536): eye =
 * -R^T t, screen-left / screen-down / forward = minus rows 0 / 1 / 2 of the
 * view's rotation.  World units map to the grid so that [-1, 1] spans its
 * longest axis around its centre, world +X = grid +x, +Y = grid +z (up),
 * +Z = grid -y; the light is fs_raymarch's (2, 5, 3) mapped the same way.  The
 * reference's initial camera (0, 0, -5) looks at the grid from 5 half-extents
 * along -y. */
int fs3d_key_camera(const int32_t dims[3], const float cam_pos[3], const float cam_rot[3], float aspect,
                    fs3d_camera *out);
/* The reference's GlobalData block (src/engine/rendering/materials.h:19-28) as
 * Material::writeGlobalAndObjectData fills it (src/engine/rendering/materials.cpp:531-545):
 * 76 floats = camPos, aspect, camDir (vec4 each), view, proj, viewproj,
 * invViewProj (column-major mat4 each). */
int fs3d_camera_globaldata(const float cam_pos[3], const float cam_rot[3], float *block);
/* One fixed 0.016 s step of the reference's key-driven camera
 * (src/engine/rendering/renderer.cpp:437-467). */
enum {
  FS3D_KEY_W = 1, FS3D_KEY_A = 2, FS3D_KEY_S = 4, FS3D_KEY_D = 8,
  FS3D_KEY_CTRL = 16, FS3D_KEY_SPACE = 32, FS3D_KEY_LEFT = 64, FS3D_KEY_RIGHT = 128
};
int fs3d_camera_integrate(float cam_pos[3], float cam_rot[3], uint32_t keys);
/* The 256-entry RGBA palette (rgba[256*4]). */
int fs3d_palette(float *rgba);

/* ---- serving ----------------------------------------------------------------- */
/* Continuous batching of simulation requests.  No reference counterpart (the
 * reference is one interactive window, src/engine/engine.cpp:54-74).  Requests
 * (dims, rain-fill seed and threshold, tick count) are admitted in arrival order
 * while their grids fit budget_bytes, filled on the device, kept resident, and
 * advanced in rounds of up to round_ticks ticks,
 * each tick one wide launch over every active world; requests arriving meanwhile
 * join at the next round.  A request's final hash is fused into its last tick.
 * Memory: grids (nx*ny*nz + 256 bytes, rounded to 256) are placed in device
 * chunks of up to 1 GiB, best fit, and coalesced when released; the budget
 * bounds the chunk bytes actually allocated.  A request that does not fit waits
 * at the head of the queue until finished grids come back; one larger than the
 * whole budget, or one the device cannot hold while no other grid is resident,
 * is rejected.  Calls on one server must be serialised. */
typedef struct fs3d_server fs3d_server; /* opaque */
enum { FS3D_REQ_QUEUED = 0, FS3D_REQ_ACTIVE = 1, FS3D_REQ_DONE = 2, FS3D_REQ_REJECTED = 3 };
typedef struct fs3d_server_config {
  int64_t budget_bytes; /* HBM budget for resident grids */
  int32_t round_ticks;  /* ticks per round (>= 1) */
  int32_t max_active;   /* concurrent requests (hash slots) */
  int32_t seg;          /* tick segment bytes: 8, 16, 32 (0 = 8; nx <= 64 * seg) */
  int32_t reserved;
} fs3d_server_config;
typedef struct fs3d_server_stats {
  int64_t submitted, completed, rejected;
  int64_t rounds, launches;
  double tick_p50_ms, tick_p99_ms, tick_max_ms; /* per launch: one tick of every world in it */
  double req_p50_ms, req_p99_ms;                /* submission -> final hash observed */
  double voxel_updates;                         /* cell updates enqueued so far */
  double busy_ms;                               /* summed device time of completed launches */
  int64_t peak_bytes, used_bytes;               /* grid bytes in use: high-water mark, now */
  int32_t peak_active;
  int32_t reserved;
  int64_t alloc_bytes, peak_alloc_bytes;        /* device bytes the grid arena holds: now, high-water mark */
} fs3d_server_stats;
int fs3d_server_create(const fs3d_server_config *cfg, fs3d_server **out);
int fs3d_server_destroy(fs3d_server *s); /* NULL is a no-op; waits for work in flight */
/* Queues a request (ticks >= 1); *id is its index, in submission order. */
int fs3d_server_submit(fs3d_server *s, const int32_t dims[3], uint64_t seed, uint64_t thr, int32_t ticks,
                       int64_t *id);
/* One scheduling step: observe finished rounds, admit, plan and enqueue a round
 * (at most two in flight).  Never blocks on the GPU. */
int fs3d_server_poll(fs3d_server *s, int32_t *n_active, int32_t *n_queued, int64_t *n_done);
/* Polls until every submitted request is done or rejected. */
int fs3d_server_drain(fs3d_server *s);
int fs3d_server_result(const fs3d_server *s, int64_t id, int32_t *status, uint64_t *hash, double *latency_ms);
/* Real-time driver: submits request i when arrival[i] seconds have passed since
 * the call began, polling in between, until all are done; ids[i] receives the
 * request ids. */
int fs3d_server_replay(fs3d_server *s, int32_t n, const int32_t *dims, const uint64_t *seeds, const uint64_t *thr,
                       const int32_t *ticks, const double *arrival, int64_t *ids);
int fs3d_server_get_stats(const fs3d_server *s, fs3d_server_stats *out);

/* ---- world files (checkpoint / resume) ----------------------------------------
 * No reference counterpart (the reference keeps no simulation state; its assets
 * are meshes and textures).  A .fs3d file is a 64-byte little-endian header and
 * the raw grid:
 *    0  char[8]  "FS3DGRID"
 *    8  u32      format version (1)
 *   12  u32      header bytes (64)
 *   16  i32[3]   nx, ny, nz
 *   28  u32      flags (0)
 *   32  u64      tick counter
 *   40  u64      rule hash (fs3d_rule_hash of the library that wrote it)
 *   48  u64      state hash of the grid (as fs3d_world_hash)
 *   56  u64      payload bytes (nx*ny*nz)
 *   64  u8[]     v[z][y][x]
 * Loading refuses a file whose rule hash differs from this library's (the same
 * bytes would evolve differently) and one whose payload does not match its state
 * hash. */
#define FS3D_FILE_VERSION 1
/* Hash of the tick's block rule: SplitMix64-mixed sum over the packed table. */
int fs3d_rule_hash(uint64_t *hash);
/* Host-only file access (no device needed). */
int fs3d_file_write(const char *path, const int32_t dims[3], uint64_t tick, uint64_t state_hash, const uint8_t *grid,
                    size_t nbytes);
/* Reads the header; dims, tick and state_hash may be NULL. */
int fs3d_file_info(const char *path, int32_t dims[3], uint64_t *tick, uint64_t *state_hash);
/* Reads the payload into grid (nbytes must equal nx*ny*nz). */
int fs3d_file_read(const char *path, uint8_t *grid, size_t nbytes);
/* Saves the world's current grid and tick (synchronises). */
int fs3d_world_save(fs3d_world *w, const char *path);
/* Loads a file into a world of the same dims: grid, tick counter, checked hashes. */
int fs3d_world_load(fs3d_world *w, const char *path);

/* ---- ticks without allocation ---------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the ticks of fs3d_world_run (dense, bit-identical results)
 * on a grid in caller-owned device memory.  a and b are two distinct device
 * buffers of at least nx*ny*nz + FS3D_GRID_GUARD_BYTES bytes (tick loads may
 * read up to one lane segment past the last row); the state is read from a, the
 * buffers alternate tick by tick, and *final_index receives the one holding the
 * result (0 = a, 1 = b).  hashes[nticks / hash_every] (host memory) receives
 * the state hash after every hash_every-th tick; at most FS3D_NOALLOC_MAX_HASHES
 * of them, or FS3D_ERR_INVALID.  The kernels' own device memory (hash partial
 * sums and hash slots, 96 KiB) is a static array per GPU, set up by fs3d_init
 * (FS3D_ERR_STATE before it): the call allocates nothing.  stream is a
 * hipStream_t (NULL: the default stream); the call synchronises it.  Calls on
 * one GPU share that array: calls that overlap there race. */
#define FS3D_GRID_GUARD_BYTES 256
#define FS3D_NOALLOC_MAX_HASHES 4096
int fs3d_grid_run_noalloc(uint8_t *a, uint8_t *b, const int32_t dims[3], uint64_t tick, int32_t nticks,
                          int32_t hash_every, uint64_t *hashes, int32_t *final_index, void *stream);

/* ---- timing ---------------------------------------------------------------- */
/* Average device time (ms) of the tick kernels launched by the last
 * fs3d_world_run on this world, measured with HIP events on the world's stream
 * (first launch to last launch, divided by the tick count). */
int fs3d_world_last_run_ms(const fs3d_world *w, double *ms_per_tick);
/* Mean device time (ms) of a plain streaming copy of nbytes (16-byte vector
 * loads and stores per lane): the practical bandwidth ceiling the tick is
 * compared with, and the known-size kernel that scales rocprof byte counters.
 * No reference counterpart (the reference has no timing at all). */
int fs3d_copy_probe(size_t nbytes, int32_t iters, double *ms_per_copy);
/* The same copy with streaming (nontemporal) loads and stores: the ceiling for
 * the passes that stream their rows (the chunked dense tick's stores, the fused
 * passes' loads and stores, the packed skipping tick's stores).  No reference
 * counterpart. */
int fs3d_copy_probe_streaming(size_t nbytes, int32_t iters, double *ms_per_copy);
/* Mean device time (ms) of one pass of the packed skipping tick's access
 * shape without its arithmetic, on a 2048^3 grid pair (16 GiB, allocated and
 * freed by the call): per_rq (1..64) scattered, distinct 32-byte segments
 * of every row quad, four rows each, copied from one grid to the other by one
 * lane per item; sector != 0 copies each segment's whole 64-byte sector
 * instead.  *item_bytes receives the bytes the items read and write.  No
 * reference counterpart. */
int fs3d_scatter_probe(int32_t per_rq, int32_t sector, int32_t iters, double *ms_per_pass, double *item_bytes);
/* Mean device time (ms) of one pass of the dense chunked tick's access shape
 * without its arithmetic, on a 2048^3 grid pair (16 GiB, allocated and freed
 * by the call): every row quad's four rows copied from one grid to the other,
 * one wave per 1 KiB x-chunk, 16 bytes per lane and row; streaming_loads /
 * streaming_stores != 0 make those accesses streaming (nontemporal).  No
 * reference counterpart. */
int fs3d_rowquad_probe(int32_t streaming_loads, int32_t streaming_stores, int32_t iters, double *ms_per_pass);

#ifdef __cplusplus
}
#endif
#endif
