"""Host-side mirror of the fs3d world interface (Python, over the C ABI).

The reference engine exposes init/run/cleanup (src/engine/engine.h:16-35) and a
Renderer with init/draw/cleanup (src/engine/rendering/renderer.h:81-98); it has
no voxel world.  `World` is the object its run() loop would step between event
handling and drawing (src/engine/engine.cpp:59-69): it owns a uint8 grid
v[z][y][x] (z up) resident on the GPU, double-buffered, and every method maps
one-to-one onto an `fs3d_*` entry point of include/fs3d.h.  Errors raise
`Fs3dError` (the reference logs and throws std::runtime_error,
src/util/debug.cpp:23-27).
"""
import ctypes

import numpy as np

from ._lib import (BoxParams, Fs3dError, PillarParams, c_double, c_i32, c_i64, c_int, c_u8, c_u64,
                   check, lib, ptr, vp)

EMPTY, STATIC, SAND, FLUID = 0, 1, 2, 3

# Scene parameters of the measured workloads.
BOXES_512 = dict(side_min=8, side_max=64, pct=(20, 10, 5), max_boxes=1 << 20)
PILLARS_2048 = (256, 4, 32, 256, 1800, 256, 682, 256, 682, 32, 128)
DENSE = 1  # world flag: settled-segment skipping off
SKIP_OFF, SKIP_AUTO, SKIP_LISTED, SKIP_WAVES, SKIP_SEGS = 0, 1, 2, 3, 4
TRACK_AUTO, TRACK_ON, TRACK_OFF = 0, 1, 2  # renderer occupancy tracking between frames


def device_count():
    n = c_int(0)
    check(lib().fs3d_device_count(ctypes.byref(n)))
    return n.value


def init(device=0):
    check(lib().fs3d_init(device))


def rule_lut():
    """The packed block rule lut[v*65536 + s] (host-side table, no GPU needed)."""
    lut = np.zeros(4 * 65536, np.uint16)
    check(lib().fs3d_rule_lut(ptr(lut, ctypes.c_uint16)))
    return lut


def copy_probe(nbytes, iters=20, streaming=False):
    """Mean ms of a plain device copy of nbytes (bandwidth calibration); with
    streaming, the copy uses nontemporal loads and stores."""
    ms = c_double(0)
    fn = lib().fs3d_copy_probe_streaming if streaming else lib().fs3d_copy_probe
    check(fn(nbytes, iters, ctypes.byref(ms)))
    return ms.value


def rowquad_probe(streaming_loads, streaming_stores, iters=10):
    """ms per pass of the dense chunked tick's access shape (every row quad's
    four rows, one wave per 1 KiB x-chunk) on a 2048^3 grid pair, with streaming
    or temporal loads and stores."""
    ms = c_double(0)
    check(lib().fs3d_rowquad_probe(1 if streaming_loads else 0, 1 if streaming_stores else 0, iters,
                                   ctypes.byref(ms)))
    return ms.value


def scatter_probe(per_rq, sector, iters=10):
    """(ms per pass, item bytes) of the packed tick's access shape on a 2048^3
    grid pair: per_rq scattered 32-byte segments of every row quad (sector:
    their 64-byte sectors) read from one grid and written to the other."""
    ms, nb = c_double(0), c_double(0)
    check(lib().fs3d_scatter_probe(per_rq, 1 if sector else 0, iters, ctypes.byref(ms), ctypes.byref(nb)))
    return ms.value, nb.value


def threshold(num, den):
    t = c_u64(0)
    check(lib().fs3d_threshold(num, den, ctypes.byref(t)))
    return t.value


def batch_specs(n, seed, lo=16, hi=128):
    dims = np.zeros(3 * n, np.int32)
    thr = np.zeros(n, np.uint64)
    seeds = np.zeros(n, np.uint64)
    check(lib().fs3d_batch_specs(n, seed, lo, hi, ptr(dims, c_i32), ptr(thr, c_u64), ptr(seeds, c_u64)))
    return dims.reshape(n, 3), thr, seeds


def serve_specs(n, seed, lo=32, hi=256, tick_lo=10, tick_hi=500, lam=2000.0):
    dims = np.zeros(3 * n, np.int32)
    thr = np.zeros(n, np.uint64)
    seeds = np.zeros(n, np.uint64)
    ticks = np.zeros(n, np.int32)
    arr = np.zeros(n, np.float64)
    check(lib().fs3d_serve_specs(n, seed, lo, hi, tick_lo, tick_hi, lam, ptr(dims, c_i32), ptr(thr, c_u64),
                                 ptr(seeds, c_u64), ptr(ticks, c_i32), ptr(arr, c_double)))
    return dims.reshape(n, 3), thr, seeds, ticks, arr


def rule_hash():
    """Hash of the tick's block rule (stored in world files)."""
    h = c_u64(0)
    check(lib().fs3d_rule_hash(ctypes.byref(h)))
    return h.value


def file_write(path, grid, tick=0, state_hash=0):
    """Writes a .fs3d world file from a host grid g[z][y][x] (no GPU needed)."""
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    nz, ny, nx = g.shape
    d = (c_i32 * 3)(nx, ny, nz)
    check(lib().fs3d_file_write(str(path).encode(), d, tick, state_hash, g.ctypes.data, g.size))


def file_info(path):
    """(dims (nx, ny, nz), tick, state hash) from a .fs3d header."""
    d = (c_i32 * 3)()
    t, h = c_u64(0), c_u64(0)
    check(lib().fs3d_file_info(str(path).encode(), d, ctypes.byref(t), ctypes.byref(h)))
    return (d[0], d[1], d[2]), t.value, h.value


def file_read(path):
    """The grid g[z][y][x] of a .fs3d file (refused if written under another rule)."""
    (nx, ny, nz), _, _ = file_info(path)
    g = np.empty((nz, ny, nx), np.uint8)
    check(lib().fs3d_file_read(str(path).encode(), g.ctypes.data, g.size))
    return g


class World:
    """A uint8 voxel grid of dims (nx, ny, nz) on the current GPU."""

    def __init__(self, dims, flags=0):
        self.dims = tuple(int(d) for d in dims)
        h = vp()
        d = (c_i32 * 3)(*self.dims)
        check(lib().fs3d_world_create(d, flags, ctypes.byref(h)))
        self._h = h

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().fs3d_world_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        if not self._h:
            raise Fs3dError(6, "world is closed")
        return self._h

    @property
    def shape(self):
        nx, ny, nz = self.dims
        return (nz, ny, nx)

    @property
    def size(self):
        nx, ny, nz = self.dims
        return nx * ny * nz

    # -- grid access ---------------------------------------------------------
    def set_voxel(self, x, y, z, v):
        check(lib().fs3d_world_set_voxel(self.handle, x, y, z, v))

    def get_voxel(self, x, y, z):
        v = c_u8(0)
        check(lib().fs3d_world_get_voxel(self.handle, x, y, z, ctypes.byref(v)))
        return v.value

    def upload(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.uint8)
        if g.shape != self.shape:
            raise Fs3dError(1, f"grid shape {g.shape} != {self.shape} (z, y, x)")
        check(lib().fs3d_world_upload(self.handle, g.ctypes.data, g.size))

    def download(self):
        g = np.empty(self.shape, np.uint8)
        check(lib().fs3d_world_download(self.handle, g.ctypes.data, g.size))
        return g

    @property
    def tick_count(self):
        t = c_u64(0)
        check(lib().fs3d_world_get_tick(self.handle, ctypes.byref(t)))
        return t.value

    @tick_count.setter
    def tick_count(self, t):
        check(lib().fs3d_world_set_tick(self.handle, t))

    # -- tick ----------------------------------------------------------------
    def step(self, nticks=1):
        check(lib().fs3d_world_step(self.handle, nticks))

    def tick(self):
        """One tick; returns the hash of the new state (fused into the kernel)."""
        h = c_u64(0)
        check(lib().fs3d_world_tick(self.handle, ctypes.byref(h)))
        return h.value

    def hash(self):
        h = c_u64(0)
        check(lib().fs3d_world_hash(self.handle, ctypes.byref(h)))
        return h.value

    def run(self, nticks, hash_every=0):
        """nticks ticks as one captured HIP graph; returns the periodic hashes."""
        nh = nticks // hash_every if hash_every > 0 else 0
        out = np.zeros(max(nh, 1), np.uint64)
        check(lib().fs3d_world_run(self.handle, nticks, hash_every, ptr(out, c_u64)))
        return out[:nh]

    def prepare_run(self, nticks, hash_every=0):
        """Builds (captures and instantiates) the graph run(nticks, hash_every)
        would launch from the current state, without running it."""
        check(lib().fs3d_world_prepare_run(self.handle, nticks, hash_every))

    def last_run_ms_per_tick(self):
        ms = c_double(0)
        check(lib().fs3d_world_last_run_ms(self.handle, ctypes.byref(ms)))
        return ms.value

    def counts(self):
        c = np.zeros(256, np.uint64)
        check(lib().fs3d_world_counts(self.handle, ptr(c, c_u64)))
        return c

    def sync(self):
        check(lib().fs3d_world_sync(self.handle))

    def device_ptr(self):
        p = vp()
        check(lib().fs3d_world_device_ptr(self.handle, ctypes.byref(p)))
        return p.value

    @property
    def segment(self):
        s = c_i32(0)
        check(lib().fs3d_world_get_segment(self.handle, ctypes.byref(s)))
        return s.value

    @segment.setter
    def segment(self, s):
        check(lib().fs3d_world_set_segment(self.handle, s))

    @property
    def skip(self):
        """Settled-segment skipping mode: SKIP_OFF, SKIP_AUTO (the default unless
        the world was created with DENSE), SKIP_LISTED, SKIP_WAVES or SKIP_SEGS.  Setting
        True / False means SKIP_AUTO / SKIP_OFF."""
        s = c_i32(0)
        check(lib().fs3d_world_get_skip(self.handle, ctypes.byref(s)))
        return s.value

    @skip.setter
    def skip(self, mode):
        check(lib().fs3d_world_set_skip(self.handle, int(mode)))

    def save(self, path):
        """Checkpoints the grid and tick counter to a .fs3d file."""
        check(lib().fs3d_world_save(self.handle, str(path).encode()))

    def load(self, path):
        """Resumes from a .fs3d file of the same dims (grid, tick; hashes checked)."""
        check(lib().fs3d_world_load(self.handle, str(path).encode()))

    @property
    def fuse(self):
        """Whether dense ticks run as fused pairs (set it to allow or forbid that)."""
        s = c_i32(0)
        check(lib().fs3d_world_get_fuse(self.handle, ctypes.byref(s)))
        return bool(s.value)

    @fuse.setter
    def fuse(self, on):
        check(lib().fs3d_world_set_fuse(self.handle, 1 if on else 0))

    @property
    def render_track(self):
        """Occupancy tracking between frames: TRACK_AUTO (default), TRACK_ON or
        TRACK_OFF.  With it, skipping ticks after a render record the 8^3 bricks
        they change and the next render reads only those bricks of the grid."""
        s = c_i32(0)
        check(lib().fs3d_world_get_render_track(self.handle, ctypes.byref(s)))
        return s.value

    @render_track.setter
    def render_track(self, mode):
        check(lib().fs3d_world_set_render_track(self.handle, int(mode)))

    def sweep_stats(self, enable=True):
        """Counters since the previous call: voxels read and written, voxels only
        read (settled segments on hash ticks), and settled segments that changed
        (always 0).  enable keeps (or starts) counting; False stops it."""
        out = (c_u64 * 3)()
        check(lib().fs3d_world_sweep_stats(self.handle, 1 if enable else 0, out))
        return dict(swept=int(out[0]), read_only=int(out[1]), violations=int(out[2]))

    # -- scenes ----------------------------------------------------------------
    def fill_rain(self, seed, thr):
        check(lib().fs3d_world_fill_rain(self.handle, seed, thr))

    def fill_boxes(self, seed, side_min=8, side_max=64, pct=(20, 10, 5), max_boxes=1 << 20):
        p = BoxParams(side_min, side_max, pct[0], pct[1], pct[2], max_boxes)
        n = c_i64(0)
        check(lib().fs3d_world_fill_boxes(self.handle, seed, ctypes.byref(p), ctypes.byref(n)))
        return n.value

    def fill_pillars(self, seed, params=PILLARS_2048):
        p = PillarParams(*params)
        check(lib().fs3d_world_fill_pillars(self.handle, seed, ctypes.byref(p)))
