"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only to check the HIP product path (or, in bench.py, to time the CPU
baseline).  The product package never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

u8p = ctypes.POINTER(ctypes.c_uint8)
u16p = ctypes.POINTER(ctypes.c_uint16)
u64p = ctypes.POINTER(ctypes.c_uint64)
i32p = ctypes.POINTER(ctypes.c_int32)
f64p = ctypes.POINTER(ctypes.c_double)


class BoxParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("side_min", "side_max", "pct_sand", "pct_fluid", "pct_static", "max_boxes")]


class PillarParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("n_pillars", "r_min", "r_max", "h_min", "h_max", "pool_w_min", "pool_w_max",
                 "pool_d_min", "pool_d_max", "pool_h_min", "pool_h_max")]


class Camera(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("basis", ctypes.c_float * 9), ("light", ctypes.c_float * 3),
                ("aspect", ctypes.c_float)]


f32p = ctypes.POINTER(ctypes.c_float)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        sig = {
            "fso_mix64": (ctypes.c_uint64, [ctypes.c_uint64]),
            "fso_sm64_next": (ctypes.c_uint64, [u64p]),
            "fso_cell_rand": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
            "fso_mulhi64": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
            "fso_uniform": (ctypes.c_uint64, [ctypes.c_uint64] * 3),
            "fso_threshold": (ctypes.c_uint64, [ctypes.c_uint64] * 2),
            "fso_fmix32": (ctypes.c_uint32, [ctypes.c_uint32]),
            "fso_group_word": (ctypes.c_uint32, [ctypes.c_uint32] * 4),
            "fso_variant": (ctypes.c_int, [ctypes.c_uint32] * 4),
            "fso_loguniform": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]),
            "fso_block_rule": (None, [u8p, ctypes.c_int, u8p]),
            "fso_rule_state": (ctypes.c_uint16, [ctypes.c_uint16, ctypes.c_int]),
            "fso_build_lut": (None, [u16p]),
            "fso_set_threads": (None, [ctypes.c_int]),
            "fso_get_threads": (ctypes.c_int, []),
            "fso_tick": (None, [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]),
            "fso_tick_lut": (None, [u8p, u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_uint64, u16p]),
            "fso_hash": (ctypes.c_uint64, [u8p, ctypes.c_uint64]),
            "fso_counts": (None, [u8p, ctypes.c_uint64, u64p]),
            "fso_run": (ctypes.c_int, [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                       ctypes.c_int, ctypes.c_int, u64p]),
            "fso_run_lut": (ctypes.c_int, [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                           ctypes.c_int, ctypes.c_int, u64p, u16p]),
            "fso_scene_rain": (None, [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.c_uint64]),
            "fso_scene_boxes": (ctypes.c_int64, [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_uint64, ctypes.POINTER(BoxParams)]),
            "fso_scene_pillars": (ctypes.c_int, [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_uint64, ctypes.POINTER(PillarParams)]),
            "fso_batch_specs": (None, [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                       i32p, u64p, u64p]),
            "fso_serve_specs": (None, [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_double, i32p, u64p, u64p,
                                       i32p, f64p]),
            "fso_pixel_dir": (None, [ctypes.c_int] * 4 + [ctypes.c_float, f32p]),
            "fso_raymarch_sdf": (None, [f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int, f32p]),
            "fso_box_vertex": (None, [f32p, f32p, f32p]),
            "fso_box_march": (None, [f32p, f32p, ctypes.c_int, f32p]),
            "fso_render_dda": (None, [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Camera), f32p,
                                      ctypes.c_int, ctypes.c_int, f32p, u64p, u8p]),
            "fso_orbit_camera": (None, [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_float,
                                                                                    ctypes.POINTER(Camera)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def mix64(z):
    return lib().fso_mix64(z)


def cell_rand(key, i):
    return lib().fso_cell_rand(key, i)


def threshold(num, den):
    return lib().fso_threshold(num, den)


def fmix32(h):
    return lib().fso_fmix32(h)


def variant(bx, by, bz, t):
    return lib().fso_variant(bx, by, bz, t & 0xFFFFFFFF)


def splitmix_stream(seed, n):
    st = ctypes.c_uint64(seed)
    return [lib().fso_sm64_next(ctypes.byref(st)) for _ in range(n)]


def block_rule(cats, v):
    a = np.ascontiguousarray(cats, dtype=np.uint8)
    out = np.zeros(8, np.uint8)
    lib().fso_block_rule(_p(a, u8p), v, _p(out, u8p))
    return out


def rule_state(s, v):
    return lib().fso_rule_state(s, v)


def build_lut():
    lut = np.zeros(4 * 65536, np.uint16)
    lib().fso_build_lut(_p(lut, u16p))
    return lut


def set_threads(n):
    lib().fso_set_threads(n)


def tick(grid, t):
    """One tick of a (nz, ny, nx) uint8 grid at tick counter t; returns a new grid."""
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    out = np.empty_like(g)
    nz, ny, nx = g.shape
    lib().fso_tick(_p(g, u8p), _p(out, u8p), nx, ny, nz, t)
    return out


def tick_lut(grid, t, lut):
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    out = np.empty_like(g)
    nz, ny, nx = g.shape
    lib().fso_tick_lut(_p(g, u8p), _p(out, u8p), nx, ny, nz, t, _p(lut, u16p))
    return out


def state_hash(grid):
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    return lib().fso_hash(_p(g, u8p), g.size)


def counts(grid):
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    c = np.zeros(256, np.uint64)
    lib().fso_counts(_p(g, u8p), g.size, _p(c, u64p))
    return c


_LUT = None


def run(grid, t0, nticks, hash_every=0, use_lut=False):
    """Runs nticks ticks on a copy; returns (grid, hashes).  use_lut evaluates
    blocks through the oracle's own rule table (faster; same results)."""
    global _LUT
    g = np.array(grid, dtype=np.uint8, copy=True, order="C")
    nz, ny, nx = g.shape
    nh = nticks // hash_every if hash_every > 0 else 0
    h = np.zeros(max(nh, 1), np.uint64)
    if use_lut:
        if _LUT is None:
            _LUT = build_lut()
        rc = lib().fso_run_lut(_p(g, u8p), nx, ny, nz, t0, nticks, hash_every, _p(h, u64p), _p(_LUT, u16p))
    else:
        rc = lib().fso_run(_p(g, u8p), nx, ny, nz, t0, nticks, hash_every, _p(h, u64p))
    if rc != 0:
        raise MemoryError("oracle scratch allocation failed")
    return g, h[:nh]


def scene_rain(dims, seed, thr):
    nx, ny, nz = dims
    g = np.empty((nz, ny, nx), np.uint8)
    lib().fso_scene_rain(_p(g, u8p), nx, ny, nz, seed, thr)
    return g


def scene_boxes(dims, seed, side_min=8, side_max=64, pct=(20, 10, 5), max_boxes=1 << 20):
    nx, ny, nz = dims
    g = np.empty((nz, ny, nx), np.uint8)
    P = BoxParams(side_min, side_max, pct[0], pct[1], pct[2], max_boxes)
    n = lib().fso_scene_boxes(_p(g, u8p), nx, ny, nz, seed, ctypes.byref(P))
    if n < 0:
        raise ValueError("invalid box parameters")
    return g, n


def scene_pillars(dims, seed, params):
    nx, ny, nz = dims
    g = np.empty((nz, ny, nx), np.uint8)
    P = PillarParams(*params)
    if lib().fso_scene_pillars(_p(g, u8p), nx, ny, nz, seed, ctypes.byref(P)) != 0:
        raise ValueError("invalid pillar parameters")
    return g


def batch_specs(n, seed, lo=16, hi=128):
    dims = np.zeros(3 * n, np.int32)
    thr = np.zeros(n, np.uint64)
    seeds = np.zeros(n, np.uint64)
    lib().fso_batch_specs(n, seed, lo, hi, _p(dims, i32p), _p(thr, u64p), _p(seeds, u64p))
    return dims.reshape(n, 3), thr, seeds


def serve_specs(n, seed, lo=32, hi=256, tick_lo=10, tick_hi=500, lam=2000.0):
    dims = np.zeros(3 * n, np.int32)
    thr = np.zeros(n, np.uint64)
    seeds = np.zeros(n, np.uint64)
    ticks = np.zeros(n, np.int32)
    arr = np.zeros(n, np.float64)
    lib().fso_serve_specs(n, seed, lo, hi, tick_lo, tick_hi, lam, _p(dims, i32p), _p(thr, u64p),
                          _p(seeds, u64p), _p(ticks, i32p), _p(arr, f64p))
    return dims.reshape(n, 3), thr, seeds, ticks, arr


def pixel_dir(px, py, W, H, aspect):
    d = np.zeros(3, np.float32)
    lib().fso_pixel_dir(px, py, W, H, aspect, _p(d, f32p))
    return d


def raymarch_sdf(cam_pos, aspect, W, H):
    """RGBA float32 image (H, W, 4) of the fs_raymarch sphere tracer."""
    cp = np.ascontiguousarray(cam_pos, np.float32)
    out = np.zeros((H, W, 4), np.float32)
    lib().fso_raymarch_sdf(_p(cp, f32p), aspect, W, H, _p(out, f32p))
    return out


def box_vertex(cam_pos, inv_model):
    """Object-space ray origin of shaders/raymarch_sphere.vert (inv_model column-major 4x4)."""
    cp = np.ascontiguousarray(cam_pos, np.float32)
    m = np.ascontiguousarray(inv_model, np.float32).reshape(16)
    out = np.zeros(3, np.float32)
    lib().fso_box_vertex(_p(cp, f32p), _p(m, f32p), _p(out, f32p))
    return out


def box_march(origins, locals_):
    """RGBA (n, 4) of shaders/raymarch_sphere.frag for n (origin, local position) pairs."""
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    l = np.ascontiguousarray(locals_, np.float32).reshape(-1, 3)
    out = np.zeros((len(o), 4), np.float32)
    lib().fso_box_march(_p(o, f32p), _p(l, f32p), len(o), _p(out, f32p))
    return out


def camera_from(origin, basis, light, aspect):
    c = Camera()
    c.origin[:] = [float(v) for v in origin]
    c.basis[:] = [float(v) for v in np.asarray(basis).reshape(9)]
    c.light[:] = [float(v) for v in light]
    c.aspect = float(aspect)
    return c


def orbit_camera(dims, angle, dist_scale, elev, aspect):
    c = Camera()
    lib().fso_orbit_camera(int(dims[0]), int(dims[1]), int(dims[2]), angle, dist_scale, elev, aspect, ctypes.byref(c))
    return c


def render_dda(grid, cam, palette, W, H):
    """(rgb (H,W,3) float32, hit (H,W) uint64, face (H,W) uint8) of the voxel DDA renderer."""
    g = np.ascontiguousarray(grid, np.uint8)
    nz, ny, nx = g.shape
    pal = np.ascontiguousarray(palette, np.float32).reshape(256 * 4)
    rgb = np.zeros((H, W, 3), np.float32)
    hit = np.zeros((H, W), np.uint64)
    face = np.zeros((H, W), np.uint8)
    lib().fso_render_dda(_p(g, u8p), nx, ny, nz, ctypes.byref(cam), _p(pal, f32p), W, H, _p(rgb, f32p),
                         _p(hit, u64p), _p(face, u8p))
    return rgb, hit, face
