"""ctypes loader for libfs3d.so (the HIP library built in-tree).

Importing this module never falls back to anything: when the library is missing
the import of `fallingsand3d_amd` still works (so CPU-only tooling can inspect
symbols), but every call that needs it raises `Fs3dError`.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfs3d.so")

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_u8 = ctypes.c_uint8
c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_i64 = ctypes.c_int64
c_size = ctypes.c_size_t
c_double = ctypes.c_double
c_float = ctypes.c_float
vp = ctypes.c_void_p
P = ctypes.POINTER

FS3D_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "NOMEM", 3: "BUDGET", 4: "HIP", 5: "NODEVICE", 6: "STATE"}


class Fs3dError(RuntimeError):
    """Raised for any non-zero fs3d_status; carries the code and library message."""

    def __init__(self, code, msg):
        super().__init__(f"fs3d error {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class BoxParams(ctypes.Structure):
    _fields_ = [(n, c_i32) for n in
                ("side_min", "side_max", "pct_sand", "pct_fluid", "pct_static", "max_boxes")]


class PillarParams(ctypes.Structure):
    _fields_ = [(n, c_i32) for n in
                ("n_pillars", "r_min", "r_max", "h_min", "h_max", "pool_w_min", "pool_w_max",
                 "pool_d_min", "pool_d_max", "pool_h_min", "pool_h_max")]


class Camera(ctypes.Structure):
    _fields_ = [("origin", c_float * 3), ("basis", c_float * 9), ("light", c_float * 3), ("aspect", c_float)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", c_i64), ("round_ticks", c_i32), ("max_active", c_i32), ("seg", c_i32),
                ("reserved", c_i32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", c_i64), ("completed", c_i64), ("rejected", c_i64), ("rounds", c_i64),
                ("launches", c_i64), ("tick_p50_ms", c_double), ("tick_p99_ms", c_double),
                ("tick_max_ms", c_double), ("req_p50_ms", c_double), ("req_p99_ms", c_double),
                ("voxel_updates", c_double), ("busy_ms", c_double), ("peak_bytes", c_i64), ("used_bytes", c_i64),
                ("peak_active", c_i32), ("reserved", c_i32), ("alloc_bytes", c_i64), ("peak_alloc_bytes", c_i64)]


# name -> (restype, argtypes); every entry point of include/fs3d.h
SIGNATURES = {
    "fs3d_last_error": (ctypes.c_char_p, []),
    "fs3d_abi_version": (c_int, []),
    "fs3d_device_count": (c_int, [P(c_int)]),
    "fs3d_init": (c_int, [c_int]),
    "fs3d_rule_lut": (c_int, [P(ctypes.c_uint16)]),
    "fs3d_world_create": (c_int, [P(c_i32), c_u32, P(vp)]),
    "fs3d_world_destroy": (c_int, [vp]),
    "fs3d_world_dims": (c_int, [vp, P(c_i32)]),
    "fs3d_world_set_voxel": (c_int, [vp, c_i32, c_i32, c_i32, c_u8]),
    "fs3d_world_get_voxel": (c_int, [vp, c_i32, c_i32, c_i32, P(c_u8)]),
    "fs3d_world_upload": (c_int, [vp, vp, c_size]),
    "fs3d_world_download": (c_int, [vp, vp, c_size]),
    "fs3d_world_get_tick": (c_int, [vp, P(c_u64)]),
    "fs3d_world_set_tick": (c_int, [vp, c_u64]),
    "fs3d_world_step": (c_int, [vp, c_i32]),
    "fs3d_world_tick": (c_int, [vp, P(c_u64)]),
    "fs3d_world_hash": (c_int, [vp, P(c_u64)]),
    "fs3d_world_run": (c_int, [vp, c_i32, c_i32, P(c_u64)]),
    "fs3d_world_prepare_run": (c_int, [vp, c_i32, c_i32]),
    "fs3d_world_counts": (c_int, [vp, P(c_u64)]),
    "fs3d_world_sync": (c_int, [vp]),
    "fs3d_world_device_ptr": (c_int, [vp, P(vp)]),
    "fs3d_world_set_segment": (c_int, [vp, c_i32]),
    "fs3d_world_get_segment": (c_int, [vp, P(c_i32)]),
    "fs3d_world_set_skip": (c_int, [vp, c_i32]),
    "fs3d_world_get_skip": (c_int, [vp, P(c_i32)]),
    "fs3d_world_set_fuse": (c_int, [vp, c_i32]),
    "fs3d_world_get_fuse": (c_int, [vp, P(c_i32)]),
    "fs3d_world_sweep_stats": (c_int, [vp, c_i32, P(c_u64)]),
    "fs3d_world_fill_rain": (c_int, [vp, c_u64, c_u64]),
    "fs3d_world_fill_boxes": (c_int, [vp, c_u64, P(BoxParams), P(c_i64)]),
    "fs3d_world_fill_pillars": (c_int, [vp, c_u64, P(PillarParams)]),
    "fs3d_batch_specs": (c_int, [c_i32, c_u64, c_i32, c_i32, P(c_i32), P(c_u64), P(c_u64)]),
    "fs3d_serve_specs": (c_int, [c_i32, c_u64, c_i32, c_i32, c_i32, c_i32, c_double, P(c_i32),
                                 P(c_u64), P(c_u64), P(c_i32), P(c_double)]),
    "fs3d_threshold": (c_int, [c_u64, c_u64, P(c_u64)]),
    "fs3d_world_last_run_ms": (c_int, [vp, P(c_double)]),
    "fs3d_copy_probe": (c_int, [c_size, c_i32, P(c_double)]),
    "fs3d_copy_probe_streaming": (c_int, [c_size, c_i32, P(c_double)]),
    "fs3d_scatter_probe": (c_int, [c_i32, c_i32, c_i32, P(c_double), P(c_double)]),
    "fs3d_rowquad_probe": (c_int, [c_i32, c_i32, c_i32, P(c_double)]),
    "fs3d_batch_create": (c_int, [c_i32, P(c_i32), P(vp)]),
    "fs3d_batch_destroy": (c_int, [vp]),
    "fs3d_batch_size": (c_int, [vp, P(c_i32)]),
    "fs3d_batch_upload": (c_int, [vp, c_i32, vp, c_size]),
    "fs3d_batch_download": (c_int, [vp, c_i32, vp, c_size]),
    "fs3d_batch_fill_rain": (c_int, [vp, P(c_u64), P(c_u64)]),
    "fs3d_batch_get_tick": (c_int, [vp, P(c_u64)]),
    "fs3d_batch_set_tick": (c_int, [vp, c_u64]),
    "fs3d_batch_set_config": (c_int, [vp, c_i32, c_i32, c_i32]),
    "fs3d_batch_get_config": (c_int, [vp, P(c_i32), P(c_i32), P(c_i32), P(c_i32)]),
    "fs3d_batch_set_skip": (c_int, [vp, c_i32]),
    "fs3d_batch_get_skip": (c_int, [vp, P(c_i32)]),
    "fs3d_batch_run": (c_int, [vp, c_i32, c_i32, P(c_u64)]),
    "fs3d_batch_last_run_ms": (c_int, [vp, P(c_double)]),
    "fs3d_batch_autotune": (c_int, [vp, c_i32, P(c_i32), P(c_i32), P(c_i32), P(c_i32), P(c_double)]),
    "fs3d_server_create": (c_int, [P(ServerConfig), P(vp)]),
    "fs3d_server_destroy": (c_int, [vp]),
    "fs3d_server_submit": (c_int, [vp, P(c_i32), c_u64, c_u64, c_i32, P(c_i64)]),
    "fs3d_server_poll": (c_int, [vp, P(c_i32), P(c_i32), P(c_i64)]),
    "fs3d_server_drain": (c_int, [vp]),
    "fs3d_server_result": (c_int, [vp, c_i64, P(c_i32), P(c_u64), P(c_double)]),
    "fs3d_server_replay": (c_int, [vp, c_i32, P(c_i32), P(c_u64), P(c_u64), P(c_i32), P(c_double), P(c_i64)]),
    "fs3d_server_get_stats": (c_int, [vp, P(ServerStats)]),
    "fs3d_render_sdf": (c_int, [P(c_float), c_float, c_i32, c_i32, P(c_float)]),
    "fs3d_box_vertex": (c_int, [P(c_float), P(c_float), P(c_float)]),
    "fs3d_render_box_march": (c_int, [P(c_float), P(c_float), c_i32, P(c_float)]),
    "fs3d_world_render": (c_int, [vp, P(Camera), c_i32, c_i32, P(c_float), P(c_u64), P(c_u8)]),
    "fs3d_world_last_render_ms": (c_int, [vp, P(c_double)]),
    "fs3d_world_set_render_track": (c_int, [vp, c_i32]),
    "fs3d_world_get_render_track": (c_int, [vp, P(c_i32)]),
    "fs3d_world_render_maps": (c_int, [vp, P(c_u64), ctypes.c_size_t, P(c_u64), ctypes.c_size_t, P(c_u64),
                                       ctypes.c_size_t]),
    "fs3d_orbit_camera": (c_int, [P(c_i32), c_double, c_double, c_double, c_float, P(Camera)]),
    "fs3d_key_camera": (c_int, [P(c_i32), P(c_float), P(c_float), c_float, P(Camera)]),
    "fs3d_camera_globaldata": (c_int, [P(c_float), P(c_float), P(c_float)]),
    "fs3d_camera_integrate": (c_int, [P(c_float), P(c_float), c_u32]),
    "fs3d_palette": (c_int, [P(c_float)]),
    "fs3d_rule_hash": (c_int, [P(c_u64)]),
    "fs3d_file_write": (c_int, [ctypes.c_char_p, P(c_i32), c_u64, c_u64, vp, c_size]),
    "fs3d_file_info": (c_int, [ctypes.c_char_p, P(c_i32), P(c_u64), P(c_u64)]),
    "fs3d_file_read": (c_int, [ctypes.c_char_p, vp, c_size]),
    "fs3d_world_save": (c_int, [vp, ctypes.c_char_p]),
    "fs3d_world_load": (c_int, [vp, ctypes.c_char_p]),
    "fs3d_grid_run_noalloc": (c_int, [vp, vp, P(c_i32), c_u64, c_i32, c_i32, P(c_u64), P(c_i32), vp]),
}

_LIB = None


def lib():
    """Loads libfs3d.so once; raises Fs3dError when it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise Fs3dError(5, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != FS3D_OK:
        msg = lib().fs3d_last_error()
        raise Fs3dError(rc, msg.decode() if msg else "")
    return rc


def ptr(a, t):
    return a.ctypes.data_as(P(t))
