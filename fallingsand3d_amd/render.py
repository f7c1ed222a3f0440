"""Rendering and camera helpers (host mirror of the fs3d render entry points).

`render_sdf` is the HIP port of the reference's full-screen sphere tracer
(shaders/fs_raymarch.frag); `render(world, camera, W, H)` ray-casts a world's
grid with the same pixel mapping and lighting, returning colour, the hit voxel
index and the entry face per pixel.
"""
import ctypes

import numpy as np

from ._lib import Camera, c_double, c_float, c_i32, c_u8, c_u64, check, lib, ptr

MISS = np.iinfo(np.uint64).max
KEY_W, KEY_A, KEY_S, KEY_D, KEY_CTRL, KEY_SPACE, KEY_LEFT, KEY_RIGHT = 1, 2, 4, 8, 16, 32, 64, 128


def palette():
    p = np.zeros((256, 4), np.float32)
    check(lib().fs3d_palette(ptr(p, c_float)))
    return p


def render_sdf(cam_pos=(0.0, 0.0, -5.0), aspect=1700.0 / 900.0, W=850, H=450):
    cp = np.ascontiguousarray(cam_pos, np.float32)
    out = np.zeros((H, W, 4), np.float32)
    check(lib().fs3d_render_sdf(ptr(cp, c_float), aspect, W, H, ptr(out, c_float)))
    return out


def box_vertex(cam_pos, inv_model):
    """Ray origin of shaders/raymarch_sphere.vert: invModel * (-camPos, 1) (inv_model
    column-major 4x4)."""
    cp = np.ascontiguousarray(cam_pos, np.float32)
    m = np.ascontiguousarray(inv_model, np.float32).reshape(16)
    out = np.zeros(3, np.float32)
    check(lib().fs3d_box_vertex(ptr(cp, c_float), ptr(m, c_float), ptr(out, c_float)))
    return out


def box_march(origins, locals_):
    """shaders/raymarch_sphere.frag on the GPU for n fragments: (n, 4) RGBA."""
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    l = np.ascontiguousarray(locals_, np.float32).reshape(-1, 3)
    out = np.zeros((len(o), 4), np.float32)
    check(lib().fs3d_render_box_march(ptr(o, c_float), ptr(l, c_float), len(o), ptr(out, c_float)))
    return out


def orbit_camera(dims, angle, dist_scale=1.6, elev=0.4, aspect=16.0 / 9.0):
    d = (c_i32 * 3)(*[int(v) for v in dims])
    cam = Camera()
    check(lib().fs3d_orbit_camera(d, angle, dist_scale, elev, aspect, ctypes.byref(cam)))
    return cam


def key_camera(dims, cam_pos, cam_rot, aspect=16.0 / 9.0):
    """Voxel camera for the reference's key-driven camera state (see camera_integrate)."""
    d = (c_i32 * 3)(*[int(v) for v in dims])
    p = np.ascontiguousarray(cam_pos, np.float32)
    r = np.ascontiguousarray(cam_rot, np.float32)
    cam = Camera()
    check(lib().fs3d_key_camera(d, ptr(p, c_float), ptr(r, c_float), aspect, ctypes.byref(cam)))
    return cam


def camera(origin, basis, light, aspect):
    cam = Camera()
    cam.origin[:] = [float(v) for v in origin]
    cam.basis[:] = [float(v) for v in np.asarray(basis, np.float64).reshape(9)]
    cam.light[:] = [float(v) for v in light]
    cam.aspect = float(aspect)
    return cam


def render(world, cam, W, H):
    """(rgb (H,W,3) float32, hit (H,W) uint64, face (H,W) uint8) of the world's grid."""
    rgb = np.zeros((H, W, 3), np.float32)
    hit = np.zeros((H, W), np.uint64)
    face = np.zeros((H, W), np.uint8)
    check(lib().fs3d_world_render(world.handle, ctypes.byref(cam), W, H, ptr(rgb, c_float), ptr(hit, c_u64),
                                  ptr(face, c_u8)))
    return rgb, hit, face


def last_render_ms(world):
    ms = c_double(0)
    check(lib().fs3d_world_last_render_ms(world.handle, ctypes.byref(ms)))
    return ms.value


def render_maps(world):
    """(coarse, fine, region) occupancy maps of the world's last render, as
    uint64 arrays in the layouts fs3d_world_render_maps documents."""
    nx, ny, nz = world.dims
    bnx, bny, bnz = (nx + 7) // 8, (ny + 7) // 8, (nz + 7) // 8
    rnx, rny, rnz = (bnx + 3) // 4, (bny + 3) // 4, (bnz + 3) // 4
    coarse = np.zeros(((bnx + 63) // 64) * bny * bnz, np.uint64)
    fine = np.zeros(bnx * bny * bnz, np.uint64)
    region = np.zeros(((rnx + 63) // 64) * rny * rnz, np.uint64)
    check(lib().fs3d_world_render_maps(world.handle, ptr(coarse, c_u64), coarse.size, ptr(fine, c_u64), fine.size,
                                       ptr(region, c_u64), region.size))
    return coarse, fine, region


def camera_globaldata(cam_pos, cam_rot):
    """The reference's 76-float GlobalData block for a camera state."""
    p = np.ascontiguousarray(cam_pos, np.float32)
    r = np.ascontiguousarray(cam_rot, np.float32)
    out = np.zeros(76, np.float32)
    check(lib().fs3d_camera_globaldata(ptr(p, c_float), ptr(r, c_float), ptr(out, c_float)))
    return out


def camera_integrate(cam_pos, cam_rot, keys):
    """One 0.016 s step of the reference's key-driven camera; returns (pos, rot)."""
    p = np.array(cam_pos, np.float32)
    r = np.array(cam_rot, np.float32)
    check(lib().fs3d_camera_integrate(ptr(p, c_float), ptr(r, c_float), keys))
    return p, r
