"""The viewer shell (fs3d_view) and the key-driven voxel camera it uses.

The camera mapping is pinned against the reference's own view matrices (the
glm fixture of tests/golden/camera_globaldata.json): the eye must be -R^T t and
the screen axes minus the rows of R, carried into grid coordinates by the
documented world -> grid transform."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import announce

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _to_grid(v):
    return np.array([v[0], -v[2], v[1]])


def test_key_camera_matches_glm_view_matrices():
    from fallingsand3d_amd import render as R
    fix = json.load(open(os.path.join(ROOT, "tests", "golden", "camera_globaldata.json")))
    dims = (96, 64, 40)
    c = np.array(dims) * 0.5
    half = 0.5 * max(dims)
    for state, block in zip(fix["states"], fix["blocks"]):
        pos, rot = state[:3], state[3:]
        view = np.array(block[12:28], np.float64).reshape(4, 4).T  # column-major -> [row][col]
        Rm, t = view[:3, :3], view[:3, 3]
        eye = -Rm.T @ t
        cam = R.key_camera(dims, pos, rot, 1.5)
        np.testing.assert_allclose(np.array(cam.origin), c + half * _to_grid(eye), rtol=1e-5, atol=1e-3)
        basis = np.array(cam.basis).reshape(3, 3)
        for k in range(3):
            np.testing.assert_allclose(basis[k], _to_grid(-Rm[k]), atol=1e-5)
        assert cam.aspect == pytest.approx(1.5)


def test_key_camera_initial_state_closed_form():
    from fallingsand3d_amd import render as R
    dims = (128, 100, 60)
    cam = R.key_camera(dims, (0.0, 0.0, -5.0), (0.0, 0.0, 0.0), 2.0)
    # 5 half-extents in front of the grid centre along -y, looking along +y
    np.testing.assert_allclose(cam.origin, (64.0, 50.0 - 5 * 64.0, 30.0), atol=1e-4)
    np.testing.assert_allclose(np.array(cam.basis).reshape(3, 3), [[-1, 0, 0], [0, 0, -1], [0, 1, 0]], atol=1e-6)
    np.testing.assert_allclose(cam.light, (64.0 + 2 * 64.0, 50.0 - 3 * 64.0, 30.0 + 5 * 64.0), atol=1e-4)
    # a quarter turn to the right (held RIGHT raises the yaw) looks along +x
    cam = R.key_camera(dims, (0.0, 0.0, -5.0), (0.0, 90.0, 0.0), 2.0)
    np.testing.assert_allclose(np.array(cam.basis).reshape(3, 3)[2], (1, 0, 0), atol=1e-6)


def build_viewer(tmp_path):
    from fallingsand3d_amd import _lib
    exe = tmp_path / "fs3d_view"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "fallingsand3d_amd", "viewer", "fs3d_view.cpp"), "-o", str(exe),
                    "-L", libdir, "-lfs3d", "-Wl,-rpath," + libdir], check=True)
    return exe


def test_viewer_reports_script_errors_before_running(tmp_path):
    exe = build_viewer(tmp_path)
    script = tmp_path / "bad.txt"
    script.write_text("down W\nframes 3\njump 2\n")
    r = subprocess.run([str(exe), "--script", str(script)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "ERROR: " in r.stderr and "bad.txt:3: unknown event 'jump'" in r.stderr
    script.write_text("down Q\n")
    r = subprocess.run([str(exe), "--script", str(script)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "bad.txt:1: expected a key" in r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_viewer_without_gpu_fails_loudly(tmp_path):
    import fallingsand3d_amd as F
    if F.device_count() > 0:
        pytest.skip("a GPU is visible")
    exe = build_viewer(tmp_path)
    script = tmp_path / "s.txt"
    script.write_text("frames 1\n")
    r = subprocess.run([str(exe), "--script", str(script)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "ERROR: " in r.stderr


def read_ppm(path):
    data = open(path, "rb").read()
    head, W, H = b"", 0, 0
    parts = data.split(b"\n", 3)
    assert parts[0] == b"P6" and parts[2] == b"255"
    W, H = map(int, parts[1].split())
    return np.frombuffer(parts[3], np.uint8).reshape(H, W, 3)


def quantize(rgb):
    """The viewer's PPM rounding: clamp to [0, 1], then (uint8)(v * 255 + 0.5) in fp32."""
    v = np.clip(rgb.astype(np.float32), np.float32(0), np.float32(1))
    return (v * np.float32(255) + np.float32(0.5)).astype(np.uint8)


def check_frames_against_oracle(out, dims, tpf, O):
    from fallingsand3d_amd import render as R
    W_KEY, RIGHT_KEY = 1, 128
    pal = np.array(json.load(open(os.path.join(ROOT, "tests", "golden", "palette.json")))["rgba"], np.float32)
    # (drawn frame, ticks before its render, keys held on each drawn frame up to it)
    held = [W_KEY] * 5 + [RIGHT_KEY] * 4 + [0]
    g0, _ = O.scene_boxes(dims, 7, side_min=8, side_max=32)
    cases = [(0, 1 * tpf), (8, 9 * tpf), (9, 13 * tpf)]
    for drawn, nticks in cases:
        announce(f"viewer frame {drawn} vs oracle render ({nticks} ticks)")
        if drawn == 9:  # the edit lands before the last frame's ticks
            g, _ = O.run(g0, 0, nticks - tpf, use_lut=True)
            g[60, 30, 20] = 2
            g, _ = O.run(g, nticks - tpf, tpf, use_lut=True)
        else:
            g, _ = O.run(g0, 0, nticks, use_lut=True)
        pos, rot = (0.0, 0.0, -5.0), (0.0, 0.0, 0.0)
        for keys in held[:drawn + 1]:
            pos, rot = R.camera_integrate(pos, rot, keys)
        img = read_ppm(os.path.join(out, f"frame_{drawn:06d}.ppm"))
        H, W = img.shape[:2]
        cam = R.key_camera(dims, pos, rot, W / H)
        ocam = O.camera_from(list(cam.origin), list(cam.basis), list(cam.light), cam.aspect)
        rgb, hit, _ = O.render_dda(g, ocam, pal, W, H)
        want = quantize(rgb)
        diff = np.abs(img.astype(np.int16) - want.astype(np.int16))
        # RGB within 1e-5 of the oracle can round to the neighbouring byte only
        # where v * 255 + 0.5 lies within 0.003 of an integer
        assert diff.max() <= 1, (drawn, int(diff.max()))
        near = np.abs((np.clip(rgb, 0, 1) * 255 + 0.5) % 1.0 - 0.5) > 0.497
        assert not (diff.astype(bool) & ~near).any(), drawn
        assert (hit != R.MISS).mean() > 0.02  # the grid is in view (5 half-extents away)


@pytest.mark.gpu
def test_viewer_session_matches_plain_world(gpu, O, tmp_path):
    exe = build_viewer(tmp_path)
    script = tmp_path / "session.txt"
    script.write_text(
        "# walk forward, turn right, minimise, drop a grain, snapshot\n"
        "down W\nframes 5\nup W\n"
        "down RIGHT\nframes 4\nup RIGHT\n"
        "minimize\nframes 3\nrestore\n"
        "set 20 30 60 2\nsnap\nframes 2\n"
        "quit\nframes 100\n")
    out = tmp_path / "frames"
    out.mkdir()
    dims, tpf = (96, 80, 64), 3
    announce(f"viewer session {dims} boxes, {tpf} ticks per frame")
    r = subprocess.run([str(exe), "--script", str(script), "--dims", *map(str, dims), "--scene", "boxes",
                        "--seed", "7", "--ticks-per-frame", str(tpf), "--size", "64", "40", "--every", "4",
                        "--out", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    got = {f[i]: f[i + 1] for i in range(0, len(f) - 1, 2) if f[i] in ("frames", "drawn", "ticks", "hash", "yaw")}
    assert int(got["frames"]) == 14 and int(got["drawn"]) == 11 and int(got["ticks"]) == 14 * tpf
    # the camera moves only on drawn frames: 5 forward steps, 4 yaw steps
    cam = f[f.index("cam") + 1:f.index("cam") + 4]
    assert float(cam[2]) == pytest.approx(-5.0 + 5 * 1.5 * 0.016, abs=1e-5)
    assert float(got["yaw"]) == pytest.approx(4 * 10.0 * 0.016, abs=1e-5)
    # frames 0, 4, 8 on the cadence, and the snapshot (drawn frame 9)
    names = sorted(os.listdir(out))
    assert names == ["frame_000000.ppm", "frame_000004.ppm", "frame_000008.ppm", "frame_000009.ppm"], names
    img = open(out / "frame_000000.ppm", "rb").read()
    assert img.startswith(b"P6\n64 40\n255\n") and len(img) == len(b"P6\n64 40\n255\n") + 64 * 40 * 3
    assert max(img[14:]) > 0  # the grid is in view
    # frames 0, 8 and 9 against the oracle's render of the oracle's grid at that
    # tick, under the camera the held keys give, after the PPM's 8-bit rounding
    check_frames_against_oracle(out, dims, tpf, O)
    # the same ticks and edit on a plain world give the same state
    with gpu.World(dims) as w:
        w.fill_boxes(7, side_min=8, side_max=32)
        w.step(12 * tpf)
        w.set_voxel(20, 30, 60, 2)
        w.step(2 * tpf)
        assert w.hash() == int(got["hash"])
