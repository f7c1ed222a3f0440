"""The ray caster's occupancy maps between frames: once a world has been
rendered, skipping ticks mark the 8^3 bricks they change and the next frame
reads only those bricks again; dense ticks and host-side edits make it read
every brick.  Every frame must equal the oracle's render of the current grid
(hit and face exact, RGB within 1e-5), whatever happened since the last one."""
import json
import os

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

RGB_TOL = 1e-5


@pytest.fixture(scope="module")
def pal():
    gold = os.path.join(os.path.dirname(__file__), "golden", "palette.json")
    return np.array(json.load(open(gold))["rgba"], np.float32)


def expected_maps(g):
    """The occupancy maps of grid g[z][y][x] (numpy, from their definition):
    coarse bits per 8^3 brick, fine words with a bit per 2x2x2 sub-brick, and
    region bits per 32^3 region."""
    nz, ny, nx = g.shape
    bnx, bny, bnz = -(-nx // 8), -(-ny // 8), -(-nz // 8)
    pad = np.zeros((bnz * 8, bny * 8, bnx * 8), bool)
    pad[:nz, :ny, :nx] = g != 0
    sub = pad.reshape(bnz * 4, 2, bny * 4, 2, bnx * 4, 2).any(axis=(1, 3, 5))  # [sz][sy][sx]
    sub = sub.reshape(bnz, 4, bny, 4, bnx, 4).transpose(0, 2, 4, 1, 3, 5)     # [bz][by][bx][z][y][x]
    weights = (np.uint64(1) << np.arange(64, dtype=np.uint64)).reshape(4, 4, 4)  # bit x + 4 y + 16 z
    fine = (sub.astype(np.uint64) * weights).sum(axis=(3, 4, 5), dtype=np.uint64).reshape(-1)
    bw = -(-bnx // 64)
    occ = (fine != 0).reshape(bnz * bny, bnx)
    coarse = np.zeros(bnz * bny * bw, np.uint64)
    for bx in range(bnx):
        coarse[np.arange(bnz * bny) * bw + bx // 64] |= occ[:, bx].astype(np.uint64) << np.uint64(bx % 64)
    # regions of 4 x 4 x 4 bricks
    rnx, rny, rnz = -(-bnx // 4), -(-bny // 4), -(-bnz // 4)
    rpad = np.zeros((rnz * 4, rny * 4, rnx * 4), bool)
    rpad[:bnz, :bny, :bnx] = (fine != 0).reshape(bnz, bny, bnx)
    rocc = rpad.reshape(rnz, 4, rny, 4, rnx, 4).any(axis=(1, 3, 5)).reshape(rnz * rny, rnx)
    rw = -(-rnx // 64)
    region = np.zeros(rnz * rny * rw, np.uint64)
    for rx in range(rnx):
        region[np.arange(rnz * rny) * rw + rx // 64] |= rocc[:, rx].astype(np.uint64) << np.uint64(rx % 64)
    return coarse, fine, region


def _frame(R, O, w, cam, pal, W, H, label):
    rgb, hit, face = R.render(w, cam, W, H)
    g = w.download()
    # the maps after a render describe the current grid exactly, however they
    # were brought up to date
    coarse, fine, region = R.render_maps(w)
    want_c, want_f, want_r = expected_maps(g)
    assert (fine == want_f).all(), (label, "fine", int((fine != want_f).sum()))
    assert (coarse == want_c).all(), (label, "coarse")
    assert (region == want_r).all(), (label, "region")
    ocam = O.camera_from(list(cam.origin), list(cam.basis), list(cam.light), cam.aspect)
    wrgb, whit, wface = O.render_dda(g, ocam, pal, W, H)
    assert (hit == whit).all(), (label, int((hit != whit).sum()))
    assert (face == wface).all(), label
    assert np.abs(rgb - wrgb).max() <= RGB_TOL, label
    return float((hit != R.MISS).mean())


CASES = [("WAVES", 16, (160, 130, 100), "ON"), ("LISTED", 16, (160, 130, 100), "ON"),
         ("SEGS", 16, (160, 130, 100), "ON"), ("SEGS", 8, (150, 70, 90), "ON"), ("WAVES", 32, (300, 41, 77), "ON"),
         ("SEGS", 64, (2100, 20, 24), "ON"), ("OFF", 16, (160, 130, 100), "ON"),
         ("SEGS", 16, (160, 130, 100), "OFF"), ("WAVES", 16, (160, 130, 100), "AUTO")]


@pytest.mark.parametrize("mode,seg,shape,track", CASES, ids=lambda c: str(c))
def test_frames_between_ticks_and_edits(gpu, O, pal, mode, seg, shape, track):
    from fallingsand3d_amd import render as R
    nx, ny, nz = shape
    with gpu.World(shape) as w:
        w.segment = seg
        w.skip = getattr(gpu, "SKIP_" + mode)
        w.render_track = getattr(gpu, "TRACK_" + track)
        assert w.render_track == getattr(gpu, "TRACK_" + track)
        if nx * ny * nz >= 1 << 20 and nx >= 64 and ny >= 64 and nz >= 64:
            w.fill_boxes(5, side_min=6, side_max=24)
        else:
            w.fill_rain(5, gpu.threshold(1, 4))
        cams = [R.orbit_camera(w.dims, a, 1.5, 0.4, 4 / 3) for a in (0.7, 3.9)]
        W, H = 160, 120
        steps = [("first frame", lambda: None),
                 ("step 3", lambda: w.step(3)),
                 ("run 8, hash every 4 (graph)", lambda: w.run(8, 4)),
                 ("tick with hash", lambda: w.tick()),
                 ("set_voxel", lambda: w.set_voxel(nx // 2, ny // 3, nz - 2, gpu.SAND)),
                 ("step 5", lambda: w.step(5)),
                 ("run 6 twice (graph replay)", lambda: (w.run(6), w.run(6))),
                 ("no change", lambda: None),
                 ("set tick", lambda: setattr(w, "tick_count", w.tick_count + 3)),
                 ("step 7", lambda: w.step(7))]
        for i, (what, op) in enumerate(steps):
            announce(f"incremental frames {mode} seg={seg} {shape} track={track}: {what}")
            op()
            _frame(R, O, w, cams[i % 2], pal, W, H, (mode, seg, shape, what))
        # a grid uploaded from the host, then ticks on top of it
        g = w.download()
        g[nz // 2:nz // 2 + 3, :, :] = gpu.SAND
        w.upload(g)
        _frame(R, O, w, cams[0], pal, W, H, (mode, seg, shape, "upload"))
        w.run(10, 5)
        assert _frame(R, O, w, cams[1], pal, W, H, (mode, seg, shape, "run after upload")) > 0


def test_fused_dense_frames(gpu, O, pal):
    """Dense fused pairs leave no change masks: every frame rebuilds the maps."""
    from fallingsand3d_amd import render as R
    with gpu.World((600, 40, 48), flags=gpu.DENSE) as w:
        w.fuse = True
        assert w.fuse
        w.fill_rain(9, gpu.threshold(1, 3))
        cam = R.orbit_camera(w.dims, 2.2, 1.3, 0.5, 2.0)
        for what in ("first", "run 10", "step 4"):
            announce(f"fused dense frames: {what}")
            if what == "run 10":
                w.run(10)
            elif what == "step 4":
                w.step(4)
            _frame(R, O, w, cam, pal, 200, 100, what)


def test_tracking_switched_between_frames(gpu, O, pal):
    """Tracking turned on after a few frames and off again: frames stay exact."""
    from fallingsand3d_amd import render as R
    with gpu.World((140, 120, 96)) as w:
        w.fill_boxes(8, side_min=6, side_max=20)
        cam = R.orbit_camera(w.dims, 1.9, 1.4, 0.45, 1.5)
        for i, track in enumerate(["AUTO", "AUTO", "ON", "ON", "ON", "OFF", "OFF", "ON", "ON"]):
            announce(f"tracking switched: frame {i} track={track}")
            w.render_track = getattr(gpu, "TRACK_" + track)
            w.run(4) if i % 2 else w.step(3)
            _frame(R, O, w, cam, pal, 120, 80, (i, track))
