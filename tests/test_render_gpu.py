"""GPU renderers against the CPU transcription: hit voxel index and face must be
bit-exact, colours fp32 within 1e-5 (the same float operations in the same
order on both sides)."""
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


@pytest.mark.parametrize("cam_pos,aspect,W,H", [((0.0, 0.0, -5.0), 1700.0 / 900.0, 850, 450),
                                                ((0.3, -0.2, -2.0), 1.6, 320, 200),
                                                ((0.0, 0.0, 0.1), 1.0, 64, 48),
                                                ((1.5, 2.0, -3.0), 1700.0 / 900.0, 257, 131)])
def test_sdf_port_matches_transcription(gpu, O, cam_pos, aspect, W, H):
    from fallingsand3d_amd import render as R
    announce(f"sdf render cam={cam_pos} {W}x{H}")
    got = R.render_sdf(cam_pos, aspect, W, H)
    want = O.raymarch_sdf(np.array(cam_pos, np.float32), aspect, W, H)
    assert np.abs(got - want).max() <= RGB_TOL
    assert (got[..., 0] > 0).any() or cam_pos[2] > -0.5


def _compare(R, O, w, cam, pal, W, H, label):
    from test_render_incremental_gpu import expected_maps
    rgb, hit, face = R.render(w, cam, W, H)
    g = w.download()
    maps = R.render_maps(w)
    want = expected_maps(g)
    assert all((a == b).all() for a, b in zip(maps, want)), (label, "occupancy maps")
    ocam = O.camera_from(list(cam.origin), list(cam.basis), list(cam.light), cam.aspect)
    wrgb, whit, wface = O.render_dda(g, ocam, pal, W, H)
    assert (hit == whit).all(), (label, int((hit != whit).sum()))
    assert (face == wface).all(), label
    assert np.abs(rgb - wrgb).max() <= RGB_TOL, label
    return (hit != R.MISS).mean()


def test_dda_scenes_and_cameras(gpu, O, pal):
    from fallingsand3d_amd import render as R
    scenes = []
    w1 = gpu.World((64, 64, 64))
    w1.fill_rain(0, gpu.threshold(3, 10))
    scenes.append(("rain64", w1))
    w2 = gpu.World((96, 80, 64))
    w2.fill_boxes(11, side_min=4, side_max=16)
    scenes.append(("boxes", w2))
    w3 = gpu.World((120, 90, 50))
    w3.fill_pillars(4, (12, 2, 9, 10, 40, 8, 30, 8, 30, 2, 6))
    scenes.append(("pillars", w3))
    rng = np.random.default_rng(2)
    g = rng.choice(np.array([0, 0, 0, 0, 1, 2, 3, 77, 200], np.uint8), size=(9, 13, 17))
    w4 = gpu.World((17, 13, 9))
    w4.upload(g)
    scenes.append(("tiny+reserved", w4))
    # sparse single voxels (most occupied bricks hold one or two), and voxels on
    # the last x, y and z planes of a grid whose sizes are not multiples of 2 or 8
    g5 = np.zeros((45, 67, 101), np.uint8)
    sparse = rng.random(g5.shape) < 0.002
    g5[sparse] = rng.choice(np.array([1, 2, 3, 9], np.uint8), size=int(sparse.sum()))
    g5[-1, :, ::7] = 2
    g5[:, -1, 5::11] = 3
    g5[::5, ::9, -1] = 1
    w5 = gpu.World((101, 67, 45))
    w5.upload(g5)
    scenes.append(("sparse+far-planes", w5))
    try:
        for name, w in scenes:
            for ang, ds, el, (W, H) in [(0.4, 1.6, 0.45, (320, 180)), (2.9, 1.3, 0.1, (257, 131)),
                                        (5.0, 2.2, -0.3, (96, 96))]:
                announce(f"dda {name} angle={ang} {W}x{H}")
                cam = R.orbit_camera(w.dims, ang, ds, el, W / H)
                frac = _compare(R, O, w, cam, pal, W, H, (name, ang))
                assert frac > (0.002 if name.startswith("sparse") else 0.01), (name, ang)
            # a camera inside the grid looking along +y, and an axis-aligned one from above
            nx, ny, nz = w.dims
            inside = R.camera([nx * 0.5 + 0.25, 0.6, nz * 0.75], [[-1, 0, 0], [0, 0, -1], [0, 1, 0]],
                              [0, -3 * ny, 0], 1.25)
            _compare(R, O, w, inside, pal, 120, 96, (name, "inside"))
            above = R.camera([nx * 0.5, ny * 0.5, nz * 3.0], [[-1, 0, 0], [0, 1, 0], [0, 0, -1]],
                             [0, 0, -5 * nz], 1.0)
            _compare(R, O, w, above, pal, 64, 64, (name, "above"))
            # rays with a subnormal direction component (1/d is +inf on a moving axis):
            # the walk takes single steps across empty bricks instead of jumping
            # screen x moves along world x only by 1e-39: every ray's x component is subnormal
            tiny = [[1e-39, 0, 0.8], [0, 0, 0.6], [0, 1, 0]]
            skew = R.camera([nx * 0.37 + 0.1, -0.5 * ny, nz * 0.41], tiny, [0, -3 * ny, 0], 1.0)
            _compare(R, O, w, skew, pal, 48, 48, (name, "subnormal"))
    finally:
        for _, w in scenes:
            w.close()


def test_render_hand_off_after_ticks(gpu, O, pal):
    from fallingsand3d_amd import render as R
    announce("dda after 40 ticks of a 128^3 box scene")
    with gpu.World((128, 128, 128)) as w:
        w.fill_boxes(3, side_min=6, side_max=24)
        w.step(40)
        cam = R.orbit_camera(w.dims, 1.1, 1.5, 0.5, 16 / 9)
        _compare(R, O, w, cam, pal, 480, 270, "ticked")
        assert R.last_render_ms(w) > 0


def test_box_march_port_matches_transcription(gpu, O):
    from fallingsand3d_amd import render as R
    announce("raymarch_sphere fragment stage: 200k fragments vs the transcription")
    rng = np.random.default_rng(13)
    n = 200000
    origins = (rng.normal(size=(n, 3)) * rng.uniform(0.2, 6, size=(n, 1))).astype(np.float32)
    locals_ = rng.uniform(-0.5, 0.5, size=(n, 3)).astype(np.float32)
    face = rng.integers(0, 3, size=n)
    locals_[np.arange(n // 2), face[: n // 2]] = np.sign(origins[np.arange(n // 2), face[: n // 2]]) * 0.5
    got = R.box_march(origins, locals_)
    want = O.box_march(origins, locals_)
    assert np.abs(got - want).max() <= RGB_TOL
    assert (got[:, 0] > 0).mean() > 0.3
