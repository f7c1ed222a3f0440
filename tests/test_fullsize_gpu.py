"""Full-size parity runs of the measured workloads against the oracle (OpenMP on
the host; tens of seconds each): the 512^3 box scene's 1000-tick hash sequence
and the 2048^3 pillar scene's fill, two ticks and one 1920x1080 frame."""
import json
import os

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def test_boxes512_1000_ticks_every_50th_hash(gpu, O):
    announce("512^3 boxes: 1000-tick graph, hash every 50, vs oracle")
    with gpu.World((512, 512, 512)) as w:
        w.fill_boxes(1)
        h = w.run(1000, 50)
    g, _ = O.scene_boxes((512, 512, 512), 1)
    _, want = O.run(g, 0, 1000, 50, use_lut=True)
    assert [int(x) for x in h] == [int(x) for x in want]


def test_pillars2048_fill_ticks_and_frame(gpu, O):
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import orbit_angles
    from fallingsand3d_amd import render as R
    params = (256, 4, 32, 256, 1800, 256, 682, 256, 682, 32, 128)
    pal = np.array(json.load(open(os.path.join(os.path.dirname(__file__), "golden", "palette.json")))["rgba"],
                   np.float32)
    announce("2048^3 pillars: device fill vs oracle fill")
    with gpu.World((2048, 2048, 2048)) as w:
        w.fill_pillars(2, params)
        g = O.scene_pillars((2048, 2048, 2048), 2, params)
        assert w.hash() == O.state_hash(g)
        announce("2048^3 pillars: two ticks (both parities) vs oracle")
        w.step(2)
        g, _ = O.run(g, 0, 2, use_lut=True)
        assert w.hash() == O.state_hash(g)
        announce("2048^3 pillars: one 1920x1080 frame vs oracle")
        cam = R.orbit_camera(w.dims, orbit_angles(2, 1)[0], 1.4, 0.35, 1920 / 1080)
        rgb, hit, face = R.render(w, cam, 1920, 1080)
        ocam = O.camera_from(list(cam.origin), list(cam.basis), list(cam.light), cam.aspect)
        wrgb, whit, wface = O.render_dda(g, ocam, pal, 1920, 1080)
        assert (hit == whit).all() and (face == wface).all()
        assert np.abs(rgb - wrgb).max() <= 1e-5
        assert (hit != R.MISS).mean() > 0.05
        announce("2048^3 pillars: four more (packed skipping) ticks, a frame from the changed bricks' maps")
        w.step(4)
        g, _ = O.run(g, 2, 4, use_lut=True)
        assert w.hash() == O.state_hash(g)
        rgb, hit, face = R.render(w, cam, 960, 540)
        ocam = O.camera_from(list(cam.origin), list(cam.basis), list(cam.light), cam.aspect)
        wrgb, whit, wface = O.render_dda(g, ocam, pal, 960, 540)
        assert (hit == whit).all() and (face == wface).all()
        assert np.abs(rgb - wrgb).max() <= 1e-5


def test_pillars2048_skipping_matches_dense_and_fused(gpu):
    # The default tick at this size packs the active segments of listed row
    # quads; past the first two ticks most segments are settled.  Dense single
    # ticks (checked against the oracle above) and fused pairs are the reference.
    params = (256, 4, 32, 256, 1800, 256, 682, 256, 682, 32, 128)
    dims = (2048, 2048, 2048)
    announce("2048^3 pillars: 40 ticks, default (packed skipping) vs dense vs fused, hash every 10")
    hashes = {}
    for kind in ("skip", "dense", "fused"):
        with gpu.World(dims, flags=0 if kind == "skip" else gpu.DENSE) as w:
            if kind != "skip":
                w.fuse = kind == "fused"
                assert w.fuse == (kind == "fused")
            else:
                assert w.skip == gpu.SKIP_AUTO
                w.sweep_stats(True)
            w.fill_pillars(2, params)
            hashes[kind] = [int(x) for x in w.run(40, 10)]
            if kind == "skip":
                st = w.sweep_stats(False)
                assert st["violations"] == 0
                assert st["swept"] < 0.5 * 40 * 2048 ** 3  # settled segments were skipped
    assert hashes["skip"] == hashes["dense"] == hashes["fused"]
