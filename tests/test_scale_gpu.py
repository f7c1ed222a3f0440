"""The measured workloads at their full sizes against the oracle's fixtures
(tests/golden/make_scale_golden.py wrote them in the build container):

  - the 512^3 box scene: the state hash after every one of 2100 ticks;
  - the 2048^3 pillar scene: 200 ticks, a 1920x1080 frame after every 10th
    (state hash, SHA-256 of hit and face, RGB of 512 hit pixels per frame; the
    full RGB of three frames against the host oracle's render of the same grid),
    with the default (packed skipping), dense and fused ticks;
  - 4096 batched worlds: every world's hash after tick 20 (persistent, wide and
    plain) and after tick 500 (persistent, with and without settled-world
    skipping);
  - the 6000-request serving stream: every request's final hash, replayed in
    real time and all at once.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PILLARS_2048 = (256, 4, 32, 256, 1800, 256, 682, 256, 682, 32, 128)


def rows(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return [[int(v) for v in line.split()] for line in f if line.strip() and not line.startswith("#")]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_boxes512_every_tick_to_2100(gpu):
    want = [r[0] for r in rows("boxes512_hashes.txt")]
    assert len(want) == 2101
    announce("512^3 boxes: 2100 ticks, hash after every tick, vs oracle fixture")
    with gpu.World((512, 512, 512)) as w:
        w.fill_boxes(1)
        assert w.hash() == want[0]
        w.sweep_stats(True)
        got = [int(x) for x in w.run(2100, 1)]
        assert w.sweep_stats(False)["violations"] == 0
    bad = [k + 1 for k, (g, h) in enumerate(zip(got, want[1:])) if g != h]
    assert not bad, f"first mismatching ticks {bad[:5]}"


def test_boxes512_graph_windows_match_fixture(gpu):
    # the timed shapes the benchmark uses: warm-up, then hashes every 50 and on
    # the last tick (here a 1000-tick window after 100 and a 20-tick one after 5)
    want = [r[0] for r in rows("boxes512_hashes.txt")]
    sys.path.insert(0, ROOT)
    from bench import hash_segments
    for warmup, steps in ((100, 1000), (5, 20), (7, 123)):
        announce(f"512^3 boxes: warm-up {warmup}, {steps} timed ticks in graph runs")
        with gpu.World((512, 512, 512)) as w:
            w.fill_boxes(1)
            w.run(warmup, 0)
            t = warmup
            for nt, he in hash_segments(steps):
                h = [int(x) for x in w.run(nt, he)]
                assert h == [want[t + he * (i + 1)] for i in range(nt // he)], (warmup, steps, nt, he)
                t += nt
            assert w.tick_count == warmup + steps


@pytest.fixture(scope="module")
def frames():
    with open(os.path.join(GOLDEN, "pillars2048_frames.json")) as f:
        meta = json.load(f)
    sample = np.load(os.path.join(GOLDEN, "pillars2048_rgb_sample.npz"))
    return meta, sample["index"], sample["rgb"]


def test_pillars2048_200_ticks_20_frames(gpu, O, frames):
    from bench import orbit_angles
    from fallingsand3d_amd import render as R
    meta, sidx, srgb = frames
    W, H = meta["W"], meta["H"]
    fx = meta["frames"]
    assert len(fx) == 20
    pal = np.array(json.load(open(os.path.join(GOLDEN, "palette.json")))["rgba"], np.float32)
    angles = orbit_angles(2, 20)
    full_rgb = {0, 9, 19}  # frames whose every pixel is checked against the host oracle
    with gpu.World((2048, 2048, 2048)) as w:
        w.fill_pillars(2, PILLARS_2048)
        w.sweep_stats(True)
        for k, ang in enumerate(angles):
            announce(f"2048^3 pillars: ticks to {10 * (k + 1)}, frame {k} (default tick, render tracking)")
            w.run(10, 0)
            assert w.hash() == fx[k]["hash"], k
            cam = R.orbit_camera(w.dims, ang, 1.4, 0.35, W / H)
            rgb, hit, face = R.render(w, cam, W, H)
            assert sha(hit.astype("<u8")) == fx[k]["hit_sha256"], k
            assert sha(face) == fx[k]["face_sha256"], k
            got = rgb.reshape(-1, 3)[sidx[k]]
            assert np.abs(got - srgb[k]).max() <= 1e-5, k
            if k in full_rgb:
                announce(f"2048^3 pillars: frame {k}, every pixel vs the host oracle's render")
                g = w.download()
                ocam = O.camera_from(list(cam.origin), list(cam.basis), list(cam.light), cam.aspect)
                wrgb, whit, wface = O.render_dda(g, ocam, pal, W, H)
                del g
                assert (hit == whit).all() and (face == wface).all(), k
                assert np.abs(rgb - wrgb).max() <= 1e-5, k
        assert w.sweep_stats(False)["violations"] == 0


@pytest.mark.parametrize("kind", ["dense", "fused"])
def test_pillars2048_dense_and_fused_match_fixture(gpu, frames, kind):
    meta, _, _ = frames
    want = [f["hash"] for f in meta["frames"]]
    announce(f"2048^3 pillars: 200 {kind} ticks, hash every 10 vs oracle fixture")
    with gpu.World((2048, 2048, 2048), flags=gpu.DENSE) as w:
        w.fuse = kind == "fused"
        assert w.fuse == (kind == "fused")
        w.fill_pillars(2, PILLARS_2048)
        got = [int(x) for x in w.run(200, 10)]
    assert got == want


def test_batch4096_against_fixture(gpu):
    from fallingsand3d_amd.batch import PERSISTENT, WIDE
    fx = rows("batch4096_hashes.txt")
    dims, thr, seeds = gpu.batch_specs(4096, 3, 16, 128)
    assert [tuple(r[:3]) for r in fx] == [tuple(int(v) for v in d) for d in dims]
    want20 = np.array([r[3] for r in fx], np.uint64)
    want500 = np.array([r[4] for r in fx], np.uint64)
    with gpu.Batch(dims) as b:
        for mode, name in ((PERSISTENT, "persistent"), (WIDE, "wide"), (None, "plain")):
            announce(f"batch 4096 worlds: 20 ticks {name} vs oracle fixture")
            b.tick_count = 0
            b.fill_rain(seeds, thr)
            h = b.run(20, plain=True) if mode is None else b.run(20, mode=mode)
            assert (h == want20).all(), (name, np.flatnonzero(h != want20)[:5])
        for skip in (True, False):
            announce(f"batch 4096 worlds: 500 ticks persistent, settled-world skipping {skip}")
            b.skip = skip
            b.tick_count = 0
            b.fill_rain(seeds, thr)
            h = b.run(500, mode=PERSISTENT)
            assert (h == want500).all(), (skip, np.flatnonzero(h != want500)[:5])
        b.skip = True
        announce("batch 4096 worlds: autotuned run of 500 ticks")
        b.autotune(trial_ticks=5, cache_file=None)
        b.tick_count = 0
        b.fill_rain(seeds, thr)
        h = b.run(500)
        assert (h == want500).all()


def test_serve6000_against_fixture(gpu):
    fx = rows("serve6000_hashes.txt")
    n = len(fx)
    dims, thr, seeds, ticks, arr = gpu.serve_specs(n, 5, 32, 256, 10, 500, 2000.0)
    assert [tuple(r[:4]) for r in fx] == [(*(int(v) for v in d), int(t)) for d, t in zip(dims, ticks)]
    want = [r[4] for r in fx]
    for name, arrival in (("real time", arr), ("all at once", np.zeros(n))):
        announce(f"serving {n} requests ({name}) vs oracle fixture")
        with gpu.Server(budget_bytes=200 << 30, round_ticks=8, max_active=1 << 16) as s:
            ids = s.replay(dims, seeds, thr, ticks, arrival)
            res = [s.result(int(i)) for i in ids]
            st = s.stats()
        assert all(r[0] == 2 for r in res)
        bad = [i for i, r in enumerate(res) if r[1] != want[i]]
        assert not bad, (name, bad[:5])
        assert st["peak_alloc_bytes"] <= 200 << 30
