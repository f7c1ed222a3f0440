"""Two-tick fusion of the dense tick: fused pairs must equal two single ticks (and
the oracle) bit for bit, for either start parity, odd sizes, partial last lanes,
tile and z-range edges, and graph runs whose hash ticks break pairs."""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

SHAPES = [(257, 9, 10), (300, 7, 66), (512, 12, 8), (513, 5, 70), (700, 13, 3), (1000, 8, 9), (1024, 2, 2),
          (1025, 11, 6), (2048, 6, 7), (2000, 3, 65), (333, 1, 1), (400, 31, 2),
          (1500, 100, 80), (2100, 5, 7), (3584, 3, 4), (3632, 4, 5)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fused_pairs_match_single_ticks(gpu, O, shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 31 + ny * 7 + nz)
    g = rng.choice(np.array([0, 0, 0, 1, 2, 2, 3, 3], np.uint8), size=(nz, ny, nx))
    g[rng.random(g.shape) < 0.02] = 230
    for t0 in (6, 11):  # both start parities
        announce(f"fused vs single {shape} t0={t0}")
        with gpu.World(shape, flags=gpu.DENSE) as f, gpu.World(shape, flags=gpu.DENSE) as s:
            f.fuse, s.fuse = True, False
            assert f.fuse and not s.fuse
            for w in (f, s):
                w.upload(g)
                w.tick_count = t0
            f.step(6)
            s.step(6)
            a, b = f.download(), s.download()
            assert (a == b).all(), (shape, t0, int((a != b).sum()))
            want, _ = O.run(g, t0, 6, use_lut=True)
            assert (a == want).all()
            # odd counts and runs with hashes: pairs around hash ticks fall back to single ticks
            assert f.tick() == s.tick()
            f.step(3)
            s.step(3)
            assert [int(x) for x in f.run(20, 3)] == [int(x) for x in s.run(20, 3)]
            assert (f.download() == s.download()).all()
            assert f.tick_count == s.tick_count == t0 + 30


def test_fused_graph_replays(gpu):
    announce("fused graph runs: 512^3 box scene, 2 x 100 ticks with hashes every 25")
    with gpu.World((512, 512, 512), flags=gpu.DENSE) as f, gpu.World((512, 512, 512), flags=gpu.DENSE) as s:
        f.fuse, s.fuse = True, False
        for w in (f, s):
            w.fill_boxes(1)
        for _ in range(2):
            assert [int(x) for x in f.run(100, 25)] == [int(x) for x in s.run(100, 25)]
        assert f.hash() == s.hash()


def test_fuse_defaults_and_limits(gpu):
    with gpu.World((600, 8, 8), flags=gpu.DENSE) as w:
        assert not w.fuse  # small grids run single ticks by default
        w.fuse = True
        assert w.fuse
    with gpu.World((100, 8, 8), flags=gpu.DENSE) as w:
        w.fuse = True
        assert not w.fuse  # 100 bytes need 13 lanes of 8 bytes: several row quads per wave
    with gpu.World((3632, 4, 4), flags=gpu.DENSE) as w:
        w.fuse = True
        assert w.fuse  # 16-byte x-chunks: 3 layers x 6 rows x 3632 bytes fit 64 KiB
    with gpu.World((3633, 4, 4), flags=gpu.DENSE) as w:
        w.fuse = True
        assert not w.fuse  # 3 layers x 6 rows x 3648 bytes: the LDS ring would exceed 64 KiB
    with gpu.World((1024, 1024, 260), flags=gpu.DENSE) as w:
        assert w.fuse  # on by default above 2^28 voxels
    with gpu.World((600, 8, 8)) as w:
        w.fuse = True
        assert not w.fuse  # skipping on: dense fusion not in effect
        w.skip = gpu.SKIP_OFF
        assert w.fuse
