"""Settled-segment skipping: the skipping tick must equal the dense tick (and the
oracle) bit for bit, through host edits, segment changes, tick resets and graph
runs; its self-check counter (settled segments that changed, checked on hash
ticks) must stay 0."""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

# the three schedules: a plan pass plus a persistent pass over the listed row
# quads, one wave per row-quad group skipping on its own, and a plan pass plus a
# persistent pass over the listed lane segments
MODES = pytest.mark.parametrize("mode", [2, 3, 4], ids=["listed", "waves", "segs"])


def _scene(rng, shape, fluid=True):
    nx, ny, nz = shape
    mats = [0, 0, 0, 0, 0, 1, 2, 2, 3, 3] if fluid else [0, 0, 0, 0, 1, 2, 2]
    g = rng.choice(np.array(mats, np.uint8), size=(nz, ny, nx))
    g[0] = 1
    g[rng.random(g.shape) < 0.01] = 200
    return g


@pytest.mark.parametrize("shape,seg", [((64, 40, 48), 8), ((100, 33, 37), 16), ((257, 12, 30), 32),
                                       ((1000, 6, 20), 16), ((3000, 4, 9), 64), ((31, 18, 70), 8)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
@MODES
def test_skip_equals_dense_single_ticks(gpu, O, shape, seg, mode):
    rng = np.random.default_rng(sum(shape) * 7 + seg)
    g = _scene(rng, shape, fluid=shape[0] != 1000)
    t0 = int(rng.integers(0, 1 << 20))
    announce(f"skip vs dense {shape} seg={seg} t0={t0} mode={mode}")
    with gpu.World(shape) as w, gpu.World(shape, flags=gpu.DENSE) as d:
        assert w.skip == gpu.SKIP_AUTO and d.skip == gpu.SKIP_OFF
        w.skip = mode
        for x in (w, d):
            x.segment = seg
            x.upload(g)
            x.tick_count = t0
        w.sweep_stats(True)
        cur = g
        for k in range(48):
            if k % 5 == 4:  # hash ticks read settled segments and self-check them
                assert w.tick() == d.tick()
            else:
                w.step(1)
                d.step(1)
            if k % 8 == 7:
                a, b = w.download(), d.download()
                assert (a == b).all(), (shape, seg, k, int((a != b).sum()))
        cur, _ = O.run(cur, t0, 48, use_lut=True)
        assert (w.download() == cur).all()
        st = w.sweep_stats(False)
    assert st["violations"] == 0
    n = shape[0] * shape[1] * shape[2]
    assert 0 < st["swept"] <= 48 * n


@pytest.mark.parametrize("seg", [8, 16, 32, 64])
@MODES
def test_skip_full_width_rows(gpu, seg, mode):
    # nx = 64 segments: the last lane's mask bit is bit 63 (regression: it was
    # dropped by a left shift, so the last lane of a full row skipped while moving)
    nx, ny, nz = 64 * seg, 24, 60
    rng = np.random.default_rng(seg)
    g = np.zeros((nz, ny, nx), np.uint8)
    g[0] = 1
    for x0 in (0, seg - 2, nx // 2 - 3, nx - seg - 3, nx - 9, nx - 4):
        y0 = int(rng.integers(0, ny - 6))
        g[1:int(rng.integers(20, nz)), y0:y0 + 5, x0:x0 + 4] = 2
    g[1:4, :, nx - seg:] = np.where(rng.random((3, ny, seg)) < 0.3, 3, 0)
    announce(f"skip vs dense, full-width rows nx={nx} seg={seg} mode={mode}")
    with gpu.World((nx, ny, nz)) as w, gpu.World((nx, ny, nz), flags=gpu.DENSE) as d:
        w.skip = mode
        for x in (w, d):
            x.segment = seg
            x.upload(g)
        w.sweep_stats(True)
        for k in range(12):
            assert [int(v) for v in w.run(10, 5)] == [int(v) for v in d.run(10, 5)], k
            if k == 6:
                for x in (w, d):
                    x.set_voxel(nx - 1, 3, nz - 1, 2)
                    x.set_voxel(nx - seg, 4, nz - 2, 3)
        assert (w.download() == d.download()).all()
        assert w.sweep_stats(False)["violations"] == 0


def _edge_scene(rng, shape, seg):
    """A settled base (static floor, a sand bed, static posts) with isolated sand
    grains and fluid drops next to segment and 64-byte sector edges: most
    segments are settled, and the active ones sit beside settled neighbours, so
    on odd ticks the blocks straddling their edges are mostly fixed points."""
    nx, ny, nz = shape
    g = np.zeros((nz, ny, nx), np.uint8)
    g[0] = 1
    g[1:3] = 2
    g[1:nz // 2, ::5, ::97] = 1
    edges = [b + d for b in range(seg, nx, seg) for d in (-2, -1, 0, 1)]
    for x in edges:
        if rng.random() < 0.6:
            y, z = int(rng.integers(0, ny)), int(rng.integers(nz // 3, nz))
            g[z, y, x] = 2 if rng.random() < 0.7 else 3
    # a few busy row quads whose active sectors spill across groups of items
    zb = nz - 3
    g[zb:zb + 2, 0:2, :] = np.where(rng.random((2, min(2, ny), nx)) < 0.4, 2, 0)
    return g


@pytest.mark.parametrize("shape,seg", [((2048, 6, 26), 32), ((1000, 8, 22), 16), ((2000, 5, 30), 32),
                                       ((512, 9, 24), 8)],
                         ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
@MODES
def test_skip_sparse_activity_at_edges(gpu, O, shape, seg, mode):
    # Sparse activity beside settled segments exercises the schedules' settled
    # neighbours: the full-grid tick's keep-input blocks and whole-sector stores,
    # and the packed tick's sector items and edge masks (a straddling block known
    # not to change keeps its input), across hash ticks, edits and tick resets.
    rng = np.random.default_rng(shape[0] + seg)
    g = _edge_scene(rng, shape, seg)
    t0 = int(rng.integers(0, 1 << 16)) | 1  # starts on an odd tick
    announce(f"skip vs dense, sparse edge activity {shape} seg={seg} t0={t0} mode={mode}")
    with gpu.World(shape) as w, gpu.World(shape, flags=gpu.DENSE) as d:
        w.skip = mode
        for x in (w, d):
            x.segment = seg
            x.upload(g)
            x.tick_count = t0
        w.sweep_stats(True)
        cur, t = g.copy(), t0
        for k in range(8):
            assert [int(v) for v in w.run(15, 7)] == [int(v) for v in d.run(15, 7)], k
            cur, _ = O.run(cur, t, 15, use_lut=True)
            t += 15
            if k in (2, 5):
                # edits on both sides of an edge, while both worlds run on
                for x, y, z, v in ((seg, 1, shape[2] - 2, 3), (seg - 1, 2, shape[2] - 3, 2),
                                   (2 * seg + 1, 0, shape[2] - 1, 2)):
                    for wd in (w, d):
                        wd.set_voxel(x, y, z, v)
                    cur[z, y, x] = v
            if k == 6:
                t += 3
                for wd in (w, d):
                    wd.tick_count = t
        a = w.download()
        assert (a == d.download()).all()
        assert (a == cur).all()
        assert w.sweep_stats(False)["violations"] == 0


@MODES
def test_skip_with_host_edits(gpu, O, mode):
    shape = (72, 30, 40)
    rng = np.random.default_rng(11)
    g = _scene(rng, shape)
    announce(f"skip with set_voxel / upload / set_tick / set_segment edits, mode={mode}")
    with gpu.World(shape) as w:
        w.skip = mode
        w.upload(g)
        w.sweep_stats(True)
        cur, t = g.copy(), 0
        for round_ in range(6):
            w.step(12)
            cur, _ = O.run(cur, t, 12, use_lut=True)
            t += 12
            # single-cell edits, including segment edges and the grid border
            for _ in range(25):
                x, y, z = int(rng.integers(0, 72)), int(rng.integers(0, 30)), int(rng.integers(1, 40))
                if _ % 5 == 0:
                    x = int(rng.choice([0, 7, 8, 15, 16, 63, 64, 71]))
                v = int(rng.choice([0, 2, 3, 1, 90]))
                w.set_voxel(x, y, z, v)
                cur[z, y, x] = v
            if round_ == 2:
                w.segment = 16
            if round_ == 3:
                t += 5
                w.tick_count = t
            if round_ == 4:
                cur[5:9, :, 10:20] = 3
                w.upload(cur)
            assert (w.download() == cur).all(), round_
        h = w.tick()
        cur, _ = O.run(cur, t, 1, use_lut=True)
        assert h == O.state_hash(cur)
        assert w.sweep_stats(False)["violations"] == 0


@MODES
def test_skip_graph_runs_match_dense(gpu, mode):
    announce(f"skip vs dense: 400-tick graphs, hashes every 20, box scene 192^3, mode={mode}")
    dims = (192, 192, 192)
    with gpu.World(dims) as w, gpu.World(dims, flags=gpu.DENSE) as d:
        w.skip = mode
        w.fill_boxes(5, side_min=6, side_max=40)
        d.fill_boxes(5, side_min=6, side_max=40)
        w.sweep_stats(True)
        hw = w.run(400, 20)
        hd = d.run(400, 20)
        assert [int(x) for x in hw] == [int(x) for x in hd]
        st = w.sweep_stats(True)
        assert st["violations"] == 0
        # replay of the cached graph continues exactly
        hw = w.run(400, 20)
        hd = d.run(400, 20)
        assert [int(x) for x in hw] == [int(x) for x in hd]
        st2 = w.sweep_stats(False)
        assert st2["violations"] == 0
        n = dims[0] * dims[1] * dims[2]
        # the scene settles: the second run sweeps less than the first
        assert st2["swept"] < st["swept"] < 400 * n
        # switching skipping off and on keeps results exact
        w.skip = gpu.SKIP_OFF
        d.skip = mode
        assert [int(x) for x in w.run(40, 10)] == [int(x) for x in d.run(40, 10)]


@MODES
def test_settled_world_sweeps_nothing(gpu, mode):
    announce(f"a settled world: ticks after settling read and write nothing, mode={mode}")
    dims = (128, 64, 32)
    g = np.zeros(dims[::-1], np.uint8)
    g[0] = 1
    g[1:4] = 2  # a flat sand layer: settled from the start
    with gpu.World(dims) as w:
        w.skip = mode
        w.upload(g)
        w.step(2)  # the first two ticks evaluate every segment
        w.sweep_stats(True)
        w.step(50)
        st = w.sweep_stats(False)
        assert st["swept"] == 0 and st["read_only"] == 0
        assert (w.download() == g).all()
        # one grain dropped in wakes only its neighbourhood
        w.set_voxel(70, 30, 20, 2)
        w.sweep_stats(True)
        w.step(40)
        st = w.sweep_stats(False)
        assert 0 < st["swept"] < 40 * (128 * 64 * 32) // 50
        out = w.download()
        assert out[4, 30, 70] == 2 and int((out == 2).sum()) == int((g == 2).sum()) + 1


@pytest.mark.parametrize("seg", [16, 32])
@MODES
def test_skip_reserved_bytes_on_segment_edges(gpu, O, seg, mode):
    # Reserved bytes (which never move) on both sides of every segment edge,
    # next to sand and fluid that keep moving: the straddling blocks read them
    # through the adjacent lane or from memory (packed schedule), and a lane that
    # stores a neighbour's cell must never overwrite one.
    nx, ny, nz = 64 * seg, 10, 40
    rng = np.random.default_rng(seg * 10 + mode)
    g = np.zeros((nz, ny, nx), np.uint8)
    g[0] = 1
    edges = np.arange(seg, nx, seg)
    for x in edges[::3]:
        g[1:nz // 2, :, x - 1] = 200
        g[1:nz // 2, ::2, x] = 77
    g[nz // 2:] = np.where(rng.random((nz - nz // 2, ny, nx)) < 0.15, rng.choice([2, 3], size=(nz - nz // 2, ny, nx)), 0)
    announce(f"reserved bytes on segment edges, nx={nx} seg={seg} mode={mode}")
    with gpu.World((nx, ny, nz)) as w, gpu.World((nx, ny, nz), flags=gpu.DENSE) as d:
        w.skip = mode
        for x in (w, d):
            x.segment = seg
            x.upload(g)
        w.sweep_stats(True)
        assert [int(v) for v in w.run(60, 6)] == [int(v) for v in d.run(60, 6)]
        out = w.download()
        assert (out == d.download()).all()
        want, _ = O.run(g, 0, 60, use_lut=True)
        assert (out == want).all()
        assert ((out == 200) == (g == 200)).all() and ((out == 77) == (g == 77)).all()
        assert w.sweep_stats(False)["violations"] == 0


@MODES
def test_skip_from_one_buffer_into_the_other(gpu, mode):
    # Grids above 2^27 bytes keep two buffers (smaller ones tick in place): the
    # full-grid and listed schedules then rely on the destination holding the
    # bytes of two ticks before.  1024 x 520 x 260 is 138 MB, just above the limit.
    shape = (1024, 520, 260)
    assert shape[0] * shape[1] * shape[2] > 1 << 27
    announce(f"skip vs dense from one buffer into the other {shape} mode={mode}")
    rng = np.random.default_rng(mode)
    with gpu.World(shape) as w, gpu.World(shape, flags=gpu.DENSE) as d:
        w.skip = mode
        d.fuse = False
        for x in (w, d):
            x.fill_boxes(3)
        w.sweep_stats(True)
        for k in range(4):
            assert [int(v) for v in w.run(12, 4)] == [int(v) for v in d.run(12, 4)], k
            if k == 1:
                for _ in range(20):
                    x, y, z = int(rng.integers(0, shape[0])), int(rng.integers(0, shape[1])), int(rng.integers(1, shape[2]))
                    for wd in (w, d):
                        wd.set_voxel(x, y, z, 2)
        assert w.hash() == d.hash()
        assert w.sweep_stats(False)["violations"] == 0
