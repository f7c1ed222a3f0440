"""Serving loop on the GPU: every request's final hash is the same alone and in
the mix, equals the oracle, and admission respects the HBM budget."""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def requests(n, seed, lo, hi, tlo, thi):
    rng = np.random.default_rng(seed)
    dims = rng.integers(lo, hi + 1, size=(n, 3)).astype(np.int32)
    seeds = rng.integers(0, 1 << 62, size=n, dtype=np.uint64)
    thr = rng.integers(1 << 60, 1 << 63, size=n, dtype=np.uint64)
    ticks = rng.integers(tlo, thi + 1, size=n).astype(np.int32)
    return dims, seeds, thr, ticks


def alone_hash(gpu, dims, seed, thr, ticks):
    with gpu.World(tuple(int(v) for v in dims)) as w:
        w.fill_rain(int(seed), int(thr))
        w.step(int(ticks))
        return w.hash()


@pytest.mark.parametrize("round_ticks", [1, 8, 64])
def test_mixed_results_match_oracle_and_alone(gpu, O, round_ticks):
    dims, seeds, thr, ticks = requests(24, 7 + round_ticks, 3, 60, 1, 40)
    announce(f"serve 24 mixed requests round_ticks={round_ticks}")
    with gpu.Server(budget_bytes=1 << 30, round_ticks=round_ticks, max_active=64) as s:
        ids = [s.submit(dims[i], seeds[i], thr[i], ticks[i]) for i in range(24)]
        s.drain()
        res = [s.result(i) for i in ids]
        st = s.stats()
    assert all(r[0] == 2 for r in res)
    assert st["completed"] == 24 and st["used_bytes"] == 0
    for i in range(24):
        g = O.scene_rain(tuple(int(v) for v in dims[i]), int(seeds[i]), int(thr[i]))
        want, _ = O.run(g, 0, int(ticks[i]))
        assert res[i][1] == O.state_hash(want), i
        if i % 6 == 0:
            assert res[i][1] == alone_hash(gpu, dims[i], seeds[i], thr[i], ticks[i])


def test_budget_admission_and_rejection(gpu, O):
    # grids of 40^3 = 64000 cells round to a 64 KiB class (plus guard); a budget of
    # 200 KiB admits at most three at a time; a 100^3 request can never fit.
    n = 9
    dims = np.array([[40, 40, 40]] * n + [[100, 100, 100]], np.int32)
    seeds = np.arange(1, n + 2, dtype=np.uint64)
    thr = np.full(n + 1, O.threshold(1, 4), np.uint64)
    ticks = np.array([5, 9, 13, 3, 7, 11, 2, 6, 10, 4], np.int32)
    announce("serve budget: 9 admissible requests under 200 KiB, one oversize")
    with gpu.Server(budget_bytes=200 << 10, round_ticks=4, max_active=16) as s:
        ids = [s.submit(dims[i], seeds[i], thr[i], ticks[i]) for i in range(n + 1)]
        s.drain()
        st = s.stats()
        res = [s.result(i) for i in ids]
    assert res[-1][0] == 3 and st["rejected"] == 1
    assert st["peak_bytes"] <= 200 << 10 and st["peak_active"] <= 3
    for i in range(n):
        assert res[i][0] == 2
        g = O.scene_rain((40, 40, 40), int(seeds[i]), int(thr[i]))
        want, _ = O.run(g, 0, int(ticks[i]))
        assert res[i][1] == O.state_hash(want)


def test_replay_poisson_stream(gpu):
    dims, thr, seeds, ticks, arr = gpu.serve_specs(400, 11, 16, 64, 10, 80, 4000.0)
    announce("serve replay: 400 requests, Poisson 4000/s, edges 16-64, 10-80 ticks")
    with gpu.Server(budget_bytes=1 << 30, round_ticks=8, max_active=1024) as s:
        ids = s.replay(dims, seeds, thr, ticks, arr)
        st = s.stats()
        res = [s.result(int(i)) for i in ids]
    assert st["completed"] == 400 and all(r[0] == 2 for r in res)
    assert st["tick_p50_ms"] > 0 and st["tick_p99_ms"] >= st["tick_p50_ms"]
    assert st["req_p99_ms"] >= st["req_p50_ms"] > 0
    for i in range(0, 400, 37):
        assert res[i][1] == alone_hash(gpu, dims[i], seeds[i], thr[i], ticks[i]), i


def test_server_errors(gpu):
    with pytest.raises(gpu.Fs3dError):
        gpu.Server(budget_bytes=0)
    with gpu.Server(budget_bytes=1 << 20, round_ticks=2, max_active=4) as s:
        with pytest.raises(gpu.Fs3dError):
            s.submit((8, 8, 8), 1, 1 << 62, 0)
        with pytest.raises(gpu.Fs3dError):
            s.submit((600, 8, 8), 1, 1 << 62, 3)
        with pytest.raises(gpu.Fs3dError):
            s.result(5)
