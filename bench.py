"""fs3d benchmark: voxel-updates/s of the tick on MI355X, against the HBM roofline.

Default workload (one JSON line on stdout): the 512^3 box scene (20% sand, 10%
fluid, 5% static over a floor), K ticks in one captured HIP graph with a fused
state hash every 50 ticks, after W warm-up ticks.  A "step" is one tick of the
whole world.  Extra measurements ride in the same line under "extra": the 2048^3
pillar scene's dense tick.

Roofline: the tick moves 2 algorithmic bytes per voxel-update (1 read + 1 write);
achieved = 2 * voxels / mean tick-kernel time, measured with HIP events on the
world's stream over a graph of plain ticks; peak = 8.0e12 B/s (MI355X HBM3E).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
BYTES_PER_UPDATE = 2


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def tick_kernel_ms(F, w, nticks):
    """Mean device time per plain tick over a graph of nticks ticks (HIP events)."""
    w.run(nticks, 0)
    return w.last_run_ms_per_tick()


def bench_boxes512(F, steps, warmup):
    n = 512
    with F.World((n, n, n)) as w:
        t0 = time.perf_counter()
        w.fill_boxes(1)
        log(f"[bench] 512^3 box scene generated in {time.perf_counter() - t0:.2f}s")
        w.run(max(warmup, 1), 0)
        w.sync()
        t0 = time.perf_counter()
        hashes = w.run(steps, 50)
        wall = time.perf_counter() - t0
        graph_ms = w.last_run_ms_per_tick()
        kern_ms = tick_kernel_ms(F, w, 200)
        vox = n ** 3
        return dict(vox=vox, wall=wall, graph_ms=graph_ms, kern_ms=kern_ms, hashes=[int(h) for h in hashes],
                    seg=w.segment)


def bench_pillars2048(F, ticks):
    n = 2048
    with F.World((n, n, n)) as w:
        t0 = time.perf_counter()
        w.fill_pillars(2)
        log(f"[bench] 2048^3 pillar scene generated in {time.perf_counter() - t0:.2f}s")
        w.run(4, 0)
        t0 = time.perf_counter()
        w.run(ticks, 0)
        wall = time.perf_counter() - t0
        kern_ms = w.last_run_ms_per_tick()
        h = w.hash()
        return dict(vox=n ** 3, wall=wall, kern_ms=kern_ms, ticks=ticks, hash=int(h), seg=w.segment)


def cpu_baseline(seconds=12.0):
    """The oracle's tick (single thread, table-driven) on a 512^3 box scene,
    timed for about `seconds` of host work."""
    from oracle import oracle as O
    g, _ = O.scene_boxes((512, 512, 512), 1)
    O.set_threads(1)
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < seconds:
        g, _ = O.run(g, k, 1, 0, use_lut=True)
        k += 1
    dt = time.perf_counter() - t0
    O.set_threads(0)
    return dict(value=k * 512 ** 3 / dt / 1e9, unit="Gvoxel-updates/s", cores=1, kind="port",
                sample=f"{k} ticks of the 512^3 box scene, oracle table-driven tick, 1 thread, {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every world lives on one device and nothing here "
              "shards across GPUs (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import fallingsand3d_amd as F
    if F.device_count() < 1:
        raise SystemExit("no GPU visible")
    F.init(0)

    r = bench_boxes512(F, args.steps, args.warmup)
    vox = r["vox"]
    value = vox * args.steps / r["wall"] / 1e9
    achieved = BYTES_PER_UPDATE * vox / (r["kern_ms"] * 1e-3) / 1e9
    out = {
        "metric": "Gvoxel-updates/s + %HBM roofline at 512^3/2048^3; batched vs plain; p99 tick ms",
        "value": round(value, 3),
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded box scene generated in-process)",
        "config": {"workload": "boxes512: 512^3 uint8 world, seeded boxes (20% sand, 10% fluid, 5% static) "
                               "over a floor; ticks in one HIP graph with a fused state hash every 50 ticks",
                   "dims": [512, 512, 512], "hash_every": 50, "segment_bytes": r["seg"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved * 1e9 / HBM_PEAK, 4), "traffic": None,
                     "kernel": f"tick_kernel<{r['seg']},false>", "kernel_ms": round(r["kern_ms"], 5)},
        "final_hash": r["hashes"][-1] if r["hashes"] else None,
    }
    extra = {}
    if not args.no_extra:
        p = bench_pillars2048(F, 200)
        ach = BYTES_PER_UPDATE * p["vox"] / (p["kern_ms"] * 1e-3) / 1e9
        extra["pillars2048"] = {
            "workload": "2048^3 uint8, 256 seeded sand pillars in a fluid pool, 200 plain ticks in one graph",
            "gvox_per_s": round(p["vox"] * p["ticks"] / p["wall"] / 1e9, 2),
            "kernel_ms": round(p["kern_ms"], 4), "achieved_GBps": round(ach, 1),
            "hbm_frac": round(ach * 1e9 / HBM_PEAK, 4), "final_hash": p["hash"], "segment_bytes": p["seg"]}
    out["extra"] = extra
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
