"""fs3d benchmark: voxel-updates/s of the tick on MI355X, against the HBM roofline.

Default run (one JSON line on stdout):
  value    the 512^3 box scene (20% sand, 10% fluid, 5% static boxes over a floor),
           K ticks in one captured HIP graph with a fused state hash every 50
           ticks, after W warm-up ticks, with the library's default tick (settled
           segments skipped, results bit-identical to the dense tick).  A "step" is
           one tick of the whole world; value counts every voxel of every tick.
  roofline the timed region: algorithmic bytes actually moved (2 per swept
           voxel-update, 1 per voxel only read on hash ticks; counted by replaying
           the same ticks with sweep counters) / device time of the timed graph
           (HIP events on the world's stream), against 8.0 TB/s; "dense" holds the
           dense tick kernel on the same world (2 bytes per voxel-update).
  extra    pillars2048: the 2048^3 pillar scene, skipping, dense and dense fused
           in pairs (the dense single-tick kernel is the HBM-bound figure: the
           512^3 pair of buffers nearly fits the 256 MB Infinity Cache);
           pillars2048_render: that scene with a
           1920x1080 voxel render every 10 ticks; copy: a plain device copy of the
           same bytes, the practical bandwidth ceiling; batched: 4096 worlds against
           the one-launch-per-world loop; serving: the continuous-batching server
           under a Poisson load.
  cpu_baseline  the oracle's tick on this host, one thread, a bounded sample; its
           "openmp" field times the same tick over the host threads the run may use.

--only {boxes512,pillars2048,copy,batched,serving,render2048} runs one workload
(for profiler runs) and prints its own JSON line.
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
# Bytes per tick of the timed region from rocprofv3 counter runs of this same
# command (profiles/tick_kernel_boxes512.md): FETCH_SIZE x 2 + WRITE_SIZE x 1 KiB
# per dispatch (factors from a copy of known size), averaged over its skipping
# and hashing ticks.
PMC_TRAFFIC_BOXES512 = 171.2e6
METRIC = "Gvoxel-updates/s + %HBM roofline at 512^3/2048^3; batched vs plain; p99 tick ms"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def roofline(vox, kern_ms):
    ach = BYTES_PER_UPDATE * vox / (kern_ms * 1e-3)
    return ach, ach / HBM_PEAK


def bench_boxes512(F, steps, warmup, kernel_ticks=200):
    """The 512^3 box scene with the library's default tick (settled segments are
    skipped): warm-up ticks, then the timed graph of `steps` ticks with a fused
    hash every 50.  The same tick sequence is replayed with sweep counters on to
    get the algorithmic bytes of the timed region (skipping is deterministic), and
    a dense world runs the plain tick kernel for the dense roofline."""
    n = 512
    with F.World((n, n, n)) as w:
        t0 = time.perf_counter()
        w.fill_boxes(1)
        log(f"[bench] boxes512: scene generated in {time.perf_counter() - t0:.2f}s")
        if warmup:
            w.run(warmup, 0)
        w.prepare_run(steps, 50)  # graph capture is set-up, not a tick
        w.sync()
        t0 = time.perf_counter()
        hashes = w.run(steps, 50)
        wall = time.perf_counter() - t0
        dev_ms = w.last_run_ms_per_tick()
        log(f"[bench] boxes512: {steps} ticks in {wall * 1e3:.2f} ms ({dev_ms * 1e3:.2f} us/tick on the device)")
        # replay with counters
        w.fill_boxes(1)
        w.tick_count = 0
        if warmup:
            w.run(warmup, 0)
        w.sweep_stats(True)
        h2 = w.run(steps, 50)
        st = w.sweep_stats(False)
        if [int(x) for x in h2] != [int(x) for x in hashes]:
            raise SystemExit("boxes512: counter replay diverged from the timed run")
        mode, seg = w.skip, w.segment
    with F.World((n, n, n), flags=F.DENSE) as d:
        d.fill_boxes(1)
        d.run(kernel_ticks, 0)
        d.run(kernel_ticks, 0)
        kern_ms = d.last_run_ms_per_tick()
    return dict(vox=n ** 3, wall=wall, dev_ms=dev_ms, kern_ms=kern_ms, hashes=[int(h) for h in hashes], seg=seg,
                swept=st["swept"], read_only=st["read_only"], violations=st["violations"], mode=mode)


def bench_pillars2048(F, ticks):
    """The 2048^3 pillar scene: `ticks` plain ticks in one graph with the default
    (skipping) tick, with the dense single-tick kernel (the HBM roofline figure)
    and with dense ticks fused in pairs (two ticks per pass over the grid)."""
    n = 2048
    out = {}
    for kind in ("skip", "dense", "fused"):
        with F.World((n, n, n), flags=0 if kind == "skip" else F.DENSE) as w:
            if kind != "skip":
                w.fuse = kind == "fused"
                if w.fuse != (kind == "fused"):
                    raise SystemExit("pillars2048: fusion not available at 2048^3")
            t0 = time.perf_counter()
            w.fill_pillars(2)
            log(f"[bench] pillars2048: scene generated in {time.perf_counter() - t0:.2f}s")
            w.run(4, 0)
            w.fill_pillars(2)
            w.tick_count = 0
            if kind == "skip":
                w.sweep_stats(True)
            w.prepare_run(ticks, 0)
            w.sync()
            t0 = time.perf_counter()
            w.run(ticks, 0)
            wall = time.perf_counter() - t0
            kern_ms = w.last_run_ms_per_tick()
            st = w.sweep_stats(False) if kind == "skip" else None
            h = w.hash()
            log(f"[bench] pillars2048 {kind}: {ticks} ticks in {wall * 1e3:.1f} ms ({kern_ms:.3f} ms/tick)")
            out[kind] = dict(wall=wall, kern_ms=kern_ms, hash=int(h), swept=st["swept"] if st else None,
                             seg=w.segment)
    if not out["dense"]["hash"] == out["skip"]["hash"] == out["fused"]["hash"]:
        raise SystemExit("pillars2048: skipping, dense and fused ticks disagree")
    return dict(vox=n ** 3, ticks=ticks, **out)


def orbit_angles(seed, n):
    """Seeded orbit-camera angles: 2*pi * (SplitMix64 output >> 11) / 2^53."""
    import math
    st, out = seed, []
    for _ in range(n):
        st = (st + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        z ^= z >> 31
        out.append(2.0 * math.pi * (z >> 11) / float(1 << 53))
    return out


def bench_pillars_render(F, ticks=200, every=10, W=1920, H=1080, seed=2):
    """The 2048^3 collapsing-column scene: `ticks` ticks in graphs of `every`,
    a W x H voxel render after each, from a seeded orbit camera."""
    from fallingsand3d_amd import render as R
    n = 2048
    frames, tick_ms, render_ms, steps = [], [], [], 0
    angles = orbit_angles(seed, ticks // every)
    with F.World((n, n, n)) as w:
        w.fill_pillars(seed)
        # a first render turns on occupancy tracking (this grid is above 2^29
        # voxels), then the graph is captured with its changed-brick marking
        R.render(w, R.orbit_camera(w.dims, angles[0], 1.4, 0.35, W / H), W, H)
        w.run(every, 0)
        w.fill_pillars(seed)
        w.tick_count = 0
        t0 = time.perf_counter()
        for k, ang in enumerate(angles):
            w.run(every, 0)
            tick_ms.append(w.last_run_ms_per_tick())
            cam = R.orbit_camera(w.dims, ang, 1.4, 0.35, W / H)
            rgb, hit, face = R.render(w, cam, W, H)
            render_ms.append(R.last_render_ms(w))
            frames.append(float((hit != R.MISS).mean()))
            t1 = time.perf_counter()
            steps += voxel_steps(cam, w.dims, hit, W, H)
            t0 += time.perf_counter() - t1  # host-side counting is not part of the loop's wall time
        wall = time.perf_counter() - t0
        h = w.hash()
    return dict(wall=wall, tick_ms=sum(tick_ms) / len(tick_ms), render_ms=sum(render_ms) / len(render_ms),
                render_ms_max=max(render_ms), hit_frac=sum(frames) / len(frames), frames=len(frames), hash=int(h),
                vox=n ** 3, ticks=ticks, rays=W * H, voxel_steps=steps / len(frames))


def voxel_steps(cam, dims, hit, W, H):
    """Cells a frame's rays walked through: per ray, the sum over axes of the
    distance between its first cell and its last one (the hit cell, or the cell
    where it leaves the grid), plus one; each axis only moves one way in a
    walk.  Computed on the host in fp32 from the camera, so a ray's count can
    be off by a cell where its entry or exit point lies on a cell face."""
    import numpy as np
    nx, ny, nz = dims
    f32 = np.float32
    px = np.arange(W, dtype=f32)[None, :]
    py = np.arange(H, dtype=f32)[:, None]
    ux = ((f32(W) - px) - f32(0.5)) / f32(W) * f32(2) - f32(1)
    uy = (((py + f32(0.5)) / f32(H)) * f32(2) - f32(1)) * (f32(1) / f32(cam.aspect))
    ux, uy = np.broadcast_to(ux, (H, W)), np.broadcast_to(uy, (H, W))
    b = np.array(cam.basis[:], f32).reshape(3, 3)
    d = ux[..., None] * b[0] + uy[..., None] * b[1] + b[2]
    d /= np.sqrt((d * d).sum(-1, keepdims=True))
    o = np.array(cam.origin[:], f32)
    n = np.array([nx, ny, nz], f32)
    with np.errstate(divide="ignore", invalid="ignore"):
        t1, t2 = (0 - o) / d, (n - o) / d
    lo = np.where(d == 0, -np.inf, np.minimum(t1, t2)).max(-1)
    hi = np.where(d == 0, np.inf, np.maximum(t1, t2)).min(-1)
    tmin = np.maximum(lo, 0)
    inside = (tmin <= hi) & ~((d == 0) & ((o < 0) | (o > n))).any(-1)
    ni = np.array([nx, ny, nz])
    start = np.clip(np.floor(o + tmin[..., None] * d), 0, ni - 1).astype(np.int64)
    end = np.clip(np.floor(o + hi[..., None] * d), 0, ni - 1).astype(np.int64)
    hit = np.asarray(hit).reshape(H, W)
    h = hit != np.iinfo(np.uint64).max
    li = hit[h].astype(np.int64)
    end[h] = np.stack([li % nx, (li // nx) % ny, li // (nx * ny)], -1)
    steps = np.abs(end - start).sum(-1) + 1
    return int(steps[inside].sum())


def bench_copy(F, nbytes, iters=20):
    ms = F.copy_probe(nbytes, iters)
    log(f"[bench] copy: {nbytes} bytes in {ms:.4f} ms ({2 * nbytes / ms / 1e6:.1f} GB/s)")
    return dict(nbytes=nbytes, ms=ms, GBps=2 * nbytes / (ms * 1e-3) / 1e9)


def bench_batched(F, nworlds=4096, ticks=500, plain_ticks=20, seed=3):
    """4096 worlds, edges log-uniform in [16, 128], rain fill with per-world
    density in [0.05, 0.5): all worlds step `ticks` ticks in one persistent
    launch (shape chosen by the autotuner) against the plain loop of one launch
    per world per tick.  The plain loop is timed on its first `plain_ticks`
    ticks (2 x 4096 x 20 launches would not change its per-tick cost) and its
    hashes are compared with the persistent kernel's over the same ticks."""
    dims, thr, seeds = F.batch_specs(nworlds, seed, 16, 128)
    vox = int(dims.astype("int64").prod(1).sum())
    with F.Batch(dims) as b:
        b.fill_rain(seeds, thr)
        tune = b.autotune(trial_ticks=10)
        log(f"[bench] batched: autotune -> {tune}")
        h_sample = b.run(plain_ticks)
        with F.Batch(dims) as p:
            p.fill_rain(seeds, thr)
            t0 = time.perf_counter()
            h_plain = p.run(plain_ticks, plain=True)
            plain_wall = time.perf_counter() - t0
        exact = bool((h_sample == h_plain).all())
        log(f"[bench] batched: plain loop {plain_ticks} ticks in {plain_wall:.3f}s, bit-exact={exact}")
        b.tick_count = 0
        b.fill_rain(seeds, thr)
        b.run(ticks)  # graph capture + warm caches
        b.tick_count = 0
        b.fill_rain(seeds, thr)
        t0 = time.perf_counter()
        h = b.run(ticks)
        wall = time.perf_counter() - t0
        dev_ms = b.last_run_ms()
        log(f"[bench] batched: persistent {ticks} ticks in {wall * 1e3:.1f} ms (device {dev_ms:.1f} ms)")
        # the same run with every tick of every world executed (settled worlds too)
        b.skip = False
        b.tick_count = 0
        b.fill_rain(seeds, thr)
        b.run(ticks)
        b.tick_count = 0
        b.fill_rain(seeds, thr)
        t0 = time.perf_counter()
        h_all = b.run(ticks)
        wall_all = time.perf_counter() - t0
        b.skip = True
        exact = exact and bool((h_all == h).all())
        log(f"[bench] batched: every tick executed: {wall_all * 1e3:.1f} ms, same hashes={bool((h_all == h).all())}")
        digest = 0
        for v in h:
            digest = (digest + int(v)) % (1 << 64)
        plain_per_tick = plain_wall / plain_ticks
        return dict(nworlds=nworlds, vox=vox, ticks=ticks, wall=wall, dev_ms=dev_ms,
                    gvox_per_s=vox * ticks / wall / 1e9, plain_ms_per_tick=plain_per_tick * 1e3,
                    plain_gvox_per_s=vox / plain_per_tick / 1e9,
                    speedup=(plain_per_tick * ticks) / wall, bit_exact_sample=exact,
                    all_ticks_wall=wall_all, all_ticks_gvox_per_s=vox * ticks / wall_all / 1e9,
                    all_ticks_speedup=(plain_per_tick * ticks) / wall_all,
                    tuned={"mode": {0: "persistent", 2: "wide"}.get(tune["mode"], tune["mode"]), "seg": tune["seg"],
                           "threads": tune["threads"], "groups": tune["groups"]}, hash_digest=digest)


def bench_serving(F, seconds=3.0, lam=2000.0, seed=5, round_ticks=8, check=8):
    """Poisson stream of requests at `lam` per second for `seconds` seconds of
    arrivals (edges log-uniform 32-256, ticks uniform 10-500, rain fill with
    density U(0.05, 0.5)) through the continuous-batching server under a
    200 GiB budget, replayed in real time; `check` requests spread over the
    stream are re-run alone and their final hashes compared."""
    n = int(seconds * lam)
    dims, thr, seeds, ticks, arr = F.serve_specs(n, seed, 32, 256, 10, 500, lam)
    with F.Server(budget_bytes=200 << 30, round_ticks=round_ticks, max_active=1 << 16) as s:
        t0 = time.perf_counter()
        ids = s.replay(dims, seeds, thr, ticks, arr)
        wall = time.perf_counter() - t0
        st = s.stats()
        res = [s.result(int(i)) for i in ids]
    exact = True
    for i in range(0, n, max(1, n // check)):
        with F.World(tuple(int(v) for v in dims[i])) as w:
            w.fill_rain(int(seeds[i]), int(thr[i]))
            w.step(int(ticks[i]))
            exact = exact and (w.hash() == res[i][1]) and res[i][0] == 2
    log(f"[bench] serving: {n} requests in {wall:.2f}s, tick p50 {st['tick_p50_ms']:.3f} ms p99 "
        f"{st['tick_p99_ms']:.3f} ms, request p50 {st['req_p50_ms']:.1f} ms p99 {st['req_p99_ms']:.1f} ms, "
        f"bit-exact={exact}")
    return dict(n=n, wall=wall, stats=st, exact=exact,
                gvox_per_s=st["voxel_updates"] / wall / 1e9, busy_frac=st["busy_ms"] / (wall * 1e3))


def cpu_baseline(seconds=12.0, omp_seconds=6.0):
    """The oracle's tick (table-driven) on the 512^3 box scene: one thread for
    about `seconds` of host work, then OpenMP over the host threads this run may
    use (OMP_NUM_THREADS, else the CPU count) for about `omp_seconds`; then two
    ticks of the 2048^3 pillar scene with OpenMP (sampled: every voxel is
    updated, as the oracle does not skip)."""
    from oracle import oracle as O

    def timed(threads, budget):
        g, _ = O.scene_boxes((512, 512, 512), 1)
        O.set_threads(threads)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < budget:
            g, _ = O.run(g, k, 1, 0, use_lut=True)
            k += 1
        return k, time.perf_counter() - t0

    k, dt = timed(1, seconds)
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    k2, dt2 = timed(threads, omp_seconds)
    # 2048^3: two ticks (one of each parity) of the pillar scene, sampled
    from fallingsand3d_amd.world import PILLARS_2048
    O.set_threads(threads)
    g = O.scene_pillars((2048, 2048, 2048), 2, PILLARS_2048)
    t0 = time.perf_counter()
    g, _ = O.run(g, 0, 2, 0, use_lut=True)
    dt3 = time.perf_counter() - t0
    del g
    O.set_threads(0)
    return dict(value=round(k * 512 ** 3 / dt / 1e9, 4), unit="Gvoxel-updates/s", cores=1, kind="port",
                sample=f"{k} ticks of the 512^3 box scene (oracle table-driven tick, 1 thread, {dt:.1f} s)",
                openmp=dict(value=round(k2 * 512 ** 3 / dt2 / 1e9, 4), cores=threads,
                            sample=f"{k2} ticks of the same scene, OpenMP over {threads} threads, {dt2:.1f} s"),
                pillars2048_sampled=dict(ms_per_tick=round(dt3 / 2 * 1e3, 1),
                                         value=round(2 * 2048 ** 3 / dt3 / 1e9, 4), cores=threads,
                                         sample="2 ticks (one of each parity) of the 2048^3 pillar scene, OpenMP over "
                                                f"{threads} threads, {dt3:.1f} s (sampled)"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--only", choices=["boxes512", "pillars2048", "copy", "batched", "serving", "render2048"],
                    default=None)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every world lives on one device and nothing here shards "
              "across GPUs (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import fallingsand3d_amd as F
    if F.device_count() < 1:
        raise SystemExit("no GPU visible")
    F.init(0)

    if args.only == "pillars2048":
        p = bench_pillars2048(F, args.steps)
        print(json.dumps({"workload": "pillars2048", "dense_kernel_ms": p["dense"]["kern_ms"],
                          "fused_ms_per_tick": p["fused"]["kern_ms"], "skip_ms_per_tick": p["skip"]["kern_ms"],
                          "hash": p["dense"]["hash"]}), flush=True)
        return
    if args.only == "copy":
        c = bench_copy(F, 8 << 30)
        print(json.dumps({"workload": "copy", **c}), flush=True)
        return
    if args.only == "batched":
        bb = bench_batched(F, ticks=args.steps)
        print(json.dumps({"workload": "batched", **bb}), flush=True)
        return
    if args.only == "render2048":
        pr = bench_pillars_render(F)
        print(json.dumps({"workload": "render2048", **pr}), flush=True)
        return
    if args.only == "serving":
        sv = bench_serving(F)
        print(json.dumps({"workload": "serving", **{k: v for k, v in sv.items() if k != "stats"},
                          **sv["stats"]}), flush=True)
        return

    r = bench_boxes512(F, args.steps, args.warmup)
    vox = r["vox"]
    value = vox * args.steps / r["wall"] / 1e9
    # the timed region: algorithmic bytes actually moved (2 per swept voxel-update,
    # 1 per voxel only read on hash ticks) over its device time (HIP events)
    alg_bytes = 2 * r["swept"] + r["read_only"]
    ach = alg_bytes / (r["dev_ms"] * 1e-3 * args.steps)
    frac = ach / HBM_PEAK
    ach_d, frac_d = roofline(vox, r["kern_ms"])
    sched = {1: "auto", 2: "listed", 3: "waves"}.get(r["mode"], str(r["mode"]))
    out = {
        "metric": METRIC,
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
        "roofline": {"bound": "hbm", "achieved": round(ach / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(frac, 4), "traffic": PMC_TRAFFIC_BOXES512,
                     "traffic_source": "profiles/tick_kernel_boxes512.md (rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                                       "bytes per tick)",
                     "kernel": f"tick_kernel<{r['seg']},*,skip> ({sched} schedule), every tick of the timed region",
                     "ms_per_tick_device": round(r["dev_ms"], 5),
                     "alg_bytes_per_tick": int(alg_bytes / args.steps),
                     "swept_fraction": round(r["swept"] / (vox * args.steps), 4),
                     "skip_violations": r["violations"],
                     "dense": {"kernel": f"tick_kernel<{r['seg']},false> (every voxel, no skipping)",
                               "kernel_ms": round(r["kern_ms"], 5), "achieved": round(ach_d / 1e9, 1),
                               "frac": round(frac_d, 4)}},
        "final_hash": r["hashes"][-1] if r["hashes"] else None,
    }
    if args.only is None:
        extra = {}
        p = bench_pillars2048(F, 200)
        ach2, frac2 = roofline(p["vox"], p["dense"]["kern_ms"])
        c = bench_copy(F, 8 << 30)
        pr = bench_pillars_render(F)
        extra["pillars2048_render"] = {
            "workload": "2048^3 pillars, 200 ticks in graphs of 10, 1920x1080 voxel render after each (20 frames), "
                        "seeded orbit camera",
            "wall_s": round(pr["wall"], 3), "tick_ms": round(pr["tick_ms"], 4),
            "render_ms_mean": round(pr["render_ms"], 3), "render_ms_max": round(pr["render_ms_max"], 3),
            "rays_per_s": round(pr["rays"] / (pr["render_ms"] * 1e-3), 1), "hit_fraction": round(pr["hit_frac"], 4),
            "voxel_steps_per_frame": round(pr["voxel_steps"]),
            "voxel_steps_per_s": round(pr["voxel_steps"] / (pr["render_ms"] * 1e-3), 1),
            "final_hash": pr["hash"]}
        ps, pd = p["skip"], p["dense"]
        extra["pillars2048"] = {
            "workload": "2048^3 uint8, 256 seeded sand pillars in a fluid pool, 200 plain ticks in one graph",
            "gvox_per_s": round(p["vox"] * p["ticks"] / ps["wall"] / 1e9, 2),
            "ms_per_tick": round(ps["kern_ms"], 4), "swept_fraction": round(ps["swept"] / (p["vox"] * p["ticks"]), 4),
            "dense_gvox_per_s": round(p["vox"] * p["ticks"] / pd["wall"] / 1e9, 2),
            "dense_kernel_ms": round(pd["kern_ms"], 4), "dense_achieved_GBps": round(ach2 / 1e9, 1),
            "dense_hbm_frac": round(frac2, 4), "dense_frac_of_copy": round(ach2 / 1e9 / c["GBps"], 4),
            "fused_pairs": {"kernel": "tick2_kernel<32> (two dense ticks per pass, first tick's rows in LDS)",
                            "gvox_per_s": round(p["vox"] * p["ticks"] / p["fused"]["wall"] / 1e9, 2),
                            "ms_per_tick": round(p["fused"]["kern_ms"], 4),
                            "achieved_GBps": round(p["vox"] * 2 / (2 * p["fused"]["kern_ms"] * 1e-3) / 1e9, 1),
                            "note": "achieved counts the 2 bytes per voxel that one pass moves for two ticks"},
            "final_hash": pd["hash"], "segment_bytes": pd["seg"]}
        extra["copy"] = {"bytes": c["nbytes"], "ms": round(c["ms"], 4), "GBps": round(c["GBps"], 1),
                         "hbm_frac": round(c["GBps"] * 1e9 / HBM_PEAK, 4)}
        bb = bench_batched(F)
        extra["batched"] = {
            "workload": "4096 worlds, edges log-uniform 16-128, rain fill, density U(0.05,0.5), 500 ticks",
            "voxels": bb["vox"], "gvox_per_s": round(bb["gvox_per_s"], 2), "wall_ms": round(bb["wall"] * 1e3, 2),
            "device_ms": round(bb["dev_ms"], 2), "plain_gvox_per_s": round(bb["plain_gvox_per_s"], 3),
            "plain_ms_per_tick_sampled": round(bb["plain_ms_per_tick"], 3),
            "speedup_vs_plain": round(bb["speedup"], 2), "bit_exact_sample": bb["bit_exact_sample"],
            "every_tick_executed": {"gvox_per_s": round(bb["all_ticks_gvox_per_s"], 2),
                                    "wall_ms": round(bb["all_ticks_wall"] * 1e3, 2),
                                    "speedup_vs_plain": round(bb["all_ticks_speedup"], 2)},
            "settled_worlds": "a world with two quiet ticks in a row is a fixed point; its remaining ticks "
                              "are not run (only its final, hashed tick)",
            "autotuned": bb["tuned"], "final_hash_digest": bb["hash_digest"]}
        sv = bench_serving(F)
        st = sv["stats"]
        extra["serving"] = {
            "workload": "Poisson 2000 req/s for 3 s, edges log-uniform 32-256, ticks U{10..500}, rain fill, "
                        "200 GiB budget, rounds of 8 ticks",
            "requests": sv["n"], "wall_s": round(sv["wall"], 3), "gvox_per_s": round(sv["gvox_per_s"], 2),
            "tick_p50_ms": round(st["tick_p50_ms"], 4), "tick_p99_ms": round(st["tick_p99_ms"], 4),
            "request_p50_ms": round(st["req_p50_ms"], 2), "request_p99_ms": round(st["req_p99_ms"], 2),
            "device_busy_frac": round(sv["busy_frac"], 3), "peak_active": st["peak_active"],
            "peak_bytes": st["peak_bytes"], "rejected": st["rejected"], "bit_exact_sample": sv["exact"]}
        out["extra"] = extra
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
