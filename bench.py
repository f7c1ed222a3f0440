"""fs3d benchmark: voxel-updates/s of the tick on MI355X, against the HBM roofline.

Default run (one JSON line on stdout):
  value    the 512^3 box scene (20% sand, 10% fluid, 5% static boxes over a floor),
           K ticks in captured HIP graphs with a fused state hash every 50 ticks
           and always on the last timed tick, after W warm-up ticks, with the
           library's default tick (settled segments skipped, results bit-identical
           to the dense tick).  A "step" is one tick of the whole world; value
           counts every voxel of every tick.
  parity   every hash of the timed region against the oracle's per-tick hashes of
           the same scene (tests/golden/boxes512_hashes.txt).
  roofline the timed region: algorithmic bytes actually moved (2 per swept
           voxel-update, 1 per voxel only read on hash ticks; counted by replaying
           the same ticks with sweep counters) / device time of the timed graphs
           (HIP events on the world's stream), against 8.0 TB/s.  "traffic" is
           measured for the same window: this script runs the window again in two
           child processes under rocprofv3 --pmc (FETCH_SIZE, then WRITE_SIZE),
           scaled by the counters of a copy of known size in the same process.
           "dense" holds the dense tick kernel on the same world; "hbm2048" the
           dense tick of the 2048^3 pillar scene, whose buffers (2 x 8 GiB) do not
           fit the 256 MB Infinity Cache: the HBM-bound figure.
  extra    pillars2048: the 2048^3 pillar scene, skipping, dense and dense fused
           in pairs; pillars2048_render: that scene with a 1920x1080 voxel render
           every 10 ticks; copy: a plain device copy of the same bytes, with
           temporal and with streaming (nontemporal) accesses, the practical
           bandwidth ceilings; batched: 4096 worlds against the
           one-launch-per-world loop; serving: the continuous-batching server
           under a Poisson load, saturated, and against one world per request;
           slab2048: one GPU's z-slab of the 2048^3 world (dense and fused ticks
           with a two-plane halo each side) and its halo exchange.
           Each workload's hashes (and frame digests) are compared with the
           oracle's fixtures in tests/golden/.
  cpu_baseline  the oracle's tick on this host, one thread, a bounded sample; its
           "openmp" field times the same tick over the host threads the run may use.

--only {boxes512,pillars2048,copy,batched,serving,render2048,slab2048} runs one workload
(for profiler runs) and prints its own JSON line; --only boxes512-window and
--only dense2048-window run just a timed window (and a copy of known size), for
the counter runs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
XGMI_LINK_BPS = 153e9  # one MI355X xGMI link, per direction (published figure; one GPU here)
BYTES_PER_UPDATE = 2
HASH_EVERY = 50
COPY_CAL_BYTES = 1 << 30  # the known-size copy that scales the byte counters
GOLDEN = os.path.join(ROOT, "tests", "golden")
METRIC = "Gvoxel-updates/s + %HBM roofline at 512^3/2048^3; batched vs plain; p99 tick ms"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def roofline(vox, kern_ms):
    ach = BYTES_PER_UPDATE * vox / (kern_ms * 1e-3)
    return ach, ach / HBM_PEAK


def golden_rows(name):
    """Rows of integers of an oracle fixture under tests/golden ('#' lines are
    comments), or None when the file is absent."""
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return [[int(v) for v in line.split()] for line in f if line.strip() and not line.startswith("#")]


def hash_segments(steps, every=HASH_EVERY):
    """The timed ticks as graph runs (nticks, hash_every): whole groups of `every`
    ticks, hashed at each group's end, then the remainder as one run hashed at
    its last tick, so the last timed tick is always hashed."""
    q, r = divmod(steps, every)
    return ([(q * every, every)] if q else []) + ([(r, r)] if r else [])


def timed_boxes_window(w, warmup, steps):
    """Warm-up ticks, then the timed graph runs of hash_segments(steps).  Both
    graphs are captured beforehand: the first run has an even tick count, so the
    second starts from the same buffer it was captured on.  The warm-up hashes
    its last tick, so the hashing kernels' first use in the process (tens of
    microseconds of one-time setup) falls outside the timed region.  Returns (wall s, device ms
    per tick, hashes, absolute tick of each hash)."""
    if warmup:
        w.run(warmup, warmup)
    segs = hash_segments(steps)
    for nt, he in segs:
        w.prepare_run(nt, he)
    w.sync()
    # each run returns when its ticks are done on the device; the wall clock
    # covers the runs, not the bookkeeping between them
    hashes, dev, wall = [], 0.0, 0.0
    for nt, he in segs:
        t0 = time.perf_counter()
        out = w.run(nt, he)
        wall += time.perf_counter() - t0
        hashes += [int(x) for x in out]
        dev += w.last_run_ms_per_tick() * nt
    ticks, t = [], warmup
    for nt, he in segs:
        ticks += [t + he * (i + 1) for i in range(nt // he)]
        t += nt
    return wall, dev / steps, hashes, ticks


def bench_boxes512(F, steps, warmup, kernel_ticks=200):
    """The 512^3 box scene with the library's default tick (settled segments are
    skipped): warm-up ticks, then the timed graphs of `steps` ticks with a fused
    hash every 50 and on the last tick.  The same tick sequence is replayed with
    sweep counters on to get the algorithmic bytes of the timed region (skipping
    is deterministic), and a dense world runs the plain tick kernel for the dense
    roofline."""
    n = 512
    with F.World((n, n, n)) as w:
        t0 = time.perf_counter()
        w.fill_boxes(1)
        log(f"[bench] boxes512: scene generated in {time.perf_counter() - t0:.2f}s")
        wall, dev_ms, hashes, hticks = timed_boxes_window(w, warmup, steps)
        log(f"[bench] boxes512: {steps} ticks in {wall * 1e3:.2f} ms ({dev_ms * 1e3:.2f} us/tick on the device)")
        # replay with counters
        w.fill_boxes(1)
        w.tick_count = 0
        if warmup:
            w.run(warmup, 0)
        w.sweep_stats(True)
        h2 = []
        for nt, he in hash_segments(steps):
            h2 += [int(x) for x in w.run(nt, he)]
        st = w.sweep_stats(False)
        if h2 != hashes:
            raise SystemExit("boxes512: counter replay diverged from the timed run")
        mode, seg = w.skip, w.segment
    with F.World((n, n, n), flags=F.DENSE) as d:
        d.fill_boxes(1)
        d.run(kernel_ticks, 0)
        d.run(kernel_ticks, 0)
        kern_ms = d.last_run_ms_per_tick()
    return dict(vox=n ** 3, wall=wall, dev_ms=dev_ms, kern_ms=kern_ms, hashes=hashes, hash_ticks=hticks, seg=seg,
                swept=st["swept"], read_only=st["read_only"], violations=st["violations"], mode=mode)


def boxes512_parity(hashes, ticks):
    """Every hash of the timed region against the oracle's per-tick hashes."""
    fx = golden_rows("boxes512_hashes.txt")
    if fx is None or not ticks or max(ticks) >= len(fx):
        return {"checked": 0, "match": None, "fixture": "tests/golden/boxes512_hashes.txt",
                "note": "fixture absent or shorter than warm-up + steps"}
    bad = [t for h, t in zip(hashes, ticks) if fx[t][0] != h]
    return {"checked": len(ticks), "match": not bad, "ticks": [ticks[0], ticks[-1]],
            "fixture": "tests/golden/boxes512_hashes.txt (oracle, hash after every tick)",
            **({"mismatched_ticks": bad[:8]} if bad else {})}


def window_main(F, mode, steps, warmup):
    """A timed window alone, then a copy of COPY_CAL_BYTES: what the counter runs
    profile (one dispatch per tick of the window's tick kernel)."""
    if mode == "boxes512-window":
        with F.World((512, 512, 512)) as w:
            w.fill_boxes(1)
            _, _, hashes, ticks = timed_boxes_window(w, warmup, steps)
        out = {"workload": mode, "hash": hashes[-1], "tick": ticks[-1]}
    else:
        with F.World((2048, 2048, 2048), flags=F.DENSE) as w:
            w.fuse = False
            w.fill_pillars(2)
            w.run(steps, 0)
            out = {"workload": mode, "hash": int(w.hash()), "tick": steps}
    F.copy_probe(COPY_CAL_BYTES, 3)
    print(json.dumps(out), flush=True)


def pmc_traffic(mode, steps, warmup, kernel_sub, last_n):
    """HBM-side bytes per launch of the last `last_n` dispatches whose kernel
    name contains one of `kernel_sub`, from two child runs of `--only mode` under
    rocprofv3 --pmc (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch).  Each
    counter is scaled by what it reads for the copy of COPY_CAL_BYTES in the
    same child.  Runs before this process touches the GPU.  Returns a dict, or
    {"error": ...}."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return {"error": "rocprofv3 not found"}
    base = tempfile.mkdtemp(prefix="fs3d_pmc_")
    vals = {}
    try:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(base, ctr)
            cmd = [prof, "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", "pmc", "--", sys.executable,
                   os.path.abspath(__file__), "--only", mode, "--steps", str(steps), "--warmup", str(warmup),
                   "--no-cpu"]
            log(f"[bench] traffic: {ctr} counter run of --only {mode} --steps {steps} --warmup {warmup}")
            t0 = time.perf_counter()
            p = subprocess.Popen(cmd, cwd=base, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                                 start_new_session=True)
            try:
                text, _ = p.communicate(timeout=240)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.communicate()
                return {"error": f"{ctr} counter run timed out"}
            done = f'"workload": "{mode}"' in text
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            # rocprofv3 may end with SIGSEGV after the program finished and its results are whole
            if not done or not files or p.returncode not in (0, 139, -11):
                tail = " | ".join(text.strip().splitlines()[-3:])
                return {"error": f"{ctr} counter run failed (rc {p.returncode}): {tail[-300:]}"}
            per = {}
            with open(files[0]) as f:
                for r in csv.DictReader(f):
                    if r.get("Counter_Name") != ctr:
                        continue
                    did = int(r["Dispatch_Id"])
                    name, v = per.get(did, (r["Kernel_Name"], 0.0))
                    per[did] = (name, v + float(r["Counter_Value"]))
            disp = [per[k] for k in sorted(per)]
            ticks = [v for name, v in disp if any(k in name for k in kernel_sub)][-last_n:]
            copies = [v for name, v in disp if "copy_kernel" in name]
            if len(ticks) != last_n or not copies:
                return {"error": f"{ctr}: {len(ticks)} of {last_n} tick dispatches, {len(copies)} copies found"}
            vals[ctr] = (sum(ticks) / len(ticks), sum(copies) / len(copies))
            log(f"[bench] traffic: {ctr} {vals[ctr][0]:.1f} KiB per dispatch ({time.perf_counter() - t0:.1f} s)")
    finally:
        shutil.rmtree(base, ignore_errors=True)
    f_read = COPY_CAL_BYTES / (vals["FETCH_SIZE"][1] * 1024.0)
    f_write = COPY_CAL_BYTES / (vals["WRITE_SIZE"][1] * 1024.0)
    rd = vals["FETCH_SIZE"][0] * 1024.0 * f_read
    wr = vals["WRITE_SIZE"][0] * 1024.0 * f_write
    return {"bytes": rd + wr, "read": rd, "written": wr, "dispatches": last_n,
            "scale": {"FETCH_SIZE": round(f_read, 4), "WRITE_SIZE": round(f_write, 4)}}


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


def bench_slab2048(F, ticks=20, slabs=8, halo=2):
    """One GPU's share of the 2048^3 world cut into `slabs` z-slabs: 2048 x 2048 x
    (2048 / slabs) planes plus `halo` planes on each side, dense single ticks and
    fused pairs over all of them, and the device copy of one boundary plane (what
    the receiving GPU writes of a halo exchange).  The link transfer itself needs
    a second GPU and is only estimated (one plane each way per link per odd tick
    at 153 GB/s per xGMI link)."""
    n, own = 2048, 2048 // slabs
    nz = own + 2 * halo
    out = {}
    for kind in ("dense", "fused"):
        with F.World((n, n, nz), flags=F.DENSE) as w:
            w.fuse = kind == "fused"
            if w.fuse != (kind == "fused"):
                raise SystemExit("slab2048: fusion not available for the slab shape")
            w.fill_pillars(2)
            w.run(4, 0)
            w.prepare_run(ticks, 0)
            w.sync()
            w.run(ticks, 0)
            out[kind] = w.last_run_ms_per_tick()
    plane = n * n
    copy_ms = F.copy_probe(plane, 50)
    link_ms = plane / XGMI_LINK_BPS * 1e3
    best = min(out["dense"], out["fused"])
    tick8 = best + link_ms / 2  # one exchange per two ticks, not overlapped
    return {"workload": f"one of {slabs} z-slabs of the 2048^3 pillar world: {n}x{n}x{own} planes plus a "
                        f"{halo}-plane halo on each side ({nz} planes), pillar fill, {ticks} dense ticks",
            "dims": [n, n, nz], "dense_ms_per_tick": round(out["dense"], 4),
            "fused_ms_per_tick": round(out["fused"], 4),
            "own_gvox_per_s": round(plane * own / (best * 1e-3) / 1e9, 1),
            "halo": {"plane_bytes": plane, "device_copy_ms": round(copy_ms, 4),
                     "xgmi_ms_per_odd_tick_estimate": round(link_ms, 4),
                     "note": "per odd tick each GPU sends its two boundary planes to its two z-neighbours over "
                             "two links at once; the estimate assumes 153 GB/s per link (not measured: one GPU)"},
            "estimate_8gpu": {"ms_per_tick": round(tick8, 4),
                              "gvox_per_s": round(n ** 3 / (tick8 * 1e-3) / 1e9, 1),
                              "note": "the faster of dense and fused slab ticks plus half an unoverlapped "
                                      "exchange per tick; interior row quads could hide it"}}


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


def golden_frames():
    path = os.path.join(GOLDEN, "pillars2048_frames.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["frames"]


def bench_pillars_render(F, ticks=200, every=10, W=1920, H=1080, seed=2):
    """The 2048^3 collapsing-column scene: `ticks` ticks in graphs of `every`,
    a W x H voxel render after each, from a seeded orbit camera.  After each
    frame (outside the timed loop) the state hash and the SHA-256 of the hit and
    face maps are compared with the oracle's (tests/golden/pillars2048_frames.json)."""
    import hashlib

    import numpy as np

    from fallingsand3d_amd import render as R
    n = 2048
    frames, tick_ms, render_ms, steps, checks = [], [], [], 0, []
    angles = orbit_angles(seed, ticks // every)
    fx = golden_frames()
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
            t1 = time.perf_counter()  # counting and checking are not part of the loop's wall time
            steps += voxel_steps(cam, w.dims, hit, W, H)
            got = dict(tick=(k + 1) * every, hash=int(w.hash()),
                       hit_sha256=hashlib.sha256(np.ascontiguousarray(hit, "<u8").tobytes()).hexdigest(),
                       face_sha256=hashlib.sha256(np.ascontiguousarray(face).tobytes()).hexdigest())
            if fx is not None and k < len(fx):
                checks.append(all(fx[k][key] == got[key] for key in got))
            t0 += time.perf_counter() - t1
        wall = time.perf_counter() - t0
        h = w.hash()
    return dict(wall=wall, tick_ms=sum(tick_ms) / len(tick_ms), render_ms=sum(render_ms) / len(render_ms),
                render_ms_max=max(render_ms), hit_frac=sum(frames) / len(frames), frames=len(frames), hash=int(h),
                vox=n ** 3, ticks=ticks, rays=W * H, voxel_steps=steps / len(frames),
                parity={"frames_checked": len(checks), "match": all(checks) if checks else None,
                        "what": "state hash after every 10th tick, SHA-256 of each frame's hit and face maps, "
                                "against the oracle (tests/golden/pillars2048_frames.json)"})


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


def bench_copy(F, nbytes, iters=20, scatter=True):
    """The plain copy (temporal and streaming loads and stores); the dense
    chunked tick's access shape without its arithmetic (every row quad's four
    rows of a 2048^3 grid pair, one wave per 1 KiB x-chunk); and the packed
    skipping tick's: 5 scattered 32-byte segments of every row quad (about the
    tick's density on the pillar scene), copied as 32-byte pieces and as whole
    64-byte sectors."""
    ms = F.copy_probe(nbytes, iters)
    log(f"[bench] copy: {nbytes} bytes in {ms:.4f} ms ({2 * nbytes / ms / 1e6:.1f} GB/s)")
    out = dict(nbytes=nbytes, ms=ms, GBps=2 * nbytes / (ms * 1e-3) / 1e9)
    # with streaming (nontemporal) loads and stores, as the large-grid passes use
    ms_nt = F.copy_probe(nbytes, iters, streaming=True)
    out["streaming"] = dict(ms=round(ms_nt, 4), GBps=round(2 * nbytes / (ms_nt * 1e-3) / 1e9, 1))
    log(f"[bench] streaming copy: {2 * nbytes / ms_nt / 1e6:.1f} GB/s")
    # the same copy with a 128 MiB working set (64 MiB read, 64 MiB written), which
    # stays in the 256 MB Infinity Cache like an in-place 512^3 tick's grid
    small = 64 << 20
    ms_s = F.copy_probe(small, 200)
    out["cache_resident"] = dict(nbytes=small, ms=round(ms_s, 5), GBps=round(2 * small / (ms_s * 1e-3) / 1e9, 1))
    log(f"[bench] copy, 128 MiB working set: {2 * small / ms_s / 1e6:.1f} GB/s")
    if scatter:
        # the dense chunked tick's loads and stores without its arithmetic, with
        # the tick's own policy (temporal loads, streaming stores)
        rms = F.rowquad_probe(False, True)
        out["rowquad_chunks"] = dict(ms=round(rms, 4), GBps=round(2 * 2048 ** 3 / (rms * 1e-3) / 1e9, 1),
                                     policy="temporal loads, streaming stores")
        log(f"[bench] rowquad_chunks: {rms:.4f} ms per pass of a 2048^3 pair")
        # 5 of the 64 segments: about the packed tick's density on the pillar scene
        for key, per_rq, sector in (("scatter5_seg32", 5, False), ("scatter5_sector64", 5, True)):
            sms, sb = F.scatter_probe(per_rq, sector)
            out[key] = dict(ms=round(sms, 4), item_bytes=sb, GBps=round(sb / (sms * 1e-3) / 1e9, 1))
            log(f"[bench] {key}: {sb / 1e9:.3f} GB of items in {sms:.4f} ms")
    return out


def bench_batched(F, nworlds=4096, ticks=500, plain_ticks=20, seed=3):
    """4096 worlds, edges log-uniform in [16, 128], rain fill with per-world
    density in [0.05, 0.5): all worlds step `ticks` ticks in one persistent
    launch (shape chosen by the autotuner) against the plain loop of one launch
    per world per tick.  The plain loop is timed on its first `plain_ticks`
    ticks (2 x 4096 x 20 launches would not change its per-tick cost).  The
    hashes after tick 20 (plain and persistent) and after tick 500 (persistent,
    with and without settled-world skipping) are compared with the oracle's
    (tests/golden/batch4096_hashes.txt)."""
    import numpy as np
    dims, thr, seeds = F.batch_specs(nworlds, seed, 16, 128)
    vox = int(dims.astype("int64").prod(1).sum())
    fx = golden_rows("batch4096_hashes.txt")
    fx_ok = fx is not None and len(fx) == nworlds and all(tuple(r[:3]) == tuple(int(v) for v in d)
                                                           for r, d in zip(fx, dims))
    want20 = np.array([r[3] for r in fx], np.uint64) if fx_ok else None
    want_end = np.array([r[4] for r in fx], np.uint64) if fx_ok and ticks == 500 else None
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
        log(f"[bench] batched: plain loop {plain_ticks} ticks in {plain_wall:.3f}s, plain = persistent: {exact}")
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
        parity = {"fixture": "tests/golden/batch4096_hashes.txt (oracle)", "match": None}
        if want20 is not None:
            m20 = bool((h_plain == want20).all() and (h_sample == want20).all())
            m_end = bool((h == want_end).all() and (h_all == want_end).all()) if want_end is not None else None
            parity = {**parity, "tick20_plain_and_persistent": m20, f"tick{ticks}_persistent": m_end,
                      "worlds": nworlds, "match": m20 and m_end is not False}
        plain_per_tick = plain_wall / plain_ticks
        return dict(nworlds=nworlds, vox=vox, ticks=ticks, wall=wall, dev_ms=dev_ms,
                    gvox_per_s=vox * ticks / wall / 1e9, plain_ms_per_tick=plain_per_tick * 1e3,
                    plain_gvox_per_s=vox / plain_per_tick / 1e9,
                    speedup=(plain_per_tick * ticks) / wall, bit_exact_sample=exact,
                    all_ticks_wall=wall_all, all_ticks_gvox_per_s=vox * ticks / wall_all / 1e9,
                    all_ticks_speedup=(plain_per_tick * ticks) / wall_all,
                    tuned={"mode": {0: "persistent", 2: "wide"}.get(tune["mode"], tune["mode"]), "seg": tune["seg"],
                           "threads": tune["threads"], "groups": tune["groups"]}, hash_digest=digest, parity=parity)


def bench_serving(F, seconds=3.0, lam=2000.0, seed=5, round_ticks=8, plain_every=10):
    """Poisson stream of requests at `lam` per second for `seconds` seconds of
    arrivals (edges log-uniform 32-256, ticks uniform 10-500, rain fill with
    density U(0.05, 0.5)) through the continuous-batching server under a
    200 GiB budget, replayed in real time; then the same requests all at once
    (the server's capacity); then the plain shape of one world per request (its
    own grid, fill, one tick launch per tick, hash), timed on every
    `plain_every`-th request.  Every final hash is compared with the oracle's
    (tests/golden/serve6000_hashes.txt); the plain sample's with the served ones."""
    import numpy as np
    n = int(seconds * lam)
    dims, thr, seeds, ticks, arr = F.serve_specs(n, seed, 32, 256, 10, 500, lam)
    cells = dims.astype(np.int64).prod(1)
    budget = 200 << 30
    with F.Server(budget_bytes=budget, round_ticks=round_ticks, max_active=1 << 16) as s:
        t0 = time.perf_counter()
        ids = s.replay(dims, seeds, thr, ticks, arr)
        wall = time.perf_counter() - t0
        st = s.stats()
        res = [s.result(int(i)) for i in ids]
    log(f"[bench] serving: {n} requests in {wall:.2f}s, tick p50 {st['tick_p50_ms']:.3f} ms p99 "
        f"{st['tick_p99_ms']:.3f} ms, request p50 {st['req_p50_ms']:.1f} ms p99 {st['req_p99_ms']:.1f} ms")
    with F.Server(budget_bytes=budget, round_ticks=round_ticks, max_active=1 << 16) as s:
        t0 = time.perf_counter()
        ids2 = s.replay(dims, seeds, thr, ticks, np.zeros(n))
        wall_sat = time.perf_counter() - t0
        st_sat = s.stats()
        res_sat = [s.result(int(i)) for i in ids2]
    log(f"[bench] serving: all {n} requests at once in {wall_sat:.2f}s "
        f"({st_sat['voxel_updates'] / wall_sat / 1e9:.1f} Gvox/s)")
    sample = list(range(0, n, plain_every))
    plain_hash = {}
    t0 = time.perf_counter()
    for i in sample:
        with F.World(tuple(int(v) for v in dims[i])) as w:
            w.fill_rain(int(seeds[i]), int(thr[i]))
            w.step(int(ticks[i]))
            plain_hash[i] = w.hash()
    plain_wall = time.perf_counter() - t0
    plain_vox = float((cells[sample] * ticks[sample]).sum())
    plain_rate = plain_vox / plain_wall
    served_ok = all(r[0] == 2 for r in res) and all(r[0] == 2 for r in res_sat)
    same = served_ok and all(res[i][1] == res_sat[i][1] for i in range(n)) and \
        all(plain_hash[i] == res[i][1] for i in sample)
    fx = golden_rows("serve6000_hashes.txt")
    parity = {"fixture": "tests/golden/serve6000_hashes.txt (oracle)", "match": None}
    if fx is not None and len(fx) >= n and all(tuple(fx[i][:4]) == (*(int(v) for v in dims[i]), int(ticks[i]))
                                               for i in range(n)):
        parity = {**parity, "requests": n, "match": served_ok and all(res[i][1] == fx[i][4] for i in range(n))}
    log(f"[bench] serving: plain loop on {len(sample)} requests in {plain_wall:.2f}s "
        f"({plain_rate / 1e9:.1f} Gvox/s); served = saturated = plain: {same}; oracle: {parity['match']}")
    rate = st["voxel_updates"] / wall
    rate_sat = st_sat["voxel_updates"] / wall_sat
    return dict(n=n, wall=wall, stats=st, exact=same, gvox_per_s=rate / 1e9, busy_frac=st["busy_ms"] / (wall * 1e3),
                saturated=dict(wall=wall_sat, gvox_per_s=rate_sat / 1e9, stats=st_sat),
                plain=dict(requests=len(sample), wall=plain_wall, gvox_per_s=plain_rate / 1e9),
                speedup_vs_plain=rate / plain_rate, saturated_speedup_vs_plain=rate_sat / plain_rate, parity=parity)


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


DENSE2048_WINDOW = 10  # dense 2048^3 ticks of the counter run for roofline.hbm2048
# the dense single tick at 2048^3: 16-byte x-chunks (rows longer than 1024 bytes)
DENSE2048_KERNELS = ("tick_chunk_kernel<", "tick_kernel<")
DENSE2048_LABEL = "tick_chunk_kernel<16,false> (dense tick in 16-byte x-chunks, 2048^3 pillar scene)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--only", choices=["boxes512", "pillars2048", "copy", "batched", "serving", "render2048",
                                       "slab2048", "boxes512-window", "dense2048-window"], default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter runs of the timed window")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every world lives on one device and nothing here shards "
              "across GPUs (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        raise SystemExit("--steps >= 1 and --warmup >= 0")

    # the counter runs are child processes started before this one touches the GPU
    traffic = traffic2048 = None
    if args.only in (None, "boxes512") and not args.no_traffic:
        traffic = pmc_traffic("boxes512-window", args.steps, args.warmup, ("tick_kernel<",), args.steps)
        if args.only is None:
            traffic2048 = pmc_traffic("dense2048-window", DENSE2048_WINDOW, 0, DENSE2048_KERNELS, DENSE2048_WINDOW)

    import fallingsand3d_amd as F
    if F.device_count() < 1:
        raise SystemExit("no GPU visible")
    F.init(0)

    if args.only in ("boxes512-window", "dense2048-window"):
        window_main(F, args.only, args.steps, args.warmup)
        return
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
    if args.only == "slab2048":
        print(json.dumps({"workload": "slab2048", **bench_slab2048(F)}), flush=True)
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
    alg_per_tick = alg_bytes / args.steps
    roof = {"bound": "hbm", "achieved": round(ach / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(frac, 4), "traffic": None,
            "kernel": f"tick_kernel<{r['seg']},*,skip> ({sched} schedule), every tick of the timed region",
            "ms_per_tick_device": round(r["dev_ms"], 5), "alg_bytes_per_tick": int(alg_per_tick),
            "swept_fraction": round(r["swept"] / (vox * args.steps), 4), "skip_violations": r["violations"],
            "note": "the 128 MiB grid (updated in place) fits the 256 MB Infinity Cache: cache-assisted; "
                    "hbm2048 is the HBM-bound figure"}
    if traffic is not None:
        if "error" in traffic:
            roof["traffic_error"] = traffic["error"]
        else:
            roof["traffic"] = round(traffic["bytes"])
            roof["traffic_detail"] = {
                "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this window (--only boxes512-window, "
                          "same --steps/--warmup), mean over its timed tick dispatches, scaled by a 1 GiB copy",
                "read": round(traffic["read"]), "written": round(traffic["written"]),
                "vs_algorithmic": round(traffic["bytes"] / alg_per_tick, 3), "scale": traffic["scale"]}
    roof["dense"] = {"kernel": f"tick_kernel<{r['seg']},false> (every voxel, no skipping)",
                     "kernel_ms": round(r["kern_ms"], 5), "achieved": round(ach_d / 1e9, 1), "frac": round(frac_d, 4)}
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
                               "over a floor; ticks in HIP graphs with a fused state hash every 50 ticks and on "
                               "the last timed tick",
                   "dims": [512, 512, 512], "hash_every": HASH_EVERY, "segment_bytes": r["seg"]},
        "roofline": roof,
        "final_hash": r["hashes"][-1],
        "final_tick": r["hash_ticks"][-1],
        "parity": boxes512_parity(r["hashes"], r["hash_ticks"]),
    }
    if args.only is None:
        extra = {}
        p = bench_pillars2048(F, 200)
        ach2, frac2 = roofline(p["vox"], p["dense"]["kern_ms"])
        c = bench_copy(F, 8 << 30)
        if c.get("cache_resident"):
            # the dense 512^3 tick against a copy of the same working set
            roof["dense"]["frac_of_cache_copy"] = round(roof["dense"]["achieved"] / c["cache_resident"]["GBps"], 4)
        hbm = {"bound": "hbm", "achieved": round(ach2 / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
               "frac": round(frac2, 4), "traffic": None,
               "kernel": DENSE2048_LABEL,
               "kernel_ms": round(p["dense"]["kern_ms"], 5), "alg_bytes_per_tick": 2 * p["vox"],
               "frac_of_copy": round(ach2 / 1e9 / c["GBps"], 4),
               "frac_of_streaming_copy": round(ach2 / 1e9 / c["streaming"]["GBps"], 4)}
        if c.get("rowquad_chunks"):
            # against its own loads and stores without the arithmetic
            hbm["frac_of_access_shape"] = round(ach2 / 1e9 / c["rowquad_chunks"]["GBps"], 4)
        if traffic2048 is not None:
            if "error" in traffic2048:
                hbm["traffic_error"] = traffic2048["error"]
            else:
                hbm["traffic"] = round(traffic2048["bytes"])
                hbm["traffic_detail"] = {
                    "source": f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of {DENSE2048_WINDOW} dense ticks "
                              "(--only dense2048-window), scaled by a 1 GiB copy",
                    "read": round(traffic2048["read"]), "written": round(traffic2048["written"]),
                    "vs_algorithmic": round(traffic2048["bytes"] / (2 * p["vox"]), 4)}
        roof["hbm2048"] = hbm
        fx_frames = golden_frames()
        want200 = fx_frames[-1]["hash"] if fx_frames and fx_frames[-1]["tick"] == 200 else None
        pr = bench_pillars_render(F)
        extra["pillars2048_render"] = {
            "workload": "2048^3 pillars, 200 ticks in graphs of 10, 1920x1080 voxel render after each (20 frames), "
                        "seeded orbit camera",
            "wall_s": round(pr["wall"], 3), "tick_ms": round(pr["tick_ms"], 4),
            "render_ms_mean": round(pr["render_ms"], 3), "render_ms_max": round(pr["render_ms_max"], 3),
            "rays_per_s": round(pr["rays"] / (pr["render_ms"] * 1e-3), 1), "hit_fraction": round(pr["hit_frac"], 4),
            "voxel_steps_per_frame": round(pr["voxel_steps"]),
            "voxel_steps_per_s": round(pr["voxel_steps"] / (pr["render_ms"] * 1e-3), 1),
            "final_hash": pr["hash"], "parity": pr["parity"]}
        ps, pd = p["skip"], p["dense"]
        extra["pillars2048"] = {
            "workload": "2048^3 uint8, 256 seeded sand pillars in a fluid pool, 200 plain ticks in one graph",
            "gvox_per_s": round(p["vox"] * p["ticks"] / ps["wall"] / 1e9, 2),
            "ms_per_tick": round(ps["kern_ms"], 4), "swept_fraction": round(ps["swept"] / (p["vox"] * p["ticks"]), 4),
            "dense_gvox_per_s": round(p["vox"] * p["ticks"] / pd["wall"] / 1e9, 2),
            "dense_kernel_ms": round(pd["kern_ms"], 4), "dense_achieved_GBps": round(ach2 / 1e9, 1),
            "dense_hbm_frac": round(frac2, 4), "dense_frac_of_copy": round(ach2 / 1e9 / c["GBps"], 4),
            "dense_frac_of_streaming_copy": round(ach2 / 1e9 / c["streaming"]["GBps"], 4),
            "fused_pairs": {"kernel": "tick2c_kernel<16> (two dense ticks per pass in 16-byte x-chunks of 1 KiB, "
                                             "first tick's rows in LDS)",
                            "gvox_per_s": round(p["vox"] * p["ticks"] / p["fused"]["wall"] / 1e9, 2),
                            "ms_per_tick": round(p["fused"]["kern_ms"], 4),
                            "achieved_GBps": round(p["vox"] * 2 / (2 * p["fused"]["kern_ms"] * 1e-3) / 1e9, 1),
                            "note": "achieved counts the 2 bytes per voxel that one pass moves for two ticks"},
            "final_hash": pd["hash"], "segment_bytes": pd["seg"], "dense_kernel": DENSE2048_LABEL,
            "parity": {"tick200_matches_oracle": (pd["hash"] == want200) if want200 is not None else None,
                       "fixture": "tests/golden/pillars2048_frames.json"}}
        extra["slab2048"] = bench_slab2048(F)
        extra["copy"] = {"bytes": c["nbytes"], "ms": round(c["ms"], 4), "GBps": round(c["GBps"], 1),
                         "hbm_frac": round(c["GBps"] * 1e9 / HBM_PEAK, 4), "streaming": c["streaming"],
                         "scatter5_seg32": c.get("scatter5_seg32"), "scatter5_sector64": c.get("scatter5_sector64"),
                         "cache_resident": c.get("cache_resident"), "rowquad_chunks": c.get("rowquad_chunks")}
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
            "autotuned": bb["tuned"], "final_hash_digest": bb["hash_digest"], "parity": bb["parity"]}
        sv = bench_serving(F)
        st = sv["stats"]
        extra["serving"] = {
            "workload": "Poisson 2000 req/s for 3 s, edges log-uniform 32-256, ticks U{10..500}, rain fill, "
                        "200 GiB budget, rounds of 8 ticks",
            "requests": sv["n"], "wall_s": round(sv["wall"], 3), "gvox_per_s": round(sv["gvox_per_s"], 2),
            "tick_p50_ms": round(st["tick_p50_ms"], 4), "tick_p99_ms": round(st["tick_p99_ms"], 4),
            "request_p50_ms": round(st["req_p50_ms"], 2), "request_p99_ms": round(st["req_p99_ms"], 2),
            "device_busy_frac": round(sv["busy_frac"], 3), "peak_active": st["peak_active"],
            "peak_bytes": st["peak_bytes"], "rejected": st["rejected"],
            "saturated": {"what": "the same requests submitted at once: the server's capacity",
                          "wall_s": round(sv["saturated"]["wall"], 3),
                          "gvox_per_s": round(sv["saturated"]["gvox_per_s"], 2),
                          "tick_p50_ms": round(sv["saturated"]["stats"]["tick_p50_ms"], 4),
                          "tick_p99_ms": round(sv["saturated"]["stats"]["tick_p99_ms"], 4),
                          "peak_active": sv["saturated"]["stats"]["peak_active"]},
            "plain": {"what": "one world per request (own grid, fill, one tick launch per tick, hash), "
                              f"every 10th request ({sv['plain']['requests']})",
                      "wall_s": round(sv["plain"]["wall"], 3), "gvox_per_s": round(sv["plain"]["gvox_per_s"], 2)},
            "speedup_vs_plain": round(sv["speedup_vs_plain"], 2),
            "saturated_speedup_vs_plain": round(sv["saturated_speedup_vs_plain"], 2),
            "served_equals_alone": sv["exact"], "parity": sv["parity"]}
        out["extra"] = extra
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
