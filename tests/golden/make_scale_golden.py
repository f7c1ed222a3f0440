"""Writes the oracle fixtures of the full-size workloads, computed by the CPU
oracle in this container (plain C + OpenMP, `oracle/`):

  boxes512   tests/golden/boxes512_hashes.txt
             512^3 box scene (seed 1: boxes of side 8-64, 20% sand, 10% fluid,
             5% static over a floor): the state hash after every tick,
             ticks 0..2100 (covers any warm-up + timed window up to 2100 ticks).
  batch4096  tests/golden/batch4096_hashes.txt
             the 4096 worlds of batch_specs(4096, seed 3, edges 16..128), rain
             fill: each world's hash after tick 20 and after tick 500.
  serve6000  tests/golden/serve6000_hashes.txt
             the 6000-request stream serve_specs(6000, seed 5, edges 32..256,
             ticks 10..500, 2000 req/s), rain fill: each request's final hash
             (after its own tick count, from tick 0).
  pillars2048 tests/golden/pillars2048_frames.json (+ pillars2048_rgb_sample.npz)
             the 2048^3 pillar scene (seed 2, 256 pillars of radius 4-32 and
             height 256-1800 in a fluid pool): the hash after every 10th tick up
             to 200, and after each of those ticks one 1920x1080 frame of the
             voxel ray caster from the seeded orbit camera (angles from a
             SplitMix64 stream of seed 2, distance 1.4, elevation 0.35): SHA-256
             of the hit and face maps, hit fraction, and the RGB of 512 hit
             pixels drawn by a fixed numpy generator.

The tick rules are this project's own (the reference snapshot has no
simulation), so these fixtures pin the GPU path to the oracle at the measured
sizes, not to reference output.  Cost in an 8-core container: boxes512 ~8 min,
batch4096 ~5 min, serve6000 ~30 min, pillars2048 ~45 min.

Usage: python tests/golden/make_scale_golden.py {boxes512,batch4096,serve6000,pillars2048} [--jobs N]
"""
import argparse
import hashlib
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

PILLARS_2048 = (256, 4, 32, 256, 1800, 256, 682, 256, 682, 32, 128)
MISS = np.iinfo(np.uint64).max


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def orbit_angles(seed, n):
    """2*pi * (SplitMix64 output >> 11) / 2^53, n outputs of the stream `seed`."""
    return [2.0 * math.pi * (int(v) >> 11) / float(1 << 53) for v in O.splitmix_stream(seed, n)]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make_boxes512(ticks=2100):
    O.set_threads(0)
    g, nboxes = O.scene_boxes((512, 512, 512), 1)
    h0 = O.state_hash(g)
    out = [h0]
    t0 = time.time()
    for k in range(0, ticks, 100):
        n = min(100, ticks - k)
        g, h = O.run(g, k, n, 1, use_lut=True)
        out += [int(x) for x in h]
        log(f"boxes512: tick {k + n} ({time.time() - t0:.0f} s)")
    with open(os.path.join(HERE, "boxes512_hashes.txt"), "w") as f:
        f.write(f"# 512^3 box scene, seed 1 ({nboxes} boxes); line k = state hash after tick k (k = 0: the fill)\n")
        for h in out:
            f.write(f"{h}\n")


def _batch_world(args):
    dims, seed, thr = args
    g = O.scene_rain(dims, seed, thr)
    g, _ = O.run(g, 0, 20, use_lut=True)
    h20 = O.state_hash(g)
    g, _ = O.run(g, 20, 480, use_lut=True)
    return h20, O.state_hash(g)


def _init_worker():
    O.set_threads(1)


def make_batch4096(jobs):
    dims, thr, seeds = O.batch_specs(4096, 3, 16, 128)
    work = [(tuple(int(v) for v in dims[i]), int(seeds[i]), int(thr[i])) for i in range(4096)]
    t0 = time.time()
    res = []
    with mp.Pool(jobs, initializer=_init_worker) as pool:
        for i, r in enumerate(pool.imap(_batch_world, work, chunksize=4)):
            res.append(r)
            if (i + 1) % 512 == 0:
                log(f"batch4096: {i + 1} worlds ({time.time() - t0:.0f} s)")
    with open(os.path.join(HERE, "batch4096_hashes.txt"), "w") as f:
        f.write("# batch_specs(4096, seed 3, edges 16..128), rain fill; per world: nx ny nz, hash after tick 20, "
                "hash after tick 500\n")
        for (d, _, _), (h20, h500) in zip(work, res):
            f.write(f"{d[0]} {d[1]} {d[2]} {h20} {h500}\n")


def _serve_request(args):
    dims, seed, thr, ticks = args
    g = O.scene_rain(dims, seed, thr)
    g, _ = O.run(g, 0, ticks, use_lut=True)
    return O.state_hash(g)


def make_serve6000(jobs, n=6000):
    dims, thr, seeds, ticks, _ = O.serve_specs(n, 5, 32, 256, 10, 500, 2000.0)
    work = [(tuple(int(v) for v in dims[i]), int(seeds[i]), int(thr[i]), int(ticks[i])) for i in range(n)]
    # largest first keeps the pool busy to the end; results are put back in order
    order = sorted(range(n), key=lambda i: -(np.prod(work[i][0], dtype=np.int64) * work[i][3]))
    t0 = time.time()
    res = [None] * n
    with mp.Pool(jobs, initializer=_init_worker) as pool:
        for k, h in enumerate(pool.imap(_serve_request, [work[i] for i in order], chunksize=1)):
            res[order[k]] = h
            if (k + 1) % 250 == 0:
                log(f"serve6000: {k + 1} requests ({time.time() - t0:.0f} s)")
    with open(os.path.join(HERE, "serve6000_hashes.txt"), "w") as f:
        f.write("# serve_specs(6000, seed 5, edges 32..256, ticks 10..500, 2000 req/s), rain fill; per request: "
                "nx ny nz ticks final hash\n")
        for (d, _, _, tk), h in zip(work, res):
            f.write(f"{d[0]} {d[1]} {d[2]} {tk} {h}\n")


def make_pillars2048(ticks=200, every=10, W=1920, H=1080, nsample=512):
    O.set_threads(0)
    dims = (2048, 2048, 2048)
    pal = np.array(json.load(open(os.path.join(HERE, "palette.json")))["rgba"], np.float32)
    angles = orbit_angles(2, ticks // every)
    t0 = time.time()
    g = O.scene_pillars(dims, 2, PILLARS_2048)
    frames = []
    sidx = np.zeros((len(angles), nsample), np.int64)
    srgb = np.zeros((len(angles), nsample, 3), np.float32)
    for k, ang in enumerate(angles):
        g, _ = O.run(g, k * every, every, use_lut=True)
        h = O.state_hash(g)
        cam = O.orbit_camera(dims, ang, 1.4, 0.35, W / H)
        rgb, hit, face = O.render_dda(g, cam, pal, W, H)
        hits = np.flatnonzero(hit.reshape(-1) != MISS)
        pick = np.sort(np.random.default_rng(1000 + k).choice(hits, nsample, replace=False))
        sidx[k] = pick
        srgb[k] = rgb.reshape(-1, 3)[pick]
        frames.append(dict(tick=(k + 1) * every, hash=int(h), angle=ang, hit_sha256=digest(hit.astype("<u8")),
                           face_sha256=digest(face), hit_fraction=float(len(hits) / hit.size),
                           rgb_sum=float(rgb.astype(np.float64).sum())))
        log(f"pillars2048: tick {(k + 1) * every}, hash {h}, hit fraction {frames[-1]['hit_fraction']:.4f} "
            f"({time.time() - t0:.0f} s)")
    meta = dict(scene="2048^3 pillars, seed 2", params=list(PILLARS_2048), every=every, ticks=ticks, W=W, H=H,
                camera=dict(kind="orbit", dist_scale=1.4, elev=0.35, aspect="W/H", angle_seed=2),
                hit_dtype="<u8 (linear voxel index; 2^64-1 = miss)", face_dtype="u1", frames=frames)
    with open(os.path.join(HERE, "pillars2048_frames.json"), "w") as f:
        json.dump(meta, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "pillars2048_rgb_sample.npz"), index=sidx, rgb=srgb)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["boxes512", "batch4096", "serve6000", "pillars2048"])
    ap.add_argument("--jobs", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    if a.which == "boxes512":
        make_boxes512()
    elif a.which == "batch4096":
        make_batch4096(a.jobs)
    elif a.which == "serve6000":
        make_serve6000(a.jobs)
    else:
        make_pillars2048()


if __name__ == "__main__":
    main()
