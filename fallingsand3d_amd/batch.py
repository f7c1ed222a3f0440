"""Batched worlds (host mirror of the fs3d_batch_* entry points).

`Batch` holds many independent worlds of uneven sizes in one device arena.
`run(nticks)` advances all of them in one persistent kernel launch (captured in
a HIP graph); `run(nticks, plain=True)` is the one-launch-per-world-per-tick
baseline.  Both produce bit-identical grids.

`autotune()` times every persistent-kernel shape on a scratch copy and keeps the
fastest; the choice is cached in a JSON file keyed by the batch's size class and
the device architecture, so a later batch of the same class skips the sweep.
"""
import ctypes
import json
import os

import numpy as np

from ._lib import Fs3dError, c_double, c_i32, c_u64, check, lib, ptr, vp

PERSISTENT, PLAIN, WIDE, TUNED = 0, 1, 2, 3
MODE_NAMES = {PERSISTENT: "persistent", PLAIN: "plain", WIDE: "wide"}

DEFAULT_TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")


def size_class(dims):
    """Key of a batch for the tuning cache: world count and x-extent/volume buckets."""
    d = np.asarray(dims, dtype=np.int64).reshape(-1, 3)
    vol = d.prod(1)
    nxb = int(np.ceil(np.log2(max(int(d[:, 0].max()), 1))))
    volb = int(np.round(np.log2(max(float(vol.mean()), 1.0))))
    nb = int(np.ceil(np.log2(max(len(d), 1))))
    return f"n2^{nb}-nx2^{nxb}-vol2^{volb}"


class Batch:
    def __init__(self, dims):
        d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1, 3))
        self.dims = d
        self.n = len(d)
        h = vp()
        check(lib().fs3d_batch_create(self.n, ptr(d, c_i32), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().fs3d_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        if not self._h:
            raise Fs3dError(6, "batch is closed")
        return self._h

    def shape(self, i):
        nx, ny, nz = (int(v) for v in self.dims[i])
        return (nz, ny, nx)

    def upload(self, i, grid):
        g = np.ascontiguousarray(grid, dtype=np.uint8)
        if g.shape != self.shape(i):
            raise Fs3dError(1, f"world {i}: grid shape {g.shape} != {self.shape(i)}")
        check(lib().fs3d_batch_upload(self.handle, i, g.ctypes.data, g.size))

    def download(self, i):
        g = np.empty(self.shape(i), np.uint8)
        check(lib().fs3d_batch_download(self.handle, i, g.ctypes.data, g.size))
        return g

    def fill_rain(self, seeds, thr):
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        t = np.ascontiguousarray(thr, dtype=np.uint64)
        if len(s) != self.n or len(t) != self.n:
            raise Fs3dError(1, "fill_rain needs one seed and threshold per world")
        check(lib().fs3d_batch_fill_rain(self.handle, ptr(s, c_u64), ptr(t, c_u64)))

    @property
    def tick_count(self):
        t = c_u64(0)
        check(lib().fs3d_batch_get_tick(self.handle, ctypes.byref(t)))
        return t.value

    @tick_count.setter
    def tick_count(self, t):
        check(lib().fs3d_batch_set_tick(self.handle, t))

    @property
    def config(self):
        """(segment bytes, workgroup threads, row-quad groups, resident workgroups)."""
        s, t, r, g = c_i32(0), c_i32(0), c_i32(0), c_i32(0)
        check(lib().fs3d_batch_get_config(self.handle, ctypes.byref(s), ctypes.byref(t), ctypes.byref(r),
                                          ctypes.byref(g)))
        return s.value, t.value, r.value, g.value

    def set_config(self, seg, threads, groups=1):
        check(lib().fs3d_batch_set_config(self.handle, seg, threads, groups))

    @property
    def skip(self):
        """Settled worlds stop ticking in persistent runs (results are identical)."""
        on = c_i32(0)
        check(lib().fs3d_batch_get_skip(self.handle, ctypes.byref(on)))
        return bool(on.value)

    @skip.setter
    def skip(self, on):
        check(lib().fs3d_batch_set_skip(self.handle, 1 if on else 0))

    def run(self, nticks, plain=False, hashes=True, mode=None):
        """Advances every world; returns the per-world hashes after the last tick.
        mode: PERSISTENT, WIDE, PLAIN or TUNED (default: TUNED, or PLAIN when
        plain=True)."""
        if mode is None:
            mode = PLAIN if plain else TUNED
        if mode == TUNED and getattr(self, "_tuned_mode", None) is not None:
            mode = self._tuned_mode
        out = np.zeros(self.n, np.uint64)
        check(lib().fs3d_batch_run(self.handle, nticks, mode, ptr(out, c_u64) if hashes else None))
        return out if hashes else None

    def hashes(self):
        return self.run(0)

    def last_run_ms(self):
        ms = c_double(0)
        check(lib().fs3d_batch_last_run_ms(self.handle, ctypes.byref(ms)))
        return ms.value

    def autotune(self, trial_ticks=10, cache_file=DEFAULT_TUNING_FILE, arch="gfx950"):
        """Picks the persistent-kernel shape, reusing a cached choice when the
        cache holds one for this size class and architecture."""
        key = f"{arch}/{size_class(self.dims)}"
        cache = {}
        if cache_file and os.path.exists(cache_file):
            try:
                with open(cache_file) as f:
                    cache = json.load(f)
            except (OSError, ValueError):
                cache = {}
        if key in cache and "mode" in cache[key]:
            c = cache[key]
            self.set_config(c["seg"], c["threads"], c["groups"])
            self._tuned_mode = c["mode"]
            return dict(seg=c["seg"], threads=c["threads"], groups=c["groups"], mode=c["mode"], cached=True,
                        ms=c.get("ms"))
        s, t, r, m, ms = c_i32(0), c_i32(0), c_i32(0), c_i32(0), c_double(0)
        check(lib().fs3d_batch_autotune(self.handle, trial_ticks, ctypes.byref(s), ctypes.byref(t),
                                        ctypes.byref(r), ctypes.byref(m), ctypes.byref(ms)))
        self._tuned_mode = m.value
        cache[key] = {"seg": s.value, "threads": t.value, "groups": r.value, "mode": m.value, "ms": ms.value,
                      "trial_ticks": trial_ticks}
        if cache_file:
            try:
                with open(cache_file, "w") as f:
                    json.dump(cache, f, indent=1, sort_keys=True)
            except OSError:
                pass
        return dict(seg=s.value, threads=t.value, groups=r.value, mode=m.value, cached=False, ms=ms.value)
