"""Serving loop (host mirror of the fs3d_server_* entry points).

`Server` admits simulation requests under an HBM budget, keeps their grids
resident and advances all active ones together, tick by tick, in wide launches
over the whole GPU; requests arriving meanwhile join at the next round.
`replay()` drives a recorded arrival schedule in real time from native code and
`stats()` reports per-tick and per-request latency percentiles and throughput.
"""
import ctypes

import numpy as np

from ._lib import Fs3dError, ServerConfig, ServerStats, c_double, c_i32, c_i64, c_u64, check, lib, ptr, vp

QUEUED, ACTIVE, DONE, REJECTED = 0, 1, 2, 3


class Server:
    def __init__(self, budget_bytes=200 << 30, round_ticks=8, max_active=16384, seg=8):
        cfg = ServerConfig(int(budget_bytes), round_ticks, max_active, seg, 0)
        h = vp()
        check(lib().fs3d_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().fs3d_server_destroy(self._h)
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
            raise Fs3dError(6, "server is closed")
        return self._h

    def submit(self, dims, seed, thr, ticks):
        d = (c_i32 * 3)(*[int(v) for v in dims])
        rid = c_i64(-1)
        check(lib().fs3d_server_submit(self.handle, d, int(seed), int(thr), int(ticks), ctypes.byref(rid)))
        return rid.value

    def poll(self):
        a, q, n = c_i32(0), c_i32(0), c_i64(0)
        check(lib().fs3d_server_poll(self.handle, ctypes.byref(a), ctypes.byref(q), ctypes.byref(n)))
        return a.value, q.value, n.value

    def drain(self):
        check(lib().fs3d_server_drain(self.handle))

    def result(self, rid):
        st, h, lat = c_i32(0), c_u64(0), c_double(0)
        check(lib().fs3d_server_result(self.handle, rid, ctypes.byref(st), ctypes.byref(h), ctypes.byref(lat)))
        return st.value, h.value, lat.value

    def replay(self, dims, seeds, thr, ticks, arrival):
        d = np.ascontiguousarray(np.asarray(dims, np.int32).reshape(-1, 3))
        n = len(d)
        s = np.ascontiguousarray(seeds, np.uint64)
        t = np.ascontiguousarray(thr, np.uint64)
        k = np.ascontiguousarray(ticks, np.int32)
        a = np.ascontiguousarray(arrival, np.float64)
        ids = np.zeros(n, np.int64)
        check(lib().fs3d_server_replay(self.handle, n, ptr(d, c_i32), ptr(s, c_u64), ptr(t, c_u64), ptr(k, c_i32),
                                       ptr(a, c_double), ptr(ids, c_i64)))
        return ids

    def stats(self):
        st = ServerStats()
        check(lib().fs3d_server_get_stats(self.handle, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_ if name != "reserved"}
