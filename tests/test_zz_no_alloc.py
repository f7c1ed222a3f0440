"""fs3d_grid_run_noalloc (ticks of a grid in caller-owned device memory, with
the kernels' own device memory in a static per-GPU array) against
fs3d_world_run on the same grid and ticks: hashes and final grid bit for bit,
on two calls in a row, the second on a smaller grid; both also against the
oracle.  The device buffers are the caller's, allocated here with hipMalloc."""
import ctypes

import numpy as np
import pytest

from conftest import announce

GUARD = 256  # FS3D_GRID_GUARD_BYTES


def hip():
    L = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return L


H2D, D2H = 1, 2


class DevGrid:
    """A caller-owned device buffer of nx*ny*nz + guard bytes."""

    def __init__(self, L, n):
        self.L, self.n, self.p = L, n, ctypes.c_void_p()
        assert L.hipMalloc(ctypes.byref(self.p), n + GUARD) == 0
        assert L.hipMemset(self.p, 0, n + GUARD) == 0

    def upload(self, g):
        g = np.ascontiguousarray(g, np.uint8)
        assert g.size == self.n
        assert self.L.hipMemcpy(self.p, g.ctypes.data, self.n, H2D) == 0

    def download(self, shape):
        out = np.empty(shape, np.uint8)
        assert self.L.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def free(self):
        assert self.L.hipFree(self.p) == 0


def random_grid(dims, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    g = rng.choice(np.array([0, 0, 0, 1, 2, 2, 3, 3], np.uint8), size=(nz, ny, nx))
    g[rng.random(g.shape) < 0.01] = 77  # reserved bytes stay put
    g[0] = 1
    return g


def noalloc_run(F, a, b, dims, tick, nticks, hash_every):
    from fallingsand3d_amd._lib import check, lib
    nh = nticks // hash_every if hash_every else 0
    hashes = np.zeros(max(nh, 1), np.uint64)
    fin = ctypes.c_int32(-1)
    d = (ctypes.c_int32 * 3)(*dims)
    check(lib().fs3d_grid_run_noalloc(a.p, b.p, d, tick, nticks, hash_every,
                                      hashes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(fin),
                                      None))
    return [int(h) for h in hashes[:nh]], fin.value


@pytest.mark.gpu
@pytest.mark.parametrize("calls", [[((96, 70, 54), 3, 13, 2, 41), ((40, 33, 17), 0, 9, 3, 42)]])
def test_noalloc_matches_world_run(gpu, O, calls):
    L = hip()
    for dims, t0, nticks, every, seed in calls:
        announce(f"noalloc run dims={dims} t0={t0} nticks={nticks} hash_every={every}")
        g = random_grid(dims, seed)
        with gpu.World(dims) as w:
            w.upload(g)
            w.tick_count = t0
            want_h = [int(x) for x in w.run(nticks, every)]
            want_g = w.download()
        n = int(np.prod(dims))
        a, b = DevGrid(L, n), DevGrid(L, n)
        try:
            a.upload(g)
            got_h, fin = noalloc_run(gpu, a, b, dims, t0, nticks, every)
            shape = (dims[2], dims[1], dims[0])
            got_g = (a if fin == 0 else b).download(shape)
        finally:
            a.free()
            b.free()
        assert fin == nticks % 2
        assert got_h == want_h
        assert (got_g == want_g).all()
        og, oh = O.run(g, t0, nticks, every)
        assert got_h == [int(x) for x in oh]
        assert (got_g == og).all()


@pytest.mark.gpu
def test_noalloc_refuses_more_hashes_than_its_scratch(gpu):
    from fallingsand3d_amd._lib import lib
    L = hip()
    dims = (16, 16, 16)
    a, b = DevGrid(L, 4096), DevGrid(L, 4096)
    try:
        hashes = np.zeros(4097, np.uint64)
        fin = ctypes.c_int32(-1)
        rc = lib().fs3d_grid_run_noalloc(a.p, b.p, (ctypes.c_int32 * 3)(*dims), 0, 4097, 1,
                                         hashes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(fin),
                                         None)
        assert rc == 1 and b"FS3D_NOALLOC_MAX_HASHES" in lib().fs3d_last_error()
        # the largest accepted count still works
        g = random_grid(dims, 5)
        a.upload(g)
        got_h, fin = noalloc_run(gpu, a, b, dims, 0, 4096, 1)
        assert len(got_h) == 4096 and fin == 0
    finally:
        a.free()
        b.free()


def test_noalloc_argument_checks_before_any_device_call():
    # no device is touched before these checks, so they run without a GPU
    from fallingsand3d_amd._lib import lib
    buf_a, buf_b = ctypes.create_string_buffer(8), ctypes.create_string_buffer(8)
    fin = ctypes.c_int32(-1)
    hashes = np.zeros(8, np.uint64)
    hp = hashes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    d = (ctypes.c_int32 * 3)(16, 16, 16)
    a, b = ctypes.addressof(buf_a), ctypes.addressof(buf_b)
    L = lib()
    assert L.fs3d_grid_run_noalloc(a, b, d, 0, 4097, 1, hp, ctypes.byref(fin), None) == 1
    assert b"FS3D_NOALLOC_MAX_HASHES" in L.fs3d_last_error()
    assert L.fs3d_grid_run_noalloc(a, a, d, 0, 2, 1, hp, ctypes.byref(fin), None) == 1
    assert L.fs3d_grid_run_noalloc(a, b, d, 0, 2, 1, None, ctypes.byref(fin), None) == 1
    assert L.fs3d_grid_run_noalloc(a, b, (ctypes.c_int32 * 3)(0, 16, 16), 0, 2, 0, hp, ctypes.byref(fin), None) == 1
    assert L.fs3d_grid_run_noalloc(a, b, (ctypes.c_int32 * 3)(5000, 16, 16), 0, 2, 0, hp, ctypes.byref(fin),
                                   None) == 1
    assert L.fs3d_grid_run_noalloc(a, b, d, 0, -1, 0, hp, ctypes.byref(fin), None) == 1
    assert fin.value == -1
