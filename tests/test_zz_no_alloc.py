"""p3_iterate_noalloc: the allocation-free dispatch against p3_iterate.

Both entry points are called through the C ABI with the same blobs; the
histogram that p3_iterate_noalloc adds into a caller-owned device buffer
must equal, bit for bit, the one p3_iterate leaves in a fresh renderer. Two
calls in a row share the device's static scratch array, the second on a
smaller grid with fewer walkers, so what the first leaves behind must not
leak into the second.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import blobs
from pyr3_amd import _lib, synth

P3_E_ARG, P3_E_STATE = -1, -3


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _blobs(g, W, H, iters, fuse, seed, walkers):
    xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=fuse, seed=seed, walkers=walkers)
    return (np.ascontiguousarray(xf, dtype=np.float32), np.ascontiguousarray(ds, dtype=np.uint32),
            np.ascontiguousarray(u, dtype=np.uint32), np.ascontiguousarray(pal, dtype=np.float32))


def _iterate_reference(lib, W, H, xf, ds, u, pal):
    """p3_iterate on a fresh W x H renderer; returns its histogram."""
    taps = np.array([0.25, 0.5, 0.25], dtype=np.float32)
    h = C.c_void_p()
    assert lib.p3_create(W, H, 1, _fp(taps), 3, C.byref(h)) == 0, lib.p3_last_error(None)
    try:
        assert lib.p3_reset(h, _fp(pal)) == 0, lib.p3_last_error(h)
        assert lib.p3_iterate(h, _fp(xf), xf.shape[0], _up(ds), ds.shape[0], _up(u)) == 0, \
            lib.p3_last_error(h)
        out = np.zeros(H * W * 4, dtype=np.uint32)
        assert lib.p3_read_histogram(h, _up(out), out.size) == 0, lib.p3_last_error(h)
        return out
    finally:
        lib.p3_destroy(h)


class DeviceBuffer:
    """A zeroed device buffer from the HIP runtime the library itself links
    (the caller-owned output; the test allocates it, the entry point never
    does)."""

    def __init__(self, n_u32):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = n_u32
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(4 * n_u32)) == 0
        zero = np.zeros(n_u32, dtype=np.uint32)
        assert self.hip.hipMemcpy(self.p, zero.ctypes.data_as(C.c_void_p), C.c_size_t(4 * n_u32),
                                  1) == 0   # hipMemcpyHostToDevice

    def read(self):
        out = np.zeros(self.n, dtype=np.uint32)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(4 * self.n),
                                  2) == 0   # hipMemcpyDeviceToHost
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _iterate_noalloc(lib, xf, ds, u, pal, W, H):
    buf = DeviceBuffer(H * W * 4)
    try:
        rc = lib.p3_iterate_noalloc(_fp(xf), xf.shape[0], _up(ds), ds.shape[0], _up(u), _fp(pal),
                                    buf.p, buf.n, None)
        assert rc == 0, lib.p3_last_error(None)
        return buf.read()
    finally:
        buf.free()


def test_noalloc_rejects_before_any_device_call():
    """Argument errors, the size limit included, return before any HIP call."""
    lib = _lib.load()
    g = synth.bench_genome(0, 64, 48)
    xf, ds, u, pal = _blobs(g, 64, 48, 10, 5, 1, 256)
    buf = C.c_void_p(1 << 40)  # never dereferenced: every case fails validation
    assert lib.p3_iterate_noalloc(_fp(xf), xf.shape[0], _up(ds), ds.shape[0], _up(u), _fp(pal),
                                  buf, 64 * 48 * 4 - 1, None) == P3_E_ARG
    assert b"histogram holds" in lib.p3_last_error(None)
    big = u.copy()
    big[0], big[1] = 2600, 2000   # 5.2 M bins x 40 bytes > 96 MiB
    assert lib.p3_iterate_noalloc(_fp(xf), xf.shape[0], _up(ds), ds.shape[0], _up(big),
                                  _fp(pal), buf, 2600 * 2000 * 4, None) == P3_E_ARG
    assert b"maximum" in lib.p3_last_error(None)
    cap = u.copy()
    cap[14] = 1
    assert lib.p3_iterate_noalloc(_fp(xf), xf.shape[0], _up(ds), ds.shape[0], _up(cap), _fp(pal),
                                  buf, 64 * 48 * 4, None) == P3_E_ARG


@pytest.mark.gpu
def test_noalloc_matches_iterate_on_two_calls():
    lib = _lib.load()
    g = synth.bench_genome(0, 320, 240)
    cases = [(320, 240, 4096, 60, 20, 0xA11C), (160, 120, 1152, 25, 10, 0xB22D)]
    for W, H, walkers, iters, fuse, seed in cases:
        print(f"noalloc case {W}x{H} walkers {walkers} iters {iters} seed {seed:#x}", flush=True)
        xf, ds, u, pal = _blobs(g, W, H, iters, fuse, seed, walkers)
        ref = _iterate_reference(lib, W, H, xf, ds, u, pal)   # also loads the shape's kernel
        got = _iterate_noalloc(lib, xf, ds, u, pal, W, H)
        assert ref[3::4].sum() > 0
        assert np.array_equal(got, ref), (W, H, int(np.abs(got.astype(np.int64) - ref).max()))


@pytest.mark.gpu
def test_noalloc_refuses_a_shape_not_loaded():
    lib = _lib.load()
    # a shape no other test dispatches: one xform of three rarely used entries
    from helpers import genome, xform
    from pyr3_amd.catalog import VARIATION_INDEX as V
    g = genome([xform(1.0, (0.5, 0, 0, 0, 0.5, 0),
                      vars=((V["hopalong"], 0.3), (V["peano"], 0.2), (V["doyle"], 0.1)))],
               scale=20.0)
    xf, ds, u, pal = _blobs(g, 32, 24, 5, 2, 3, 64)
    buf = DeviceBuffer(32 * 24 * 4)
    try:
        rc = lib.p3_iterate_noalloc(_fp(xf), xf.shape[0], _up(ds), ds.shape[0], _up(u), _fp(pal),
                                    buf.p, buf.n, None)
        assert rc == P3_E_STATE and b"not loaded" in lib.p3_last_error(None)
        assert int(buf.read().sum()) == 0
    finally:
        buf.free()
