"""GPU: fs_band_noalloc, the band construction that allocates nothing (its
distance scratch is a fixed device array inside the library), against fs_band
(which allocates its scratch per call) on the same solid masks: two calls in a
row, the second on a smaller grid, bit-exact; the int-distance path
(d0 >= 255); and the size limit, rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from myfluidsimulation_amd import capi

pytestmark = pytest.mark.gpu


def _lib():
    return capi.lib()


def blobs(nx, ny, nz, seed):
    """seeded solid mask: random boxes plus y/z wall planes"""
    rng = np.random.default_rng(seed)
    s = np.zeros((nz, ny, nx), np.uint8)
    for _ in range(12):
        x0, y0, z0 = rng.integers(0, nx - 4), rng.integers(0, ny - 4), rng.integers(0, nz - 4)
        s[z0:z0 + rng.integers(2, 9), y0:y0 + rng.integers(2, 9), x0:x0 + rng.integers(2, 9)] = 1
    s[:, 0, :] = s[:, -1, :] = 1
    return s.reshape(-1)


def noalloc_band(fs, solid, nx, ny, nz, d0):
    L = _lib()
    n = nx * ny * nz
    d_solid = fs.DeviceBuffer.from_host(solid)
    d_band = fs.DeviceBuffer(n)
    try:
        capi.check(L.fs_band_noalloc(C.c_void_p(d_solid.ptr), nx, ny, nz, d0,
                                     C.c_void_p(d_band.ptr), None), "fs_band_noalloc")
        capi.call("fs_device_synchronize")
        return d_band.to_host(np.uint8, n)
    finally:
        d_solid.free()
        d_band.free()


def test_noalloc_band_matches_fs_band(fs):
    cases = [(200, 150, 120, 3, 1), (64, 48, 40, 2, 2), (40, 24, 20, 300, 3)]
    for nx, ny, nz, d0, seed in cases:  # second call smaller than the first
        print(f"noalloc band {nx}x{ny}x{nz} d0 {d0}", flush=True)
        solid = blobs(nx, ny, nz, seed)
        got = noalloc_band(fs, solid, nx, ny, nz, d0)
        want = fs.band(solid, nx, ny, nz, d0)
        assert want.sum() > 0
        assert np.array_equal(got, want), (nx, ny, nz, d0)


def test_noalloc_band_rejects_oversize(fs):
    L = _lib()
    # 4096 x 4096 x 8 = 2^27 cells: 128 MiB of byte distances, above the limit;
    # checked before any pointer is used
    rc = L.fs_band_noalloc(C.c_void_p(16), 4096, 4096, 8, 2, C.c_void_p(16), None)
    assert rc == capi.FS_ERR_INVALID
    assert b"scratch" in L.fs_last_error()
