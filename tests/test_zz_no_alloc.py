"""ds2d_laplace_band_noalloc (device buffers, library-owned scratch, no
allocation) against ds2d_laplace_band on the same images: bit-identical, on
two calls in a row, the second with a smaller image, and in place; an image
above the scratch limit is rejected before any device work.

The caller's device buffers come from the same HIP runtime libds2d.so links
(hipMalloc / hipMemcpy through ctypes)."""
import ctypes as C

import numpy as np
import pytest

from ds2d_amd import laplace_band
from ds2d_amd.lib import load

pytestmark = pytest.mark.gpu
LIMIT = 1 << 22
H2D, D2H = 1, 2


def _hip():
    lib = load()  # loads libds2d.so and with it the HIP runtime it links
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return lib, hip


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.nbytes <= self.n and self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape):
        out = np.empty(shape, np.float32)
        assert out.nbytes <= self.n and self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def test_noalloc_band_filter_matches_allocating_entry_point():
    lib, hip = _hip()
    rng = np.random.default_rng(5)
    big = 512 * 256 * 4
    d_in, d_out = DevBuf(hip, big), DevBuf(hip, big)
    try:
        for shape in ((512, 256), (251, 101)):  # the second call is smaller: stale scratch must not leak
            img = rng.standard_normal(shape).astype(np.float32)
            d_in.put(img)
            d_out.put(np.full(shape, np.nan, np.float32))
            rc = lib.ds2d_laplace_band_noalloc(shape[1], shape[0], d_in.p, d_out.p, None)
            assert rc == 0, lib.ds2d_last_error()
            assert hip.hipDeviceSynchronize() == 0
            got, ref = d_out.get(shape), laplace_band(img)
            print(f"noalloc {shape}: max |diff| {np.nanmax(np.abs(got - ref)):.3e}")
            assert np.array_equal(got, ref), shape
        # in place (in == out)
        img = rng.standard_normal((96, 40)).astype(np.float32)
        d_in.put(img)
        assert lib.ds2d_laplace_band_noalloc(40, 96, d_in.p, d_in.p, None) == 0
        assert hip.hipDeviceSynchronize() == 0
        assert np.array_equal(d_in.get((96, 40)), laplace_band(img))
        # above the limit: rejected on the host, the buffers are never touched
        rc = lib.ds2d_laplace_band_noalloc(LIMIT + 1, 1, d_in.p, d_out.p, None)
        assert rc < 0 and b"limit" in lib.ds2d_last_error()
    finally:
        d_in.free()
        d_out.free()
