"""hos_mlp_forward_noalloc (the allocation-free inference entry point) against
hos_mlp_forward, through the C-ABI: two calls in a row, the second with fewer rows, and
the size limit.  The test allocates its device arrays itself (HIP runtime via ctypes);
the entry point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from hosnwt import learned as hl
from hosnwt.lib import check, load


def _hip():
    return C.CDLL("libamdhip64.so")


def _dev_alloc(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
    return p


@pytest.mark.gpu
def test_mlp_forward_noalloc_matches_forward():
    lib = load()
    hip = _hip()
    net = hl.OrderNet(max_rows=256, seed=3)
    rng = np.random.default_rng(4)
    xd = _dev_alloc(hip, 256 * 512 * 4)
    yd = _dev_alloc(hip, 256 * 64 * 4)
    try:
        for rows in (200, 37):  # the second call smaller than the first
            x = rng.standard_normal((rows, 512)).astype(np.float32)
            ref = net.forward(x)
            assert hip.hipMemcpy(xd, x.ctypes.data_as(C.c_void_p), C.c_size_t(x.nbytes), 1) == 0  # H2D
            check(lib.hos_mlp_forward_noalloc(net.h, rows, xd, yd, None))
            assert hip.hipDeviceSynchronize() == 0
            y = np.zeros((rows, 64), dtype=np.float32)
            assert hip.hipMemcpy(y.ctypes.data_as(C.c_void_p), yd, C.c_size_t(y.nbytes), 2) == 0  # D2H
            print(f"rows {rows}: max |noalloc - forward| = {np.abs(y - ref).max():.3e}", flush=True)
            assert np.array_equal(y, ref)
        with pytest.raises(RuntimeError):
            check(lib.hos_mlp_forward_noalloc(net.h, 257, xd, yd, None))
    finally:
        hip.hipFree(xd)
        hip.hipFree(yd)
        net.close()
