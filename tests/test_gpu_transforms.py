"""GPU r2r transforms vs the oracle's (FFTW-definition) transforms.

FP64: relative L2 <= 1e-13; FP32: relative L2 <= 2e-6.  Sizes cover smooth
lengths (Stockham radix 2/3/4/5/8), every compiled register-engine length
(512 ... 12288) and non-smooth ones (Bluestein):
2(n-1) = 16, 32, 510 (2.3.5.17), 2046 (2.3.11.31), 6144, 1536, 4094 (2.23.89),
and x lines longer than both in-workgroup engines (multi-launch HBM passes):
14400 (2^6 3^2 5^2), 16384, 20480, 28800.
"""
import numpy as np
import pytest

from helpers import rel
from hosnwt import lib
from oracle import hos_oracle as ho

SIZES_1D = [5, 9, 17, 33, 129, 256, 257, 385, 513, 769, 1024, 1025, 1537, 2048, 2049, 3073, 4097, 6145,
            7201, 8193, 10241, 14401]
SIZES_2D = [(9, 5), (33, 17), (256, 64), (1024, 256), (3073, 769), (2049, 513), (8193, 5)]


@pytest.mark.gpu
@pytest.mark.parametrize("n1", SIZES_1D)
@pytest.mark.parametrize("kind", ["cos", "sin"])
def test_r2r_1d(n1, kind):
    rng = np.random.default_rng(n1)
    a = rng.standard_normal((n1, 1))
    F = ho.Fourier(n1, 1)
    for fwd in (True, False):
        ref = F.space_2_fourier(a, kind, "cos") if fwd else F.fourier_2_space(a, kind, "cos")
        got = lib.r2r(a, kind, "cos", fwd, 0)
        assert rel(got, ref) < 1e-13, (n1, kind, fwd, rel(got, ref))
        got32 = lib.r2r(a, kind, "cos", fwd, 1)
        assert rel(got32, ref) < 2e-6, (n1, kind, fwd, rel(got32, ref))


@pytest.mark.gpu
@pytest.mark.parametrize("n1,n2", SIZES_2D)
def test_r2r_2d(n1, n2):
    rng = np.random.default_rng(n1 + n2)
    a = rng.standard_normal((n1, n2))
    F = ho.Fourier(n1, n2)
    for cx in ("cos", "sin"):
        for cy in ("cos", "sin"):
            for fwd in (True, False):
                ref = F.space_2_fourier(a, cx, cy) if fwd else F.fourier_2_space(a, cx, cy)
                got = lib.r2r(a, cx, cy, fwd, 0)
                assert rel(got, ref) < 1e-13, (n1, n2, cx, cy, fwd, rel(got, ref))
    ref = F.space_2_fourier(a)
    assert rel(lib.r2r(a, "cos", "cos", True, 1), ref) < 2e-6
