"""Pins the oracle's transforms to FFTW's documented r2r definitions.

The closed forms below are FFTW's definitions of REDFT00/RODFT00/REDFT01/
REDFT10/RODFT01/RODFT10 (FFTW 3 manual, "1d Real-even DFTs (DCTs)" and
"1d Real-odd DFTs (DSTs)"), evaluated as O(N^2) sums; the normalisations are
fourier_FFTW.f90:178-229.
"""
import math

import numpy as np
import pytest

from oracle import hos_oracle as ho


def fftw_def(x, kind):
    n = len(x)
    j = np.arange(n)
    y = np.zeros(n)
    for k in range(n):
        if kind == "REDFT00":
            y[k] = x[0] + (-1) ** k * x[n - 1] + 2 * np.sum(x[1:n - 1] * np.cos(math.pi * j[1:n - 1] * k / (n - 1)))
        elif kind == "RODFT00":
            y[k] = 2 * np.sum(x * np.sin(math.pi * (j + 1) * (k + 1) / (n + 1)))
        elif kind == "REDFT10":
            y[k] = 2 * np.sum(x * np.cos(math.pi * (j + 0.5) * k / n))
        elif kind == "REDFT01":
            y[k] = x[0] + 2 * np.sum(x[1:] * np.cos(math.pi * j[1:] * (k + 0.5) / n))
        elif kind == "RODFT10":
            y[k] = 2 * np.sum(x * np.sin(math.pi * (j + 0.5) * (k + 1) / n))
        elif kind == "RODFT01":
            y[k] = (-1) ** k * x[n - 1] + 2 * np.sum(x[:n - 1] * np.sin(math.pi * (j[:n - 1] + 1) * (k + 0.5) / n))
    return y


@pytest.mark.parametrize("kind", ["REDFT00", "RODFT00", "REDFT10", "REDFT01", "RODFT10", "RODFT01"])
@pytest.mark.parametrize("n", [2, 3, 5, 8, 17])
def test_r2r_matches_fftw_definition(kind, n):
    if kind == "REDFT00" and n < 2:
        pytest.skip()
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    np.testing.assert_allclose(ho.r2r(x, kind, 0), fftw_def(x, kind), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n1,n2", [(9, 1), (12, 1), (9, 7), (16, 10)])
def test_f2s_is_cosine_series_at_nodes(n1, n2):
    """F2S(a)(x_i) = sum_k a_k cos(k pi x_i / L) (cos) and sum a_k sin (sin), node by node."""
    rng = np.random.default_rng(0)
    F = ho.Fourier(n1, n2)
    a = rng.standard_normal((n1, n2))
    i1 = np.arange(n1)
    i2 = np.arange(n2)
    for cx in ("cos", "sin"):
        for cy in (("cos", "sin") if n2 > 1 else ("cos",)):
            bx = (np.cos if cx == "cos" else np.sin)(math.pi * np.outer(i1, i1) / (n1 - 1))
            if n2 > 1:
                by = (np.cos if cy == "cos" else np.sin)(math.pi * np.outer(i2, i2) / (n2 - 1))
            else:
                by = np.ones((1, 1))
            aa = a.copy()
            if cx == "sin":
                aa[[0, n1 - 1], :] = 0.0
            if n2 > 1 and cy == "sin":
                aa[:, [0, n2 - 1]] = 0.0
            expect = bx @ aa @ by.T
            np.testing.assert_allclose(F.fourier_2_space(a, cx, cy), expect, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("n1,n2", [(9, 1), (33, 1), (9, 7), (17, 12)])
def test_s2f_inverts_f2s(n1, n2):
    rng = np.random.default_rng(1)
    F = ho.Fourier(n1, n2)
    a = rng.standard_normal((n1, n2))
    np.testing.assert_allclose(F.space_2_fourier(F.fourier_2_space(a)), a, rtol=1e-12, atol=1e-12)
    # sine series on interior modes
    a[[0, n1 - 1], :] = 0
    np.testing.assert_allclose(F.space_2_fourier(F.fourier_2_space(a, "sin", "cos"), "sin", "cos"), a,
                               rtol=1e-12, atol=1e-12)


def test_single_mode_derivative():
    """eta_x of cos(k x) is -k sin(k x) through the sin-cos inverse (resol_HOS.f90:985-1000)."""
    n1, xlen = 33, 3.0
    F = ho.Fourier(n1, 1)
    x = np.arange(n1) * xlen / (n1 - 1)
    m = 5
    k = m * math.pi / xlen
    a = np.zeros((n1, 1))
    a[m, 0] = 1.0
    kx = np.arange(n1) * math.pi / xlen
    d = F.fourier_2_space(-kx[:, None] * a, "sin", "cos")
    np.testing.assert_allclose(d[:, 0], -k * np.sin(k * x), atol=1e-12)


def test_paddle_transforms_round_trip_and_basis():
    """S2F_add inverts F2S_add; F2S_add of mode (k3, k2) is cos(kx_add z') cos(ky y) on paddle nodes."""
    n3a, n2 = 8, 6
    FA = ho.FourierAdd(n3a, n2)
    rng = np.random.default_rng(2)
    a = rng.standard_normal((n3a, n2))
    np.testing.assert_allclose(FA.space_2_fourier(FA.fourier_2_space(a)), a, rtol=1e-12, atol=1e-12)
    # basis: the REDFT10 inverse evaluates sum_m a_m cos(pi (m+1/2) j / n) at paddle node j, i.e.
    # cos(kx_add(m) z_j) with kx_add = (2m+1) pi / 4 and z_j = j / (n3-1) (f2s = 1/2 cancels the 2)
    e = np.zeros((n3a, n2))
    e[3, 2] = 1.0
    j = np.arange(n3a)
    i2 = np.arange(n2)
    expect = np.outer(np.cos(math.pi * (3 + 0.5) * j / n3a), np.cos(math.pi * i2 * 2 / (n2 - 1)))
    np.testing.assert_allclose(FA.fourier_2_space(e), expect, atol=1e-13)


def test_y_inverse_matches_base_transform():
    n1, n2 = 7, 9
    F = ho.Fourier(n1, n2)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((n1, n2))
    # x transform identity (only y): compare against the 2-D transform of an x-delta
    y = ho.fourier_2_space_y_axis(a, "cos", 1)
    i2 = np.arange(n2)
    np.testing.assert_allclose(y, a @ np.cos(math.pi * np.outer(i2, i2) / (n2 - 1)).T, atol=1e-12)
    ys = ho.fourier_2_space_y_axis(a, "sin", 1)
    np.testing.assert_allclose(ys, a @ np.sin(math.pi * np.outer(i2, i2) / (n2 - 1)).T, atol=1e-12)
