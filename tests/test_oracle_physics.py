"""Pins the oracle's physics against what the mathematics fixes.

Each check is chosen so that a plausible slip (dropped term, wrong sign,
transposed operand, wrong index) fails it:
  * the separable wavemaker sum equals the reference's literal quadruple loop;
  * the continuation polynomial meets its four defining conditions;
  * HOS W at order M converges to the exact Laplace solution like (kA)^(M+1);
  * with eta = 0 every order above 1 vanishes: W = F2S(phis_hat k tanh k);
  * GtoF inverts expon; linear sloshing (M=1) is integrated exactly;
  * the nonlinear (M=3) sloshing step conserves energy and volume;
  * partial dealiasing with p = M makes cos-cos products grid independent,
    and the sine-to-cosine conversion reproduces the reference grid exactly.
"""
import math

import numpy as np
import pytest

from oracle import hos_oracle as ho


def tiny_params(**kw):
    base = dict(n1=6, n2=5, mHOS=3, xlen=2.0, ylen=1.5, n3=5, icase=2, iramp=0)
    base.update(kw)
    return ho.Params(**base)


def test_phiadd_separable_equals_quadruple_loop():
    p = tiny_params()
    tb = ho.Tables(p)
    rng = np.random.default_rng(0)
    a = rng.standard_normal((p.n3_add, p.n2))
    da = rng.standard_normal((p.n3_add, p.n2))
    eta = 0.05 * rng.standard_normal((p.n1, p.n2))
    for e in (None, eta):
        sep = ho.phiadd_sl(a, da, e, tb)
        bf = ho.phiadd_sl_bruteforce(a, da, e, tb)
        for k in sep:
            np.testing.assert_allclose(sep[k], bf[k], rtol=1e-11, atol=1e-12, err_msg=k)


def test_phiadd_separable_2d():
    p = tiny_params(n1=9, n2=1)
    tb = ho.Tables(p)
    rng = np.random.default_rng(1)
    a = rng.standard_normal((p.n3_add, 1))
    da = rng.standard_normal((p.n3_add, 1))
    eta = 0.05 * rng.standard_normal((p.n1, 1))
    sep = ho.phiadd_sl(a, da, eta, tb)
    bf = ho.phiadd_sl_bruteforce(a, da, eta, tb)
    for k in sep:
        np.testing.assert_allclose(sep[k], bf[k], rtol=1e-11, atol=1e-12, err_msg=k)


def test_continuation_conditions():
    f, df, d2f = 0.3, -0.7, 1.9
    c = ho.match_surface_coeffs(f, df, d2f)[:, 0]

    def P(x, d=0):
        # derivatives of (c1 x^7 + c2 x^5 + c3 x^3 + c4 x) e^{-6(x+1)} by finite differences of high order
        fx = lambda t: (((c[0] * t * t + c[1]) * t * t + c[2]) * t * t + c[3]) * t * math.exp(-6 * (t + 1))
        h = 1e-3
        if d == 0:
            return fx(x)
        if d == 1:
            return (-fx(x + 2 * h) + 8 * fx(x + h) - 8 * fx(x - h) + fx(x - 2 * h)) / (12 * h)
        return (-fx(x + 2 * h) + 16 * fx(x + h) - 30 * fx(x) + 16 * fx(x - h) - fx(x - 2 * h)) / (12 * h * h)
    assert abs(P(-1.0) - f) < 1e-12
    assert abs(P(-1.0, 1) - df) < 1e-7
    assert abs(P(-1.0, 2) - d2f) < 1e-5
    assert abs(P(-2.0 / 3.0)) < 1e-12


def test_match_surface_uses_one_sided_stencils():
    """For quadratic paddle data the reference stencils are exact, so the continuation starts from f, f', f''."""
    n3 = 9
    delz = 1.0 / (n3 - 1)
    z = np.arange(2 * (n3 - 1)) * delz
    q = lambda t: 0.2 - 0.7 * t + 1.1 * t * t
    col = np.zeros((2 * (n3 - 1), 1))
    col[:n3, 0] = q(z[:n3])
    out = ho.match_surface_wmk(col, n3, delz)
    c = ho.match_surface_coeffs(q(1.0), -0.7 + 2.2, 2.2)[:, 0]
    xx = np.arange(n3 + 1, 2 * (n3 - 1) + 1 - 0) - 1
    xx = xx / (n3 - 1) - 2.0
    P = (((c[0] * xx ** 2 + c[1]) * xx ** 2 + c[2]) * xx ** 2 + c[3]) * xx * np.exp(-6 * (xx + 1))
    np.testing.assert_allclose(out[n3:, 0], P, rtol=1e-10, atol=1e-13)
    np.testing.assert_array_equal(out[:n3, 0], col[:n3, 0])


def _hos_w(M, A, n1=128, xlen=2 * math.pi, m=4):
    """W from hosphis_modes for eta = A cos(kx), phis = exact Laplace potential at z = eta."""
    p = ho.Params(n1=n1, n2=1, mHOS=M, xlen=xlen, icase=1, iabsnb=0)
    o = ho.Oracle(p)
    x = np.arange(n1) * xlen / (n1 - 1)
    k = m * math.pi / xlen
    eta = (A * np.cos(k * x))[:, None]
    phis = (np.cosh(k * (eta[:, 0] + 1)) / np.cosh(k) * np.cos(k * x))[:, None]
    w_exact = k * np.sinh(k * (eta[:, 0] + 1)) / np.cosh(k) * np.cos(k * x)
    etapm, etax, etay, g2, px, py = o.phisxy_etaxy(phis, eta)
    h = o.hosphis_modes(phis, etapm, g2)
    return np.max(np.abs(h["phiz"][:, 0] - w_exact)) / np.max(np.abs(w_exact))


def test_hos_w_converges_with_order():
    """|W_M - W_exact| ~ (kA)^(M+1): each order gains about a factor kA."""
    A = 0.02
    errs = [_hos_w(M, A) for M in (1, 2, 3, 4, 5)]
    k = 4 * math.pi / (2 * math.pi)
    for a, b in zip(errs[:-1], errs[1:]):
        assert b < a * 3 * k * A + 1e-13, errs
    assert errs[-1] < 1e-7


def test_w_at_flat_surface_is_linear():
    p = ho.Params(n1=33, n2=9, mHOS=5, xlen=3.0, ylen=2.0, icase=1, iabsnb=0)
    o = ho.Oracle(p)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((33, 9)) * np.exp(-0.5 * o.tb.k)
    phis = o.F.fourier_2_space(a)
    eta = np.zeros((33, 9))
    etapm, etax, etay, g2, px, py = o.phisxy_etaxy(phis, eta)
    h = o.hosphis_modes(phis, etapm, g2)
    w = o.F.fourier_2_space(a * o.tb.k * np.tanh(o.tb.k))
    np.testing.assert_allclose(h["phiz"], w, atol=1e-12)
    np.testing.assert_allclose(h["W1"], w, atol=1e-12)


def test_gtof_inverts_expon():
    p = ho.Params(n1=17, n2=1, mHOS=1, xlen=3.0, icase=1)
    o = ho.Oracle(p)
    rng = np.random.default_rng(4)
    a, b = rng.standard_normal((17, 1)), rng.standard_normal((17, 1))
    f1, f2 = o.expon(a, b, 0.37)
    g1, g2 = o.gtof(f1, f2, 0.37)
    np.testing.assert_allclose(g1, a, atol=1e-14)
    np.testing.assert_allclose(g2, b, atol=1e-14)


def test_linear_sloshing_is_exact():
    """M=1 without wavemaker: eta = a cos(kx) cos(w t), w^2 = k tanh k, to rounding."""
    n1, xlen, m = 33, 4.0, 3
    p = ho.Params(n1=n1, n2=1, mHOS=1, xlen=xlen, icase=1, iabsnb=0)
    o = ho.Oracle(p)
    x = np.arange(n1) * xlen / (n1 - 1)
    k = m * math.pi / xlen
    w = math.sqrt(k * math.tanh(k))
    a = 0.01
    st = o.initialize(eta0=(a * np.cos(k * x))[:, None])
    t = 0.0
    for h in (0.05, 0.11, 0.07):
        st, err, _ = o.runge4_adapt2(t, st, h)
        t += h
    np.testing.assert_allclose(st.eta[:, 0], a * np.cos(k * x) * math.cos(w * t), atol=1e-15)
    assert err < 1e-13


def test_nonlinear_sloshing_conserves_energy_and_volume():
    n1, xlen, m = 64, 4.0, 2
    p = ho.Params(n1=n1, n2=1, mHOS=3, xlen=xlen, icase=1, iabsnb=0)
    o = ho.Oracle(p)
    x = np.arange(n1) * xlen / (n1 - 1)
    k = m * math.pi / xlen
    st = o.initialize(eta0=(0.05 * np.cos(k * x))[:, None])
    e0 = o.energy(st.eta, st.phis, np.zeros_like(st.eta))[2]
    t = 0.0
    for _ in range(30):
        st, err, deta = o.runge4_adapt2(t, st, 0.05)
        t += 0.05
    e1 = o.energy(st.eta, st.phis, deta)[2]
    # HOS at order M conserves energy to O((kA)^(M+1)) ~ 4e-5 here; a dropped term gives O(kA)
    assert abs(e1 - e0) / e0 < 1e-4
    assert abs(o.volume(st.eta)) < 1e-15


def test_volume_energy_closed_forms():
    p = ho.Params(n1=41, n2=21, mHOS=1, xlen=4.0, ylen=2.0, icase=1)
    o = ho.Oracle(p)
    one = np.ones((41, 21))
    assert abs(o.volume(one) - 1.0) < 1e-14
    x = np.arange(41) * 4.0 / 40
    c = np.repeat(np.cos(math.pi * x / 4.0)[:, None], 21, axis=1)
    assert abs(o.volume(c)) < 1e-15
    # potential energy of eta = 1: 1/2 * area (trapezoid of a constant is exact)
    assert abs(o.energy(one, 0 * one, 0 * one)[0] - 0.5 * 4.0 * 2.0) < 1e-13


def test_padded_grid_is_exact_for_cos_products():
    """Design basis: with p = M, retained modes of the HOS RHS do not depend on the extended grid."""
    for n2 in (1, 8):
        outs = []
        for pad in (0, 1, 3):
            p = ho.Params(n1=24, n2=n2, mHOS=5, xlen=4.0, ylen=2.0, icase=1, iabsnb=0)
            p.Nd1 += pad
            if n2 > 1:
                p.Nd2 += pad
            o = ho.Oracle(p)
            rng = np.random.default_rng(5)
            G1 = 0.2 * rng.standard_normal((24, n2)) * np.exp(-0.4 * o.tb.k)
            G2 = 0.2 * rng.standard_normal((24, n2)) * np.exp(-0.4 * o.tb.k) * o.tb.omega
            z = np.zeros((24, n2))
            outs.append(o.solve_hos2(z, z, G2, G1, np.zeros((p.n3_add, n2)), 0.02))
        for o2 in outs[1:]:
            for a, b in zip(outs[0][:2], o2[:2]):
                assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-12


def sin_to_cos(N, n, K):
    M = N - 1
    s2f = ho._s2f_vec(N)[:n]
    q = np.arange(K)[None, :] + np.arange(n)[:, None]
    r = np.arange(K)[None, :] - np.arange(n)[:, None]

    def C(v):
        out = np.zeros(v.shape)
        odd = v % 2 != 0
        out[odd] = 1.0 / np.tan(math.pi * v[odd] / (2 * M))
        return out
    return s2f[:, None] * (C(q) + C(r))


@pytest.mark.parametrize("n,M", [(32, 3), (48, 4), (40, 5)])
def test_sine_to_cosine_conversion_is_exact(n, M):
    """filter_ext on the reference grid of (sine series) x (cosine series) from any finer grid."""
    Nd = ((M + 1) * n) // 2
    rng = np.random.default_rng(n)
    e = rng.standard_normal(n)
    a = rng.standard_normal(n)

    def ext(v, N):
        z = np.zeros((N, 1))
        z[:n, 0] = v
        return z
    F = ho.Fourier(Nd, 1)
    ref = F.space_2_fourier(F.fourier_2_space(ext(e, Nd), "sin") * F.fourier_2_space(ext(a, Nd)))[:n, 0]
    Np = Nd + 7
    Fp = ho.Fourier(Np, 1)
    K = min(2 * n, Np)
    P = Fp.space_2_fourier(Fp.fourier_2_space(ext(e, Np), "sin") * Fp.fourier_2_space(ext(a, Np)), "sin")[:K, 0]
    got = sin_to_cos(Nd, n, K) @ P
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13
    naive = Fp.space_2_fourier(Fp.fourier_2_space(ext(e, Np), "sin") * Fp.fourier_2_space(ext(a, Np)))[:n, 0]
    assert np.linalg.norm(naive - ref) / np.linalg.norm(ref) > 1e-8  # the conversion is needed


def test_ramp_closed_forms():
    p = ho.Params(n1=8, n2=1, mHOS=1, xlen=1.0, iramp=2, Tramp=3.0, T_stop=100.0)
    a, at, att = ho.ramp(1.5, p)
    assert (a, at, att) == (0.25, 2 * 0.5 / 3.0, 2 / 9.0)
    assert ho.ramp(5.0, p) == (1.0, 0.0, 0.0)
    assert ho.ramp(104.0, p) == (0.0, 0.0, 0.0)
    p.iramp = 1
    assert ho.ramp(1.5, p) == (0.5, 1.0 / 3.0, 0.0)


def test_rk_tableau_on_linear_rhs():
    """runge4_adapt2 with a diagonal linear RHS f(G) = lam * G (lam per mode and per field).

    The classical RK4 step multiplies each mode by 1 + z + z^2/2 + z^3/6 + z^4/24
    (z = lam h) and Kutta's third-order companion by 1 + z + z^2/2 + z^3/6, so the
    embedded error of every mode is |z^4/24 G| over max|G_new|.  A wrong A, b or e
    entry, or a stage fed the wrong predecessor, changes one of the polynomials.
    """
    n1, n2 = 9, 6
    p = ho.Params(n1=n1, n2=n2, mHOS=1, xlen=3.0, ylen=2.0, icase=1, iabsnb=0)
    o = ho.Oracle(p)
    rng = np.random.default_rng(21)
    lam1 = -0.8 + 1.3 * rng.standard_normal((n1, n2))
    lam2 = 0.4 * rng.standard_normal((n1, n2))

    def linear_rhs(a1st, eta1st, G2, G1, a_add, delta_t, *second_order_state):
        o.second = None
        za = np.zeros_like(a_add)
        return lam1 * G1, lam2 * G2, np.zeros_like(a1st), np.zeros_like(eta1st), za, za.copy()

    o.solve_hos2 = linear_rhs
    eta = rng.standard_normal((n1, n2)) * 0.01
    phis = rng.standard_normal((n1, n2)) * 0.01
    st = o.initialize(eta0=eta, phis0=phis)
    h = 0.23
    new, err, _ = o.runge4_adapt2(0.0, st, h)

    def p4(z):
        return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24

    G1 = o.s2f(eta)
    G2 = o.s2f(phis) * o.tb.omega
    G1n, G2n = p4(lam1 * h) * G1, p4(lam2 * h) * G2
    G2n[0, 0] = 0.0  # k_G2(1,1) = 0 and omega(1,1) = 0
    e1 = np.abs((lam1 * h) ** 4 / 24 * G1).max() / (np.abs(G1n).max() + ho.TINY)
    e2 = np.abs((lam2 * h) ** 4 / 24 * G2).max() / (np.abs(G2n).max() + ho.TINY)
    assert abs(err - max(e1, e2)) <= 1e-12 * max(e1, e2)
    etark, _ = o.gtof(G1n, G2n, h)
    etark[0, 0] = 0.0  # no wavemaker: the mean-volume mode is reset to 0
    np.testing.assert_allclose(new.eta, o.f2s(etark), atol=1e-15, rtol=0)


def test_filter_weights_properties():
    """filtering_base weights, pinned by identities rather than by their formulas.

    * c >= 0 keeps exactly the first INT(n c) modes (a projection);
    * -3 is cos^2(pi i / n);
    * -4 is the degree-7 smoothstep p(s) = I_s(4, 4): p(s) + p(1 - s) = 1 with
      s = cos^2(pi i / n), and p = 1/2 at s = 1/2;
    * -2 is sinc(2 pi i / n), vanishing at i = n/2;
    * -5 (erfc profile) decreases from 1 to 0 and passes 1/2 at i = 3n/4;
    * -1 zeroes only the last mode (the reference's integer quotient i/n).
    """
    n = 16
    i = np.arange(1, n + 1)
    w = ho.filter_weights(n, 0.6)
    assert np.array_equal(w, (i <= int(16 * 0.6)).astype(float))
    np.testing.assert_allclose(ho.filter_weights(n, -3), np.cos(np.pi * i / n) ** 2, atol=1e-15)
    s = np.cos(np.pi * i / n) ** 2
    w4 = ho.filter_weights(n, -4)
    # 1 - s(i) = sin^2(pi i/n) = s(n/2 - i)
    ii = np.arange(1, n // 2)
    np.testing.assert_allclose(w4[ii - 1] + w4[n // 2 - ii - 1], 1.0, atol=1e-14)
    assert abs(w4[n // 4 - 1] - 0.5) < 1e-14 and abs(s[n // 4 - 1] - 0.5) < 1e-14
    w2 = ho.filter_weights(n, -2)
    assert abs(w2[n // 2 - 1]) < 1e-15 and abs(w2[0] - np.sinc(2.0 / n)) < 1e-15
    w5 = ho.filter_weights(n, -5)
    assert np.all(np.diff(w5) < 0) and abs(w5[3 * n // 4 - 1] - 0.5) < 1e-15 and w5[0] > 0.999
    w1 = ho.filter_weights(n, -1)
    assert np.array_equal(w1, np.r_[np.ones(n - 1), 0.0])


def test_filtering_acts_on_modes_and_is_placed_before_expon():
    """filtering_x of a single cosine mode scales it by the mode weight; in solveHOS2
    the filter multiplies dG1, dG2 (and the frozen mean term) by the 2-D weight."""
    p = ho.Params(n1=24, n2=8, mHOS=3, xlen=4.0, ylen=2.0, icase=1, iabsnb=0)
    o = ho.Oracle(p)
    x = np.arange(24) * 4.0 / 23
    for k in (3, 20):
        f = np.repeat(np.cos(k * np.pi * x / 4.0)[:, None], 8, axis=1)
        g = o.filtering(f, 0.5, 0)
        np.testing.assert_allclose(g, f * (1.0 if k < 12 else 0.0), atol=1e-13)
    rng = np.random.default_rng(5)
    kx = np.arange(24) * np.pi / 4.0
    ky = np.arange(8) * np.pi / 2.0
    dec = np.exp(-0.3 * np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2))
    G1 = 0.01 * rng.standard_normal((24, 8)) * dec
    G2 = 0.01 * rng.standard_normal((24, 8)) * dec
    z = np.zeros((24, 8))
    za = np.zeros((p.n3_add, 8))
    ref = o.solve_hos2(z, z, G2, G1, za, 0.02)
    pf = ho.Params(n1=24, n2=8, mHOS=3, xlen=4.0, ylen=2.0, icase=1, iabsnb=0, coeffilt=(-3.0, 0.5, 1.0))
    of = ho.Oracle(pf)
    got = of.solve_hos2(z, z, G2, G1, za, 0.02)
    W = ho.filter_weights(24, -3.0)[:, None] * ho.filter_weights(8, 0.5)[None, :]
    np.testing.assert_allclose(got[0], W * ref[0], atol=1e-16)
    np.testing.assert_allclose(got[1], W * ref[1], atol=1e-16)


def test_check_slope_filters_only_steep_states():
    p = ho.Params(n1=32, n2=1, mHOS=3, xlen=4.0, icase=1, iabsnb=0)
    o = ho.Oracle(p)
    x = np.arange(32) * 4.0 / 31
    eta = (0.01 * np.cos(np.pi * x / 4.0) + 0.001 * np.cos(30 * np.pi * x / 4.0))[:, None]
    st = o.initialize(eta0=eta, phis0=0.5 * eta)
    same, fx, fy = o.check_slope(st, np.array([1.0]), None)
    assert not fx and not fy and np.array_equal(same.eta, st.eta)
    filt, fx, _ = o.check_slope(st, np.array([9.0]), None)
    assert fx
    np.testing.assert_allclose(filt.eta[:, 0], 0.01 * np.cos(np.pi * x / 4.0), atol=1e-14)
    np.testing.assert_allclose(filt.phis, 0.5 * filt.eta, atol=1e-14)


def test_second_order_paddle_fields_are_consistent():
    """d1st_for_2nd pinned by identities between independent code paths.

    * at the paddle top (z = 0) the free-surface-mode sums equal the x = 0 column of the
      base-grid transforms: phi_xx = F2S(-a kx^2), phi_y = F2S(-a ky) (cos-sin),
      phi_z = F2S(a k tanh k);
    * at the bottom (z = -1) the vertical derivative vanishes (no flux);
    * phi_z of the sums is the z derivative of the potential sum a cosh(k(z+1))/cosh k
      (checked by centred differences in z);
    * the additional-potential part on the paddle is F2S_add of the modal coefficients
      times k_add^2, -ky, -kx_add (cos/sin kinds), i.e. its own z derivative again.
    """
    p = ho.Params(n1=20, n2=9, mHOS=3, xlen=5.0, ylen=3.0, n3=9, icase=2, i_wmk=2, iabsnb=0)
    o = ho.Oracle(p)
    tb = o.tb
    rng = np.random.default_rng(8)
    dec = np.exp(-0.4 * tb.k)
    a1 = 0.01 * rng.standard_normal((20, 9)) * dec
    da1 = 0.01 * rng.standard_normal((20, 9)) * dec
    zadd = np.zeros((p.n3_add, 9))
    sl = ho.phiadd_sl(zadd, zadd, None, tb, need=o.SECOND_FIELDS)
    f, _ = o.d1st_for_2nd(a1, da1, sl, zadd, zadd, np.zeros((20, 9)))
    top = p.n3 - 1
    np.testing.assert_allclose(f["axx"][top], o.f2s(-a1 * (tb.kx ** 2)[:, None])[0], atol=1e-14)
    np.testing.assert_allclose(f["ay"][top], o.f2s(-a1 * tb.ky[None, :], "cos", "sin")[0], atol=1e-14)
    np.testing.assert_allclose(f["az"][top], o.f2s(a1 * tb.kth[:, :, 0])[0], atol=1e-14)
    np.testing.assert_allclose(f["az"][0], 0.0, atol=1e-15)
    # potential sum and its z derivative by centred differences
    def phi_at(z):
        ch = np.cosh(tb.k * (z + 1.0)) / np.cosh(tb.k)
        return ho.fourier_2_space_y_axis((a1 * ch).sum(axis=0)[None, :], "cos", 1)[0]
    for i3 in (2, 5):
        z = tb.cz[i3]
        dz = 1e-5
        fd = (phi_at(z + dz) - phi_at(z - dz)) / (2 * dz)
        np.testing.assert_allclose(f["az"][i3], fd, atol=1e-9)
    # additional-potential part: with FS modes zero, axx / az follow F2S_add of the coefficients
    aadd = 0.01 * rng.standard_normal((p.n3_add, 9)) * np.exp(-0.3 * np.arange(p.n3_add))[:, None]
    g, _ = o.d1st_for_2nd(np.zeros((20, 9)), np.zeros((20, 9)), sl, aadd, zadd, np.zeros((20, 9)))
    phi_add = o.FA.fourier_2_space(aadd, "cos", "cos")[:p.n3]
    # z derivative of phi_add on the paddle by centred differences of the cos(kx_add (z+1)) series
    def phiadd_at(z):
        c = np.cos(tb.kx_add[:, None] * (z + 1.0)) * aadd
        return ho.fourier_2_space_y_axis(c.sum(axis=0)[None, :], "cos", 1)[0]
    # the inverse sine transform along z is FFTW's RODFT10, Y_k = 2 sum_j X_j sin(pi (j+1/2)(k+1)/n):
    # entry k holds node k+1, so the reference's paddle z derivative at index i3 is the one at
    # node i3+1 (resol_wmkr.f90:548-566 keeps that index), whereas the cosine fields are at node i3
    for i3 in (1, 6):
        z = tb.cz[i3 + 1]
        fd = (phiadd_at(z + 1e-5) - phiadd_at(z - 1e-5)) / 2e-5
        np.testing.assert_allclose(g["az"][i3], fd, atol=1e-8)
    np.testing.assert_allclose(g["axx"], o.FA.fourier_2_space(aadd * tb.k_add ** 2, "cos", "cos")[:p.n3],
                               atol=1e-15)
    assert np.abs(phi_add).max() > 0


@pytest.mark.parametrize("n2", [1, 6])
def test_second_order_z0_fields_are_derivatives(n2):
    """The extra z = 0 wavemaker fields of phiadd_1st_SWEET (resol_wmkr.f90:500-524) are
    derivatives of the brute-force-pinned fields: z derivatives by centred differences of
    phiadd_sl at z = +-d, time derivatives as the same sums of da, and (2-D) phi_xx = -phi_zz
    from Laplace's equation."""
    p = tiny_params(n1=9, n2=n2)
    tb = ho.Tables(p)
    rng = np.random.default_rng(21)
    a = rng.standard_normal((p.n3_add, n2)) * np.exp(-0.3 * np.arange(p.n3_add))[:, None]
    da = rng.standard_normal((p.n3_add, n2)) * np.exp(-0.3 * np.arange(p.n3_add))[:, None]
    z0 = ho.phiadd_sl(a, da, None, tb, need=ho.Oracle.SECOND_FIELDS)
    d = 1e-5
    up = ho.phiadd_sl(a, da, np.full((p.n1, n2), d), tb)
    dn = ho.phiadd_sl(a, da, np.full((p.n1, n2), -d), tb)
    scale = max(np.abs(z0[k]).max() for k in z0)
    for fld, base in (("phizz", "phiz"), ("phixz", "phix")):
        np.testing.assert_allclose(z0[fld], (up[base] - dn[base]) / (2 * d), atol=1e-7 * scale, err_msg=fld)
    # the reference's phi_yz sum has no kx_add factor (resol_wmkr.f90:522-523): it is the z
    # derivative of phi_y for the coefficients a / kx_add
    ak = a / tb.kx_add[:, None]
    upk = ho.phiadd_sl(ak, da, np.full((p.n1, n2), d), tb, need=("phiy",))
    dnk = ho.phiadd_sl(ak, da, np.full((p.n1, n2), -d), tb, need=("phiy",))
    np.testing.assert_allclose(z0["phiyz"], (upk["phiy"] - dnk["phiy"]) / (2 * d), atol=1e-7 * scale)
    zt_up = ho.phiadd_sl(da, da, np.full((p.n1, n2), d), tb, need=("phiz",))
    zt_dn = ho.phiadd_sl(da, da, np.full((p.n1, n2), -d), tb, need=("phiz",))
    np.testing.assert_allclose(z0["phizt"], ho.phiadd_sl(da, da, None, tb, need=("phiz",))["phiz"], atol=1e-14)
    np.testing.assert_allclose(z0["phizt"], (zt_up["phiz"] + zt_dn["phiz"]) / 2, atol=1e-8 * scale)
    dd = ho.phiadd_sl(da, da, None, tb)
    np.testing.assert_allclose(z0["phixt"], dd["phix"], atol=1e-14)
    np.testing.assert_allclose(z0["phiyt"], dd["phiy"], atol=1e-14)
    if n2 == 1:
        np.testing.assert_allclose(z0["phixx"], -z0["phizz"], atol=1e-12 * scale)
    assert np.abs(z0["phizz"]).max() > 0 and np.abs(z0["phixz"]).max() > 0


@pytest.mark.parametrize("n2", [1, 7])
def test_second_order_free_surface_conditions_closed_form(n2):
    """solveHOS_SWEET_2's free-surface conditions (resol_wmkr.f90:302-318) on a single
    first-order mode phi1 = a cosh(k(z+1))/cosh k cos(kx x) cos(ky y), eta1 = e cos cos, with a
    still paddle (no additional potential): the Taylor-expanded second-order forcing
      deta2 = -phi1_x eta1_x - phi1_y eta1_y + eta1 phi1_zz,
      da2   = S2F(-(phi1_x^2 + phi1_y^2 + phi1_z^2)/2 - eta1 phi1_zt),  phi1_zt = -e k tanh k cc
    (da1 = -eta1 without beach) in closed form at the nodes."""
    n1, m, nm = 16, 3, (2 if n2 > 1 else 0)
    p = ho.Params(n1=n1, n2=n2, mHOS=3, xlen=6.0, ylen=4.0, n3=9, icase=2, i_wmk=2, iabsnb=0,
                  pos_ini=np.zeros((n2, 0), complex), freq=np.zeros(0))
    o = ho.Oracle(p)
    tb = o.tb
    a, e = 0.013, 0.021
    a1 = np.zeros((n1, n2))
    a1[m, nm] = a
    kx, ky = tb.kx[m], tb.ky[nm]
    k = math.hypot(kx, ky)
    X, Y = np.meshgrid(tb.cx, tb.cy if n2 > 1 else np.zeros(1), indexing="ij")
    cx, sx = np.cos(kx * X), np.sin(kx * X)
    cy, sy = np.cos(ky * Y), np.sin(ky * Y)
    e1 = e * cx * cy
    z1, za = np.zeros((n1, n2)), np.zeros((p.n3_add, n2))
    o.time = 0.3
    o.solve_hos2(a1, e1, z1, z1, za, 0.01, a2nd=z1, eta2nd=z1, a_add_2=za)
    da2, deta2, d2 = o.second
    th = math.tanh(k)
    px, py, pz = -a * kx * sx * cy, -a * ky * cx * sy, a * k * th * cx * cy
    ex, ey = -e * kx * sx * cy, -e * ky * cx * sy
    pzz, pzt = a * k * k * cx * cy, -e * k * th * cx * cy
    want_deta = -px * ex - py * ey + e1 * pzz
    want_tmp = -0.5 * (px ** 2 + py ** 2 + pz ** 2) - e1 * pzt
    np.testing.assert_allclose(d2, 0.0, atol=1e-18)
    np.testing.assert_allclose(deta2, want_deta, atol=1e-15)
    np.testing.assert_allclose(o.f2s(da2), want_tmp, atol=1e-15)
    assert np.abs(want_deta).max() > 1e-5 and np.abs(want_tmp).max() > 1e-5


@pytest.mark.parametrize("n2", [1, 6])
def test_second_order_paddle_condition_is_a_time_derivative(n2):
    """solveHOS_SWEET_2's paddle condition (resol_wmkr.f90:264-277) is
    d2 = -d/dt[alpha (-X phi_xx + X_y phi_y + X_z phi_z)] on the paddle, the time derivative
    taken along the first-order solution (a1 + t da1, a_add_1 + t da_add_1, X(t)).  Checked by
    centred differences in t of that bracket, evaluated by solveHOS_SWEET_1 / d1st_for_2nd at
    shifted times and states, then continued, transformed and scaled as the reference does."""
    n1 = 14
    rng = np.random.default_rng(31)
    nf = 2
    pos = (0.02 * rng.standard_normal((n2, nf)) + 0.02j * rng.standard_normal((n2, nf)))
    dpos = (0.05 * rng.standard_normal((n2, nf)) + 0.05j * rng.standard_normal((n2, nf))) if n2 > 1 else \
        np.zeros((1, nf), complex)
    p = ho.Params(n1=n1, n2=n2, mHOS=3, xlen=5.0, ylen=3.0, n3=9, icase=3, i_wmk=2, igeom=2, d_hinge=0.3,
                  iramp=2, Tramp=4.0, iabsnb=0, freq=np.array([0.31, 0.47]), pos_ini=pos, dposdy_ini=dpos)
    o = ho.Oracle(p)
    tb = o.tb
    dec = np.exp(-0.4 * tb.k)
    a1 = 0.01 * rng.standard_normal((n1, n2)) * dec
    e1 = 0.01 * rng.standard_normal((n1, n2)) * dec
    aad = 0.003 * rng.standard_normal((p.n3_add, n2)) * np.exp(-0.3 * np.arange(p.n3_add))[:, None]
    z1, za = np.zeros((n1, n2)), np.zeros((p.n3_add, n2))
    t0 = 1.3
    o.time = t0
    o.solve_hos2(a1, e1, z1, z1, aad, 0.01, a2nd=z1, eta2nd=z1, a_add_2=za)
    d2 = o.second[2]
    o.time = t0
    da1, _, d1, _ = o.solve_sweet_1(a1, e1, aad)

    def bracket(dt):
        o.time = t0 + dt
        o.solve_sweet_1(a1 + dt * da1, e1, aad + dt * d1)
        f = o.first
        dy, _, dz, _ = ho.dpos_1st(o.time, p, tb)
        return o.alpha * (-o.pos1st * f["axx"] + dy * f["ay"] + dz * f["az"])

    h = 1e-4
    fd = np.zeros((p.n3_add, n2))
    fd[:p.n3] = -(bracket(h) - bracket(-h)) / (2 * h)
    fd = ho.match_surface_wmk(fd, p.n3, tb.delz)
    fd = o.FA.space_2_fourier(fd) / tb.k_add_thk_add
    scale = np.abs(d2).max()
    assert scale > 1e-6
    np.testing.assert_allclose(d2, fd, atol=1e-6 * scale)


@pytest.mark.parametrize("igeom,d", [(1, 0.0), (2, 0.0), (2, 0.3), (2, 0.55)])
def test_dgzeddz_is_the_derivative_of_gzed(igeom, d):
    """dgzeddz (wavemaking.f90:219-250) against centred differences of gzed, including the
    smoothed hinge region."""
    h = 1e-6
    for z in np.linspace(-0.999, -0.001, 37):
        fd = (ho.gzed(z + h, igeom, d) - ho.gzed(z - h, igeom, d)) / (2 * h)
        assert abs(ho.dgzeddz(z, igeom, d) - fd) < 1e-6, (z, igeom, d)


def _wmk3_params(n2, **kw):
    rng = np.random.default_rng(41)
    nf = 2
    pos = 0.02 * rng.standard_normal((n2, nf)) + 0.02j * rng.standard_normal((n2, nf))
    dpos = (0.05 * rng.standard_normal((n2, nf)) + 0.05j * rng.standard_normal((n2, nf))) if n2 > 1 else \
        np.zeros((1, nf), complex)
    base = dict(n1=14, n2=n2, mHOS=3, xlen=5.0, ylen=3.0, n3=9, icase=3, i_wmk=3, igeom=2, d_hinge=0.3,
                iramp=2, Tramp=4.0, iabsnb=0, freq=np.array([0.31, 0.47]), pos_ini=pos, dposdy_ini=dpos)
    base.update(kw)
    return ho.Params(**base)


@pytest.mark.parametrize("n2", [1, 6])
def test_third_order_first_order_paddle_derivatives_by_direct_sums(n2):
    """wmk_3rdorder's first-order third derivatives on the paddle (resol_wmkr.f90:982-1012)
    against direct sums of the additional potential
    phi = sum a cos(kx_add (z+1)) cosh(k (x - xlen))/cosh(k xlen) cos(ky y)
    differentiated analytically at x = 0 and evaluated at the physical nodes (no transforms);
    the z-sine field xz is read at node i3+1 (the DST-II index shift)."""
    p = _wmk3_params(n2)
    o = ho.Oracle(p)
    tb = o.tb
    rng = np.random.default_rng(43)
    aad = rng.standard_normal((p.n3_add, n2)) * np.exp(-0.5 * np.arange(p.n3_add))[:, None]
    z1, za = np.zeros((p.n1, n2)), np.zeros((p.n3_add, n2))
    f1, _ = o.wmk_3rdorder(z1, z1, aad, za, za, za)
    kt = np.tanh(tb.k_add * p.xlen)  # cosh'(.)/cosh at x = 0 is -k tanh(k xlen)
    for i3 in range(p.n3 - 1):
        z = tb.cz[i3]
        zs = tb.cz[i3 + 1]
        for i2 in range(n2):
            y = tb.cy[i2]
            xxx = xy = xz = 0.0
            for m in range(p.n3_add):
                for j in range(n2):
                    k, kxa, ky = tb.k_add[m, j], tb.kx_add[m], tb.ky[j]
                    xxx += aad[m, j] * math.cos(kxa * (z + 1.0)) * (-k ** 3 * kt[m, j]) * math.cos(ky * y)
                    xy += aad[m, j] * math.cos(kxa * (z + 1.0)) * (-k * kt[m, j]) * (-ky * math.sin(ky * y))
                    xz += aad[m, j] * (-kxa * math.sin(kxa * (zs + 1.0))) * (-k * kt[m, j]) * math.cos(ky * y)
            assert abs(f1["xxx"][i3, i2] - xxx) < 1e-9 * max(1.0, abs(xxx)), (i3, i2)
            assert abs(f1["xy"][i3, i2] - xy) < 1e-9 * max(1.0, abs(xy)), (i3, i2)
            assert abs(f1["xz"][i3, i2] - xz) < 1e-9 * max(1.0, abs(xz)), (i3, i2)


@pytest.mark.parametrize("n2", [1, 6])
def test_third_order_paddle_condition_is_a_time_derivative(n2):
    """solveHOS_SWEET_3's paddle condition (resol_wmkr.f90:366-396) is
    d3 = -d/dt[alpha (-X phi2_xx + X_y phi2_y + X_z phi2_z)
               + alpha^2 (-X^2/2 phi1_xxx + X X_y phi1_xy + X X_z phi1_xz)]
    along the solution (a2nd + t da2nd, a_add_2 + t da_add_2, a_add_1 + t da_add_1, X(t)),
    checked by centred differences in t through wmk_3rdorder at shifted states."""
    p = _wmk3_params(n2)
    o = ho.Oracle(p)
    tb = o.tb
    rng = np.random.default_rng(44)
    dec = np.exp(-0.4 * tb.k)
    a1 = 0.01 * rng.standard_normal((p.n1, n2)) * dec
    e1 = 0.01 * rng.standard_normal((p.n1, n2)) * dec
    a2 = 0.004 * rng.standard_normal((p.n1, n2)) * dec
    e2 = 0.004 * rng.standard_normal((p.n1, n2)) * dec
    w = np.exp(-0.3 * np.arange(p.n3_add))[:, None]
    aad = 0.003 * rng.standard_normal((p.n3_add, n2)) * w
    aad2 = 0.002 * rng.standard_normal((p.n3_add, n2)) * w
    aad3 = 0.001 * rng.standard_normal((p.n3_add, n2)) * w
    z1 = np.zeros((p.n1, n2))
    t0 = 1.3
    o.time = t0
    r = o.solve_hos2(a1, e1, z1, z1, aad, 0.01, a2nd=a2, eta2nd=e2, a_add_2=aad2, a_add_3=aad3)
    d1 = r[4]
    da2, _, d2 = o.second
    d3 = o.third

    def bracket(dt):
        t = t0 + dt
        al, _, _ = ho.ramp(t, p)
        X, _, _ = ho.pos_1st(t, p, tb)
        Xy, _, Xz, _ = ho.dpos_1st(t, p, tb)
        f1, f2 = o.wmk_3rdorder(a2 + dt * da2, da2, aad + dt * d1, d1, aad2 + dt * d2, d2)
        return (al * (-X * f2["axx"] + Xy * f2["ay"] + Xz * f2["az"])
                + al ** 2 * (-(X ** 2) / 2.0 * f1["xxx"] + X * Xy * f1["xy"] + X * Xz * f1["xz"]))

    h = 1e-4
    fd = np.zeros((p.n3_add, n2))
    fd[:p.n3] = -(bracket(h) - bracket(-h)) / (2 * h)
    fd = ho.match_surface_wmk(fd, p.n3, tb.delz)
    fd = o.FA.space_2_fourier(fd) / tb.k_add_thk_add
    scale = np.abs(d3).max()
    assert scale > 1e-7
    np.testing.assert_allclose(d3, fd, atol=1e-6 * scale)


def test_third_order_rejects_filters():
    with pytest.raises(ValueError):
        _wmk3_params(1, coeffilt=(0.8, 1.0, 1.0))
