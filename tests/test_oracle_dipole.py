"""Pins the oracle's dipole wavemaker (igeom = 3) against what the mathematics fixes.

phi_dipoles sums, for each wave component, a line dipole below the free surface, its image
below the bottom and pairs of wall images.  A slip in a sign, an image position or a
derivative term fails one of:
  * each derivative field equals the centred difference of the potential (x, z = eta, t);
  * the potential is harmonic: phi_xx + phi_zz = 0 away from the dipoles;
  * the bottom image makes phi_z vanish at z = -1 exactly;
  * the wall images make phi_x vanish at x = 0 and x = xlen, better with more image pairs;
  * dipoles_ini's depth classes and moments (mu_x = -i mu_z tanh(k (z + 1)), image depth
    -2 - z) and the transfer function's two branches meeting at k = 20.
"""
import math

import numpy as np

from oracle import hos_oracle as ho


def dip_params(n_side=6, **kw):
    base = dict(n1=9, n2=1, mHOS=3, xlen=12.0, n3=5, icase=3, igeom=3, iramp=0)
    base.update(kw)
    p = ho.Params(**base)
    freq = np.array([0.45, 0.55, 0.8, 1.2])
    k = np.array([ho_k(f) for f in freq])
    rng = np.random.default_rng(5)
    p.dip_freq = freq
    p.dip_mu = ho.dipoles_ini(freq, k, 0.01 * (1 + rng.random(4)), rng.uniform(0, 2 * math.pi, 4), p.icase,
                              p.x_dip)
    p.n_side = n_side
    return p


def ho_k(f):
    w2 = (2 * math.pi * f) ** 2
    lo, hi = 0.0, 50.0
    for _ in range(200):
        m = 0.5 * (lo + hi)
        if m * math.tanh(m) > w2:
            hi = m
        else:
            lo = m
    return 0.5 * (lo + hi)


def field(p, x, z, t=0.37, alpha=1.0, alphat=0.0):
    x = np.asarray(x, float).reshape(-1, 1)
    z = np.broadcast_to(np.asarray(z, float).reshape(-1, 1), x.shape).copy()
    return ho.phi_dipoles(z, t, alpha, alphat, p, x)


def test_dipole_derivatives_are_consistent():
    p = dip_params()
    x = np.linspace(0.4, 11.0, 23)
    z = -0.1 + 0.05 * np.sin(x)
    h = 1e-5
    phi, phit, phix, phiz = field(p, x, z)
    # x derivative
    dpx = (field(p, x + h, z)[0] - field(p, x - h, z)[0]) / (2 * h)
    np.testing.assert_allclose(phix, dpx, rtol=1e-6, atol=1e-9 * np.abs(phix).max())
    # z derivative (eta moves the evaluation point)
    dpz = (field(p, x, z + h)[0] - field(p, x, z - h)[0]) / (2 * h)
    np.testing.assert_allclose(phiz, dpz, rtol=1e-6, atol=1e-9 * np.abs(phiz).max())
    # t derivative with alpha = 1, alpha' = 0
    dpt = (field(p, x, z, t=0.37 + h)[0] - field(p, x, z, t=0.37 - h)[0]) / (2 * h)
    np.testing.assert_allclose(phit, dpt, rtol=1e-6, atol=1e-9 * np.abs(phit).max())
    # the ramp: phi_t = alpha' phi + alpha d phi / dt
    _, phit2, _, _ = field(p, x, z, alpha=0.6, alphat=0.3)
    np.testing.assert_allclose(phit2, 0.3 * phi + 0.6 * phit, rtol=1e-12, atol=1e-15)


def test_dipole_potential_is_harmonic():
    p = dip_params()
    x = np.linspace(0.5, 11.5, 31)
    x = x[np.abs(x - p.x_dip) > 0.6]
    z = np.full_like(x, -0.05)
    h = 1e-3
    lap = ((field(p, x + h, z)[0] - 2 * field(p, x, z)[0] + field(p, x - h, z)[0])
           + (field(p, x, z + h)[0] - 2 * field(p, x, z)[0] + field(p, x, z - h)[0])) / (h * h)
    scale = np.abs(field(p, x, z)[0]).max()
    assert np.abs(lap).max() < 1e-4 * scale


def test_dipole_bottom_condition_is_exact():
    p = dip_params()
    x = np.linspace(0.0, p.xlen, 17)
    _, _, _, phiz = field(p, x, -1.0)
    _, _, _, phiz_up = field(p, x, -0.5)
    assert np.abs(phiz).max() < 1e-13 * np.abs(phiz_up).max()


def test_dipole_wall_images_reduce_wall_flux():
    res = []
    for n_side in (2, 6, 30):
        p = dip_params(n_side=n_side)
        z = np.linspace(-0.95, -0.05, 7)
        r = []
        for xw in (0.0, p.xlen):
            _, _, phix, _ = field(p, np.full_like(z, xw), z)
            r.append(np.abs(phix).max())
        _, _, phix_mid, _ = field(p, np.full_like(z, 0.5 * p.xlen), z)
        res.append(max(r) / np.abs(phix_mid).max())
    assert res[0] > res[1] > res[2]
    assert res[2] < 0.1 * res[0]


def test_dipoles_ini_moments_and_depths():
    freq = np.array([0.3, 0.55, 0.65, 1.1])
    k = np.array([ho_k(f) for f in freq])
    amp = np.array([0.01, 0.02, 0.03, 0.04])
    ph = np.array([0.1, 0.2, 0.3, 0.4])
    mu = ho.dipoles_ini(freq, k, amp, ph, 3, 2.5)
    np.testing.assert_array_equal(mu[:, 2].real, [-0.6, -0.4, -0.2, -0.35])
    np.testing.assert_array_equal(mu[:, 3].real, -2.0 - mu[:, 2].real)
    for f in range(4):
        z = mu[f, 2].real
        tf = ho.dipole_transfer(k[f], z)
        np.testing.assert_allclose(abs(mu[f, 1]), amp[f] * abs(tf), rtol=1e-14)
        np.testing.assert_allclose(np.angle(mu[f, 1] / tf), (k[f] * 2.5 + ph[f] + math.pi) % (2 * math.pi)
                                   - math.pi, atol=1e-12)
        np.testing.assert_allclose(mu[f, 0], -1j * mu[f, 1] * math.tanh(k[f] * (z + 1.0)), rtol=1e-15)
    mono = ho.dipoles_ini(freq, k, amp, ph, 2, 2.5, z_dipmono=-0.25)
    np.testing.assert_array_equal(mono[:, 2].real, -0.25)


def test_dipole_transfer_overflows_where_tanh_rounds_to_one():
    # the reference's finite-depth formula divides by k^2 - (k tanh k)^2, which is 0 in
    # FP64 once tanh k = 1 (18.7 < k <= 20): such components are refused by the host
    assert math.isinf(ho.dipole_transfer(19.5, -0.2).real)
    assert math.isfinite(ho.dipole_transfer(18.0, -0.2).real)
    assert math.isfinite(ho.dipole_transfer(20.5, -0.2).real)


def test_dipole_transfer_branches_meet():
    # the deep-water branch is the finite-depth formula's large-k limit: with
    # 1 - tanh^2 k = 1 / cosh^2 k evaluated without cancellation, the two agree at k = 20
    for z in (-0.2, -0.35, -0.6):
        k = 20.0 + 1e-9
        w2 = k * math.tanh(k)
        k2 = k * k
        exact = (k2 / math.cosh(k) ** 2 + w2) / (2.0 * math.sqrt(w2) * (k2 / math.cosh(k) ** 2) * math.cosh(k)
                                                 * math.sinh(k * (z + 1)))
        b = ho.dipole_transfer(k, z).real
        assert abs(b - exact) / abs(exact) < 1e-6
        # and the finite-depth formula, free of cancellation, at a moderate k
        k = 3.0
        w2 = k * math.tanh(k)
        exact = (k * k / math.cosh(k) ** 2 + w2) / (2.0 * math.sqrt(w2) * (k * k / math.cosh(k) ** 2) * math.cosh(k)
                                                     * math.sinh(k * (z + 1)))
        assert abs(ho.dipole_transfer(k, z).real - exact) / exact < 1e-12
