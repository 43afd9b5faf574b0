"""Pins the oracle's SWENSE output (init_velocity_comp + HOSvel2_SWEET + HOSvel_SWEET,
velocities.f90:40-784) against what the mathematics fixes.

  * at a flat surface every re-expansion vanishes: the velocity modes at z = 0 are the
    first-order ones, -kx phis_hat (x) and k tanh k phis_hat (z), and phi_t's are those of
    d phis / dt;
  * for an exact Laplace solution phi = B cosh(k (z + 1)) / cosh k cos(k x) sampled on a
    curved surface eta, the x and z velocities at z = 0 approach the exact B k sin, B k tanh k
    cos like (kA)^(M+1): a dropped order, a wrong sign of a term or a transposed kind of the
    z multiplier (k^j / tanh k) spoils the convergence;
  * the H2 levels are a numerical device: 1 and 10 levels agree to within that truncation.
"""
import math

import numpy as np
import pytest

from oracle import hos_oracle as ho


def params(M, n1=48, xlen=2 * math.pi, **kw):
    base = dict(n1=n1, n2=1, mHOS=M, xlen=xlen, n3=9, icase=2, igeom=1, iramp=1, Tramp=10.0)
    base.update(kw)
    return ho.Params(**base)


def run(orc, eta, phis, meta2=10, t=0.0):
    p = orc.p
    G1 = orc.s2f(eta)
    G2 = orc.s2f(phis) * orc.tb.omega
    z = np.zeros((p.n1, p.n2))
    za = np.zeros((p.n3_add, p.n2))
    orc.time = t
    orc.solve_hos2(z, z, G2, G1, za, 0.0)  # wavemaker globals of the previous evaluation
    return orc.hosvel2_sweet(meta2, t, 0.0, G1, G2, eta, phis, z, z, (za, None, None))


def test_flat_surface_gives_first_order_velocities():
    orc = ho.Oracle(params(3, n1=32))
    rng = np.random.default_rng(2)
    kx = orc.tb.kx
    a = 0.01 * rng.standard_normal((32, 1)) * np.exp(-0.3 * kx)[:, None]
    a[0, 0] = 0.0
    phis = orc.f2s(a)
    eta = np.zeros_like(phis)
    sx, sy, sz, st = run(orc, eta, phis)
    ex = -kx[:, None] * a
    ex[0] = ex[-1] = 0.0  # a sine series has no end modes
    np.testing.assert_allclose(sx, ex, atol=1e-14)
    np.testing.assert_allclose(sz, orc.tb.kth[:, :, 0] * a, atol=1e-14)
    # phi_t is not moved either: the modes of its value at z = eta, and there it is
    # d phis / dt - W d eta / dt (chain rule on phis = phi(eta)), d eta / dt = W
    G1, G2 = orc.s2f(eta), orc.s2f(phis) * orc.tb.omega
    z, za = np.zeros_like(eta), np.zeros((orc.p.n3_add, 1))
    phim = orc.init_velocity_comp(0.0, 0.0, G1, G2, eta, phis, z, z, (za, None, None))
    np.testing.assert_allclose(st, orc.s2f(phim[3]), atol=1e-14)
    W = orc.f2s(orc.tb.kth[:, :, 0] * a)
    np.testing.assert_allclose(orc.f2s(orc.modesFSt), W, atol=1e-13)


@pytest.mark.parametrize("kA", [0.05, 0.1])
def test_velocities_at_z0_converge_to_the_laplace_solution(kA):
    k, B = 3.0, 0.02
    errs = {}
    for M in (2, 3, 4, 5):
        orc = ho.Oracle(params(M, n1=64))
        x = orc.tb.cx[:, None]
        A = kA / k
        eta = A * np.cos(k * x)
        phis = B * np.cosh(k * (eta + 1.0)) / np.cosh(k) * np.cos(k * x)
        sx, _, sz, _ = run(orc, eta, phis, meta2=10)
        ux = orc.f2s(sx, "sin", "cos")
        uz = orc.f2s(sz)
        ex = -B * k * np.sin(k * x)
        ez = B * k * math.tanh(k) * np.cos(k * x)
        errs[M] = (np.linalg.norm(ux - ex) / np.linalg.norm(ex), np.linalg.norm(uz - ez) / np.linalg.norm(ez))
    for f in (0, 1):
        seq = [errs[M][f] for M in (2, 3, 4, 5)]
        assert seq[-1] < 5e-3
        # each order gains about a factor kA (at least a third of it here)
        for a, b in zip(seq, seq[1:]):
            assert b < a * max(3 * kA, 0.5), (f, seq)


def test_levels_agree_within_the_truncation():
    orc = ho.Oracle(params(4, n1=48))
    x = orc.tb.cx[:, None]
    # two modes of a Laplace solution on a two-mode surface
    eta = 0.02 * np.cos(2.0 * x) + 0.01 * np.cos(4.0 * x)
    phis = sum(b * np.cosh(k * (eta + 1.0)) / np.cosh(k) * np.cos(k * x) for b, k in ((0.01, 2.0), (0.004, 4.0)))
    r10 = run(orc, eta, phis, meta2=10)
    r1 = run(orc, eta, phis, meta2=1)
    # (phi_t carries the right-hand side's truncation too: a looser bound)
    for f, bound in ((0, 1e-4), (2, 1e-4), (3, 1e-3)):
        d = np.linalg.norm(r10[f] - r1[f]) / np.linalg.norm(r10[f])
        assert d < bound, (f, d)
