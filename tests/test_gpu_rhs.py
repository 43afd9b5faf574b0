"""One solveHOS2 evaluation on the GPU (FP64) vs the oracle: relative L2 <= 1e-10.

Covers the 2-D piston tank (M=3, M=5), the 3-D tank with a directional
snake wavemaker (M=5), a hinged flap, sloshing without wavemaker (M=5), and
the spectral filters of filtering.f90 on the free-surface and paddle modes.
Each case calls the RHS twice so that the mean-volume fix, which reads the
wavemaker state of the previous evaluation, is exercised with nonzero values.
"""
import numpy as np
import pytest

from helpers import make_setup, oracle_params, random_state, rel
from hosnwt import config as hc
from hosnwt.tank import Tank
from oracle import hos_oracle as ho

TOL = 1e-10

CASES = {
    "2d_piston_M3": dict(n1=64, n2=1, M=3, n3=9),
    "2d_piston_M5": dict(n1=48, n2=1, M=5, n3=9),
    "2d_flap_M4": dict(n1=40, n2=1, M=4, n3=9, igeom=2, d_hinge=0.3),
    "3d_snake_M5": dict(n1=24, n2=12, M=5, n3=9, xlen=8.0, ylen=4.0),
    "3d_slosh_M5": dict(n1=20, n2=10, M=5, n3=9, xlen=4.0, ylen=2.0, icase=1),
    "2d_M1": dict(n1=32, n2=1, M=1, n3=9),
    # spectral filters (filtering.f90): shaped x filter, cut-off y filter, paddle z filter
    "3d_filtered_M5": dict(n1=24, n2=12, M=5, n3=9, xlen=8.0, ylen=4.0, coeffilt=(-3.0, 0.7, 0.8)),
    "2d_filtered_M3": dict(n1=40, n2=1, M=3, n3=9, coeffilt=(-4.0, 1.0, -2.0)),
}


def directional(kw):
    if kw.get("n2", 1) > 1 and kw.get("icase", 2) == 2:
        comps = hc.jonswap_components(0.05, 1.4, 3.3, n_freq=6, seed=3, rnum=6, clock=8.0, n_dir=3)
        kw = dict(kw, icase=3, components=comps)
    return kw


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_rhs_parity(name):
    kw = directional(CASES[name])
    s = make_setup(**kw)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        for call, (t, dt, seed) in enumerate([(0.7, 0.05, 1), (1.9, 0.02, 2)]):
            G1, G2, a1, e1, aad = random_state(s, seed)
            orc.time = t
            r = orc.solve_hos2(a1, e1, G2, G1, aad, dt)
            g = tank.solve_hos2(t, dt, G1, G2, a1, e1, aad)
            ref_dG2 = r[1].copy()
            ref_dG2[0, 0] = 0.0  # k_G2(1,1) = 0 (runge_kutta.f90:150)
            errs = {
                "dG1": rel(g["dG1"], r[0]), "dG2": rel(g["dG2"], ref_dG2),
                "da1st": rel(g["da1st"], r[2]), "deta1st": rel(g["deta1st"], r[3]),
                "da_add_1": rel(g["da_add_1"], r[4]), "da_add_ramp": rel(g["da_add_ramp"], r[5]),
            }
            print(name, call, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
            for k, v in errs.items():
                assert v < TOL, (name, call, k, v)
    finally:
        tank.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n2", [1, 10])
def test_check_slope_filters_committed_state(n2):
    """hos_check_slope with a threshold below the slopes filters eta and phis exactly
    like the oracle's check_slope (filtering_x, then filtering_y, coefficient 0.7)."""
    kw = dict(n1=32, n2=n2, M=3, n3=9, xlen=6.0, ylen=3.0, icase=1)
    s = make_setup(**kw)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        rng = np.random.default_rng(3)
        eta = 0.01 * rng.standard_normal((32, n2))
        phis = 0.01 * rng.standard_normal((32, n2))
        z = np.zeros((32, n2))
        za = np.zeros((2 * (9 - 1), n2))
        tank.set_state(eta, phis, za, za, z, z)
        tank.runge4_adapt2(0.0, 0.01)  # any right-hand side fills the slope arrays
        sx, sy = tank.slope_max()
        assert sx > 0.0
        fx, fy = tank.check_slope(threshold=1e9)
        assert not fx and not fy
        fx, fy = tank.check_slope(threshold=0.0)
        assert fx and fy == (n2 > 1)
        st = tank.get_state()
        e_ref = orc.filtering(eta, 0.7, 0)
        p_ref = orc.filtering(phis, 0.7, 0)
        if n2 > 1:
            e_ref = orc.filtering(e_ref, 0.7, 1)
            p_ref = orc.filtering(p_ref, 0.7, 1)
        assert rel(st["eta"], e_ref) < TOL and rel(st["phis"], p_ref) < TOL
    finally:
        tank.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n2", [1, 10])
def test_slope_max_against_the_reference_grid(n2):
    """The reference takes MAXVAL(ABS(etax)) over its extended grid Nd = (p1+1) n1 / 2
    (initial_condition.f90:694-696); the GPU samples the same trigonometric polynomial on
    its padded grid N1b x N2b (a smooth length).  Both maxima lie within
    (hx^2 max|f_xx| + 2 hx hy max|f_xy| + hy^2 max|f_yy|) / 8 of the true maximum (at an
    interior maximum the gradient vanishes and a node lies within half a cell per axis),
    with the second derivatives bounded by the sums of |modes| times kx^2, kx ky, ky^2.
    The test checks the two maxima against the sum of both grids' bounds."""
    kw = dict(n1=32, n2=n2, M=3, n3=9, xlen=6.0, ylen=3.0)
    s = make_setup(**kw)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        G1, G2, a1, e1, aad = random_state(s, 11, amp=0.05)
        orc.time = 0.4
        orc.solve_hos2(a1, e1, G2, G1, aad, 0.02)
        tank.solve_hos2(0.4, 0.02, G1, G2, a1, e1, aad)
        sx, sy = tank.slope_max()
        st = tank.stats()
        Nb1, Nb2 = st["big_n1"], st["big_n2"]
        p = orc.p
        assert (Nb1, Nb2) != (p.Nd1, p.Nd2)  # the grids differ
        kx = np.arange(p.Nd1) * np.pi / s.xlen
        ky = np.arange(p.Nd2) * np.pi / s.ylen if n2 > 1 else np.zeros(1)

        def bound(hx, hy, b):
            ab = np.abs(b)
            return (hx * hx * (ab * kx[:, None] ** 2).sum() + 2 * hx * hy * (ab * np.outer(kx, ky)).sum()
                    + hy * hy * (ab * ky[None, :] ** 2).sum()) / 8.0

        hx_r, hx_g = s.xlen / (p.Nd1 - 1), s.xlen / (Nb1 - 1)
        hy_r = s.ylen / (p.Nd2 - 1) if n2 > 1 else 0.0
        hy_g = s.ylen / (Nb2 - 1) if n2 > 1 else 0.0
        bx = orc.s2f_big(orc.etax, "sin", "cos")
        sx_ref = float(np.abs(orc.etax).max())
        lim_x = bound(hx_r, hy_r, bx) + bound(hx_g, hy_g, bx)
        print(f"n2={n2} grids {p.Nd1}x{p.Nd2} / {Nb1}x{Nb2}: max|eta_x| {sx:.6e} vs {sx_ref:.6e} "
              f"(diff {abs(sx - sx_ref):.2e}, bound {lim_x:.2e})", flush=True)
        assert abs(sx - sx_ref) <= lim_x + 1e-12 * sx_ref
        if n2 > 1:
            by = orc.s2f_big(orc.etay, "cos", "sin")
            sy_ref = float(np.abs(orc.etay).max())
            lim_y = bound(hx_r, hy_r, by) + bound(hx_g, hy_g, by)
            print(f"   max|eta_y| {sy:.6e} vs {sy_ref:.6e} (diff {abs(sy - sy_ref):.2e}, bound {lim_y:.2e})",
                  flush=True)
            assert abs(sy - sy_ref) <= lim_y + 1e-12 * sy_ref
    finally:
        tank.close()
