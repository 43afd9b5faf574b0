"""Second-order wavemaker (i_wmk = 2) on the GPU vs the oracle.

solveHOS2 with solveHOS_SWEET_1 + d1st_for_2nd + solveHOS_SWEET_2
(resol_wmkr.f90:109-329, :645-864) on seeded random first- and second-order
states: the HOS derivatives, the first-order wavemaker outputs and the
second-order ones (da2nd modal, deta2nd physical, da2nd_add paddle modes) must
agree with the oracle to relative L2 <= 1e-10 in FP64.  Flap and piston
paddles, 2-D and directional 3-D tanks (nonzero dposdy_ini), with and without
the x / y / z spectral filters.  RK replay: the GPU replays the oracle's
(h, accept) decisions and the first- and second-order states must agree.
"""
import numpy as np
import pytest

from helpers import make_setup, oracle_params, random_state, rel
from hosnwt import config as hc
from hosnwt.tank import Tank
from oracle import hos_oracle as ho

TOL = 1e-10

CASES = {
    "2d_piston_M3": dict(n1=64, n2=1, M=3, n3=9, xlen=20.0),
    "2d_flap_M5_filtered": dict(n1=48, n2=1, M=5, n3=9, xlen=12.0, igeom=2, d_hinge=0.3, T=1.2, Tramp=2.0,
                                coeffilt=(0.8, 1.0, 0.9)),
    "3d_directional_M5": dict(n1=24, n2=12, M=5, n3=9, xlen=8.0, ylen=4.0, Tramp=1.5),
    "3d_directional_M3_filtered": dict(n1=32, n2=16, M=3, n3=9, xlen=10.0, ylen=5.0, Tramp=1.5,
                                       coeffilt=(0.7, 0.8, 0.9)),
    # odd n2: the wavemaker y sums pair node i2 with n2-1-i2; the middle node is its own mirror
    "3d_directional_M3_odd_n2": dict(n1=24, n2=13, M=3, n3=9, xlen=8.0, ylen=4.0, Tramp=1.5),
}


def setup_for(name):
    kw = dict(CASES[name])
    if kw.get("n2", 1) > 1:
        kw["components"] = hc.jonswap_components(0.06, 1.4, 3.3, n_freq=5, seed=11, rnum=6, clock=8.0, n_dir=3,
                                                 spread_s=10)
        kw["icase"] = 3
    return make_setup(i_wmk=2, **kw)


def second_state(s, seed):
    G1, G2, a2, e2, aa2 = random_state(s, seed, amp=0.004)
    return a2, e2, aa2


def test_setup_has_transverse_paddle_derivative():
    s = setup_for("3d_directional_M5")
    assert s.dposdy_ini.shape == s.pos_ini.shape
    assert np.abs(s.dposdy_ini).max() > 0.0
    # central difference of pos_ini(y) with a small shift of the paddle nodes
    y = np.arange(s.cfg.n2) * s.ylen / (s.cfg.n2 - 1)
    d = 1e-5
    _, pp, _ = hc.wmk_ini(s.waves, y + d)
    _, pm, _ = hc.wmk_ini(s.waves, y - d)
    assert rel((pp - pm) / (2 * d), s.dposdy_ini) < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_rhs2_parity(name):
    s = setup_for(name)
    print(name, flush=True)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        t = 0.6 * s.Tramp
        G1, G2, a1, e1, aad = random_state(s, 3, amp=0.01)
        a2, e2, aa2 = second_state(s, 4)
        orc.time = t
        r = orc.solve_hos2(a1, e1, G2, G1, aad, 0.01, a2nd=a2, eta2nd=e2, a_add_2=aa2)
        da2, deta2, d2 = orc.second
        g = tank.solve_hos2_second(t, 0.01, G1, G2, a1, e1, aad, a2, e2, aa2)
        ref_dG2 = r[1].copy()
        ref_dG2[0, 0] = 0.0  # k_G2(1,1) = 0 (runge_kutta.f90:150)
        errs = {"dG1": rel(g["dG1"], r[0]), "dG2": rel(g["dG2"], ref_dG2), "da1st": rel(g["da1st"], r[2]),
                "deta1st": rel(g["deta1st"], r[3]), "da_add_1": rel(g["da_add_1"], r[4]),
                "da_add_ramp": rel(g["da_add_ramp"], r[5]), "da2nd": rel(g["da2nd"], da2),
                "deta2nd": rel(g["deta2nd"], deta2), "da_add_2": rel(g["da_add_2"], d2)}
        print(name, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert np.abs(d2).max() > 0.0 and np.abs(deta2).max() > 0.0
        for k, v in errs.items():
            assert v < TOL, (name, k, v)
        # hos_rhs on an i_wmk = 2 context: the second-order state taken as zero
        r0 = orc.solve_hos2(a1, e1, G2, G1, aad, 0.01)
        g0 = tank.solve_hos2(t, 0.01, G1, G2, a1, e1, aad)
        ref0 = r0[1].copy()
        ref0[0, 0] = 0.0
        assert rel(g0["dG2"], ref0) < TOL
        assert rel(g0["da1st"], r0[2]) < TOL
    finally:
        tank.close()


@pytest.mark.gpu
def test_rhs2_mixed_precision():
    """FP32 transforms: the wavemaker chain stays FP64, so the second-order outputs
    keep FP64 accuracy except through the FP32 surface elevation of phi_add."""
    name = "3d_directional_M5"
    s = setup_for(name)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=1)
    try:
        t = 0.6 * s.Tramp
        G1, G2, a1, e1, aad = random_state(s, 3, amp=0.01)
        a2, e2, aa2 = second_state(s, 4)
        orc.time = t
        r = orc.solve_hos2(a1, e1, G2, G1, aad, 0.01, a2nd=a2, eta2nd=e2, a_add_2=aa2)
        da2, deta2, d2 = orc.second
        g = tank.solve_hos2_second(t, 0.01, G1, G2, a1, e1, aad, a2, e2, aa2)
        errs = {"dG1": rel(g["dG1"], r[0]), "da2nd": rel(g["da2nd"], da2), "deta2nd": rel(g["deta2nd"], deta2),
                "da_add_2": rel(g["da_add_2"], d2)}
        print(name, "mixed", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert errs["dG1"] < 1e-5
        for k in ("da2nd", "deta2nd", "da_add_2"):
            assert errs[k] < 1e-9, (k, errs[k])
    finally:
        tank.close()


def replay(tank, log, Tramp, iramp=1):
    t = 0.0
    errs = []
    endramp = 0
    for h, acc, _ in log:
        errs.append(tank.runge4_adapt2(t, h))
        if acc:
            tank.commit()
            if iramp == 1 and Tramp - (t + h) <= 2.2e-16 and Tramp > t and endramp != 1:
                tank.end_ramp()
                endramp = 1
            t += h
    return t, errs


# Replays end away from the half periods of the monochromatic flap, where the
# 2 omega part of the second-order paddle modes has integrated back to zero.
REPLAY = {"2d_flap_M5": (dict(n1=48, n2=1, M=5, n3=9, xlen=12.0, igeom=2, d_hinge=0.3, T=1.2, Tramp=2.0), 0.4),
          "3d_directional_M5": (dict(CASES["3d_directional_M5"]), 0.3)}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(REPLAY))
def test_step_replay_second_order(name):
    kw, frac = REPLAY[name]
    kw = dict(kw, toler=1e-6)
    if kw.get("n2", 1) > 1:
        kw["components"] = hc.jonswap_components(0.06, 1.4, 3.3, n_freq=5, seed=11, rnum=6, clock=8.0, n_dir=3,
                                                 spread_s=10)
        kw["icase"] = 3
    s = make_setup(i_wmk=2, **kw)
    print(name, "replay", flush=True)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        st = orc.initialize()
        tank.initialize()
        log = []
        t_next = frac * s.delt_out
        h0 = min(0.95 * s.dt_stab, s.delt_out)
        st, t_end, _, na, nr, deta = ho.advance(orc, st, 0.0, t_next, h0, s.cfg.toler, s.dt_stab, s.delt_out, log=log)
        t_gpu, errs = replay(tank, log, s.Tramp, s.cfg.iramp)
        assert abs(t_gpu - t_end) < 1e-12
        e_err = max(abs(a - b[2]) / max(abs(b[2]), 1e-300) for a, b in zip(errs, log))
        g = tank.get_state()
        g2 = tank.get_state2()
        res = {"eta": rel(g["eta"], st.eta), "phis": rel(g["phis"], st.phis),
               "a1st_add": rel(g["a1st_add"], st.a1st_add), "a1st": rel(g["a1st"], st.a1st),
               "eta1st": rel(g["eta1st"], st.eta1st), "a2nd_add": rel(g2["a2nd_add"], st.a2nd_add),
               "a2nd": rel(g2["a2nd"], st.a2nd), "eta2nd": rel(g2["eta2nd"], st.eta2nd)}
        print(name, f"attempts={len(log)} acc={na} rej={nr} err_rel={e_err:.2e}",
              {k: f"{v:.2e}" for k, v in res.items()}, flush=True)
        assert np.abs(st.a2nd_add).max() > 1e-3 * np.abs(st.a1st_add).max()
        assert e_err < 1e-6
        for k, v in res.items():
            assert v < TOL, (name, k, v)
    finally:
        tank.close()
