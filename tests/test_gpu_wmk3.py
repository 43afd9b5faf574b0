"""Third-order wavemaker (i_wmk = 3) on the GPU vs the oracle.

solveHOS2 with solveHOS_SWEET_1, _2 and _3 (resol_wmkr.f90:109-437, :942-1115)
on seeded random first-, second- and third-order states: every output,
including da_add_3, within relative L2 1e-10 in FP64 (2-D flap, 3-D
directional).  RK replay of the oracle's (h, accept) decisions with the
third-order paddle modes in the state and in the error norm.  Contexts with
filters are refused for i_wmk = 3.
"""
import numpy as np
import pytest

from helpers import make_setup, oracle_params, random_state, rel
from hosnwt import config as hc
from hosnwt.tank import Tank
from oracle import hos_oracle as ho

TOL = 1e-10

CASES = {
    "2d_flap_M5": dict(n1=48, n2=1, M=5, n3=9, xlen=12.0, igeom=2, d_hinge=0.3, T=1.2, Tramp=2.0),
    "2d_piston_M3": dict(n1=64, n2=1, M=3, n3=9, xlen=20.0),
    "3d_directional_M5": dict(n1=24, n2=12, M=5, n3=9, xlen=8.0, ylen=4.0, Tramp=1.5),
}


def setup_for(name, **extra):
    kw = dict(CASES[name], **extra)
    if kw.get("n2", 1) > 1:
        kw["components"] = hc.jonswap_components(0.06, 1.4, 3.3, n_freq=5, seed=11, rnum=6, clock=8.0, n_dir=3,
                                                 spread_s=10)
        kw["icase"] = 3
    return make_setup(i_wmk=3, **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_rhs3_parity(name):
    s = setup_for(name)
    print(name, flush=True)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        t = 0.6 * s.Tramp
        G1, G2, a1, e1, aad = random_state(s, 3, amp=0.01)
        _, _, a2, e2, aad2 = random_state(s, 4, amp=0.004)
        _, _, _, _, aad3 = random_state(s, 5, amp=0.002)
        orc.time = t
        r = orc.solve_hos2(a1, e1, G2, G1, aad, 0.01, a2nd=a2, eta2nd=e2, a_add_2=aad2, a_add_3=aad3)
        da2, deta2, d2 = orc.second
        d3 = orc.third
        g = tank.solve_hos2_third(t, 0.01, G1, G2, a1, e1, aad, a2, e2, aad2, aad3)
        ref_dG2 = r[1].copy()
        ref_dG2[0, 0] = 0.0  # k_G2(1,1) = 0 (runge_kutta.f90:150)
        errs = {"dG1": rel(g["dG1"], r[0]), "dG2": rel(g["dG2"], ref_dG2), "da1st": rel(g["da1st"], r[2]),
                "deta1st": rel(g["deta1st"], r[3]), "da_add_1": rel(g["da_add_1"], r[4]),
                "da_add_ramp": rel(g["da_add_ramp"], r[5]), "da2nd": rel(g["da2nd"], da2),
                "deta2nd": rel(g["deta2nd"], deta2), "da_add_2": rel(g["da_add_2"], d2),
                "da_add_3": rel(g["da_add_3"], d3)}
        print(name, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert np.abs(d3).max() > 0.0
        for k, v in errs.items():
            assert v < TOL, (name, k, v)
        # hos_rhs2 on an i_wmk = 3 context: the third-order modes taken as zero
        orc.time = t
        r0 = orc.solve_hos2(a1, e1, G2, G1, aad, 0.01, a2nd=a2, eta2nd=e2, a_add_2=aad2)
        g0 = tank.solve_hos2_second(t, 0.01, G1, G2, a1, e1, aad, a2, e2, aad2)
        assert rel(g0["dG1"], r0[0]) < TOL
        assert rel(g0["da_add_2"], orc.second[2]) < TOL
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


# tolerances above the h-independent start-up ratio of the a_add_3 error term
REPLAY = {"2d_flap_M5": (3e-6, 0.4), "3d_directional_M5": (1e-6, 0.3)}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(REPLAY))
def test_step_replay_third_order(name):
    toler, frac = REPLAY[name]
    s = setup_for(name, toler=toler)
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
        a3 = tank.get_state3()
        res = {"eta": rel(g["eta"], st.eta), "phis": rel(g["phis"], st.phis),
               "a1st_add": rel(g["a1st_add"], st.a1st_add), "a1st": rel(g["a1st"], st.a1st),
               "a2nd_add": rel(g2["a2nd_add"], st.a2nd_add), "eta2nd": rel(g2["eta2nd"], st.eta2nd),
               "a3rd_add": rel(a3, st.a3rd_add)}
        print(name, f"attempts={len(log)} acc={na} rej={nr} err_rel={e_err:.2e}",
              {k: f"{v:.2e}" for k, v in res.items()}, flush=True)
        assert np.abs(st.a3rd_add).max() > 0.0
        assert e_err < 1e-6
        for k, v in res.items():
            assert v < TOL, (name, k, v)
    finally:
        tank.close()


@pytest.mark.gpu
def test_third_order_with_filters_is_refused():
    s = setup_for("2d_piston_M3", coeffilt=(0.8, 1.0, 1.0))
    with pytest.raises(RuntimeError, match="without filters"):
        Tank(s, precision=0)
