"""RK4(3) steps on the GPU (FP64) vs the oracle.

Parity procedure: the oracle runs the adaptive loop and logs its (h, accept)
decisions; the GPU replays exactly those attempts (so a 1e-14 difference in
the error estimate cannot fork the trajectory), and eta, phis (physical),
the wavemaker modes, volume and energy must agree to relative L2 <= 1e-10.
The embedded error estimate of every attempt (a max-norm of a difference of
nearly equal stage sums) must agree to 1e-6 relative.  Volume is compared
absolutely, scaled by max|eta|, since it is near zero.
A free-running GPU loop (hos_advance) must take the same decisions.
"""
import numpy as np
import pytest

from helpers import make_setup, oracle_params, rel
from hosnwt import config as hc
from hosnwt.tank import Tank
from oracle import hos_oracle as ho

TOL = 1e-10

CASES = {
    "2d_regular_M3": dict(n1=64, n2=1, M=3, n3=9, xlen=20.0),
    "2d_regular_M5_flap": dict(n1=48, n2=1, M=5, n3=9, xlen=12.0, igeom=2, d_hinge=0.3, T=1.2, Tramp=2.0),
    "3d_directional_M5": dict(n1=24, n2=12, M=5, n3=9, xlen=8.0, ylen=4.0, Tramp=1.5),
}


def setup_for(name):
    kw = dict(CASES[name])
    if kw.get("n2", 1) > 1:
        kw["components"] = hc.jonswap_components(0.06, 1.4, 3.3, n_freq=5, seed=11, rnum=6, clock=8.0, n_dir=3)
        kw["icase"] = 3
    return make_setup(**kw)


def replay(tank, log, t0=0.0, Tramp=0.0, iramp=1):
    t = t0
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


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_step_replay_parity(name):
    s = setup_for(name)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        st = orc.initialize()
        tank.initialize()
        g0 = tank.get_state()
        assert rel(g0["a1st_add"], st.a1st_add) < TOL
        log = []
        t_next = 2.5 * s.delt_out if s.cfg.n2 == 1 else 1.2 * s.delt_out
        h0 = min(0.95 * s.dt_stab, s.delt_out)
        st, t_end, _, na, nr, deta = ho.advance(orc, st, 0.0, t_next, h0, s.cfg.toler, s.dt_stab, s.delt_out, log=log)
        t_gpu, errs = replay(tank, log, Tramp=s.Tramp, iramp=s.cfg.iramp)
        assert abs(t_gpu - t_end) < 1e-12
        e_err = max(abs(a - b[2]) / max(abs(b[2]), 1e-300) for a, b in zip(errs, log))
        g = tank.get_state()
        ph = tank.physical()
        vol_o = orc.volume(st.eta)
        en_o = orc.energy(st.eta, st.phis + orc.phi_add, deta)
        vol_g = tank.volume(ph["eta"])
        en_g = tank.energy(ph["eta"], ph["phis"] + ph["phi_add"], ph["deta"])
        res = {"eta": rel(g["eta"], st.eta), "phis": rel(g["phis"], st.phis),
               "a1st_add": rel(g["a1st_add"], st.a1st_add),
               # a_add_ramp passes through zero for a purely imaginary paddle transfer function:
               # measured against the paddle-mode scale
               "a_add_ramp": np.linalg.norm(g["a_add_ramp"] - st.a_add_ramp)
               / max(np.linalg.norm(st.a_add_ramp), np.linalg.norm(st.a1st_add), 1e-300),
               "a1st": rel(g["a1st"], st.a1st), "eta1st": rel(g["eta1st"], st.eta1st),
               "deta": rel(ph["deta"], deta), "phi_add": rel(ph["phi_add"], orc.phi_add),
               "volume": abs(vol_g - vol_o) / max(np.abs(st.eta).max(), 1e-300),
               "energy": abs(en_g[2] - en_o[2]) / max(abs(en_o[2]), 1e-300)}
        print(name, f"attempts={len(log)} acc={na} rej={nr} err_rel={e_err:.2e}",
              {k: f"{v:.2e}" for k, v in res.items()}, flush=True)
        assert e_err < 1e-6
        for k, v in res.items():
            assert v < TOL, (name, k, v)
        # free-running GPU controller over the same interval takes the same decisions
        tank2 = Tank(s, precision=0)
        tank2.initialize()
        h1, na2, nr2 = tank2.advance(t_next, h0)
        assert (na2, nr2) == (na, nr)
        assert rel(tank2.get_state()["eta"], st.eta) < TOL
        tank2.close()
    finally:
        tank.close()
