"""SWENSE output on the GPU (hos_swense_modes) against the oracle.

The reference writes, at each output time, the modes of the velocities and of phi_t at z = 0
obtained from the free-surface state by the H2 operator (HOSvel2_SWEET with 10 levels, then
HOSvel_SWEET; IO/velocities.f90:40-784), with the modal eta and d eta / dt and the paddle
modes (IO/output.f90:417-467).  Its re-expansions keep grid-dependent (aliased) content, so
the GPU runs them on the reference's own extended grid; the bar is relative L2 <= 1e-10 on
every record, in 2-D and 3-D, first- and second-order wavemaker.
"""
import numpy as np
import pytest

from helpers import make_setup, oracle_params, random_state, rel
from hosnwt import config as hc
from hosnwt.tank import Tank
from oracle import hos_oracle as ho

TOL = 1e-10

CASES = {
    "2d_piston_M3": dict(n1=32, n2=1, M=3, n3=9, xlen=10.0),
    "2d_flap_M5": dict(n1=24, n2=1, M=5, n3=9, xlen=8.0, igeom=2, d_hinge=0.3),
    "2d_piston_M4_wmk2": dict(n1=24, n2=1, M=4, n3=9, xlen=8.0, i_wmk=2),
    "3d_snake_M4": dict(n1=16, n2=8, M=4, n3=9, xlen=6.0, ylen=3.0),
}


def setup_for(name):
    kw = dict(CASES[name])
    if kw.get("n2", 1) > 1:
        kw["components"] = hc.jonswap_components(0.05, 1.4, 3.3, n_freq=5, seed=3, rnum=6, clock=8.0, n_dir=3)
        kw["icase"] = 3
    return make_setup(**kw)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("meta2", [10, 2])
def test_swense_modes_match_oracle(name, meta2):
    s = setup_for(name)
    c = s.cfg
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        G1, G2, a1, e1, aad = random_state(s, 11, amp=0.01)
        t = 1.7
        # one right-hand side first on both sides: the volume fix of the next one reads the
        # wavemaker globals it leaves (resol_HOS.f90:353-366)
        orc.time = t
        orc.solve_hos2(a1, e1, G2, G1, aad, 0.0)
        if c.i_wmk >= 2:
            z, za = np.zeros_like(a1), np.zeros_like(aad)
            tank.solve_hos2_second(t, 0.0, G1, G2, a1, e1, aad, z, z, za)
        else:
            tank.solve_hos2(t, 0.0, G1, G2, a1, e1, aad)
        eta = orc.f2s(G1)
        om = orc.tb.omega.copy()
        om[0, 0] = 1.0
        phis_m = G2 / om
        phis_m[0, 0] = 0.0
        phis = orc.f2s(phis_m)
        e1_phys = e1
        tank.set_state(eta, phis, aad, aad, a1, e1_phys)
        a2 = None
        if c.i_wmk >= 2:
            rng = np.random.default_rng(5)
            a2 = 0.1 * aad * rng.standard_normal(aad.shape)
            tank.set_state2(a2, np.zeros_like(a1), np.zeros_like(e1))
        got = tank.swense_modes(t, meta2)
        G1o = orc.s2f(eta)
        G2o = orc.s2f(phis) * orc.tb.omega
        ref = orc.hosvel2_sweet(meta2, t, 0.0, G1o, G2o, eta, phis, a1, e1_phys, (aad, a2, None),
                                a2nd=np.zeros_like(a1) if a2 is not None else None,
                                eta2nd=np.zeros_like(e1) if a2 is not None else None)
        errs = {"specx": rel(got["specx"], ref[0]), "specz": rel(got["specz"], ref[2]),
                "spect": rel(got["spect"], ref[3]), "FSt": rel(got["FSt"], orc.modesFSt),
                "FS": rel(got["FS"], G1o), "add": rel(got["add"], orc.modesadd),
                "addt": rel(got["addt"], orc.modesaddt)}
        if c.n2 > 1:
            errs["specy"] = rel(got["specy"], ref[1])
        else:
            assert not np.any(got["specy"])
        print(name, meta2, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert np.linalg.norm(ref[0]) > 0 and np.linalg.norm(ref[3]) > 0
        for k, v in errs.items():
            assert v < TOL, (name, meta2, k, v)
    finally:
        tank.close()
