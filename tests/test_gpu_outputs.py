"""Output getters behind the reference's output files (output.f90:232-482), 3-D.

* hos_probes: eta at wave probes = sum_{i1,i2} eta_hat cos(kx x) cos(ky y)
  (mat_coscos, initial_condition.f90:502-510), against the modal transform of
  the physical eta the context returns;
* hos_get_modes: the committed modal eta, phis equal space_2_Fourier of the
  physical fields;
* hos_get_wmk_motion: alpha pos1st(n3, i2) and alpha' pos1st + alpha dpos1stdt
  per y node at the last wavemaker evaluation (t_end of the last accepted
  attempt), against the oracle's ramp and pos_1st at that time.
"""
import numpy as np
import pytest

from helpers import make_setup, oracle_params, rel
from hosnwt import config as hc
from hosnwt.tank import Tank
from oracle import hos_oracle as ho


@pytest.mark.gpu
def test_probes_modes_and_wavemaker_motion_3d():
    comps = hc.jonswap_components(0.06, 1.4, 3.3, n_freq=5, seed=11, rnum=6, clock=8.0, n_dir=3)
    s = make_setup(n1=24, n2=12, M=3, n3=9, xlen=8.0, ylen=4.0, Tramp=1.5, components=comps, icase=3)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        tank.initialize()
        h, na, nr = tank.advance(0.5 * s.delt_out, min(0.95 * s.dt_stab, s.delt_out))
        print("advanced to", tank.time, "attempts", na, nr, flush=True)
        phys = tank.physical()
        F = ho.Fourier(s.cfg.n1, s.cfg.n2)
        eta_hat = F.space_2_fourier(phys["eta"])
        phis_hat = F.space_2_fourier(phys["phis"])
        m = tank.modes()
        assert rel(m["eta"], eta_hat) < 1e-12
        assert rel(m["phis"], phis_hat) < 1e-12
        xp = np.array([0.5, 2.25, 4.0, 7.9])
        yp = np.array([0.0, 1.3, 2.0, 3.7])
        tb = orc.tb
        want = np.array([np.sum(eta_hat * np.outer(np.cos(tb.kx * x), np.cos(tb.ky * y))) for x, y in zip(xp, yp)])
        got = tank.probes(xp, yp)
        print("probes", got, want, flush=True)
        assert np.abs(got - want).max() < 1e-12 * max(1.0, np.abs(want).max())
        assert np.abs(want).max() > 1e-5
        al, alt, x_top, v_top = tank.wmk_motion()
        a_ref, at_ref, _ = ho.ramp(tank.time, orc.p)
        X, Xt, _ = ho.pos_1st(tank.time, orc.p, tb)
        assert abs(al - a_ref) < 1e-14 and abs(alt - at_ref) < 1e-14
        assert rel(x_top, a_ref * X[-1, :]) < 1e-12
        assert rel(v_top, at_ref * X[-1, :] + a_ref * Xt[-1, :]) < 1e-12
    finally:
        tank.close()
