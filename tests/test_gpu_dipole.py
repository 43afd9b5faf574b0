"""Dipole wavemaker (igeom = 3) on the GPU against the oracle.

The reference generates waves with line dipoles below the free surface (phi_dipoles,
resol_wmkr.f90:1492-1647) instead of a paddle; solveHOS2 takes their potential at z = eta
(resol_HOS.f90:494-496).  Checked here:
  * one solveHOS2 evaluation (twice, the second after the first's wavemaker globals) at
    relative L2 <= 1e-10, monochromatic (icase 2) and file components over several dipole
    depths (icase 3), 2-D and 3-D;
  * RK attempts with the oracle's (h, accept) log replayed, and the free-running controller
    taking the same decisions;
  * the mixed path's right-hand side against the FP64 one;
  * the transfer function of the dipole (linear_wave.f90:316-327) by what it is for: a
    linear (M = 1) run radiates a wave of the requested amplitude away from the dipole.
"""
import math

import numpy as np
import pytest

from helpers import make_setup, oracle_params, random_state, rel
from hosnwt import config as hc
from hosnwt.tank import Tank
from oracle import hos_oracle as ho

TOL = 1e-10

CASES = {
    "2d_mono_M3": dict(n1=64, n2=1, M=3, n3=9, xlen=20.0),
    "2d_files_M4": dict(n1=48, n2=1, M=4, n3=9, xlen=16.0, files=True),
    "3d_files_M3": dict(n1=24, n2=12, M=3, n3=9, xlen=10.0, ylen=4.0, files=True),
}


def setup_for(name, **over):
    kw = dict(CASES[name], igeom=3)
    if kw.pop("files", False):
        # components spread over the four depth classes of dipoles_ini (nondimensional
        # frequency 0.3-1.3, without 18.7 < k <= 20 where the transfer function overflows)
        f = np.array([0.30, 0.42, 0.52, 0.56, 0.62, 0.66, 0.8, 1.02])
        rng = np.random.default_rng(17)
        # amplitudes scaled so that every dipole moment stays of order 1e-2 (the deep
        # dipoles of the highest frequencies have transfer functions of order 1e3)
        z = np.where(f > 1.0, -0.35, np.where(f > 0.6, -0.2, np.where(f > 0.5, -0.4, -0.6)))
        tf = np.array([abs(ho.dipole_transfer(hc.wave_number(fi), zi)) for fi, zi in zip(f, z)])
        kw["components"] = dict(harmo=np.arange(1, len(f) + 1), freq=f * math.sqrt(9.81),
                                ampli=0.004 * (1 + rng.random(len(f))) / np.maximum(1.0, tf),
                                angle=np.zeros(len(f)), phase=rng.uniform(0, 2 * math.pi, len(f)), rnum=6,
                                clock=8.0)
        kw["icase"] = 3
    kw.update(over)
    return make_setup(**kw)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_dipole_rhs_parity(name):
    s = setup_for(name)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        for call, (t, dt, seed) in enumerate([(0.7, 0.05, 1), (3.9, 0.02, 2)]):
            G1, G2, a1, e1, aad = random_state(s, seed)
            aad[:] = 0.0  # no paddle modes with dipoles
            orc.time = t
            r = orc.solve_hos2(a1, e1, G2, G1, aad, dt)
            g = tank.solve_hos2(t, dt, G1, G2, a1, e1, aad)
            ref_dG2 = r[1].copy()
            ref_dG2[0, 0] = 0.0  # k_G2(1,1) = 0 (runge_kutta.f90:150)
            errs = {"dG1": rel(g["dG1"], r[0]), "dG2": rel(g["dG2"], ref_dG2),
                    "phi_add": rel(tank.physical()["phi_add"], orc.phi_add)}
            print(name, call, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
            assert np.linalg.norm(orc.phi_add) > 0.0
            for k, v in errs.items():
                assert v < TOL, (name, call, k, v)
            for k in ("da1st", "deta1st", "da_add_1", "da_add_ramp"):
                assert not np.any(g[k]), k
    finally:
        tank.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["2d_mono_M3", "3d_files_M3"])
def test_dipole_step_replay_parity(name):
    s = setup_for(name, Tramp=2.0)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        st = orc.initialize()
        tank.initialize()
        log = []
        t_next = 1.5 * s.delt_out
        h0 = min(0.95 * s.dt_stab, s.delt_out)
        st, t_end, _, na, nr, deta = ho.advance(orc, st, 0.0, t_next, h0, s.cfg.toler, s.dt_stab, s.delt_out, log=log)
        t = 0.0
        errs = []
        for h, acc, _ in log:
            errs.append(tank.runge4_adapt2(t, h))
            if acc:
                tank.commit()
                t += h
        assert abs(t - t_end) < 1e-12
        e_err = max(abs(a - b[2]) / max(abs(b[2]), 1e-300) for a, b in zip(errs, log))
        g = tank.get_state()
        ph = tank.physical()
        res = {"eta": rel(g["eta"], st.eta), "phis": rel(g["phis"], st.phis), "deta": rel(ph["deta"], deta),
               "phi_add": rel(ph["phi_add"], orc.phi_add),
               "energy": abs(tank.energy(ph["eta"], ph["phis"] + ph["phi_add"], ph["deta"])[2]
                             - orc.energy(st.eta, st.phis + orc.phi_add, deta)[2])
               / max(abs(orc.energy(st.eta, st.phis + orc.phi_add, deta)[2]), 1e-300)}
        print(name, f"attempts={len(log)} acc={na} rej={nr} err_rel={e_err:.2e}",
              {k: f"{v:.2e}" for k, v in res.items()}, flush=True)
        assert np.abs(st.eta).max() > 0.0
        assert e_err < 1e-6
        for k, v in res.items():
            assert v < TOL, (name, k, v)
        tank2 = Tank(s, precision=0)
        tank2.initialize()
        _, na2, nr2 = tank2.advance(t_next, h0)
        assert (na2, nr2) == (na, nr)
        assert rel(tank2.get_state()["eta"], st.eta) < TOL
        tank2.close()
    finally:
        tank.close()


@pytest.mark.gpu
def test_dipole_mixed_rhs_close_to_fp64():
    s = setup_for("3d_files_M3")
    t64, t32 = Tank(s, precision=0), Tank(s, precision=1)
    try:
        G1, G2, a1, e1, aad = random_state(s, 4, amp=0.01)
        aad[:] = 0.0
        g64 = t64.solve_hos2(1.3, 0.03, G1, G2, a1, e1, aad)
        g32 = t32.solve_hos2(1.3, 0.03, G1, G2, a1, e1, aad)
        e = max(rel(g32["dG1"], g64["dG1"]), rel(g32["dG2"], g64["dG2"]))
        print(f"mixed vs FP64 dipole RHS rel L2 {e:.2e}", flush=True)
        assert e < 2e-5
    finally:
        t64.close()
        t32.close()


@pytest.mark.gpu
def test_dipole_radiates_the_requested_amplitude():
    """Linear (M = 1) 2-D tank, 20 h long, monochromatic dipole at x = 2.5 h (kh = 1.91,
    kA = 0.02, beach from 16 h): once the ramp is over and the wave front has passed, the
    envelope of eta between 1 and 3 wavelengths from the dipole is the requested amplitude
    (group velocity 0.41 sqrt(g h): the full amplitude reaches 3 wavelengths at t ~ 8 T)."""
    T = 1.5
    s = make_setup(n1=256, n2=1, M=1, n3=9, xlen=20.0, igeom=3, kA=0.02, T=T, Tramp=3 * T, xabsf=0.8, toler=1e-6)
    amp = s.cfg.amp_mono
    k = s.waves.k[0]
    tank = Tank(s, precision=0)
    try:
        tank.initialize()
        x = np.arange(s.cfg.n1) * s.xlen / (s.cfg.n1 - 1)
        sel = (x > s.x_dip + 2 * math.pi / k) & (x < s.x_dip + 3 * 2 * math.pi / k)
        Tn = s.T_mono
        t, h = 0.0, 0.5 * s.dt_stab
        t_meas = 10 * Tn
        env = np.zeros(sel.sum())
        n_out = 40
        for i in range(1, n_out + 1):
            t_next = t_meas + i * Tn / n_out * 2
            h, _, _ = tank.advance(t_next, h)
            eta = tank.get_state()["eta"][:, 0]
            env = np.maximum(env, np.abs(eta[sel]))
        ratio = env.mean() / amp
        print(f"dipole far-field amplitude / requested = {ratio:.4f} (spread {env.std() / amp:.4f})", flush=True)
        assert abs(ratio - 1.0) < 0.05
    finally:
        tank.close()
