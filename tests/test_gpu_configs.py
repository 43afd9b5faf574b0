"""solveHOS2 parity at the full sizes of the benchmark configurations (FP64 GPU vs the oracle, 1e-10;
the mixed path on the 2-D tanks with a wavemaker, 5e-5 relative L2 at amplitude 1e-2: the FP32
transforms of 6144- to 16384-point lines give 0.8-2.2e-5 on these states, the same to three digits
with the matrix-core product and with the split-K kernel for the mixed-parity filter).

* 2-D regular wave tank: n1=256, M=3, piston, monochromatic wave (T=1.5, kA=0.08),
  beach on the last 20 %.
* 2-D irregular tank: length 60, n1=2048, M=5, flap, JONSWAP (Hs 0.08, Tp 1.6,
  gamma 3.3) with 512 seeded components; its extended x lines have 12288 points.
* 3-D directional basin: 40 x 10, 1024 x 256, M=5, JONSWAP 256 frequencies x 32
  directions (cos^2s, s = 10), beach on the last 15 %; extended grid 3073 x 769.
* 3-D member of the learned-order ensemble: 512 x 128 at M_max = 7.
* 2-D ensemble members: (512, M=3), (1024, M=4), (2048, M=3), and the members whose extended
  lines exceed the in-workgroup engines: (2048, M=6), (4096, M=3), (4096, M=6).

The random modal state has a decaying spectrum so that the products stay well
conditioned; the wavemaker runs at a mid-ramp time.
"""
import numpy as np
import pytest

from helpers import make_setup, oracle_params, random_state, rel
from hosnwt import config as hc
from hosnwt.tank import Tank
from oracle import hos_oracle as ho

TOL = 1e-10

CASES = {
    "regular_2d_256_M3": dict(n1=256, n2=1, M=3, n3=33, xlen=20.0, T=1.5, kA=0.08, Tramp=4.5, xabsf=0.8),
    "irregular_2d_2048_M5": dict(n1=2048, n2=1, M=5, n3=33, xlen=60.0, igeom=2, d_hinge=0.0, Tramp=3.2,
                                 spec=dict(Hs=0.08, Tp=1.6, n_freq=512, n_dir=1, seed=7)),
    "directional_3d_1024x256_M5": dict(n1=1024, n2=256, M=5, n3=33, xlen=40.0, ylen=10.0, Tramp=2.8, xabsf=0.85,
                                       spec=dict(Hs=0.06, Tp=1.4, n_freq=256, n_dir=32, seed=2024)),
    "member_3d_512x128_M7": dict(n1=512, n2=128, M=7, n3=33, xlen=20.0, ylen=5.0, Tramp=2.8,
                                 spec=dict(Hs=0.06, Tp=1.4, n_freq=32, n_dir=8, seed=5)),
    "member_2d_512_M3": dict(n1=512, n2=1, M=3, n3=33, xlen=30.0, Tramp=3.0,
                             spec=dict(Hs=0.05, Tp=1.5, n_freq=64, n_dir=1, seed=11)),
    "member_2d_1024_M4": dict(n1=1024, n2=1, M=4, n3=33, xlen=40.0, Tramp=3.0,
                              spec=dict(Hs=0.07, Tp=1.8, n_freq=64, n_dir=1, seed=12)),
    "member_2d_2048_M3": dict(n1=2048, n2=1, M=3, n3=33, xlen=60.0, Tramp=3.0,
                              spec=dict(Hs=0.09, Tp=1.3, n_freq=64, n_dir=1, seed=13)),
    # extended lines beyond the in-workgroup engines: 14400, 16384 and 28800 points
    "member_2d_2048_M6": dict(n1=2048, n2=1, M=6, n3=33, xlen=60.0, Tramp=3.0,
                              spec=dict(Hs=0.06, Tp=1.6, n_freq=64, n_dir=1, seed=14)),
    "member_2d_4096_M3": dict(n1=4096, n2=1, M=3, n3=33, xlen=80.0, Tramp=3.0,
                              spec=dict(Hs=0.05, Tp=1.9, n_freq=64, n_dir=1, seed=15)),
    "member_2d_4096_M6": dict(n1=4096, n2=1, M=6, n3=33, xlen=80.0, Tramp=3.0,
                              spec=dict(Hs=0.08, Tp=1.4, n_freq=64, n_dir=1, seed=16)),
}


def setup_for(name):
    kw = dict(CASES[name])
    spec = kw.pop("spec", None)
    if spec is not None:
        kw["components"] = hc.jonswap_components(spec["Hs"], spec["Tp"], 3.3, n_freq=spec["n_freq"],
                                                 seed=spec["seed"], rnum=12, clock=32.0, n_dir=spec["n_dir"],
                                                 spread_s=10)
        kw["icase"] = 3
    return make_setup(**kw)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_rhs_parity_at_config_size(name):
    s = setup_for(name)
    print(name, "extended grid from", ((s.cfg.mHOS + 1) * s.cfg.n1) // 2, flush=True)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        st = tank.stats()
        print(name, "GPU extended grid", st["big_n1"], st["big_n2"], flush=True)
        t = 0.6 * s.Tramp
        G1, G2, a1, e1, aad = random_state(s, 3, amp=0.01)
        orc.time = t
        r = orc.solve_hos2(a1, e1, G2, G1, aad, 0.01)
        g = tank.solve_hos2(t, 0.01, G1, G2, a1, e1, aad)
        ref_dG2 = r[1].copy()
        ref_dG2[0, 0] = 0.0  # k_G2(1,1) = 0 (runge_kutta.f90:150)
        errs = {"dG1": rel(g["dG1"], r[0]), "dG2": rel(g["dG2"], ref_dG2), "da1st": rel(g["da1st"], r[2]),
                "deta1st": rel(g["deta1st"], r[3]), "da_add_1": rel(g["da_add_1"], r[4]),
                "da_add_ramp": rel(g["da_add_ramp"], r[5])}
        print(name, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        for k, v in errs.items():
            assert v < TOL, (name, k, v)
    finally:
        tank.close()


# 2-D tanks on both sides of the n1 = 2048 switch between the matrix-core product and the
# split-K kernel for the mixed-parity filter on the mixed path
MIXED_2D = ["member_2d_1024_M4", "irregular_2d_2048_M5", "member_2d_2048_M3", "member_2d_4096_M3"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", MIXED_2D)
def test_mixed_rhs_at_config_size(name):
    s = setup_for(name)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=1)
    try:
        t = 0.6 * s.Tramp
        G1, G2, a1, e1, aad = random_state(s, 3, amp=0.01)
        orc.time = t
        r = orc.solve_hos2(a1, e1, G2, G1, aad, 0.01)
        g = tank.solve_hos2(t, 0.01, G1, G2, a1, e1, aad)
        ref_dG2 = r[1].copy()
        ref_dG2[0, 0] = 0.0
        errs = {"dG1": rel(g["dG1"], r[0]), "dG2": rel(g["dG2"], ref_dG2), "da1st": rel(g["da1st"], r[2]),
                "deta1st": rel(g["deta1st"], r[3]), "da_add_1": rel(g["da_add_1"], r[4])}
        print(name, "mixed", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        for k, v in errs.items():
            assert v < 5e-5, (name, k, v)
    finally:
        tank.close()
