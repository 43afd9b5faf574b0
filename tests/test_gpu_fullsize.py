"""Parity at the headline and ensemble sizes.

* FP64: two runge4_adapt2 attempts of the 1024 x 256, M = 5 directional-sea workload from
  its initial state, the oracle's (h, accept) decisions replayed: eta, phis, the wavemaker
  modes, deta, volume and energy within 1e-10 (runge_kutta.f90:81-360).
* Mixed: one solveHOS2 of the same workload against the oracle, at the mixed path's bar.
* The ragged 64-member ensemble (n1 512-4096, M 3-6, mixed): every member run together
  equals itself run alone bit for bit (eight of them checked, two per n1), and one FP64
  member per n1 matches the oracle's right-hand side within 1e-10."""
import numpy as np
import pytest
import scipy.fft as sfft

from helpers import oracle_params, random_state, rel
from hosnwt.tank import Tank, advance_ensemble
from oracle import hos_oracle as ho
from test_gpu_step import replay

pytestmark = pytest.mark.gpu
TOL = 1e-10


def test_fp64_attempts_at_1024x256_match_oracle():
    import bench
    s = bench.make_setup(0)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        st = orc.initialize()
        tank.initialize()
        h0 = min(0.95 * s.dt_stab, s.delt_out)
        log = []
        with sfft.set_workers(16):
            st, t_end, _, na, nr, deta = ho.advance(orc, st, 0.0, 1.9 * h0, h0, s.cfg.toler, s.dt_stab,
                                                    s.delt_out, log=log)
        print(f"oracle: {len(log)} attempts to t = {t_end:.5f}", flush=True)
        t_gpu, errs = replay(tank, log, Tramp=s.Tramp, iramp=s.cfg.iramp)
        assert abs(t_gpu - t_end) < 1e-12 and len(log) >= 2
        e_err = max(abs(a - b[2]) / max(abs(b[2]), 1e-300) for a, b in zip(errs, log))
        g = tank.get_state()
        ph = tank.physical()
        res = {"eta": rel(g["eta"], st.eta), "phis": rel(g["phis"], st.phis),
               "a1st_add": rel(g["a1st_add"], st.a1st_add), "a1st": rel(g["a1st"], st.a1st),
               "eta1st": rel(g["eta1st"], st.eta1st), "deta": rel(ph["deta"], deta),
               "volume": abs(tank.volume(ph["eta"]) - orc.volume(st.eta)) / max(np.abs(st.eta).max(), 1e-300),
               "energy": abs(tank.energy(ph["eta"], ph["phis"] + ph["phi_add"], ph["deta"])[2]
                             - orc.energy(st.eta, st.phis + orc.phi_add, deta)[2])
               / max(abs(orc.energy(st.eta, st.phis + orc.phi_add, deta)[2]), 1e-300)}
        print(f"1024x256 M=5 FP64, {len(log)} attempts: err rel {e_err:.2e}",
              {k: f"{v:.2e}" for k, v in res.items()}, flush=True)
        assert e_err < 1e-6
        for k, v in res.items():
            assert v < TOL, (k, v)
    finally:
        tank.close()


def test_mixed_rhs_at_1024x256_vs_oracle():
    import bench
    s = bench.make_setup(1)
    orc = ho.Oracle(oracle_params(s))
    G1, G2, a1, e1, aad = random_state(s, 5, amp=0.01)
    orc.time = 3.0
    with sfft.set_workers(16):
        r = orc.solve_hos2(a1, e1, G2, G1, aad, 0.01)
    tank = Tank(s, precision=1)
    try:
        g = tank.solve_hos2(3.0, 0.01, G1, G2, a1, e1, aad)
        # the mean mode of dG2 is zero on the GPU: the step restores the frozen mean
        # potential phis_hat(1,1) (runge_kutta.f90:281-299), as in the mixed RHS tests
        ref_dG2 = r[1].copy()
        ref_dG2[0, 0] = 0.0
        errs = {"dG1": rel(g["dG1"], r[0]), "dG2": rel(g["dG2"], ref_dG2), "da1st": rel(g["da1st"], r[2]),
                "deta1st": rel(g["deta1st"], r[3]), "da_add_1": rel(g["da_add_1"], r[4])}
        print("mixed 1024x256 M=5 RHS:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        # FP32 products of orders >= 2; the first-order terms stay FP64
        assert errs["dG1"] < 2e-5 and errs["dG2"] < 2e-5
        assert errs["da1st"] < 1e-12 and errs["deta1st"] < 1e-12 and errs["da_add_1"] < 1e-12
    finally:
        tank.close()


def test_ragged_64_member_ensemble():
    import bench
    members = bench.ensemble_members(64)
    h0 = [min(0.95 * s.dt_stab, s.delt_out) for s in members]
    tanks = [Tank(s) for s in members]
    try:
        for tk in tanks:
            tk.initialize()
        _, na, nr = advance_ensemble(tanks, 0.3, h0)
        print(f"64 members to t = 0.3: {sum(na) + sum(nr)} attempts", flush=True)
        picked = []
        for n1 in (512, 1024, 2048, 4096):
            picked += [i for i, s in enumerate(members) if s.cfg.n1 == n1][:2]
        for i in picked:
            alone = Tank(members[i])
            try:
                alone.initialize()
                alone.advance(0.3, h0[i])
                assert np.array_equal(alone.get_state()["eta"], tanks[i].get_state()["eta"]), i
            finally:
                alone.close()
    finally:
        for tk in tanks:
            tk.close()
    # one FP64 member per n1 against the oracle's right-hand side
    members64 = bench.ensemble_members(64, precision=0)
    for n1 in (512, 1024, 2048, 4096):
        i = [k for k, s in enumerate(members64) if s.cfg.n1 == n1][0]
        s = members64[i]
        orc = ho.Oracle(oracle_params(s))
        G1, G2, a1, e1, aad = random_state(s, 11, amp=0.01)
        orc.time = 1.0
        r = orc.solve_hos2(a1, e1, G2, G1, aad, 0.02)
        tk = Tank(s, precision=0)
        try:
            g = tk.solve_hos2(1.0, 0.02, G1, G2, a1, e1, aad)
        finally:
            tk.close()
        ref_dG2 = r[1].copy()
        ref_dG2[0, 0] = 0.0  # the frozen mean potential (see the mixed test above)
        e = max(rel(g["dG1"], r[0]), rel(g["dG2"], ref_dG2))
        print(f"member {i}: n1 = {n1}, M = {s.cfg.mHOS}: RHS rel {e:.2e}", flush=True)
        assert e < TOL
