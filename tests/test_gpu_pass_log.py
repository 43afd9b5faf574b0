"""Per-launch timing of the extended-grid passes (hos_pass_log_begin / hos_pass_log_end),
the measurement behind the bench's roofline figure.

* Its algorithmic bytes per launch equal the timed full-transform batches' own count
  (2 x items x grid points x element size per pass), which the library derives separately,
  and its transforms equal the attempts' count of extended-grid transforms.
* Recording changes no result: attempts with and without the log give bit-identical states.
* Capacity, nesting and graph-capture refusals.
"""
import numpy as np
import pytest

from hosnwt.tank import Tank
from test_gpu_configs import setup_for


def step(tank, s, n, t=0.0):
    h = min(0.95 * s.dt_stab, s.delt_out)
    for _ in range(n):
        tank.runge4_adapt2(t, h)
        tank.commit()
        t += h
    return t


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_logged_bytes_match_the_timed_batches(precision):
    s = setup_for("member_3d_512x128_M7")
    tank = Tank(s, precision=precision)
    try:
        tank.pass_log_begin(64)
        out = tank.time_big_batch(reps=3, items=4)
        lg = tank.pass_log_end()
        print(precision, lg, out, flush=True)
        assert lg["x"]["launches"] == 3 and lg["y"]["launches"] == 3 and not lg["full"]
        assert lg["x"]["transforms"] == 12 and lg["y"]["transforms"] == 12
        assert lg["x"]["bytes"] == 3 * out["x"]["bytes"]
        assert lg["y"]["bytes"] == 3 * out["y"]["bytes"]
        assert lg["x"]["ms"] > 0.0 and lg["y"]["ms"] > 0.0
    finally:
        tank.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_logging_changes_no_result(precision):
    s = setup_for("member_3d_512x128_M7")
    states = []
    for logged in (False, True):
        tank = Tank(s, precision=precision)
        try:
            tank.initialize()
            n_big0 = tank.stats()["n_big"]
            if logged:
                tank.pass_log_begin(4096)
            step(tank, s, 3)
            if logged:
                lg = tank.pass_log_end()
                st = tank.stats()
                print(precision, lg, st["n_big"] - n_big0, flush=True)
                assert lg["x"]["launches"] > 0 and lg["y"]["launches"] > 0 and not lg["full"]
                # the extended-grid transforms the attempts counted, each with one x and one y pass
                assert lg["x"]["transforms"] == lg["y"]["transforms"] == st["n_big"] - n_big0
                # every launch carries at least one transform's outputs (n1 x n2 modes)
                assert lg["x"]["bytes"] >= lg["x"]["launches"] * s.cfg.n1 * s.cfg.n2 * 4
                assert lg["y"]["bytes"] >= lg["y"]["launches"] * s.cfg.n1 * s.cfg.n2 * 4
                # and at most a full batch of 8 accumulating full-grid transforms
                full = 8 * 3 * st["big_n1"] * st["big_n2"] * 8
                assert lg["x"]["bytes"] <= lg["x"]["launches"] * full
            states.append(tank.get_state())
        finally:
            tank.close()
    for k in states[0]:
        np.testing.assert_array_equal(states[0][k], states[1][k], err_msg=k)


@pytest.mark.gpu
def test_capacity_nesting_and_capture_refusals():
    s = setup_for("member_3d_512x128_M7")
    tank = Tank(s, precision=1)
    try:
        tank.initialize()
        with pytest.raises(RuntimeError):
            tank.pass_log_end()
        tank.pass_log_begin(1)
        with pytest.raises(RuntimeError):
            tank.pass_log_begin(1)
        with pytest.raises(RuntimeError):
            tank.advance_graph(0.05, 0.01)
        step(tank, s, 1)
        lg = tank.pass_log_end()
        assert lg["full"] and lg["x"]["launches"] + lg["y"]["launches"] == 1
    finally:
        tank.close()
