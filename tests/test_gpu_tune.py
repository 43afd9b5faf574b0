"""Extended-grid schedule autotuner (the counterpart of the reference's FFTW_PATIENT
planning in fourier_ini, fourier_FFTW.f90:123-170).

* Every compiled schedule of a grid (x / y radix sequences, y tile width) keeps RHS parity:
  FP64 against the oracle within 1e-10, FP32 against the FP64 result within 5e-5 (the
  mixed-path bound of the config-size tests).
* A tuned grid's schedule is shared by every context of that shape in the process, and two
  such contexts give bit-identical right-hand sides.
* HOSNWT_TUNE_CACHE: a second process reads the first one's choice instead of timing.
"""
import json
import os
import subprocess
import sys
from contextlib import contextmanager

import numpy as np
import pytest

from helpers import oracle_params, random_state, rel
from hosnwt.tank import Tank
from oracle import hos_oracle as ho
from test_gpu_configs import setup_for

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS = ("dG1", "dG2", "da1st", "deta1st", "da_add_1")


@contextmanager
def env(**kw):
    old = {k: os.environ.get(k) for k in kw}
    try:
        for k, v in kw.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def rhs(s, precision, state, t):
    tank = Tank(s, precision=precision)
    try:
        return tank.schedule(), tank.solve_hos2(t, 0.01, *state)
    finally:
        tank.close()


# (x sequence, y sequence, y tile): the defaults, both second sequences, and mixed choices
SCHEDULES = ["0,0,8", "1,1,8", "1,0,16", "0,1,4"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["directional_3d_1024x256_M5", "member_3d_512x128_M7", "member_2d_1024_M4",
                                  "regular_2d_256_M3"])
def test_every_schedule_keeps_parity(name):
    s = setup_for(name)
    orc = ho.Oracle(oracle_params(s))
    t = 0.6 * s.Tramp
    state = random_state(s, 3, amp=0.01)
    G1, G2, a1, e1, aad = state
    orc.time = t
    r = orc.solve_hos2(a1, e1, G2, G1, aad, 0.01)
    ref = dict(zip(FIELDS, (r[0], r[1].copy(), r[2], r[3], r[4])))
    ref["dG2"][0, 0] = 0.0  # k_G2(1,1) = 0 (runge_kutta.f90:150)
    seen = set()
    for sc in SCHEDULES:
        with env(HOSNWT_FFT_SCHED=sc):
            sch64, g64 = rhs(s, 0, state, t)
            sch32, g32 = rhs(s, 1, state, t)
        key = (sch64["x_seq"], sch64["y_seq"], sch64["y_tile"])
        print(name, sc, "FP64", sch64, "FP32", sch32, flush=True)
        if key in seen:  # the schedule clamps to one already checked (2-D: no y lines)
            continue
        seen.add(key)
        for k in FIELDS:
            e64 = rel(g64[k], ref[k])
            e32 = rel(g32[k], g64[k])
            print(f"  {k}: FP64 vs oracle {e64:.2e}, FP32 vs FP64 {e32:.2e}", flush=True)
            assert e64 < 1e-10, (name, sc, k, e64)
            assert e32 < 5e-5, (name, sc, k, e32)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_tuned_schedule_is_shared_by_contexts_of_one_shape(precision):
    s = setup_for("directional_3d_1024x256_M5")
    state = random_state(s, 5, amp=0.01)
    with env(HOSNWT_FFT_SCHED=None, HOSNWT_TUNE=None):
        a_sched, a = rhs(s, precision, state, 1.0)
        b_sched, b = rhs(s, precision, state, 1.0)
    print("precision", precision, a_sched, flush=True)
    assert a_sched == b_sched
    assert a_sched["trials"] >= 2, a_sched  # x and y sequences and tile widths were timed
    assert a_sched["us_tuned"] <= a_sched["us_default"]
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


SCRIPT = """
import json, sys
sys.path[:0] = {paths!r}
from test_gpu_configs import setup_for
from hosnwt.tank import Tank
t = Tank(setup_for("member_3d_512x128_M7"), precision=1)
print(json.dumps(t.schedule()))
t.close()
"""


@pytest.mark.gpu
def test_cache_file_carries_the_choice_to_a_new_process(tmp_path):
    cache = tmp_path / "tune.txt"
    code = SCRIPT.format(paths=[ROOT, os.path.join(ROOT, "hos-nwt_amd"), os.path.join(ROOT, "tests")])
    e = dict(os.environ, HOSNWT_TUNE_CACHE=str(cache))
    e.pop("HOSNWT_FFT_SCHED", None)
    e.pop("HOSNWT_TUNE", None)
    runs = []
    for i in range(2):
        print("process", i, flush=True)
        out = subprocess.run([sys.executable, "-c", code], env=e, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr[-2000:]
        runs.append(json.loads(out.stdout.strip().splitlines()[-1]))
        print(runs[-1], flush=True)
    assert runs[0]["trials"] >= 2
    assert runs[1]["trials"] == 0  # read from the file
    for k in ("x_seq", "y_seq", "y_tile"):
        assert runs[0][k] == runs[1][k]
    line = cache.read_text().split()
    assert [int(v) for v in line[:3]] == [runs[0]["x_length"] // 2 + 1, runs[0]["y_length"] // 2 + 1, 4]
