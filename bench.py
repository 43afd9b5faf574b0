#!/usr/bin/env python3
"""RK4(3) steps/s of the HOS-NWT time-stepping core on one MI355X.

Workload: 3-D directional sea, 40 x 10 basin (h = 1 m), n1 = 1024 x n2 = 256 modes,
HOS order M = 5, n3 = 33 paddle nodes, snake wavemaker (one paddle per y-node)
fed a seeded JONSWAP spectrum (Hs 0.06 m, Tp 1.4 s, gamma 3.3) with cos^2s
spreading (s = 10) over 256 frequencies x 32 directions, beach on the last 15 %.

A step is one runge4_adapt2 attempt (5 solveHOS2 evaluations + combination +
embedded error), accepted or rejected by the reference controller; the state
stays resident in HBM for the whole timed region.  Both precisions are timed:
FP64 everywhere, and FP32 transforms/products with FP64 RK state ("mixed").
Each path first runs untimed from rest through the wavemaker ramp (2 Tp), so
the timed attempts see a developed sea; the mixed path runs free, with its own
error estimates and accept/reject decisions, to the FP64 run's two times (end
of warm-up, end of the timed attempts).
`value` is the mixed path's steps/s when its eta stays within 1e-5 relative L2
of the FP64 path at the end of the timed attempts, otherwise the FP64 path's.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "hos-nwt_amd")]

import numpy as np  # noqa: E402

from hosnwt import config as hc  # noqa: E402

WORKLOAD = dict(n1=1024, n2=256, M=5, n3=33, xlen=40.0, ylen=10.0, Hs=0.06, Tp=1.4, gamma=3.3, n_freq=256,
                n_dir=32, s=10, xabsf=0.85, seed=2024)


def make_setup(precision: int, wl=WORKLOAD) -> hc.Setup:
    comps = hc.jonswap_components(wl["Hs"], wl["Tp"], wl["gamma"], n_freq=wl["n_freq"], seed=wl["seed"], rnum=12,
                                  clock=32.0, n_dir=wl["n_dir"], spread_s=wl["s"])
    cfg = hc.TankConfig(n1=wl["n1"], n2=wl["n2"], mHOS=wl["M"], xlen=wl["xlen"], ylen=wl["ylen"], n3=wl["n3"],
                        icase=3, igeom=1, iramp=1, Tramp=2.0 * wl["Tp"], xabsf=wl["xabsf"], T_stop=50 * wl["Tp"],
                        toler=1e-7, f_out=1.0 / wl["Tp"], precision=precision, components=comps)
    return hc.build(cfg)


def run_steps(tank, s, n, t0, h0, t_stop=None):
    """At most n attempts with the reference accept/reject rule (HOS-NWT.f90:518-563).

    ``t_stop``: stop there (the last step shortened to land on it, as the inner
    loop does at an output time).
    """
    t, h = t0, h0
    toler, stab = s.cfg.toler, s.dt_stab
    n_acc = n_rej = 0
    for _ in range(n):
        if t_stop is not None and t >= t_stop:
            break
        hl = h if t_stop is None else min(h, t_stop - t)
        err = tank.runge4_adapt2(t, hl)
        acc = not (err > toler and h > 0.001 * 0.95 * stab)
        if not acc:
            corr = 0.9 * (err / toler) ** -0.25
            n_rej += 1
        else:
            tank.commit()
            # end-of-ramp correction of the paddle modes (HOS-NWT.f90:546-550)
            if s.cfg.iramp == 1 and s.Tramp - (t + hl) <= 2.220446049250313e-16 and s.Tramp > t:
                tank.end_ramp()
            t += hl
            corr = 2.0 if err < 2.2e-16 else 0.9 * (err / toler) ** -0.25
            n_acc += 1
        h = max(min(h * corr, 0.95 * stab, s.delt_out), 1e-4 * 0.95 * stab, 1e-4 * s.delt_out)
    return t, h, n_acc, n_rej


def time_path(precision, steps, warmup, spin_up, follow=None):
    """Untimed: from rest to nondimensional time ``spin_up``, then ``warmup`` attempts; timed:
    ``steps`` attempts.  ``follow`` = (t_warm, t_end) of an earlier run: run freely (own
    controller) to t_warm untimed, then time the attempts that reach t_end."""
    from hosnwt.tank import Tank
    s = make_setup(precision)
    tank = Tank(s)
    tank.initialize()
    h0 = min(0.95 * s.dt_stab, s.delt_out)
    cap = 100000
    if follow:
        # the same landing times as the run followed: the ramp's end, then its warm-up end
        t, h, _, _ = run_steps(tank, s, cap, 0.0, h0, t_stop=min(spin_up, follow[0])) if spin_up > 0 else (0.0, h0, 0, 0)
        t, h, _, _ = run_steps(tank, s, cap, t, h, t_stop=follow[0])
    else:
        t, h, _, _ = run_steps(tank, s, cap, 0.0, h0, t_stop=spin_up) if spin_up > 0 else (0.0, h0, 0, 0)
        t, h, _, _ = run_steps(tank, s, warmup, t, h)
    t_warm = t
    tank.synchronize()
    st0 = tank.stats()
    t_start = time.perf_counter()
    t, h, na, nr = run_steps(tank, s, cap if follow else steps, t, h, t_stop=follow[1] if follow else None)
    tank.synchronize()
    tank.synchronize()
    dt = time.perf_counter() - t_start
    steps = na + nr
    st1 = tank.stats()
    eta = tank.get_state()["eta"]
    out = dict(steps_per_s=steps / dt, ms_per_step=1e3 * dt / steps, n_acc=na, n_rej=nr, t_warm=t_warm, t_end=t,
               h_end=h, eta=eta,
               big_per_step=(st1["n_big"] - st0["n_big"]) / steps, base_per_step=(st1["n_base"] - st0["n_base"]) / steps,
               big_grid=(st1["big_n1"], st1["big_n2"]), tank=tank, setup=s)
    return out


def log_passes(r, steps):
    """Per-launch HIP-event timing of the extended-grid passes as the solver launches them:
    ``steps`` further attempts of a timed run (same tank, state and controller), every pass
    launch bracketed by two events on the stream it runs on.  Returns the log and the
    attempts/s of the logged interval (the events' own cost included)."""
    tank, s = r["tank"], r["setup"]
    tank.synchronize()
    tank.pass_log_begin(32768)
    t0 = time.perf_counter()
    _, _, na, nr = run_steps(tank, s, steps, r["t_end"], r["h_end"])
    tank.synchronize()
    dt = time.perf_counter() - t0
    lg = tank.pass_log_end()
    return lg, (na + nr) / dt


def pass_figure(lg):
    ach = lg["bytes"] / (lg["ms"] * 1e-3) / 1e9
    n = lg["launches"]
    return {"achieved": round(ach, 1), "frac": round(ach / 8000.0, 4), "launches": n,
            "transforms_per_launch": round(lg["transforms"] / n, 3), "bytes_per_launch": round(lg["bytes"] / n),
            "avg_ms": round(lg["ms"] / n, 5)}


def roofline(r, precision, steps):
    """Live HIP-event timing of the dominant kernel family (extended-grid r2r passes) in the
    precision of the headline path.  Top level: the slower pass of a lone full inverse
    transform on the context's stream (full-grid bytes, 20 launches between two events).
    ``batched4``: four such transforms per launch.  ``in_step``: the solver's own launches over
    ``steps`` attempts that continue the timed run, each bracketed by two events on its stream
    (the pairs' own cost included; the logged interval's attempts/s shows it), with the
    measured HBM traffic over algorithmic bytes of those launches."""
    tank = r["tank"]
    st = tank.stats()
    grid = [st["big_n1"], st["big_n2"]]
    lone = tank.time_big_transform(reps=20)
    achieved = lone["bytes"] / (lone["ms"] * 1e-3) / 1e9
    traffic = None
    t = traffic_table(precision)
    if t is not None and list(t["grid"]) == grid:
        traffic = t["traffic_bytes"][lone["pass"]]
    b = tank.time_big_batch(reps=20, items=4)
    bp = b[lone["pass"]]
    ach4 = bp["bytes"] / (bp["ms"] * 1e-3) / 1e9
    lg, logged_rate = log_passes(r, steps)
    ti = in_step_traffic_table(precision)
    in_step = {"attempts": steps, "logged_steps_per_s": round(logged_rate, 2), "full": lg["full"]}
    for p in ("x", "y"):
        if lg[p]["launches"] > 0:
            f = pass_figure(lg[p])
            if ti is not None and list(ti["grid"]) == grid:
                f["traffic_over_bytes"] = ti["traffic_ratio"][p]
            in_step[p] = f
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 4), "traffic": traffic, "kernel": lone["kernel"],
            "bytes_per_launch": lone["bytes"], "avg_ms": round(lone["ms"], 5),
            "batched4": {"achieved": round(ach4, 1), "frac": round(ach4 / 8000.0, 4),
                         "bytes_per_launch": bp["bytes"], "avg_ms": round(bp["ms"], 5)},
            "in_step": in_step}


def traffic_table(precision=0):
    """Scaled FETCH_SIZE + WRITE_SIZE per launch of the extended-grid passes, from the committed
    rocprofv3 summary (profiles/extended_pass_1024x256_M5_traffic*.md); None when absent."""
    name = "extended_pass_1024x256_M5_traffic.md" if precision == 0 else "extended_pass_1024x256_M5_traffic_fp32.md"
    path = os.path.join(ROOT, "profiles", name)
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)
    except (OSError, IndexError, ValueError):
        return None


def in_step_traffic_table(precision=0):
    """Measured HBM traffic over algorithmic bytes of the solver's own pass launches (scaled
    FETCH_SIZE + WRITE_SIZE summed over a short run's launches), from the committed summary
    profiles/extended_pass_1024x256_M5_in_step.md; None when absent."""
    path = os.path.join(ROOT, "profiles", "extended_pass_1024x256_M5_in_step.md")
    try:
        text = open(path).read()
        blocks = [b.split("```", 1)[0] for b in text.split("```json")[1:]]
        for b in blocks:
            d = json.loads(b)
            if d.get("precision") == precision:
                return d
    except (OSError, IndexError, ValueError):
        return None
    return None


def pass_log_probe(steps, precision):
    """The solver's own extended-grid pass launches over ``steps`` attempts from rest (after
    two untimed ones), each timed with HIP events: for counter and kernel-trace runs."""
    from hosnwt.tank import Tank
    s = make_setup(precision)
    tank = Tank(s)
    tank.initialize()
    h0 = min(0.95 * s.dt_stab, s.delt_out)
    t, h, _, _ = run_steps(tank, s, 2, 0.0, h0)
    lg, rate = log_passes(dict(tank=tank, setup=s, t_end=t, h_end=h), steps)
    print(json.dumps({"precision": precision, "attempts": steps, "logged_steps_per_s": round(rate, 2),
                      "x": lg["x"], "y": lg["y"], "full": lg["full"]}), flush=True)
    tank.close()
    calibration_copy()


def calibration_copy():
    """Three 64 MiB device-to-device copies: their blit kernel moves known bytes, against which
    rocprofv3's FETCH_SIZE / WRITE_SIZE counters are scaled."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = 64 * 1024 * 1024
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, C.c_size_t(nbytes)):
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print("copied 64 MiB device-to-device 3 times", flush=True)


def pass_probe(reps, precision=0):
    """Extended-grid pass timing (the lone roofline launches) plus the counter-calibration copy."""
    from hosnwt.tank import Tank
    s = make_setup(precision)
    tank = Tank(s)
    tank.time_big_transform(reps=2)
    r = tank.time_big_transform(reps=reps)
    print(json.dumps({"pass_ms": {r["kernel"]: r["ms"], "other": r["other"]["ms"]}}), flush=True)
    tank.close()
    calibration_copy()


def mlp_probe(epochs):
    """Trains the order-prediction network on 32768 synthetic rows (a fixed random teacher
    network's outputs) for `epochs` epochs: the matrix-core kernels under the profiler."""
    from hosnwt import learned as hl
    rng = np.random.default_rng(0)
    x = rng.standard_normal((32768, 512)).astype(np.float32)
    teacher = hl.OrderNet(max_rows=1024, seed=11)
    t = np.concatenate([teacher.forward(x[i:i + 1024]) for i in range(0, len(x), 1024)])
    teacher.close()
    net = hl.OrderNet(max_rows=1024, seed=0)
    for batch in (256, 1024):
        losses, ms = net.train(x, t, epochs=epochs, batch=batch, lr=1e-3, seed=1)
        flops = hl.OrderNet.train_flops(len(x), epochs)
        print(json.dumps({"batch": batch, "epochs": epochs, "ms": round(ms, 2), "tflops": round(flops / ms / 1e9, 3),
                          "loss": [float(losses[0]), float(losses[-1])]}), flush=True)
    net.close()


def cpu_baseline(wl=WORKLOAD, rows=8):
    """Host-CPU reference-mode cost of one runge4_adapt2 attempt of the same workload.

    * FFT / HOS part: the FP64 oracle (numpy/scipy, reference transform sequence, scipy.fft
      on `cores` threads) runs one full attempt from the initialised state (wavemaker modes of
      t = 0 active).  The oracle sums the wavemaker potential in the separable form; the time
      it spends there is measured and replaced by:
    * the reference's direct sums (resol_wmkr.f90:485-532 and :1149-1175, 12 + 5 fields of
      n1 n2^2 n3_add iterations with trigonometric calls inside, tables of HOS-NWT.f90:364-405),
      restated in C (oracle/wmk_direct.c, OpenMP over `cores` threads), timed on `rows` of the
      n1 x-rows of one right-hand side and scaled to n1 rows x 5 right-hand sides."""
    import scipy.fft as sfft
    from oracle import hos_oracle as ho
    from oracle import wmk_direct as wd
    cores = min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    s = make_setup(0, wl)
    c = s.cfg
    p = ho.Params(n1=c.n1, n2=c.n2, mHOS=c.mHOS, xlen=s.xlen, ylen=s.ylen, n3=c.n3, icase=c.icase, igeom=c.igeom,
                  d_hinge=0.0, iramp=c.iramp, Tramp=s.Tramp, T_stop=s.T_stop, T_mono=s.T_mono, iabsnb=c.iabsnb,
                  xabsf=c.xabsf, coeffabsf=c.coeffabsf, freq=s.freq, pos_ini=s.pos_ini)
    orc = ho.Oracle(p)
    st = orc.initialize()
    h = min(0.95 * s.dt_stab, s.delt_out)
    sep = [0.0]
    plain = ho.phiadd_sl

    def timed_phiadd(*a, **k):
        t0 = time.perf_counter()
        r = plain(*a, **k)
        sep[0] += time.perf_counter() - t0
        return r

    ho.phiadd_sl = timed_phiadd
    try:
        with sfft.set_workers(cores):
            t0 = time.perf_counter()
            orc.runge4_adapt2(0.0, st, h)
            t_oracle = time.perf_counter() - t0
    finally:
        ho.phiadd_sl = plain
    # the direct sums on a sample of rows at the initial state's paddle modes and a sea of
    # 0.01 amplitude (the iteration count, not the values, sets the cost)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    wd.load()
    n1, n2, n3a = c.n1, c.n2, 2 * (c.n3 - 1)
    a = np.asfortranarray(st.a1st_add if hasattr(st, "a1st_add") else np.ones((n3a, n2)))
    da = np.asfortranarray(np.roll(a, 1, axis=0))
    eta = 0.01 * np.ones((n1, n2))
    lo = n1 // 2
    t0 = time.perf_counter()
    wd.phiadd_sl(n1, n2, s.xlen, s.ylen, a, da, eta, rows=(lo, lo + rows))
    t_sl = time.perf_counter() - t0
    t0 = time.perf_counter()
    wd.phiadd_1st(n1, n2, s.xlen, s.ylen, a, da, rows=(lo, lo + rows))
    t_1st = time.perf_counter() - t0
    direct = 5.0 * (n1 / rows) * (t_sl + t_1st)
    t_attempt = t_oracle - sep[0] + direct
    return {"value": round(1.0 / t_attempt, 6), "unit": "steps/s", "cores": cores, "kind": "port",
            "cpu": model,
            "sample": (f"one runge4_adapt2 attempt of the 1024x256 M=5 workload from its initial state in reference "
                       f"mode: oracle FFT/HOS part {t_oracle - sep[0]:.2f} s (scipy.fft on {cores} threads; its "
                       f"separable wavemaker sums, {sep[0]:.2f} s, left out) + the reference's direct wavemaker sums "
                       f"(17 fields, C restatement on {cores} threads) timed on {rows} of {n1} x-rows of one "
                       f"right-hand side ({t_sl + t_1st:.3f} s) and scaled x{n1 / rows:.0f} rows x 5 right-hand "
                       f"sides = {direct:.1f} s"),
            "seconds_per_attempt": round(t_attempt, 2), "direct_sums_seconds": round(direct, 2),
            "separable_sums_seconds": round(sep[0], 3),
            "direct_over_separable": round(direct / max(sep[0], 1e-9), 1)}


def small_tank_probe():
    """Attempts/s of single small tanks, stream path (hos_advance, a host wait per attempt)
    against graph-captured attempts (hos_advance_graph): cfg1 (2-D n1 = 256, M = 3, regular
    waves) and cfg2 (2-D n1 = 2048, M = 5, JONSWAP), FP64 and mixed."""
    from hosnwt.tank import Tank
    out = {}
    cases = {"cfg1 2-D 256 M=3": dict(n1=256, M=3, xlen=20.0, regular=True),
             "cfg2 2-D 2048 M=5": dict(n1=2048, M=5, xlen=60.0, regular=False)}
    for name, c in cases.items():
        for prec in (0, 1):
            if c["regular"]:
                k = hc.wave_number(1.0 / 1.5 / np.sqrt(9.81))
                cfg = hc.TankConfig(n1=c["n1"], n2=1, mHOS=c["M"], xlen=c["xlen"], n3=33, icase=2,
                                    amp_mono=0.08 / k, nu_mono=1.0 / 1.5, igeom=1, iramp=1, Tramp=4.5, xabsf=0.8,
                                    T_stop=45.0, toler=1e-7, f_out=1.0 / 1.5, precision=prec)
            else:
                comps = hc.jonswap_components(0.08, 1.6, 3.3, n_freq=512, seed=7, rnum=12, clock=32.0)
                cfg = hc.TankConfig(n1=c["n1"], n2=1, mHOS=c["M"], xlen=c["xlen"], n3=33, icase=3, igeom=2,
                                    iramp=1, Tramp=3.2, xabsf=0.8, T_stop=160.0, toler=1e-7, f_out=1.0 / 1.6,
                                    precision=prec, components=comps)
            s = hc.build(cfg)
            res = {}
            for path in ("stream", "graph"):
                tk = Tank(s)
                tk.initialize()
                h = min(0.95 * s.dt_stab, s.delt_out)
                adv = tk.advance if path == "stream" else tk.advance_graph
                h, _, _ = adv(2.0 * s.delt_out, h)  # untimed: set-up, capture, spin-up
                t0 = time.perf_counter()
                h, na, nr = adv(4.0 * s.delt_out, h)
                dt = time.perf_counter() - t0
                res[path] = round((na + nr) / dt, 1)
                tk.close()
            res["graph_vs_stream"] = round(res["graph"] / res["stream"], 3)
            out[f"{name} {'fp64' if prec == 0 else 'mixed'}"] = res
            print(f"[small] {name} prec {prec}: {res}", file=sys.stderr, flush=True)
    return out


def config_shapes():
    """The BASELINE configurations' shapes at full size (cfg2, cfg3, cfg4 members, a cfg5 member):
    name -> TankConfig keyword arguments (precision set by the caller)."""
    def js(Hs, Tp, n_freq, n_dir, seed):
        return hc.jonswap_components(Hs, Tp, 3.3, n_freq=n_freq, seed=seed, rnum=12, clock=32.0, n_dir=n_dir,
                                     spread_s=10)
    base = dict(n3=33, icase=3, iramp=1, toler=1e-7, T_stop=100.0, f_out=1.0)
    return {
        "cfg2 2-D 2048 M=5": dict(base, n1=2048, n2=1, mHOS=5, xlen=60.0, igeom=2, d_hinge=0.0, Tramp=3.2,
                                  xabsf=0.8, components=js(0.08, 1.6, 512, 1, 7)),
        "cfg3 3-D 1024x256 M=5": dict(base, n1=1024, n2=256, mHOS=5, xlen=40.0, ylen=10.0, igeom=1, Tramp=2.8,
                                      xabsf=0.85, components=js(0.06, 1.4, 256, 32, 2024)),
        "cfg4 member 2-D 512 M=3": dict(base, n1=512, n2=1, mHOS=3, xlen=30.0, igeom=1, Tramp=3.0, xabsf=0.8,
                                        components=js(0.05, 1.5, 64, 1, 11)),
        "cfg4 member 2-D 1024 M=4": dict(base, n1=1024, n2=1, mHOS=4, xlen=40.0, igeom=1, Tramp=3.0, xabsf=0.8,
                                         components=js(0.07, 1.8, 64, 1, 12)),
        "cfg4 member 2-D 2048 M=3": dict(base, n1=2048, n2=1, mHOS=3, xlen=60.0, igeom=1, Tramp=3.0, xabsf=0.8,
                                         components=js(0.09, 1.3, 64, 1, 13)),
        "cfg5 member 3-D 512x128 M=7": dict(base, n1=512, n2=128, mHOS=7, xlen=20.0, ylen=5.0, igeom=1, Tramp=2.8,
                                            xabsf=0.85, components=js(0.06, 1.4, 32, 8, 5)),
    }


def tune_report(steps=20):
    """Schedules the autotuner picks on the configuration shapes, with fixed-h attempts/s
    (h = 0.5 dt_stab, 3 warm-up attempts) under the tuned and under the default schedule
    (forced with HOSNWT_FFT_SCHED=0,0,8), one context at a time, FP64 and mixed."""
    import os
    from hosnwt.tank import Tank
    out = {}
    for name, kw in config_shapes().items():
        for prec in (0, 1):
            s = hc.build(hc.TankConfig(precision=prec, **kw))
            n = steps if kw["n2"] > 1 else 3 * steps
            res = {}
            for mode in ("default", "tuned"):
                if mode == "default":
                    os.environ["HOSNWT_FFT_SCHED"] = "0,0,8"
                else:
                    os.environ.pop("HOSNWT_FFT_SCHED", None)
                tk = Tank(s)
                tk.initialize()
                h = 0.5 * s.dt_stab
                t = 0.0
                for _ in range(3):
                    tk.runge4_adapt2(t, h)
                    tk.commit()
                    t += h
                t0 = time.perf_counter()
                for _ in range(n):
                    tk.runge4_adapt2(t, h)
                    tk.commit()
                    t += h
                dt = time.perf_counter() - t0
                res[mode] = round(n / dt, 1)
                if mode == "tuned":
                    res["schedule"] = tk.schedule()
                tk.close()
            os.environ.pop("HOSNWT_FFT_SCHED", None)
            res["tuned_vs_default"] = round(res["tuned"] / res["default"], 3)
            out[f"{name} {'fp64' if prec == 0 else 'mixed'}"] = res
            print(f"[tune] {name} prec {prec}: {res}", file=sys.stderr, flush=True)
    return out


def ensemble_members(n_members=64, precision=1, seed=2024):
    """The batched-ensemble workload: 2-D irregular realisations, length 60, n1 drawn from
    {512, 1024, 2048, 4096}, M from {3, 4, 5, 6}, JONSWAP Hs ~ U[0.03, 0.10], Tp ~ U[1.2, 2.0],
    gamma 3.3, 512 components with seeded phases, piston wavemaker, ramp 2 Tp,
    RK4(5) tolerance 1e-7."""
    rng = np.random.default_rng(seed)
    members = []
    for _ in range(n_members):
        n1 = int(rng.choice([512, 1024, 2048, 4096]))
        M = int(rng.choice([3, 4, 5, 6]))
        Hs, Tp = float(rng.uniform(0.03, 0.10)), float(rng.uniform(1.2, 2.0))
        comps = hc.jonswap_components(Hs, Tp, 3.3, n_freq=512, seed=int(rng.integers(1 << 30)), rnum=12,
                                      clock=32.0)
        cfg = hc.TankConfig(n1=n1, n2=1, mHOS=M, xlen=60.0, n3=33, icase=3, igeom=1, iramp=1, Tramp=2.0 * Tp,
                            xabsf=0.8, T_stop=100 * Tp, toler=1e-7, f_out=1.0 / Tp, precision=precision,
                            components=comps)
        members.append(hc.build(cfg))
    return members


def ensemble_bench(t_next, n_members=64):
    """The batched-ensemble workload: n_members independent 2-D irregular realisations, n1 drawn
    from {512, 1024, 2048, 4096} and M from {3, 4, 5, 6}, JONSWAP Hs ~ U[0.03, 0.10],
    Tp ~ U[1.2, 2.0], gamma 3.3, 512 components with seeded phases, length 60, ramp 2 Tp, mixed
    precision, RK4(5) tolerance 1e-7; every member advanced from rest to t_next (nondimensional),
    once member after member (hos_advance) and once together (hos_advance_ensemble)."""
    from hosnwt.tank import Tank, advance_ensemble
    members = ensemble_members(n_members)
    h0 = [min(0.95 * s.dt_stab, s.delt_out) for s in members]

    def make():
        t0 = time.perf_counter()
        tanks = [Tank(s) for s in members]
        for tk in tanks:
            tk.initialize()
        for tk in tanks:
            tk.synchronize()
        print(f"[ensemble] {len(tanks)} contexts created in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
              flush=True)
        return tanks

    tanks = make()
    t0 = time.perf_counter()
    seq = []
    for tk, h in zip(tanks, h0):
        seq.append(tk.advance(t_next, h))
    dt_seq = time.perf_counter() - t0
    seq_att = sum(na + nr for _, na, nr in seq)
    print(f"[ensemble] member after member: {seq_att} attempts in {dt_seq:.2f} s", file=sys.stderr, flush=True)
    for tk in tanks:
        tk.close()
    tanks = make()
    t0 = time.perf_counter()
    _, nas, nrs = advance_ensemble(tanks, t_next, h0)
    dt_ens = time.perf_counter() - t0
    ens_att = sum(nas) + sum(nrs)
    print(f"[ensemble] together: {ens_att} attempts in {dt_ens:.2f} s", file=sys.stderr, flush=True)
    for tk in tanks:
        tk.close()
    assert ens_att == seq_att, (ens_att, seq_att)
    # graph-captured attempts, controllers on the GPU, one host wait per round (the graphs are
    # captured in a first, untimed interval of one output period)
    from hosnwt.tank import advance_ensemble_graph
    tanks = make()
    t_warm = min(1.0, 0.5 * t_next)
    hg, _, _ = advance_ensemble_graph(tanks, t_warm, h0)
    t0 = time.perf_counter()
    _, nag, nrg = advance_ensemble_graph(tanks, t_next, hg)
    dt_graph = time.perf_counter() - t0
    graph_att = sum(nag) + sum(nrg)
    print(f"[ensemble] graphs: {graph_att} attempts from t = {t_warm} in {dt_graph:.2f} s", file=sys.stderr,
          flush=True)
    for tk in tanks:
        tk.close()
    return {"metric": f"ragged 2-D ensemble RK4(5) steps/s ({n_members} members, n1 512-4096, M 3-6)",
            "value": round(ens_att / dt_ens, 2), "unit": "attempts/s summed over members", "n_gpus": 1,
            "higher_is_better": True, "dtype": "fp32 transforms, fp64 RK", "data": "synthetic (seeded JONSWAP)",
            "t_next": t_next, "attempts": ens_att, "seconds": round(dt_ens, 3),
            "sequential_value": round(seq_att / dt_seq, 2), "sequential_seconds": round(dt_seq, 3),
            "ensemble_vs_sequential": round(dt_seq / dt_ens, 3),
            "graph_value": round(graph_att / dt_graph, 2), "graph_attempts": graph_att,
            "graph_seconds": round(dt_graph, 3), "graph_from_t": t_warm,
            "members": {f"n1={n1}": sum(s.cfg.n1 == n1 for s in members) for n1 in (512, 1024, 2048, 4096)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--spin-up", type=float, default=None, metavar="T",
                    help="untimed run from rest to nondimensional time T before the warm-up attempts "
                         "(default: the end of the wavemaker ramp, 2 Tp)")
    ap.add_argument("--precision", type=int, default=0, help="precision of --pass-probe (0 FP64, 1 FP32)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--pass-probe", type=int, default=0, metavar="REPS",
                    help="only time the roofline passes REPS times and run the counter-calibration copy")
    ap.add_argument("--pass-log-probe", type=int, default=0, metavar="STEPS",
                    help="only run STEPS attempts with every extended-grid pass launch timed (profiling)")
    ap.add_argument("--learned", type=float, default=0.0, metavar="T",
                    help="only run the learned-order ensemble (16 x 512x128, M_max = 7) to time T against "
                         "the fixed-order solver")
    ap.add_argument("--small-tanks", action="store_true",
                    help="only time the small 2-D tanks, stream path against graph-captured attempts")
    ap.add_argument("--mlp-probe", type=int, default=0, metavar="EPOCHS",
                    help="only train the order-prediction network on synthetic rows (profiling)")
    ap.add_argument("--tune-report", action="store_true",
                    help="only report the autotuned schedules of the configuration shapes and their "
                         "attempts/s against the default schedules")
    ap.add_argument("--ensemble", type=float, default=0.0, metavar="T",
                    help="only run the ragged 2-D ensemble to nondimensional time T (sequential vs together)")
    args = ap.parse_args()
    if args.pass_probe:
        pass_probe(args.pass_probe, args.precision)
        return
    if args.pass_log_probe:
        pass_log_probe(args.pass_log_probe, args.precision)
        return
    if args.ensemble:
        print(json.dumps(ensemble_bench(args.ensemble)), flush=True)
        return
    if args.tune_report:
        print(json.dumps(tune_report(args.steps)), flush=True)
        return
    if args.learned:
        from hosnwt import learned as hl
        r = hl.experiment(n_members=16, n_train=4, t_end=args.learned,
                          log=lambda m: print(m, file=sys.stderr, flush=True))
        r["metric"] = "learned-order ensemble attempts/s (16 x 3D 512x128, M_max=7) vs fixed order"
        print(json.dumps(r), flush=True)
        return
    if args.small_tanks:
        print(json.dumps(small_tank_probe()), flush=True)
        return
    if args.mlp_probe:
        mlp_probe(args.mlp_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported in this build "
              "(ensembles and a single field both fit one MI355X).", file=sys.stderr)
        sys.exit(2)
    spin = make_setup(0).Tramp if args.spin_up is None else args.spin_up
    print(f"[bench] FP64 path: spin-up to t = {spin:.3f}, {args.warmup} warmup + {args.steps} timed attempts",
          file=sys.stderr, flush=True)
    r64 = time_path(0, args.steps, args.warmup, spin)
    print(f"[bench] FP64 {r64['steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)
    roof64 = roofline(r64, 0, args.steps)
    sched = {"fp64": r64["tank"].schedule()}
    # one live context at a time, the deployment shape: with the FP64 context still alive the
    # mixed path measured 114 instead of 135 steps/s (its two streams then share the process's
    # four hardware queues with the idle context's streams)
    r64["tank"].close()
    r64["tank"] = None
    r32 = None
    if not args.fp64_only:
        print("[bench] mixed path", file=sys.stderr, flush=True)
        # free-running with its own error estimates and decisions, to the FP64 run's times
        r32 = time_path(1, args.steps, args.warmup, spin, follow=(r64["t_warm"], r64["t_end"]))
        print(f"[bench] mixed {r32['steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)
    eta_err = None
    if r32 is not None:
        d = np.linalg.norm(r32["eta"] - r64["eta"])
        eta_err = float(d / max(np.linalg.norm(r64["eta"]), 1e-300))
    use_mixed = r32 is not None and eta_err is not None and eta_err <= 1e-5
    best = r32 if use_mixed else r64
    # the roofline object describes the headline path's dominant kernel (the FP32 passes when
    # the mixed path is the headline)
    roof = roofline(r32, 1, args.steps) if use_mixed else roof64
    if r32 is not None:
        sched["mixed"] = r32["tank"].schedule()
    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline (reference-mode attempt on the host cores)", file=sys.stderr, flush=True)
        cpu = cpu_baseline()
    for r in (r64, r32):
        if r is not None and r["tank"] is not None:
            r["tank"].close()
    line = {
        "metric": "RK4(5) steps/s (3D 1024x256, M=5) mixed/learned vs FP64 GPU at eta err <=1e-5",
        "value": round(best["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transforms + f64 RK state" if use_mixed else "f64",
        "data": "synthetic (seeded JONSWAP directional sea)",
        "config": {"workload": "3D directional sea 40x10, n1=1024 x n2=256, M=5, n3=33, 256 freq x 32 dir",
                   "path": "mixed" if use_mixed else "fp64", "big_grid": list(r64["big_grid"]),
                   "big_transforms_per_step": r64["big_per_step"], "base_transforms_per_step": r64["base_per_step"]},
        "fp64_steps_per_s": round(r64["steps_per_s"], 3),
        "mixed_steps_per_s": round(r32["steps_per_s"], 3) if r32 else None,
        "mixed_vs_fp64": round(r32["steps_per_s"] / r64["steps_per_s"], 3) if r32 else None,
        "mixed_eta_rel_err": eta_err,
        "accepted_rejected": [r64["n_acc"], r64["n_rej"]],
        "mixed_accepted_rejected": [r32["n_acc"], r32["n_rej"]] if r32 else None,
        "simulated_time": [r64["t_warm"], r64["t_end"]],
        "roofline": roof,
        "roofline_fp64": roof64 if use_mixed else None,
        "fft_schedule": sched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
