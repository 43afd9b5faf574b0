"""``Tank``: the reference driver loop around the GPU time-stepping core.

Mirrors HOS-NWT.f90:446-581: per output interval, the adaptive RK4(3) loop
(on the GPU, through hos_advance), then volume and energy.  The operator names
follow the reference: ``runge4_adapt2`` (one attempt), ``solve_hos2`` (one
RHS), ``space_2_fourier`` / ``fourier_2_space``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import config as cfgmod
from .lib import HosParams, HosStats, check, fortran, load, ptr


class Tank:
    """One GPU context holding the solver state of a tank."""

    def __init__(self, setup: cfgmod.Setup, precision: int | None = None):
        self.s = setup
        cfg = setup.cfg
        self.cfg = cfg
        self.lib = load()
        self.n1, self.n2, self.n3 = cfg.n1, cfg.n2, cfg.n3
        self.n3a = 2 * (cfg.n3 - 1)
        prec = cfg.precision if precision is None else precision
        self._keep = []
        nu = fortran(setup.nu)
        pr = fortran(setup.pos_ini.real)
        pi = fortran(setup.pos_ini.imag)
        fr = np.ascontiguousarray(setup.freq, dtype=np.float64)
        self._keep += [nu, pr, pi, fr]
        dy = setup.dposdy_ini if setup.dposdy_ini is not None else np.zeros_like(setup.pos_ini)
        dyr, dyi = fortran(dy.real), fortran(dy.imag)
        nux = fortran(setup.nu_x) if setup.nu_x is not None else None
        self._keep += [dyr, dyi, nux]
        p = HosParams()
        p.n1, p.n2, p.n3, p.mHOS = cfg.n1, cfg.n2, cfg.n3, cfg.mHOS
        p.p1 = cfg.mHOS
        p.p2 = cfg.mHOS if cfg.n2 != 1 else 1
        p.icase, p.igeom, p.i_wmk, p.iramp, p.iabsnb = cfg.icase, cfg.igeom, cfg.i_wmk, cfg.iramp, cfg.iabsnb
        p.precision = prec
        p.xlen, p.ylen, p.d_hinge = setup.xlen, setup.ylen, cfg.d_hinge / cfg.h
        p.Tramp, p.T_stop, p.T_mono = setup.Tramp, setup.T_stop, setup.T_mono
        p.nu = ptr(nu)
        p.n_freq = len(fr)
        p.freq = ptr(fr) if len(fr) else None
        p.pos_ini_re = ptr(pr) if len(fr) else None
        p.pos_ini_im = ptr(pi) if len(fr) else None
        for i in range(3):
            p.coeffilt[i] = float(cfg.coeffilt[i])
        p.dposdy_ini_re = ptr(dyr) if len(fr) else None
        p.dposdy_ini_im = ptr(dyi) if len(fr) else None
        p.nu_x = ptr(nux) if nux is not None else None
        p.nu_y = None  # no side beach
        if setup.dip_mu is not None and len(setup.dip_mu):
            mu = np.asarray(setup.dip_mu, complex)
            dip = [np.ascontiguousarray(a, dtype=np.float64) for a in
                   (setup.dip_freq, mu[:, 0].real, mu[:, 0].imag, mu[:, 1].real, mu[:, 1].imag, mu[:, 2].real)]
            self._keep += dip
            p.n_dip = len(mu)
            p.dip_freq, p.dip_mux_re, p.dip_mux_im, p.dip_muz_re, p.dip_muz_im, p.dip_z = (ptr(a) for a in dip)
            p.x_dip, p.n_side = setup.x_dip, setup.n_side
        h = C.c_void_p()
        check(self.lib.hos_create(C.byref(p), C.byref(h)))
        self.h = h
        self.time = 0.0
        self.n_rk_tot = 0
        self.n_er_tot = 0

    def close(self):
        if self.h:
            check(self.lib.hos_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state -----------------------------------------------------------
    def initialize(self):
        """Initial wavemaker modes from rest (initial_condition.f90:314-466)."""
        check(self.lib.hos_initialize(self.h))
        self.time = 0.0

    def set_state(self, eta, phis, a1st_add=None, a_add_ramp=None, a1st=None, eta1st=None):
        arrs = [fortran(eta), fortran(phis)] + [None if a is None else fortran(a)
                                                 for a in (a1st_add, a_add_ramp, a1st, eta1st)]
        check(self.lib.hos_set_state(self.h, *[ptr(a) for a in arrs]))

    def get_state(self):
        n1, n2, na = self.n1, self.n2, self.n3a
        out = [np.zeros((n1, n2), order="F"), np.zeros((n1, n2), order="F"), np.zeros((na, n2), order="F"),
               np.zeros((na, n2), order="F"), np.zeros((n1, n2), order="F"), np.zeros((n1, n2), order="F")]
        check(self.lib.hos_get_state(self.h, *[ptr(a) for a in out]))
        return dict(zip(("eta", "phis", "a1st_add", "a_add_ramp", "a1st", "eta1st"), out))

    def set_state2(self, a2nd_add=None, a2nd=None, eta2nd=None):
        """Second-order wavemaker state (i_wmk = 2): paddle modes, modal a2nd, physical eta2nd."""
        arrs = [None if a is None else fortran(a) for a in (a2nd_add, a2nd, eta2nd)]
        check(self.lib.hos_set_state2(self.h, *[ptr(a) for a in arrs]))

    def get_state2(self):
        n1, n2, na = self.n1, self.n2, self.n3a
        out = [np.zeros((na, n2), order="F"), np.zeros((n1, n2), order="F"), np.zeros((n1, n2), order="F")]
        check(self.lib.hos_get_state2(self.h, *[ptr(a) for a in out]))
        return dict(zip(("a2nd_add", "a2nd", "eta2nd"), out))

    def set_state3(self, a3rd_add):
        """Third-order wavemaker paddle modes (i_wmk = 3)."""
        check(self.lib.hos_set_state3(self.h, ptr(fortran(a3rd_add))))

    def get_state3(self):
        out = np.zeros((self.n3a, self.n2), order="F")
        check(self.lib.hos_get_state3(self.h, ptr(out)))
        return out

    def physical(self):
        n1, n2 = self.n1, self.n2
        out = [np.zeros((n1, n2), order="F") for _ in range(4)]
        check(self.lib.hos_get_physical(self.h, *[ptr(a) for a in out]))
        return dict(zip(("eta", "phis", "deta", "phi_add"), out))

    # ---- operators ---------------------------------------------------------
    def runge4_adapt2(self, t: float, h: float) -> float:
        """One attempt; the trial state stays on the GPU until commit()."""
        err = C.c_double()
        check(self.lib.hos_rk_step(self.h, t, h, C.byref(err)))
        return err.value

    def commit(self):
        check(self.lib.hos_commit(self.h))

    def end_ramp(self):
        check(self.lib.hos_end_ramp(self.h))

    def solve_hos2(self, t, delta_t, G1, G2, a1st, eta1st, a_add_1):
        n1, n2, na = self.n1, self.n2, self.n3a
        ins = [fortran(x) for x in (G1, G2, a1st, eta1st, a_add_1)]
        outs = [np.zeros((n1, n2), order="F") for _ in range(4)] + [np.zeros((na, n2), order="F") for _ in range(2)]
        check(self.lib.hos_rhs(self.h, t, delta_t, *[ptr(a) for a in ins], *[ptr(a) for a in outs]))
        return dict(zip(("dG1", "dG2", "da1st", "deta1st", "da_add_1", "da_add_ramp"), outs))

    def solve_hos2_second(self, t, delta_t, G1, G2, a1st, eta1st, a_add_1, a2nd, eta2nd, a_add_2):
        """One RHS with the second-order wavemaker state (i_wmk = 2)."""
        n1, n2, na = self.n1, self.n2, self.n3a
        ins = [fortran(x) for x in (G1, G2, a1st, eta1st, a_add_1, a2nd, eta2nd, a_add_2)]
        outs = ([np.zeros((n1, n2), order="F") for _ in range(4)] + [np.zeros((na, n2), order="F") for _ in range(2)]
                + [np.zeros((n1, n2), order="F") for _ in range(2)] + [np.zeros((na, n2), order="F")])
        check(self.lib.hos_rhs2(self.h, t, delta_t, *[ptr(a) for a in ins], *[ptr(a) for a in outs]))
        return dict(zip(("dG1", "dG2", "da1st", "deta1st", "da_add_1", "da_add_ramp", "da2nd", "deta2nd",
                         "da_add_2"), outs))

    def solve_hos2_third(self, t, delta_t, G1, G2, a1st, eta1st, a_add_1, a2nd, eta2nd, a_add_2, a_add_3):
        """One RHS with the second- and third-order wavemaker state (i_wmk = 3)."""
        n1, n2, na = self.n1, self.n2, self.n3a
        ins = [fortran(x) for x in (G1, G2, a1st, eta1st, a_add_1, a2nd, eta2nd, a_add_2, a_add_3)]
        outs = ([np.zeros((n1, n2), order="F") for _ in range(4)] + [np.zeros((na, n2), order="F") for _ in range(2)]
                + [np.zeros((n1, n2), order="F") for _ in range(2)] + [np.zeros((na, n2), order="F") for _ in range(2)])
        check(self.lib.hos_rhs3(self.h, t, delta_t, *[ptr(a) for a in ins], *[ptr(a) for a in outs]))
        return dict(zip(("dG1", "dG2", "da1st", "deta1st", "da_add_1", "da_add_ramp", "da2nd", "deta2nd",
                         "da_add_2", "da_add_3"), outs))

    def advance(self, t_next: float, h_rk: float):
        """Adaptive loop to t_next (HOS-NWT.f90:498-571); returns the next step proposal."""
        t = C.c_double(self.time)
        h = C.c_double(h_rk)
        na, nr = C.c_int(), C.c_int()
        check(self.lib.hos_advance(self.h, C.byref(t), t_next, C.byref(h), self.cfg.toler, self.s.dt_stab,
                                   self.s.delt_out, C.byref(na), C.byref(nr)))
        self.time = t.value
        self.n_rk_tot += na.value
        self.n_er_tot += nr.value
        return h.value, na.value, nr.value

    def advance_graph(self, t_next: float, h_rk: float):
        """advance() with the attempt captured in a HIP graph and the controller on the GPU
        (hos_advance_graph): no host synchronisation per attempt."""
        t = C.c_double(self.time)
        h = C.c_double(h_rk)
        na, nr = C.c_int(), C.c_int()
        check(self.lib.hos_advance_graph(self.h, C.byref(t), t_next, C.byref(h), self.cfg.toler, self.s.dt_stab,
                                         self.s.delt_out, C.byref(na), C.byref(nr)))
        self.time = t.value
        self.n_rk_tot += na.value
        self.n_er_tot += nr.value
        return h.value, na.value, nr.value

    def modes(self):
        """Modal eta and phis of the committed state (HOS_modes.dat, output.f90:380-383)."""
        out = [np.zeros((self.n1, self.n2), order="F") for _ in range(2)]
        check(self.lib.hos_get_modes(self.h, *[ptr(a) for a in out]))
        return dict(zip(("eta", "phis"), out))

    def swense_modes(self, t: float, meta2: int = 10):
        """SWENSE output of the committed state at time t (velocities.f90:223-784): the modes of
        the velocities and of phi_t at z = 0 (specx, specy, specz, spect), modesFS, modesFSt
        (n1 x n2) and modesadd, modesaddt (n3_add x n2), as modes_HOS_SWENSE.dat holds them."""
        n1, n2, na = self.n1, self.n2, self.n3a
        out = [np.zeros((n1, n2), order="F") for _ in range(6)] + [np.zeros((na, n2), order="F") for _ in range(2)]
        check(self.lib.hos_swense_modes(self.h, t, meta2, *[ptr(a) for a in out]))
        return dict(zip(("specx", "specy", "specz", "spect", "FS", "FSt", "add", "addt"), out))

    def probes(self, xp, yp=None):
        """Elevation at wave probes (nondimensional positions; output.f90:266-276)."""
        xp = np.ascontiguousarray(xp, dtype=np.float64)
        yp = None if yp is None else np.ascontiguousarray(yp, dtype=np.float64)
        out = np.zeros(len(xp))
        check(self.lib.hos_probes(self.h, len(xp), ptr(xp), ptr(yp), ptr(out)))
        return out

    def wmk_motion(self):
        """alpha, alpha', paddle-top displacement and velocity per y node (wmk_motion.dat)."""
        a, at = C.c_double(), C.c_double()
        x = np.zeros(self.n2)
        v = np.zeros(self.n2)
        check(self.lib.hos_get_wmk_motion(self.h, C.byref(a), C.byref(at), ptr(x), ptr(v)))
        return a.value, at.value, x, v

    def slope_max(self):
        sx, sy = C.c_double(), C.c_double()
        check(self.lib.hos_get_slope_max(self.h, C.byref(sx), C.byref(sy)))
        return sx.value, sy.value

    def check_slope(self, threshold: float = 8.0, coeff: float = 0.7):
        """check_slope (initial_condition.f90:687-716): filter eta, phis on steep slopes."""
        fx, fy = C.c_int(), C.c_int()
        check(self.lib.hos_check_slope(self.h, threshold, coeff, C.byref(fx), C.byref(fy)))
        return bool(fx.value), bool(fy.value)

    # ---- truncation order ------------------------------------------------
    def set_order(self, m: int):
        """Runs the following attempts at HOS order m <= mHOS (the fixed-M solver at M = m)."""
        check(self.lib.hos_set_order(self.h, int(m)))

    def order(self) -> int:
        m = C.c_int()
        check(self.lib.hos_get_order(self.h, C.byref(m)))
        return m.value

    def record_features(self, mode: int):
        """0 off, 1 the first right-hand side of each attempt, 2 every right-hand side."""
        check(self.lib.hos_record_features(self.h, int(mode)))

    def take_features(self, max_rec: int = 512) -> np.ndarray:
        """Band-energy records since the last take (oldest first), shape (n, FEAT_REC)."""
        from .lib import FEAT_REC
        out = np.zeros((max_rec, FEAT_REC))
        n = C.c_int()
        check(self.lib.hos_take_features(self.h, max_rec, ptr(out), C.byref(n)))
        return out[: n.value].copy()

    def learned_last(self):
        """(record, predictions) of the member's last learned-order attempt."""
        from .lib import FEAT_REC, fptr
        rec = np.zeros(FEAT_REC)
        pred = np.zeros((64, 64), dtype=np.float32)
        check(self.lib.hos_learned_last(self.h, ptr(rec), fptr(pred)))
        return rec, pred

    def stats(self):
        s = HosStats()
        check(self.lib.hos_get_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in HosStats._fields_}

    def time_big_transform(self, reps: int = 20):
        """HIP-event timing of one extended-grid x pass and y pass; returns the slower one."""
        out = np.zeros(4)
        check(self.lib.hos_time_big_transform(self.h, reps, ptr(out)))
        if out[2] > out[0]:
            return dict(kernel="r2r_seq_kernel y pass (extended grid, inverse cos)", ms=out[2],
                        bytes=out[3], other=dict(ms=out[0], bytes=out[1]), **{"pass": "y"})
        return dict(kernel="r2r_seq_kernel x pass (extended grid, inverse cos)", ms=out[0], bytes=out[1],
                    other=dict(ms=out[2], bytes=out[3]), **{"pass": "x"})

    def time_big_batch(self, reps: int = 20, items: int = 4):
        """HIP-event timing of the extended-grid passes with `items` full inverse cos-cos
        transforms per launch (the batched form of a right-hand side): per pass the
        average launch duration and the items' algorithmic bytes."""
        out = np.zeros(4)
        check(self.lib.hos_time_big_batch(self.h, reps, items, ptr(out)))
        return dict(x=dict(ms=out[0], bytes=out[1]), y=dict(ms=out[2], bytes=out[3]), items=items)

    def pass_log_begin(self, max_launches: int = 8192):
        """Start timing every extended-grid pass launch this thread enqueues for the tank
        (two HIP events per launch on its stream; see hos_pass_log_begin)."""
        check(self.lib.hos_pass_log_begin(self.h, max_launches))

    def pass_log_end(self):
        """Stop timing: per pass the launches, their algorithmic bytes, summed durations (ms)
        and transforms, and whether launches beyond the log's capacity went unrecorded."""
        out = np.zeros(9)
        check(self.lib.hos_pass_log_end(self.h, ptr(out)))
        return dict(x=dict(launches=int(out[0]), bytes=out[1], ms=out[2], transforms=int(out[3])),
                    y=dict(launches=int(out[4]), bytes=out[5], ms=out[6], transforms=int(out[7])),
                    full=bool(out[8]))

    def schedule(self):
        """The active engine's extended-grid transform schedule (see hos_get_schedule)."""
        s8 = (C.c_int * 8)()
        us = np.zeros(2)
        check(self.lib.hos_get_schedule(self.h, s8, ptr(us)))
        keys = ("x_length", "x_values", "x_seq", "y_length", "y_values", "y_seq", "y_tile", "trials")
        out = {k: int(v) for k, v in zip(keys, s8)}
        out.update(us_default=float(us[0]), us_tuned=float(us[1]))
        return out

    def synchronize(self):
        check(self.lib.hos_synchronize(self.h))

    # ---- diagnostics (vol_energy.f90:36-119) --------------------------------
    def volume(self, eta):
        n1, n2 = self.n1, self.n2
        delx = self.s.xlen / (n1 - 1)
        w1 = np.ones(n1)
        w1[[0, -1]] = 0.5
        if n2 == 1:
            return float((eta[:, 0] * w1).sum() * delx / self.s.xlen)
        w2 = np.ones(n2)
        w2[[0, -1]] = 0.5
        dely = self.s.ylen / (n2 - 1)
        return float((eta * w2[None, :]).sum(axis=1) @ w1 * delx / self.s.xlen / self.s.ylen * dely)

    def energy(self, eta, phis, deta):
        n1, n2 = self.n1, self.n2
        delx = self.s.xlen / (n1 - 1)
        w1 = np.ones(n1)
        w1[[0, -1]] = 0.5

        def integ(f):
            if n2 == 1:
                return float((f[:, 0] * w1).sum() * delx / 2.0)
            w2 = np.ones(n2)
            w2[[0, -1]] = 0.5
            dely = self.s.ylen / (n2 - 1)
            return float((f * w2[None, :]).sum(axis=1) @ w1 * delx / 2.0 * dely)
        ep, ek = integ(eta * eta), integ(phis * deta)
        return ep, ek, ep + ek

    # ---- the driver loop -------------------------------------------------------
    def run(self, T_stop: float | None = None, on_output=None):
        """Time loop with outputs every 1/f_out (HOS-NWT.f90:446-581)."""
        s = self.s
        T_stop = s.T_stop if T_stop is None else T_stop
        delt_out = s.delt_out
        delt = min(0.95 * s.dt_stab, delt_out)
        records = []
        while self.time < T_stop + delt_out:
            ph = self.physical()
            vol = self.volume(ph["eta"])
            en = self.energy(ph["eta"], ph["phis"] + ph["phi_add"], ph["deta"])
            rec = dict(t=self.time, volume=vol, energy=en, n_rk=self.n_rk_tot, n_er=self.n_er_tot)
            records.append(rec)
            if on_output:
                on_output(rec, ph)
            if abs(self.time - T_stop) <= np.finfo(float).eps:
                break
            t_next = min(self.time + delt_out, T_stop + delt_out)
            delt, _, _ = self.advance(t_next, delt)
        return records


def advance_ensemble_graph(tanks, t_next: float, h_rk):
    """hos_advance_ensemble_graph: every tank to t_next by replays of its captured attempt,
    the controllers on the GPU, one host wait per round of replays."""
    n = len(tanks)
    lib = tanks[0].lib
    ctx = (C.c_void_p * n)(*[t.h.value for t in tanks])
    t = np.array([tk.time for tk in tanks], dtype=np.float64)
    h = np.array(h_rk, dtype=np.float64)
    stab = np.array([tk.s.dt_stab for tk in tanks], dtype=np.float64)
    dout = np.array([tk.s.delt_out for tk in tanks], dtype=np.float64)
    na = (C.c_int * n)()
    nr = (C.c_int * n)()
    toler = tanks[0].cfg.toler
    check(lib.hos_advance_ensemble_graph(n, ctx, ptr(t), t_next, ptr(h), toler, ptr(stab), ptr(dout), na, nr))
    for i, tk in enumerate(tanks):
        tk.time = float(t[i])
        tk.n_rk_tot += na[i]
        tk.n_er_tot += nr[i]
    return h, list(na), list(nr)


def advance_ensemble(tanks, t_next: float, h_rk):
    """hos_advance_ensemble: every tank to t_next, the members' attempts overlapping on the GPU.
    Returns (h proposals, accepted counts, rejected counts); each tank's time is updated."""
    n = len(tanks)
    if n == 0:
        return np.zeros(0), [], []
    if len(h_rk) != n:
        raise ValueError("one step proposal per tank")
    lib = tanks[0].lib
    ctx = (C.c_void_p * n)(*[t.h.value for t in tanks])
    t = np.array([tk.time for tk in tanks], dtype=np.float64)
    h = np.array(h_rk, dtype=np.float64)
    stab = np.array([tk.s.dt_stab for tk in tanks], dtype=np.float64)
    dout = np.array([tk.s.delt_out for tk in tanks], dtype=np.float64)
    na = (C.c_int * n)()
    nr = (C.c_int * n)()
    toler = min(tk.cfg.toler for tk in tanks)
    if any(tk.cfg.toler != toler for tk in tanks):
        raise ValueError("the ensemble shares one RK tolerance")
    rc = lib.hos_advance_ensemble(n, ctx, ptr(t), t_next, ptr(h), toler, ptr(stab), ptr(dout), na, nr)
    # the library reports every member's committed progress, also on an error
    for i, tk in enumerate(tanks):
        tk.time = float(t[i])
        tk.n_rk_tot += na[i]
        tk.n_er_tot += nr[i]
    check(rc)
    return h, list(na), list(nr)
