"""Learned truncation order: band-energy features, the order-prediction network
and the ensemble driver that picks the HOS order per attempt and per member.

The reference fixes the HOS order M at compile time (the loop bound of
HOSphis_modes, resol_HOS.f90:720-782; the dealiasing grid of
common_vars.f90:38-39).  Here a context runs any order m <= mHOS on the grid of
a context created with mHOS = m, the network predicts the band energies of the
first omitted order from those of the orders evaluated, and the fixed-M solver
(m = mHOS) is the judge: every few attempts it re-runs an attempt, and a member
whose eta drifts faster than the tolerance returns to its last judged state
(hos_advance_learned, include/hosnwt.h).

Features of one right-hand side (hos_record_features): the sums of eta_hat^2
and of (k tanh k a_j)^2 (a_j the modal potential of order j) over 64 radial
wavenumber bands.  Rows and targets of the network follow ``feature_rows``, the
same formula as the device kernel k_feat_rows.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .lib import (FEAT_REC, HosLearnedStats, HosOrderPolicy, check, fptr, load, ptr)

N_IN, WIDTH, N_OUT = 512, 256, 64
BANDS = 64
FEAT_M = 9 * BANDS        # record index of the order evaluated
FEAT_ETA = FEAT_M + 1     # record index of sum eta_hat^2
PRED_ORDERS = 6           # rows predict orders 2..7
DIMS = (N_IN, WIDTH, WIDTH, WIDTH, N_OUT)


def scale(x):
    """s(x) = (log10(max(x, 1e-30)) + 15) / 5: relative band energies to O(1) inputs."""
    return (np.log10(np.maximum(x, 1e-30)) + 15.0) / 5.0


def unscale(s):
    return 10.0 ** (5.0 * np.asarray(s, dtype=np.float64) - 15.0)


def feature_rows(rec: np.ndarray, with_targets: bool = True):
    """Network rows of one record: row for m' = 1..min(m_run - 1, 6) (targets need order
    m' + 1 in the record), or m' = 1..min(m_run, 6) without targets.  Returns
    (x [k, 512] float32, t [k, 64] float32 or None, m' list)."""
    m_run = int(rec[FEAT_M])
    et = rec[FEAT_ETA]
    inv = 1.0 / et if et > 0.0 else 0.0
    top = min(m_run - 1 if with_targets else m_run, PRED_ORDERS)
    xs, ts, ms = [], [], []
    for mp in range(1, top + 1):
        x = np.zeros(N_IN)
        x[0:BANDS] = scale(rec[0:BANDS] * inv)
        for blk in range(1, mp + 1):
            x[blk * BANDS:(blk + 1) * BANDS] = scale(rec[blk * BANDS:(blk + 1) * BANDS] * inv)
        x[7 * BANDS + mp] = 1.0
        x[7 * BANDS + 8] = (np.log10(max(et, 1e-30)) + 10.0) / 10.0
        xs.append(x.astype(np.float32))
        ms.append(mp)
        if with_targets:
            ts.append(scale(rec[(mp + 1) * BANDS:(mp + 2) * BANDS] * inv).astype(np.float32))
    x = np.array(xs, dtype=np.float32).reshape(-1, N_IN)
    t = np.array(ts, dtype=np.float32).reshape(-1, N_OUT) if with_targets else None
    return x, t, ms


def dataset(records) -> tuple[np.ndarray, np.ndarray]:
    """Training rows and targets of every record."""
    xs, ts = [], []
    for rec in records:
        x, t, _ = feature_rows(rec)
        if len(x):
            xs.append(x)
            ts.append(t)
    if not xs:
        return np.zeros((0, N_IN), np.float32), np.zeros((0, N_OUT), np.float32)
    return np.concatenate(xs), np.concatenate(ts)


def rate_of(rec: np.ndarray, j: int) -> float:
    """eta drift rate sqrt(E_j / E_eta) of order j (per unit nondimensional time)."""
    et = rec[FEAT_ETA]
    if not et > 0.0:
        return 0.0
    return float(np.sqrt(rec[j * BANDS:(j + 1) * BANDS].sum() / et))


def predicted_rate(y: np.ndarray) -> float:
    """The rate of the network's predicted band energies (one output row)."""
    return float(np.sqrt(unscale(y).sum()))


class OrderNet:
    """The order-prediction network on the GPU (hos_mlp): 512 -> 256 x 3 -> 64."""

    def __init__(self, max_rows: int = 1024, seed: int = 0):
        self.lib = load()
        n = C.c_longlong()
        check(self.lib.hos_mlp_info(None, None, None, C.byref(n)))
        self.n_params = n.value
        h = C.c_void_p()
        check(self.lib.hos_mlp_create(max_rows, C.byref(h)))
        self.h = h
        self.max_rows = max_rows
        self.set_params(self.initial_params(seed))

    @staticmethod
    def initial_params(seed: int, out_bias: np.ndarray | None = None) -> np.ndarray:
        """He-normal weights (std sqrt(2 / fan_in), the output layer sqrt(1 / fan_in)), zero
        biases (the output bias optionally set, e.g. to the mean target)."""
        rng = np.random.default_rng(seed)
        parts = []
        for l in range(4):
            fin, fout = DIMS[l], DIMS[l + 1]
            std = np.sqrt((2.0 if l < 3 else 1.0) / fin)
            parts.append(rng.standard_normal(fout * fin) * std)
            b = np.zeros(fout)
            if l == 3 and out_bias is not None:
                b = np.asarray(out_bias, dtype=np.float64)
            parts.append(b)
        return np.concatenate(parts).astype(np.float32)

    def close(self):
        if self.h:
            check(self.lib.hos_mlp_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        check(self.lib.hos_mlp_set_params(self.h, fptr(p)))

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        check(self.lib.hos_mlp_get_params(self.h, fptr(p)))
        return p

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], N_OUT), dtype=np.float32)
        check(self.lib.hos_mlp_forward(self.h, x.shape[0], fptr(x), fptr(y)))
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double()
        check(self.lib.hos_mlp_loss_grad(self.h, x.shape[0], fptr(x), fptr(t), C.byref(loss), fptr(g)))
        return loss.value, g

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int = 20, batch: int = 256, lr: float = 1e-3,
              seed: int = 0):
        """Adam on the GPU over (x, t); returns (per-epoch mean squared error, GPU ms)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        check(self.lib.hos_mlp_set_data(self.h, x.shape[0], fptr(x), fptr(t)))
        losses = np.zeros(epochs, dtype=np.float32)
        ms = C.c_double()
        check(self.lib.hos_mlp_train(self.h, epochs, batch, lr, seed, fptr(losses), C.byref(ms)))
        return losses, ms.value

    @staticmethod
    def train_flops(n_samples: int, epochs: int) -> float:
        """Matrix-core FLOPs of training: forward 2 sum(in*out), backward weight and data
        gradients 4 sum(in*out) minus the first layer's data gradient, per sample."""
        macs = sum(DIMS[l] * DIMS[l + 1] for l in range(4))
        per = 2 * macs + 2 * macs + 2 * (macs - DIMS[0] * DIMS[1])
        return float(per) * n_samples * epochs


def collect(tanks, t_next: float, h_rk, chunk: float | None = None, max_records: int | None = None):
    """Advances the tanks at their full order to t_next with every right-hand side's band
    energies recorded; returns the records (oldest first) and the step proposals."""
    h = list(h_rk)
    recs = []
    for tk in tanks:
        tk.record_features(2)
    try:
        for i, tk in enumerate(tanks):
            t = tk.time
            # chunks short enough for the 512-record ring (5 records per attempt)
            step = chunk if chunk is not None else 60.0 * h[i]
            while t < t_next - 1e-14:
                t1 = min(t + step, t_next)
                h[i], _, _ = tk.advance(t1, h[i])
                recs.append(tk.take_features())
                t = tk.time
                if max_records is not None and sum(len(r) for r in recs) >= max_records:
                    break
    finally:
        for tk in tanks:
            tk.record_features(0)
    out = np.concatenate(recs) if recs else np.zeros((0, FEAT_REC))
    return out, h


def policy(m_min: int = 3, judge_every: int = 20, rate_tol: float = 1e-6, safety: float = 0.5) -> HosOrderPolicy:
    p = HosOrderPolicy()
    p.m_min, p.judge_every, p.rate_tol, p.safety = m_min, judge_every, rate_tol, safety
    return p


def advance_learned(tanks, t_next: float, h_rk, net: OrderNet | None, pol: HosOrderPolicy):
    """hos_advance_learned over the tanks to t_next.  Returns (h_rk, n_acc, n_rej, stats)."""
    lib = load()
    n = len(tanks)
    hs = (C.c_void_p * n)(*[tk.h.value for tk in tanks])
    t = np.array([tk.time for tk in tanks], dtype=np.float64)
    h = np.array(h_rk, dtype=np.float64)
    stab = np.array([tk.s.dt_stab for tk in tanks], dtype=np.float64)
    dout = np.array([tk.s.delt_out for tk in tanks], dtype=np.float64)
    na = (C.c_int * n)()
    nr = (C.c_int * n)()
    st = (HosLearnedStats * n)()
    toler = tanks[0].s.cfg.toler
    rc = lib.hos_advance_learned(n, hs, ptr(t), t_next, ptr(h), toler, ptr(stab), ptr(dout),
                                 net.h if net is not None else None, C.byref(pol), na, nr, st)
    for i, tk in enumerate(tanks):
        tk.time = float(t[i])
        tk.n_rk_tot += na[i]
        tk.n_er_tot += nr[i]
    check(rc)
    stats = [dict(attempts=list(s.attempts), judges=s.judges, reverts=s.reverts,
                  max_judged_rate=s.max_judged_rate, max_pred_rate=s.max_pred_rate) for s in st]
    return list(h), list(na), list(nr), stats


def reset(tanks):
    for tk in tanks:
        check(tk.lib.hos_learned_reset(tk.h))


def timed(f, *a, **k):
    t0 = time.perf_counter()
    r = f(*a, **k)
    return r, time.perf_counter() - t0


# ---------------------------------------------------------------------------
# the learned-order ensemble workload and its measurement
# ---------------------------------------------------------------------------
def ensemble_members(n: int = 16, seed: int = 2025, n1: int = 512, n2: int = 128, M: int = 7,
                     precision: int = 1, n_freq: int = 128, n_dir: int = 16, toler: float = 1e-7):
    """3-D directional seas: 40 x 10 basin (h = 1 m), n1 x n2 modes, HOS order M, snake
    wavemaker (one paddle per y node), JONSWAP Hs ~ U[0.04, 0.09] m, Tp ~ U[1.2, 1.8] s,
    gamma 3.3, cos^2s spreading (s = 10) over n_freq x n_dir seeded components, ramp 2 Tp,
    beach on the last 15 %."""
    from . import config as hc
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        Hs, Tp = float(rng.uniform(0.04, 0.09)), float(rng.uniform(1.2, 1.8))
        comps = hc.jonswap_components(Hs, Tp, 3.3, n_freq=n_freq, seed=int(rng.integers(1 << 30)), rnum=12,
                                      clock=32.0, n_dir=n_dir, spread_s=10)
        cfg = hc.TankConfig(n1=n1, n2=n2, mHOS=M, xlen=40.0, ylen=10.0, n3=33, icase=3, igeom=1, iramp=1,
                            Tramp=2.0 * Tp, xabsf=0.85, T_stop=50 * Tp, toler=toler, f_out=1.0 / Tp,
                            precision=precision, components=comps)
        out.append(hc.build(cfg))
    return out


def _run_fixed(setups, outs, log):
    """Fixed order (mHOS) through hos_advance_ensemble; eta at every output time, wall time."""
    from .tank import Tank, advance_ensemble
    tanks = [Tank(s) for s in setups]
    try:
        for tk in tanks:
            tk.initialize()
        h = [min(0.95 * s.dt_stab, s.delt_out) for s in setups]
        etas, wall, att = [], 0.0, 0
        for t_out in outs:
            t0 = time.perf_counter()
            h, na, nr = advance_ensemble(tanks, t_out, h)
            wall += time.perf_counter() - t0
            att += sum(na) + sum(nr)
            etas.append([tk.physical()["eta"] for tk in tanks])
            log(f"  fixed t={t_out:.2f}: {att} attempts, {wall:.2f} s")
        return etas, att, wall
    finally:
        for tk in tanks:
            tk.close()


def _run_learned(setups, outs, net, pol, log):
    from .tank import Tank
    tanks = [Tank(s) for s in setups]
    try:
        for tk in tanks:
            tk.initialize()
        h = [min(0.95 * s.dt_stab, s.delt_out) for s in setups]
        etas, wall, att, stats = [], 0.0, 0, None
        for t_out in outs:
            t0 = time.perf_counter()
            h, na, nr, stats = advance_learned(tanks, t_out, h, net, pol)
            wall += time.perf_counter() - t0
            att += sum(na) + sum(nr)
            etas.append([tk.physical()["eta"] for tk in tanks])
            log(f"  learned t={t_out:.2f}: {att} attempts, {wall:.2f} s")
        return etas, att, wall, stats
    finally:
        for tk in tanks:
            tk.close()


def experiment(n_members: int = 16, n_train: int = 4, t_end: float = 12.0, d_out: float = 2.0,
               n1: int = 512, n2: int = 128, M: int = 7, epochs: int = 20, eta_tols=(1e-5, 1e-4),
               judge_every: int = 20, safety: float = 0.5, target_samples: int | None = None,
               log=print):
    """The learned-order ensemble measured against the fixed-order solver.

    1. Training data: n_train members (their own seeds) at the fixed order M (FP64), every
       right-hand side's band energies recorded, to t_end.
    2. The network is trained on the GPU (Adam, BF16 matrix cores).
    3. n_members evaluation members to t_end with outputs every d_out: fixed order FP64 (the
       reference) and mixed precision; then, for each eta tolerance, the learned order (mixed
       precision) with the network and with the geometric extrapolation instead of it.
    A tolerance's drift-rate budget is eta_tol / t_end.  Returns a dict of rates (attempts/s,
    all members), eta errors against the FP64 reference at every output time, and counters."""
    from .tank import Tank
    outs = [d_out * k for k in range(1, int(round(t_end / d_out)) + 1)]
    # 1. training data from the plain solver
    train = ensemble_members(n_train, seed=99, n1=n1, n2=n2, M=M, precision=0)
    tanks = [Tank(s) for s in train]
    t0 = time.perf_counter()
    try:
        for tk in tanks:
            tk.initialize()
        recs, _ = collect(tanks, t_end, [min(0.95 * s.dt_stab, s.delt_out) for s in train])
    finally:
        for tk in tanks:
            tk.close()
    x, t = dataset(recs)
    t_collect = time.perf_counter() - t0
    log(f"[learned] {len(recs)} records -> {len(x)} samples in {t_collect:.1f} s")
    if target_samples and len(x) > target_samples:
        sel = np.random.default_rng(0).choice(len(x), target_samples, replace=False)
        x, t = x[sel], t[sel]
    # held-out tenth: the next order's rate, network against geometric extrapolation
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(x))
    n_hold = max(1, len(x) // 10)
    hold, fit = perm[:n_hold], perm[n_hold:]
    # 2. training on the GPU
    net = OrderNet(max_rows=256, seed=0)
    net.set_params(OrderNet.initial_params(0, out_bias=t[fit].mean(axis=0)))
    losses, ms = net.train(x[fit], t[fit], epochs=epochs, batch=256, lr=1e-3, seed=1)
    tflops = OrderNet.train_flops(len(fit), epochs) / (ms * 1e-3) / 1e12
    pred = np.concatenate([net.forward(x[hold[i:i + 256]]) for i in range(0, n_hold, 256)])
    true_rate = np.sqrt(unscale(t[hold]).sum(axis=1))
    net_rate = np.sqrt(unscale(pred).sum(axis=1))
    geo = []
    for r in hold:
        mp = int(np.argmax(x[r, 448:456]))
        rm = np.sqrt(unscale(x[r, 64 * mp:64 * (mp + 1)]).sum())
        rm1 = np.sqrt(unscale(x[r, 64 * (mp - 1):64 * mp]).sum()) if mp >= 2 else 0.0
        geo.append(rm * rm / rm1 if rm1 > 0 else np.inf)
    geo = np.array(geo)
    lerr = lambda a: float(np.median(np.abs(np.log10(np.maximum(a, 1e-30) / np.maximum(true_rate, 1e-30)))))
    log(f"[learned] trained {epochs} epochs on {len(fit)} samples: loss {losses[0]:.3e} -> {losses[-1]:.3e}, "
        f"{ms:.0f} ms GPU, {tflops:.2f} TFLOP/s; held-out median |log10 rate error| network {lerr(net_rate):.3f}, "
        f"geometric {lerr(geo):.3f}")
    # 3. evaluation
    evals = ensemble_members(n_members, seed=2025, n1=n1, n2=n2, M=M, precision=0)
    evals32 = ensemble_members(n_members, seed=2025, n1=n1, n2=n2, M=M, precision=1)
    eta64, att64, w64 = _run_fixed(evals, outs, log)
    log(f"[learned] fixed M={M} FP64: {att64} attempts in {w64:.2f} s")
    eta32, att32, w32 = _run_fixed(evals32, outs, log)
    log(f"[learned] fixed M={M} mixed: {att32} attempts in {w32:.2f} s")

    def errs(etas):
        return [max(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(e, r))
                for e, r in zip(etas, eta64)]

    def orders(stats):
        tot = np.sum([s["attempts"] for s in stats], axis=0)
        return {int(m): int(tot[m]) for m in range(9) if tot[m]}

    res = dict(
        members=n_members, grid=[n1, n2], M=M, t_end=t_end, outputs=outs,
        samples=int(len(fit)), held_out=int(n_hold), records=int(len(recs)), epochs=epochs,
        train_ms=round(ms, 1), train_tflops=round(tflops, 3), loss_first=float(losses[0]),
        loss_last=float(losses[-1]), rate_log10_err_network=lerr(net_rate), rate_log10_err_geometric=lerr(geo),
        fixed64=dict(attempts=att64, seconds=round(w64, 3), rate=round(att64 / w64, 2)),
        fixed32=dict(attempts=att32, seconds=round(w32, 3), rate=round(att32 / w32, 2), eta_err=errs(eta32)),
        learned={}, judge_every=judge_every, safety=safety)
    for tol in eta_tols:
        pol = policy(m_min=3, judge_every=judge_every, rate_tol=tol / t_end, safety=safety)
        out = {}
        for name, use in (("network", net), ("geometric", None)):
            etaL, attL, wL, stL = _run_learned(evals32, outs, use, pol, log)
            out[name] = dict(attempts=attL, seconds=round(wL, 3), rate=round(attL / wL, 2), eta_err=errs(etaL),
                             orders=orders(stL), judges=int(sum(s["judges"] for s in stL)),
                             reverts=int(sum(s["reverts"] for s in stL)),
                             gain_vs_fixed64=round(w64 / wL, 3), gain_vs_fixed32=round(w32 / wL, 3))
            log(f"[learned] eta tol {tol:g}, {name}: {attL} attempts in {wL:.2f} s, orders {out[name]['orders']}, "
                f"reverts {out[name]['reverts']}, max eta err {max(out[name]['eta_err']):.2e}")
        out["network_vs_geometric"] = round(out["geometric"]["seconds"] / out["network"]["seconds"], 3)
        res["learned"][f"{tol:g}"] = out
    net.close()
    return res
