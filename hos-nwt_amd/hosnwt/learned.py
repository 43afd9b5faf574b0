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
