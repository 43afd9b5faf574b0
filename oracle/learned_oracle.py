"""FP64 CPU oracle of the learned-order components.

TEST INFRASTRUCTURE ONLY: only the test suite imports it, as the checker of the
GPU band-energy records and of the GPU network's forward pass and gradients.
The product path (``hos-nwt_amd``) never imports it.

* ``band_record``: the feature record of one right-hand side, restated from its
  definition: radial bands of |k| (64 bands up to |k(n1-1, n2-1)|), sums of
  eta_hat^2 and of (k tanh k a_j)^2 with a_j the modal potential of order j
  (the oracle's aHOS of HOSphis_modes, resol_HOS.f90:720-782; order 1 is
  phis_hat).  Pinned by tests/test_oracle_learned.py: a single mode lands in
  its band with its closed-form energy, and the record sums to the Parseval
  totals.
* ``mlp_forward`` / ``mlp_loss_grad``: the 512-256-256-256-64 ReLU network and
  its mean-squared-error gradient by back-propagation in FP64, optionally
  rounding operands and activations to BF16 (round to nearest even) where the
  GPU kernels store BF16.  Pinned by central finite differences of the loss.
"""
from __future__ import annotations

import math

import numpy as np

BANDS = 64
ORDERS = 8
REC = (ORDERS + 1) * BANDS + 2
DIMS = (512, 256, 256, 256, 64)


def band_record(n1: int, n2: int, xlen: float, ylen: float, eta_modes: np.ndarray, a_modes: list,
                order: int) -> np.ndarray:
    """a_modes[j] (j = 1..order): modal potentials (at least n1 x n2; extra modes ignored)."""
    pix = math.pi / xlen
    piy = math.pi / ylen if n2 > 1 else 0.0
    kx = np.arange(n1)[:, None] * pix
    ky = np.arange(n2)[None, :] * piy
    k = np.sqrt(kx * kx + ky * ky)
    kmax = math.sqrt(((n1 - 1) * pix) ** 2 + ((n2 - 1) * piy) ** 2)
    band = np.minimum((BANDS * k / kmax).astype(np.int64), BANDS - 1) if kmax > 0 else np.zeros_like(k, int)
    rec = np.zeros(REC)
    e = np.asarray(eta_modes)[:n1, :n2] ** 2
    rec[0:BANDS] = np.bincount(band.ravel(), weights=e.ravel(), minlength=BANDS)
    kth = k * np.tanh(k)
    for j in range(1, min(order, ORDERS) + 1):
        w = (kth * np.asarray(a_modes[j])[:n1, :n2]) ** 2
        rec[j * BANDS:(j + 1) * BANDS] = np.bincount(band.ravel(), weights=w.ravel(), minlength=BANDS)
    rec[ORDERS * BANDS + BANDS] = order
    rec[ORDERS * BANDS + BANDS + 1] = e.sum()
    return rec


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (f >> 16) & 1
    r = ((f + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray):
    p = np.asarray(params, dtype=np.float64)
    out, o = [], 0
    for l in range(4):
        fin, fout = DIMS[l], DIMS[l + 1]
        W = p[o:o + fin * fout].reshape(fout, fin)
        o += fin * fout
        b = p[o:o + fout]
        o += fout
        out.append((W, b))
    return out


def mlp_forward(params, x, mimic_bf16: bool = False):
    layers = unpack(params)
    h = np.asarray(x, dtype=np.float64)
    if mimic_bf16:
        h = bf16(h)
    acts = [h]
    for l, (W, b) in enumerate(layers):
        Wl = bf16(W) if mimic_bf16 else W
        z = h @ Wl.T + b
        if l < 3:
            h = np.maximum(z, 0.0)
            if mimic_bf16:
                h = bf16(h)
            acts.append(h)
        else:
            h = z
    return h, acts


def mlp_loss_grad(params, x, t, mimic_bf16: bool = False):
    """Mean squared error over rows x 64 outputs and its gradient (parameter layout)."""
    layers = unpack(params)
    y, acts = mlp_forward(params, x, mimic_bf16)
    t = np.asarray(t, dtype=np.float64)
    rows = y.shape[0]
    d = y - t
    loss = float((d * d).sum() / (rows * DIMS[4]))
    dz = 2.0 * d / (rows * DIMS[4])
    if mimic_bf16:
        dz = bf16(dz)
    grads = [None] * 4
    for l in range(3, -1, -1):
        W, _ = layers[l]
        xin = acts[l]
        gW = dz.T @ xin
        gb = dz.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            Wl = bf16(W) if mimic_bf16 else W
            dh = (dz @ Wl) * (acts[l] > 0.0)
            if mimic_bf16:
                dh = bf16(dh)
            dz = dh
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
