"""CPU pins of the learned-order oracle (oracle/learned_oracle.py) and of the host-side row
builder (hosnwt/learned.py): closed-form band records, finite-difference gradients of the
FP64 network, BF16 rounding, and the row/target format."""
import math

import numpy as np

from hosnwt import learned as hl
from oracle import learned_oracle as lo


def test_band_record_single_mode_closed_form():
    n1, n2, xlen, ylen = 32, 8, 20.0, 5.0
    i1, i2 = 9, 3
    eta = np.zeros((n1, n2))
    eta[i1, i2] = 0.3
    a = [None, np.zeros((n1, n2)), np.zeros((40, 12))]
    a[1][i1, i2] = 2.0
    a[2][i1, i2] = -0.5
    a[2][35, 10] = 9.0  # beyond the n1 x n2 modes: ignored
    rec = lo.band_record(n1, n2, xlen, ylen, eta, a, 2)
    k = math.hypot(i1 * math.pi / xlen, i2 * math.pi / ylen)
    kmax = math.hypot((n1 - 1) * math.pi / xlen, (n2 - 1) * math.pi / ylen)
    b = int(64 * k / kmax)
    kth = k * math.tanh(k)
    assert rec[b] == 0.3 ** 2 and rec[:64].sum() == rec[b]
    assert math.isclose(rec[64 + b], (kth * 2.0) ** 2, rel_tol=1e-15)
    assert math.isclose(rec[128 + b], (kth * 0.5) ** 2, rel_tol=1e-15)
    assert rec[192:576].sum() == 0.0
    assert rec[576] == 2 and rec[577] == 0.3 ** 2


def test_band_record_sums_are_parseval_totals_and_last_band_closed():
    rng = np.random.default_rng(3)
    n1, n2, xlen, ylen = 24, 6, 10.0, 3.0
    eta = rng.standard_normal((n1, n2))
    a = [None] + [rng.standard_normal((n1, n2)) for _ in range(3)]
    rec = lo.band_record(n1, n2, xlen, ylen, eta, a, 3)
    kx = np.arange(n1)[:, None] * math.pi / xlen
    ky = np.arange(n2)[None, :] * math.pi / ylen
    k = np.sqrt(kx * kx + ky * ky)
    assert math.isclose(rec[:64].sum(), (eta ** 2).sum(), rel_tol=1e-13)
    for j in (1, 2, 3):
        assert math.isclose(rec[64 * j:64 * (j + 1)].sum(), ((k * np.tanh(k) * a[j]) ** 2).sum(), rel_tol=1e-13)
    # the corner mode |k| = kmax falls in the last band
    e = np.zeros((n1, n2))
    e[-1, -1] = 1.0
    assert lo.band_record(n1, n2, xlen, ylen, e, [None, e], 1)[63] == 1.0
    # 2-D tank: bands by kx alone
    e2 = np.zeros((n1, 1))
    e2[n1 // 2, 0] = 1.0
    r2 = lo.band_record(n1, 1, xlen, 1.0, e2, [None, e2], 1)
    assert r2[int(64 * (n1 // 2) / (n1 - 1))] == 1.0


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = hl.OrderNet.initial_params(1).astype(np.float64)
    x = rng.standard_normal((5, 512))
    t = rng.standard_normal((5, 64))
    loss, g = lo.mlp_loss_grad(p, x, t)
    idx = list(rng.choice(p.size, 40, replace=False))
    # include biases of every layer and the last layer's weights
    idx += [512 * 256, 512 * 256 + 256 + 5, p.size - 1, p.size - 64 - 3]
    eps = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = eps
        lp, _ = lo.mlp_loss_grad(p + e, x, t)
        lm, _ = lo.mlp_loss_grad(p - e, x, t)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_round_to_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 3 * 2 ** -8, -2.0 - 2 ** -7, 1e-30, 0.0])
    r = lo.bf16(v)
    assert list(r[:5]) == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.0]
    assert r[5] != 0.0 and abs(r[5] - 1e-30) / 1e-30 < 2 ** -8 and r[6] == 0.0


def test_feature_rows_format():
    rec = np.zeros(hl.FEAT_REC)
    rng = np.random.default_rng(1)
    rec[:64 * 5] = rng.uniform(0.0, 1.0, 64 * 5)
    rec[hl.FEAT_M] = 4
    rec[hl.FEAT_ETA] = rec[:64].sum()
    x, t, ms = hl.feature_rows(rec)
    assert ms == [1, 2, 3] and x.shape == (3, 512) and t.shape == (3, 64)
    inv = 1.0 / rec[hl.FEAT_ETA]
    for r, mp in enumerate(ms):
        assert np.allclose(x[r, :64], hl.scale(rec[:64] * inv).astype(np.float32))
        assert np.allclose(x[r, 64:64 * (mp + 1)], hl.scale(rec[64:64 * (mp + 1)] * inv).astype(np.float32))
        assert not x[r, 64 * (mp + 1):448].any()
        assert x[r, 448 + mp] == 1.0 and np.count_nonzero(x[r, 448:456]) == 1
        assert np.allclose(t[r], hl.scale(rec[64 * (mp + 1):64 * (mp + 2)] * inv))
    xi, ti, mi = hl.feature_rows(rec, with_targets=False)
    assert mi == [1, 2, 3, 4] and ti is None
    # the rate of a record's order and of a prediction of exactly that order agree
    assert math.isclose(hl.predicted_rate(t[2]), hl.rate_of(rec, 4), rel_tol=1e-5)


def test_direct_wavemaker_sums_in_c_match_the_oracle():
    """oracle/wmk_direct.c (the reference's literal loops of phiadd_SL) against the oracle's
    brute-force loop and separable form, on small 2-D and 3-D grids; phiadd_1st_SWEET's
    z = 0 fields against phiadd_SL at eta = 0 where they coincide."""
    from helpers import make_setup, oracle_params
    from oracle import hos_oracle as ho
    from oracle import wmk_direct as wd

    for n1, n2 in ((12, 1), (10, 6)):
        s = make_setup(n1=n1, n2=n2, M=3, n3=5, ylen=3.0)
        p = oracle_params(s)
        tb = ho.Tables(p)
        rng = np.random.default_rng(n1)
        a = rng.standard_normal((p.n3_add, n2)) * np.exp(-0.5 * np.arange(p.n3_add))[:, None]
        da = rng.standard_normal((p.n3_add, n2)) * np.exp(-0.5 * np.arange(p.n3_add))[:, None]
        eta = 0.05 * rng.standard_normal((n1, n2))
        ref = ho.phiadd_sl_bruteforce(a, da, eta, tb)
        got = wd.phiadd_sl(n1, n2, s.xlen, s.ylen, a, da, eta)
        sep = ho.phiadd_sl(a, da, eta, tb)
        for k in ref:
            scale = max(np.abs(ref[k]).max(), 1e-300)
            assert np.abs(got[k] - ref[k]).max() <= 1e-12 * scale, k
            assert np.abs(sep[k] - got[k]).max() <= 1e-11 * scale, k
        # rows computed separately equal the full computation
        part = wd.phiadd_sl(n1, n2, s.xlen, s.ylen, a, da, eta, rows=(2, 5))
        assert np.array_equal(part["phi"][2:5], got["phi"][2:5])
        z0 = wd.phiadd_1st(n1, n2, s.xlen, s.ylen, a, da)
        at0 = wd.phiadd_sl(n1, n2, s.xlen, s.ylen, a, da, np.zeros((n1, n2)))
        for f, k in enumerate(("phi", "phit", "phix", "phiy", "phiz")):
            assert np.allclose(z0[f], at0[k], rtol=1e-13, atol=1e-13 * np.abs(at0[k]).max()), k
