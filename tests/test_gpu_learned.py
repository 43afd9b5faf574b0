"""GPU tests of the learned truncation order: the BF16-MFMA network (forward, gradients and
training against the FP64 CPU network of oracle/learned_oracle.py), the band-energy records
(against the oracle's per-order modal potentials), order switching (an order-m attempt of an
mHOS context equals a context created with mHOS = m) and the judge / reversion logic of
hos_advance_learned."""
import numpy as np
import pytest

from helpers import make_setup, oracle_params, random_state, rel
from hosnwt import learned as hl
from hosnwt.tank import Tank
from oracle import hos_oracle as ho
from oracle import learned_oracle as lo

pytestmark = pytest.mark.gpu


def _layer_slices():
    out, o = [], 0
    for l in range(4):
        fin, fout = hl.DIMS[l], hl.DIMS[l + 1]
        out.append((slice(o, o + fin * fout), slice(o + fin * fout, o + fin * fout + fout)))
        o += fin * fout + fout
    return out


def test_mlp_forward_matches_cpu():
    rng = np.random.default_rng(0)
    net = hl.OrderNet(max_rows=256, seed=0)
    try:
        p = net.params()
        for rows in (64, 100, 256):
            x = rng.standard_normal((rows, 512)).astype(np.float32)
            y = net.forward(x)
            ym, _ = lo.mlp_forward(p, x, mimic_bf16=True)
            yp, _ = lo.mlp_forward(p, x)
            print(f"rows {rows}: rel vs BF16-rounded FP64 {rel(y, ym):.2e}, vs FP64 {rel(y, yp):.2e}", flush=True)
            assert rel(y, ym) < 1e-3 and rel(y, yp) < 5e-2
    finally:
        net.close()


def test_mlp_gradients_match_cpu():
    rng = np.random.default_rng(1)
    net = hl.OrderNet(max_rows=256, seed=2)
    try:
        p = net.params()
        x = rng.standard_normal((150, 512)).astype(np.float32)
        t = rng.standard_normal((150, 64)).astype(np.float32)
        loss, g = net.loss_grad(x, t)
        lm, gm = lo.mlp_loss_grad(p, x, t, mimic_bf16=True)
        lp, gp = lo.mlp_loss_grad(p, x, t)
        print(f"loss GPU {loss:.6f} BF16-rounded FP64 {lm:.6f} FP64 {lp:.6f}", flush=True)
        assert abs(loss - lm) <= 1e-4 * lm
        for l, (ws, bs) in enumerate(_layer_slices()):
            for name, sl in (("W", ws), ("b", bs)):
                em, ep = rel(g[sl], gm[sl]), rel(g[sl], gp[sl])
                print(f"layer {l} d{name}: rel vs BF16-rounded FP64 {em:.2e}, vs FP64 {ep:.2e}", flush=True)
                # (BF16 operands: the first layers differ from pure FP64 by several per cent)
                assert em < 2e-3 and ep < 0.15
    finally:
        net.close()


def test_mlp_training_reduces_loss():
    """Adam on the GPU fits a fixed random teacher network (FP64 oracle outputs)."""
    rng = np.random.default_rng(3)
    teacher = hl.OrderNet.initial_params(11)
    x = rng.standard_normal((4096, 512)).astype(np.float32)
    t, _ = lo.mlp_forward(teacher, x)
    t = (t / np.abs(t).max()).astype(np.float32)
    net = hl.OrderNet(max_rows=256, seed=4)
    try:
        losses, ms = net.train(x, t, epochs=30, batch=256, lr=1e-3, seed=5)
        flops = hl.OrderNet.train_flops(len(x), 30)
        print(f"epoch losses {losses[0]:.4e} -> {losses[-1]:.4e}; {ms:.1f} ms, "
              f"{flops / (ms * 1e-3) / 1e12:.2f} TFLOP/s on the matrix cores", flush=True)
        assert np.all(np.isfinite(losses))
        assert losses[-1] < 0.2 * losses[0]
        # the trained parameters still match the CPU network's forward pass
        y = net.forward(x[:128])
        ym, _ = lo.mlp_forward(net.params(), x[:128], mimic_bf16=True)
        assert rel(y, ym) < 1e-3
    finally:
        net.close()


@pytest.mark.parametrize("n2", [1, 16])
def test_feature_record_matches_oracle(n2):
    s = make_setup(n1=64, n2=n2, M=5, n3=9, ylen=5.0)
    orc = ho.Oracle(oracle_params(s))
    tank = Tank(s, precision=0)
    try:
        G1, G2, a1, e1, aad = random_state(s, 7, amp=0.03)
        orc.time = 0.5
        orc.solve_hos2(a1, e1, G2, G1, aad, 0.02)
        tank.record_features(2)
        tank.solve_hos2(0.5, 0.02, G1, G2, a1, e1, aad)
        recs = tank.take_features()
        tank.record_features(0)
        assert recs.shape == (1, hl.FEAT_REC)
        r = recs[0]
        a_modes = [None, orc.last_phis_modes] + [orc.last_aHOS[j] for j in range(2, 6)]
        ref = lo.band_record(64, n2, s.xlen, s.ylen, orc.last_eta_modes, a_modes, 5)
        assert r[hl.FEAT_M] == 5 and abs(r[hl.FEAT_ETA] - ref[hl.FEAT_ETA]) <= 1e-12 * ref[hl.FEAT_ETA]
        for j in range(0, 6):
            e = rel(r[64 * j:64 * (j + 1)], ref[64 * j:64 * (j + 1)])
            print(f"n2={n2} block {j}: energy {ref[64 * j:64 * (j + 1)].sum():.3e}, rel {e:.2e}", flush=True)
            assert e < 1e-9
        assert not r[6 * 64:9 * 64].any()
    finally:
        tank.close()


def test_order_switch_equals_fixed_order_context():
    s5 = make_setup(n1=64, n2=16, M=5, n3=9, ylen=5.0)
    s3 = make_setup(n1=64, n2=16, M=3, n3=9, ylen=5.0)
    args = random_state(s5, 9, amp=0.03)
    G1, G2, a1, e1, aad = args
    tk5, tk3, tk5b = Tank(s5), Tank(s3), Tank(s5)
    try:
        tk5.set_order(3)
        assert tk5.order() == 3
        r53 = tk5.solve_hos2(0.4, 0.01, G1, G2, a1, e1, aad)
        r3 = tk3.solve_hos2(0.4, 0.01, G1, G2, a1, e1, aad)
        for k in r3:
            assert np.array_equal(r53[k], r3[k]), k
        tk5.set_order(5)
        r55 = tk5.solve_hos2(0.4, 0.01, G1, G2, a1, e1, aad)
        # (the volume fix reads the previous evaluation's wavemaker globals, resol_HOS.f90:353-366:
        # the fresh context evaluates once before the compared evaluation, as tk5 did)
        tk5b.solve_hos2(0.4, 0.01, G1, G2, a1, e1, aad)
        r5 = tk5b.solve_hos2(0.4, 0.01, G1, G2, a1, e1, aad)
        for k in r5:
            assert np.array_equal(r55[k], r5[k]), k
        assert rel(r5["dG1"], r3["dG1"]) > 0  # the orders differ
        with pytest.raises(RuntimeError):
            tk5.set_order(2)  # the wavemaker coupling needs order >= 3
        with pytest.raises(RuntimeError):
            tk5.set_order(6)
    finally:
        for t in (tk5, tk3, tk5b):
            t.close()


def _tank_2d(M=6, precision=0):
    s = make_setup(n1=128, n2=1, M=M, n3=9, xlen=20.0, Tramp=3.0, T=1.5, kA=0.1, precision=precision)
    tk = Tank(s)
    tk.initialize()
    return s, tk


def test_learned_reversion_reaches_the_fixed_order_run():
    """rate_tol tiny: the selection picks the lowest order its floor allows (safety huge), every
    judge fails and every window below mHOS is rolled back to the checkpoint, which only
    full-order progress moves: the run ends bit-identical to the fixed-order run, with the same
    step proposal.  The first failures raise the floor 3 -> 4 -> 5 -> 6 one order at a time."""
    s, ref = _tank_2d()
    _, tk = _tank_2d()
    try:
        h0 = min(0.95 * s.dt_stab, s.delt_out)
        t1 = 3.0
        h_ref, na_ref, nr_ref = ref.advance(t1, h0)
        pol = hl.policy(m_min=3, judge_every=5, rate_tol=1e-30, safety=1e30)
        h, na, nr, st = hl.advance_learned([tk], t1, [h0], None, pol)
        print(st[0], flush=True)
        assert tk.time == ref.time
        assert st[0]["reverts"] == st[0]["judges"] >= 3
        assert st[0]["attempts"][3] > 0 and st[0]["attempts"][4] > 0 and st[0]["attempts"][5] > 0
        a, b = tk.get_state(), ref.get_state()
        for k in a:
            assert np.array_equal(a[k], b[k]), k
        assert h[0] == h_ref
    finally:
        tk.close()
        ref.close()


def test_learned_selection_and_prediction_path():
    """A tolerant policy: judges pass (no reversion), attempts run mostly below mHOS, and the
    prediction the driver made on the GPU from its last record equals the network's forward
    pass on the host-built rows of that record."""
    s, tk = _tank_2d(precision=1)
    net = hl.OrderNet(max_rows=128, seed=0)
    try:
        h0 = min(0.95 * s.dt_stab, s.delt_out)
        pol = hl.policy(m_min=3, judge_every=4, rate_tol=1e30, safety=1.0)
        h, na, nr, st = hl.advance_learned([tk], 2.0, [h0], net, pol)
        print(st[0], flush=True)
        assert abs(tk.time - 2.0) < 1e-12
        assert st[0]["reverts"] == 0 and st[0]["judges"] > 0
        assert sum(st[0]["attempts"][3:6]) > st[0]["attempts"][6] - st[0]["judges"]
        rec, pred = tk.learned_last()
        x, _, ms = hl.feature_rows(rec, with_targets=False)
        y = net.forward(x)
        assert np.max(np.abs(y - pred[:len(ms)])) <= 1e-5 * max(1.0, np.abs(y).max())
    finally:
        net.close()
        tk.close()
