"""The learned-order ensemble against the fixed-order solver: 16 directional seas of
512 x 128 modes, M_max = 7 (JONSWAP Hs ~ U[0.04, 0.09], Tp ~ U[1.2, 1.8], s = 10).

The network is trained on the GPU from band energies of four other realisations run by the
plain solver at M = 7; the evaluation members run at fixed M = 7 in FP64 (the reference) and
in mixed precision, and at the learned order in mixed precision (with the network, and with
the geometric extrapolation in its place).  Every output time's eta stays within 1e-5 of the
FP64 fixed-order run; the step-rate gains are measured and printed (at this accuracy the
judges keep these seas close to M = 7, so the gain over the fixed mixed run is below one)."""
import json

import pytest

from hosnwt import learned as hl


@pytest.mark.gpu
def test_learned_ensemble_16_members():
    r = hl.experiment(n_members=16, n_train=4, t_end=8.0, d_out=2.0, eta_tols=(1e-5,), epochs=15,
                      log=lambda m: print(m, flush=True))
    print(json.dumps(r), flush=True)
    lr = r["learned"]["1e-05"]
    assert max(lr["network"]["eta_err"]) <= 1e-5
    assert max(lr["geometric"]["eta_err"]) <= 1e-5
    assert max(r["fixed32"]["eta_err"]) <= 1e-5
    # the trained network predicts the next order's band-energy rate better than the
    # geometric extrapolation on held-out samples
    assert r["rate_log10_err_network"] < r["rate_log10_err_geometric"]
    assert r["loss_last"] < 0.2 * r["loss_first"]
    assert lr["network"]["judges"] > 0
