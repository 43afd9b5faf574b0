"""hos_advance_ensemble: independent realisations of different sizes and orders
advanced together, their attempts overlapping on the GPU.

Every member must take exactly the decisions and reach exactly the state that
hos_advance reaches for it alone (same kernels in the same order on its own
streams), here checked to relative L2 1e-14 on eta and phis with equal
accepted / rejected counts and step proposals; FP64 and mixed members mixed in
one ensemble, driven from one host thread and from several
(HOSNWT_ENSEMBLE_THREADS).  hos_initialize on a used context restarts from rest:
the same trajectory as a fresh context.
"""
import numpy as np
import pytest

from helpers import make_setup, rel
from hosnwt import config as hc
from hosnwt.tank import Tank, advance_ensemble

MEMBERS = [  # (n1, M, Hs, Tp, seed, precision): 2-D realisations as in a ragged ensemble
    (128, 3, 0.05, 1.4, 31, 0),
    (256, 4, 0.07, 1.8, 32, 1),
    (512, 5, 0.04, 1.3, 33, 0),
    (256, 6, 0.06, 1.6, 34, 1),
]


def setups():
    out = []
    for n1, M, Hs, Tp, seed, prec in MEMBERS:
        comps = hc.jonswap_components(Hs, Tp, 3.3, n_freq=16, seed=seed, rnum=8, clock=16.0)
        out.append((make_setup(n1=n1, n2=1, M=M, n3=9, xlen=n1 / 16.0, Tramp=2.0 * Tp, components=comps, icase=3,
                               f_out=1.0 / Tp, toler=1e-6), prec))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("threads", [1, 3])
def test_ensemble_matches_members_run_alone(threads, monkeypatch):
    monkeypatch.setenv("HOSNWT_ENSEMBLE_THREADS", str(threads))
    ss = setups()
    t_next = min(s.delt_out for s, _ in ss) * 0.6
    h0 = [min(0.95 * s.dt_stab, s.delt_out) for s, _ in ss]
    alone = []
    for (s, prec), h in zip(ss, h0):
        tk = Tank(s, precision=prec)
        tk.initialize()
        hn, na, nr = tk.advance(t_next, h)
        st = tk.get_state()
        alone.append((hn, na, nr, st["eta"], st["phis"], tk.time))
        tk.close()
    tanks = [Tank(s, precision=prec) for s, prec in ss]
    try:
        for tk in tanks:
            tk.initialize()
        hs, nas, nrs = advance_ensemble(tanks, t_next, h0)
        for i, tk in enumerate(tanks):
            st = tk.get_state()
            hn, na, nr, eta, phis, t_end = alone[i]
            print(f"member {i}: acc {nas[i]}/{na} rej {nrs[i]}/{nr} eta rel {rel(st['eta'], eta):.1e}", flush=True)
            assert (nas[i], nrs[i]) == (na, nr)
            assert hs[i] == hn and tk.time == t_end
            assert rel(st["eta"], eta) < 1e-14 and rel(st["phis"], phis) < 1e-14
            assert np.abs(eta).max() > 0
    finally:
        for tk in tanks:
            tk.close()


@pytest.mark.gpu
def test_ensemble_rejects_a_repeated_context():
    s, prec = setups()[0]
    tk = Tank(s, precision=prec)
    try:
        tk.initialize()
        with pytest.raises(RuntimeError, match="twice"):
            advance_ensemble([tk, tk], 0.1, [0.01, 0.01])
    finally:
        tk.close()


@pytest.mark.gpu
def test_reinitialize_restarts_from_rest():
    s, prec = setups()[1]
    tk = Tank(s, precision=prec)
    try:
        runs = []
        for _ in range(2):
            tk.initialize()
            h, na, nr = tk.advance(0.3 * s.delt_out, min(0.95 * s.dt_stab, s.delt_out))
            st = tk.get_state()
            runs.append((h, na, nr, st["eta"], st["phis"], st["a1st_add"]))
        print("attempts", runs[0][1:3], runs[1][1:3], flush=True)
        assert runs[0][:3] == runs[1][:3]
        for a, b in zip(runs[0][3:], runs[1][3:]):
            assert np.array_equal(a, b)
        assert np.abs(runs[0][3]).max() > 0
    finally:
        tk.close()
