"""Host setup (dispersion, transfer functions, paddle amplitudes, beach) and the C-ABI surface.

No GPU: the library is loaded (it must export every symbol include/hosnwt.h
declares) but no compute entry point is called.
"""
import math
import os
import re

import numpy as np
import pytest

from hosnwt import config as hc
from hosnwt import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dispersion_relation():
    for f in (0.05, 0.2, 0.6, 1.3):
        k = hc.wave_number(f)
        assert abs(k * math.tanh(k) - (2 * math.pi * f) ** 2) < 1e-12 * max(1.0, k)
    assert hc.wave_number(0.0) == 0.0


def test_piston_transfer_function_limits():
    # deep water: b/a -> -i/2 ; shallow water: |b/a| ~ 1/k (Biesel)
    assert abs(hc.transfer_function("piston", 60.0) - (-0.5j)) < 1e-15
    tf = hc.transfer_function("piston", 30.0)
    assert abs(tf.imag + 0.5) < 1e-12 and tf.real == 0.0
    k = 1e-3
    assert abs(abs(hc.transfer_function("piston", k)) * k - 1.0) < 1e-5


def test_flap_transfer_function_matches_piston_for_zero_hinge_limit():
    # a flap hinged at the bottom (d = 0) generates less than a piston at the same stroke
    for k in (0.5, 2.0, 5.0):
        assert abs(hc.transfer_function("hinged", k, 0.0)) > abs(hc.transfer_function("piston", k))


def test_monochromatic_paddle_amplitude():
    cfg = hc.TankConfig(n1=64, n2=1, mHOS=3, xlen=20.0, icase=2, amp_mono=0.01, nu_mono=1 / 1.5, ph_mono=0.3,
                        Tramp=4.5)
    s = hc.build(cfg)
    k = hc.wave_number(s.freq[0])
    expect = hc.transfer_function("piston", k) * 0.01 * np.exp(0.3j)
    assert s.pos_ini.shape == (1, 1)
    assert abs(s.pos_ini[0, 0] - expect) < 1e-15
    assert abs(s.freq[0] - (1 / 1.5) * math.sqrt(1 / 9.81)) < 1e-15


def test_directional_components_and_hs():
    c = hc.jonswap_components(0.06, 1.4, 3.3, n_freq=64, seed=1, rnum=10, clock=32.0, n_dir=8)
    assert len(c["freq"]) == 64 * 8
    # total variance 2 * sum(a^2/2) matches Hs^2/16
    assert abs(4 * math.sqrt(0.5 * np.sum(c["ampli"] ** 2)) - 0.06) < 1e-12
    c2 = hc.jonswap_components(0.06, 1.4, 3.3, n_freq=64, seed=1, rnum=10, clock=32.0, n_dir=8)
    np.testing.assert_array_equal(c["phase"], c2["phase"])


def test_beach_and_stability_bound():
    cfg = hc.TankConfig(n1=101, n2=1, mHOS=3, xlen=20.0, xabsf=0.8, coeffabsf=1.0)
    nu = hc.absorb(cfg, 20.0)
    x = np.arange(101) * 0.2
    assert np.all(nu[x < 16.0 - 1e-12] == 0.0)
    assert abs(nu[-1, 0] - 1.0) < 1e-15
    assert abs(hc.dt_stab(101, 1, 33, 20.0, 1.0, 2) - min(math.sqrt(8) / math.pi * 0.2, math.sqrt(8) / math.pi * 4 / 64)) < 1e-15


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "hosnwt.h")).read()
    names = set(re.findall(r"^\s*(?:int|const char\*)\s+(hos_\w+)\s*\(", header, flags=re.M))
    assert len(names) >= 18
    L = lib.load()
    for n in names:
        assert hasattr(L, n), n
    assert names <= set(lib.EXPORTS) | {"hos_last_error"}


def test_ensemble_wrapper_argument_checks():
    """advance_ensemble: an empty ensemble is a no-op; proposals must match the tanks;
    the members share one tolerance (checked before the library is called)."""
    from types import SimpleNamespace

    from hosnwt.tank import advance_ensemble

    h, na, nr = advance_ensemble([], 1.0, [])
    assert len(h) == 0 and na == [] and nr == []

    def stub(toler):
        return SimpleNamespace(lib=None, h=SimpleNamespace(value=1), time=0.0,
                               s=SimpleNamespace(dt_stab=0.1, delt_out=1.0), cfg=SimpleNamespace(toler=toler))

    with pytest.raises(ValueError, match="proposal"):
        advance_ensemble([stub(1e-7), stub(1e-7)], 1.0, [0.01])
    with pytest.raises(ValueError, match="tolerance"):
        advance_ensemble([stub(1e-7), stub(1e-6)], 1.0, [0.01, 0.01])
