"""icase 4 / 41 spectra (initialize_wav_cfg, initial_condition.f90:212-308) on the host:
the discretisation, the Hs normalisation and the phase generator, pinned by closed forms and
the generator's published check value; the GPU driver test runs the same components."""
import math

import numpy as np
import pytest

from hosnwt import config as hc


def test_park_miller_published_check_value():
    # Park & Miller (CACM 31, 1988): from x_0 = 1, x_10000 = 1043618065
    u = hc.park_miller(1, 10000)
    assert round(u[-1] * 2147483646) + 1 == 1043618065
    assert np.all((u >= 0) & (u < 1))


@pytest.mark.parametrize("icase", [4, 41])
@pytest.mark.parametrize("T_stop", [2.0, 50.0, 160.0])
def test_spectrum_grid_and_hs(icase, T_stop):
    Hs, Tp, gamma = 0.08, 1.6, 3.3
    c = hc.spectrum_components(icase, Hs, Tp, gamma, 513, T_stop)
    rnum = c["rnum"]
    assert 2 ** (rnum - 1) < T_stop * 32.0 <= 2 ** rnum  # ceil(log2(T_stop clock))
    n = len(c["freq"])
    assert n == 2 * 2 ** rnum // 32 + 1
    f_base = 32.0 / 2 ** rnum
    np.testing.assert_allclose(c["freq"], (np.arange(n) + 1) * f_base, rtol=0, atol=0)
    assert list(c["harmo"]) == list(range(n))  # harmonic i - 1 at frequency i f_base, as the reference
    # Hs of the discrete spectrum: 4 sqrt(sum a^2 / 2) = Hs
    assert abs(4.0 * math.sqrt(np.sum(c["ampli"] ** 2) / 2.0) - Hs) < 1e-14
    assert np.all((c["phase"] >= 0) & (c["phase"] < 2 * math.pi))


def test_jonswap_peak_and_shape():
    Hs, Tp, gamma = 0.08, 1.6, 3.3
    c = hc.spectrum_components(4, Hs, Tp, gamma, 7, 160.0)
    S = c["ampli"] ** 2 / 2.0
    f = c["freq"]
    # the discrete peak sits at the grid frequency nearest the JONSWAP peak 1/Tp
    assert abs(f[np.argmax(S)] - 1.0 / Tp) <= f[1] - f[0]
    # gamma = 1 reduces JONSWAP to the Pierson-Moskowitz shape f^-5 exp(-5/4 (f Tp)^-4)
    c1 = hc.spectrum_components(4, Hs, Tp, 1.0, 7, 160.0)
    pm = f ** -5.0 * np.exp(-1.25 * (f * Tp) ** -4.0)
    live = pm > 1e-200  # below, exp underflows to 0 in both
    ratio = (c1["ampli"][live] ** 2) / pm[live]
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-12)
    # the peak enhancement gamma^exp(-(f - fp)^2 / (2 (sigma / Tp)^2)) with the reference's
    # single-precision sigma literals, relative to the gamma = 1 spectrum
    fl = f[live]
    sig = np.where(fl < 1.0 / Tp, float(np.float32(0.07)), float(np.float32(0.09)))
    enh = gamma ** np.exp(-((fl - 1.0 / Tp) ** 2) / (2.0 * (sig / Tp) ** 2))
    r = (c["ampli"][live] ** 2) / (c1["ampli"][live] ** 2)
    np.testing.assert_allclose(r / r[0], enh / enh[0], rtol=1e-12)


def test_bretschneider_is_pierson_moskowitz_shape():
    c = hc.spectrum_components(41, 0.05, 1.2, 3.3, 3, 50.0)
    f = c["freq"]
    pm = f ** -5.0 * np.exp(-1.25 * (f * 1.2) ** -4.0)
    live = pm > 1e-200
    ratio = (c["ampli"][live] ** 2) / pm[live]
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-12)


def test_build_accepts_icase_4_and_41():
    for icase in (4, 41):
        cfg = hc.TankConfig(n1=64, n2=1, mHOS=3, xlen=20.0, icase=icase, igeom=2, d_hinge=0.3, Tramp=1.0,
                            T_stop=2.0, Hs=0.05, Tp=1.0, gamma=3.3, iseed=513)
        s = hc.build(cfg)
        assert s.pos_ini.shape[1] == len(hc.spectrum_components(icase, 0.05, 1.0, 3.3, 513, 2.0)["freq"])
        assert np.all(np.isfinite(s.pos_ini))
