"""Shared builders for the parity tests: one Setup feeds both the GPU context and the oracle."""
import numpy as np

from hosnwt import config as hc
from oracle import hos_oracle as ho


def make_setup(n1=32, n2=1, M=3, n3=9, xlen=20.0, ylen=10.0, icase=2, igeom=1, T=1.5, kA=0.08,
               Tramp=4.5, iramp=1, d_hinge=0.0, xabsf=0.8, precision=0, components=None, T_stop=45.0,
               toler=1e-7, f_out=None, coeffilt=(1.0, 1.0, 1.0), i_wmk=1):
    """Nondimensional setup; default is the 2-D regular-wave tank (monochromatic piston)."""
    nu_mono = 1.0 / T
    k = hc.wave_number(nu_mono / np.sqrt(9.81))  # h = 1, so f* = f sqrt(h/g)
    amp = kA / k
    cfg = hc.TankConfig(n1=n1, n2=n2, mHOS=M, xlen=xlen, ylen=ylen, n3=n3, icase=icase, amp_mono=amp,
                        nu_mono=nu_mono, igeom=igeom, d_hinge=d_hinge, iramp=iramp, Tramp=Tramp, xabsf=xabsf,
                        T_stop=T_stop, toler=toler, f_out=(f_out if f_out else 1.0 / T), precision=precision,
                        components=components, coeffilt=tuple(coeffilt), i_wmk=i_wmk)
    return hc.build(cfg)


def oracle_params(s):
    c = s.cfg
    return ho.Params(n1=c.n1, n2=c.n2, mHOS=c.mHOS, xlen=s.xlen, ylen=s.ylen, n3=c.n3, icase=c.icase,
                     igeom=c.igeom, d_hinge=c.d_hinge / c.h, iramp=c.iramp, Tramp=s.Tramp, T_stop=s.T_stop,
                     T_mono=s.T_mono, iabsnb=c.iabsnb, xabsf=c.xabsf, coeffabsf=c.coeffabsf, freq=s.freq,
                     pos_ini=s.pos_ini, coeffilt=tuple(c.coeffilt), i_wmk=c.i_wmk,
                     dposdy_ini=s.dposdy_ini, **dipole_params(s))


def dipole_params(s):
    """The oracle's own dipoles_ini (resol_wmkr.f90:1404-1455) from the setup's wave components."""
    if s.cfg.igeom != 3 or s.waves is None:
        return {}
    w = s.waves
    mu = ho.dipoles_ini(w.freq, w.k, w.ampli, w.phase, s.cfg.icase, s.x_dip)
    return dict(dip_mu=mu, dip_freq=np.array(w.freq, float), x_dip=s.x_dip, n_side=s.n_side)


def random_state(s, seed, amp=0.02):
    """Smooth random modal state (decaying spectrum) for RHS parity."""
    c = s.cfg
    n1, n2, n3a = c.n1, c.n2, 2 * (c.n3 - 1)
    rng = np.random.default_rng(seed)
    kx = np.arange(n1) * np.pi / s.xlen
    ky = np.arange(n2) * np.pi / s.ylen if n2 > 1 else np.zeros(1)
    k = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    w = np.sqrt(k * np.tanh(k))
    dec = np.exp(-0.35 * k)
    G1 = amp * rng.standard_normal((n1, n2)) * dec
    G2 = amp * rng.standard_normal((n1, n2)) * dec * w
    a1 = 0.3 * amp * rng.standard_normal((n1, n2)) * dec
    e1 = 0.3 * amp * rng.standard_normal((n1, n2)) * dec
    aad = 0.1 * amp * rng.standard_normal((n3a, n2)) * np.exp(-0.3 * np.arange(n3a))[:, None]
    return G1, G2, a1, e1, aad


def rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
