"""Driver setup on the host: input parameters, wave components, paddle motion
amplitudes, numerical beach and the RK4 stability bound.

Restates the reference's setup (HOS-NWT.f90:95-255, initial_condition.f90:128-310,
wavemaking.f90:37-175, wave/wave_def.f90:54-329, wave/linear_wave.f90:119-334,
initial_condition.f90:724-794) so that the GPU context receives the same
pos_ini / freq / nu arrays the reference builds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

G = 9.81
TWOPI = 2.0 * math.pi
L_ADD = 2


def wave_number(f: float, prec: float = 1e-15) -> float:
    """Nondimensional k with (2 pi f)^2 = k tanh k (linear_wave.f90:137-151).

    Bisection on [0, 1e6] to relative accuracy ``prec`` like the reference's
    dichoto (linear_wave.f90:427-466); returns 0 for f = 0.
    """
    w2 = (TWOPI * f) ** 2
    if w2 == 0.0:
        return 0.0
    lo, hi = 0.0, 1.0e6
    x3 = hi
    while abs(lo - hi) > prec * abs(hi):
        x3 = 0.5 * (lo + hi)
        v = x3 * math.tanh(x3) - w2
        if v == 0.0:
            return x3
        if (0.0 - w2) * v < 0.0:
            hi = x3
        else:
            lo = x3
    return x3


def transfer_function(typ: str, k: float, hinge: float = 0.0) -> complex:
    """Paddle transfer function b = TF a (linear_wave.f90:277-334)."""
    if typ == "piston":
        if k < 50.0:
            return complex(0.0, -(2.0 * k + math.sinh(2.0 * k)) / (4.0 * math.sinh(k) * math.sinh(k)))
        return complex(0.0, -0.5)
    if typ == "hinged":
        d = hinge
        if k < 50 and k * d < 50:
            num = -1j * k * (1.0 - d) * (2.0 * k + math.sinh(2.0 * k)) / (4.0 * math.sinh(k))
            return num / (k * (1.0 - d) * math.sinh(k) + math.cosh(k * d) - math.cosh(k))
        if k >= 50 and k * d < 50:
            return -1j * k * (1.0 - d) / (2.0 * k * (1.0 - d) - 2.0 + 4.0 * math.cosh(k * d) * math.exp(-k))
        if k * (1.0 - d) >= 20:
            return -0.5j
        return -1j * k * (1.0 - d) / (2.0 * k * (1.0 - d) - 2.0 + 2.0 * math.exp(k * (d - 1.0)))
    if typ == "dipole":
        d = hinge
        # d is the dipole depth; IEEE arithmetic as the reference's: k^2 - w^4 is 0 in FP64
        # once tanh k rounds to 1 (18.7 < k <= 20) and the result is infinite there
        kk = np.float64(k)
        k2 = kk * kk
        w2 = kk * np.tanh(kk)
        with np.errstate(divide="ignore", invalid="ignore"):
            if kk <= 20.0:
                w4 = w2 * w2
                return complex((k2 - w4 + w2) / (2.0 * np.sqrt(w2) * (k2 - w4) * np.cosh(kk) * np.sinh(kk * (d + 1.0))))
            return complex(np.sqrt(w2) / (2.0 * np.exp(kk * d) * k2))
    return 1.0 + 0j


@dataclass
class WaveComponents:
    """Nondimensional wave components (the reference's TYPE(wave), wave_def.f90)."""

    harmo: np.ndarray
    freq: np.ndarray
    ampli: np.ndarray
    angle: np.ndarray  # radians
    phase: np.ndarray
    k: np.ndarray = field(default=None)
    TF: np.ndarray = field(default=None)
    rnum: int = 1
    clock: float = 1.0

    def init(self, typ: str, hinge: float):
        """wave_init (wave_def.f90:282-329).  For dipoles the transfer function is dipoles_ini's
        (each component's own depth), so TF stays 1 here."""
        n = len(self.freq)
        self.k = np.zeros(n)
        self.TF = np.ones(n, complex)
        for i in range(n):
            kk = wave_number(self.freq[i])
            self.k[i] = kk
            if typ != "dipole":
                self.TF[i] = transfer_function(typ, kk, hinge) if abs(kk) > np.finfo(float).eps else 1.0
            if self.freq[i] < 1e-8:
                self.k[i] = 0.0
                self.TF[i] = 1.0
        return self


@dataclass
class TankConfig:
    """Dimensional inputs in the order of input_HOS-NWT.dat (input.f90:62-133)."""

    n1: int
    n2: int
    mHOS: int
    xlen: float
    ylen: float = 1.0
    h: float = 1.0
    n3: int = 33
    icase: int = 2
    islosh: int = 1
    aslosh: float = 0.0
    amp_mono: float = 0.0
    nu_mono: float = 1.0
    theta_mono: float = 0.0
    ph_mono: float = 0.0
    ibat: int = 2
    i_wmk: int = 1
    igeom: int = 1
    d_hinge: float = 0.0
    iramp: int = 1
    Tramp: float = 0.0
    iabsnb: int = 1
    xabsf: float = 0.8
    coeffabsf: float = 1.0
    T_stop: float = 10.0
    toler: float = 1e-7
    f_out: float = 1.0
    precision: int = 0
    # spectral filters (common_vars.f90:43): x, y free-surface modes, z paddle modes
    coeffilt: tuple = (1.0, 1.0, 1.0)
    # icase = 3: components in SI (freq Hz, ampli m, angle deg, phase rad)
    components: dict | None = None
    # icase = 4 (JONSWAP) / 41 (Bretschneider): spectrum built in memory (input.f90:62-133 order)
    Hs: float = 0.05
    Tp: float = 1.0
    gamma: float = 3.3
    iseed: int = 1


@dataclass
class Setup:
    """Nondimensional run description shared by the GPU context and the oracle."""

    cfg: TankConfig
    xlen: float
    ylen: float
    Tscale: float
    Tramp: float
    T_stop: float
    T_mono: float
    f_out: float
    freq: np.ndarray
    pos_ini: np.ndarray  # (n2, n_freq) complex
    nu: np.ndarray  # (n1, n2)
    dt_stab: float
    delt_out: float
    waves: WaveComponents | None
    # y derivative of pos_ini (n2, n_freq) complex, read with i_wmk = 2
    dposdy_ini: np.ndarray | None = None
    # x derivative of the beach coefficient as the reference builds it (n1, n2)
    nu_x: np.ndarray | None = None
    # dipole wavemaker (igeom = 3): per component frequency, mu(:, 1..4), abscissa, image pairs
    dip_freq: np.ndarray | None = None
    dip_mu: np.ndarray | None = None
    x_dip: float = 2.5
    n_side: int = 6


def gzed(z: float, igeom: int, d: float) -> float:
    """Paddle shape (wavemaking.f90:184-215)."""
    delta = min(d, (1.0 - d) / 10.0)
    if igeom == 1:
        return 1.0
    if z >= d - 1.0 + delta:
        return (z + 1.0 - d) / (1.0 - d)
    if z <= d - 1.0 - delta:
        return 0.0
    u = (z - (d - 1.0)) / delta
    return delta / (2.0 * (1.0 - d)) * (5.0 / 16.0 + u + 15.0 / 16.0 * (u ** 2 - u ** 4 / 3.0 + u ** 6 / 15.0))


def build(cfg: TankConfig) -> Setup:
    """Nondimensionalise and build the wavemaker and beach inputs."""
    h = cfg.h
    Tscale = 1.0 / math.sqrt(G / h)
    xlen = cfg.xlen / h
    ylen = cfg.ylen / h if cfg.n2 != 1 else 1.0
    Tramp = (cfg.Tramp if cfg.iramp != 0 else 0.0) / Tscale
    T_stop = cfg.T_stop / Tscale
    f_out = cfg.f_out * Tscale
    nu_mono = cfg.nu_mono * Tscale
    T_mono = 1.0 / nu_mono
    typ = {1: "piston", 2: "hinged", 3: "dipole"}[cfg.igeom]
    d_hinge = cfg.d_hinge / h
    waves = None
    if cfg.icase == 2:
        # front(): harmo 1 on clock 2 nu, rnum 1 (initial_condition.f90:148-166)
        theta = 0.0 if cfg.n2 == 1 else cfg.theta_mono * math.pi / 180.0
        waves = WaveComponents(harmo=np.array([1]), freq=np.array([nu_mono]),
                               ampli=np.array([cfg.amp_mono / h]), angle=np.array([theta]),
                               phase=np.array([cfg.ph_mono]), rnum=1, clock=2.0 * nu_mono).init(typ, d_hinge)
    elif cfg.icase in (3, 4, 41):
        c = cfg.components if cfg.icase == 3 else spectrum_components(
            cfg.icase, cfg.Hs, cfg.Tp, cfg.gamma, cfg.iseed, cfg.T_stop)
        waves = WaveComponents(harmo=np.asarray(c["harmo"], int), freq=np.asarray(c["freq"], float) * Tscale,
                               ampli=np.asarray(c["ampli"], float) / h,
                               angle=np.asarray(c["angle"], float) * math.pi / 180.0,
                               phase=np.asarray(c["phase"], float), rnum=int(c["rnum"]),
                               clock=float(c["clock"]) * Tscale).init(typ, d_hinge)
    n1, n2 = cfg.n1, cfg.n2
    dely = ylen / (n2 - 1) if n2 != 1 else 0.0
    cy = np.arange(n2) * dely
    dip_freq = dip_mu = None
    x_dip = X_DIP / h
    if waves is not None and cfg.igeom == 3:
        dip_freq = np.array(waves.freq, float)
        dip_mu = dipoles_ini(waves, cfg.icase, x_dip)
        freq, pos_ini = np.zeros(0), np.zeros((n2, 0), complex)
        dposdy_ini = np.zeros((n2, 0), complex)
    elif waves is not None:
        hmax = None
        if cfg.icase in (3, 4, 41):
            # 4 n_max, n_max = 2^rnum / floor(clock sqrt(g / h)) with the nondimensional clock
            hmax = 4 * (2 ** int(waves.rnum) // max(1, math.floor(waves.clock * math.sqrt(G / h))))
        freq, pos_ini, dposdy_ini = wmk_ini(waves, cy, hmax)
    else:
        freq, pos_ini = np.zeros(0), np.zeros((n2, 0), complex)
        dposdy_ini = np.zeros((n2, 0), complex)
    nu = absorb(cfg, xlen)
    nu_x = absorb_dx(cfg, xlen)
    stab = dt_stab(n1, n2, cfg.n3, xlen, ylen, cfg.icase, cfg.igeom)
    return Setup(cfg=cfg, xlen=xlen, ylen=ylen, Tscale=Tscale, Tramp=Tramp, T_stop=T_stop, T_mono=T_mono,
                 f_out=f_out, freq=freq, pos_ini=pos_ini, nu=nu, dt_stab=stab, delt_out=1.0 / f_out,
                 waves=waves, dposdy_ini=dposdy_ini, nu_x=nu_x, dip_freq=dip_freq, dip_mu=dip_mu, x_dip=x_dip,
                 n_side=N_SIDE)


# dipole wavemaker constants (input.f90:137-139): depth of the monochromatic dipole
# (nondimensional), abscissa (m, divided by h at HOS-NWT.f90:111), wall-image pairs
Z_DIPMONO = -0.25
X_DIP = 2.5
N_SIDE = 6


def dipoles_ini(w: WaveComponents, icase: int, x_dip: float) -> np.ndarray:
    """mu(:, 1..4) of the dipole wavemaker, one row per wave component (resol_wmkr.f90:1404-1455):
    the depth from the component's nondimensional frequency (icase 2: Z_DIPMONO), the dipole
    transfer function at that depth, mu_z = a TF e^{i (k x_dip + phase)}, mu_x = -i mu_z
    tanh(k (z + 1)) and the bottom image's depth -2 - z.  The wavenumbers are the dispersion
    relation's (the dipole branch of wave_init computes them into its argument and returns its
    uninitialised local array instead, wave_def.f90:295-323)."""
    f = np.asarray(w.freq, float)
    z = np.where(f > 1.0, -0.35, np.where(f > 0.6, -0.2, np.where(f > 0.5, -0.4, -0.6)))
    if icase == 2:
        z = np.full_like(f, Z_DIPMONO)
    mu = np.zeros((len(f), 4), complex)
    for i in range(len(f)):
        tf = transfer_function("dipole", float(w.k[i]), float(z[i]))
        mu[i, 2] = z[i]
        mu[i, 3] = -2.0 - z[i]
        mu[i, 1] = w.ampli[i] * tf * np.exp(1j * (w.k[i] * x_dip + w.phase[i]))
        mu[i, 0] = -1j * mu[i, 1] * math.tanh(w.k[i] * (z[i] + 1.0))
    return mu


def wmk_ini(w: WaveComponents, cy: np.ndarray, hmax: int | None = None):
    """Per-frequency paddle amplitudes pos_ini(y, f) and their y derivatives dposdy_ini(y, f)
    (wavemaking.f90:104-155), full-width paddle (no lateral ramp).  Harmonics above ``hmax``
    (4 n_max, n_max = 2^rnum / floor(clock in s): the 4 Hz limit of wavemaking.f90:104-118)
    are left out."""
    harms = sorted(set(int(x) for x in w.harmo if hmax is None or 0 <= int(x) <= hmax))
    freq = []
    cols, dcols = [], []
    for hm in harms:
        sel = np.nonzero(w.harmo == hm)[0]
        freq.append(w.freq[sel[0]])
        col = np.zeros(len(cy), complex)
        dcol = np.zeros(len(cy), complex)
        for i in sel:
            ky = w.k[i] * math.sin(w.angle[i])
            term = w.TF[i] * math.cos(w.angle[i]) * w.ampli[i] * np.exp(1j * (-ky * cy + w.phase[i]))
            col += term
            dcol += -1j * ky * term
        cols.append(col)
        dcols.append(dcol)
    empty = np.zeros((len(cy), 0), complex)
    return (np.array(freq), (np.array(cols).T if cols else empty), (np.array(dcols).T if dcols else empty))


def absorb(cfg: TankConfig, xlen: float) -> np.ndarray:
    """Parabolic front beach nu(x) (initial_condition.f90:724-794; no back or side beach)."""
    nu = np.zeros((cfg.n1, cfg.n2))
    if cfg.iabsnb == 0:
        return nu
    x0 = xlen * cfg.xabsf
    if abs(cfg.xabsf - 1.0) > np.finfo(float).eps:
        x2 = xlen - x0
        cx = np.arange(cfg.n1) * xlen / (cfg.n1 - 1)
        sel = cx >= x0
        x3 = (cx[sel] - x0) / x2
        nu[sel, :] = (cfg.coeffabsf * x3 * x3)[:, None]
    return nu


def absorb_dx(cfg: TankConfig, xlen: float) -> np.ndarray:
    """x derivative of the front beach as initial_condition.f90:768 builds it: 2 c x3, without
    the 1/(xlen - x0) factor of the chain rule (read by the second-order wavemaker)."""
    nux = np.zeros((cfg.n1, cfg.n2))
    if cfg.iabsnb == 0:
        return nux
    x0 = xlen * cfg.xabsf
    if abs(cfg.xabsf - 1.0) > np.finfo(float).eps:
        x2 = xlen - x0
        cx = np.arange(cfg.n1) * xlen / (cfg.n1 - 1)
        sel = cx >= x0
        x3 = (cx[sel] - x0) / x2
        nux[sel, :] = (cfg.coeffabsf * 2.0 * x3)[:, None]
    return nux


def dt_stab(n1, n2, n3, xlen, ylen, icase, igeom=1) -> float:
    """RK4 stability bound (HOS-NWT.f90:132-176); dipoles have no paddle grid (:145-153)."""
    c = math.sqrt(8.0 / math.pi ** 2)
    d1 = c * xlen / (n1 - 1)
    d2 = c * ylen / (n2 - 1) if n2 != 1 else d1 + 1.0
    d3 = c * (2.0 * L_ADD) / (L_ADD * (n3 - 1))
    if icase == 1 or igeom == 3:
        return min(d1, d2)
    if n2 == 1:
        return min(d1, d3)
    return min(d1, d2, d3)


def park_miller(seed: int, n: int) -> np.ndarray:
    """n uniforms (x - 1) / (2^31 - 2) of the Park-Miller minimal-standard sequence
    x <- 16807 x mod (2^31 - 1) (Park & Miller, CACM 31, 1988), x_0 = seed (1 when
    seed mod (2^31 - 1) is 0)."""
    m = 2147483647
    x = seed % m or 1
    out = np.empty(n)
    for i in range(n):
        x = (16807 * x) % m
        out[i] = (x - 1) / (m - 1)
    return out


def spectrum_components(icase: int, Hs: float, Tp: float, gamma: float, iseed: int, T_stop: float,
                        clock: float = 32.0):
    """Long-crested irregular components of icase 4 (JONSWAP) / 41 (Bretschneider) as
    initialize_wav_cfg builds them (initial_condition.f90:212-308), in SI units for
    ``build``: rnum = ceil(log2(T_stop clock)) with T_stop in seconds, components
    i = 1 .. round(2 2^rnum / clock + 1) with harmonic i - 1 and frequency i clock / 2^rnum,
    S(f) df with the single-precision sigma literals 0.07 / 0.09, amplitudes sqrt(2 S df)
    rescaled to Hs.  Phases 2 pi u from ``park_miller(iseed)``: the reference draws them with
    gfortran's SRAND / RAND; the Fortran driver draws this same sequence."""
    if icase not in (4, 41):
        raise ValueError("icase 4 (JONSWAP) or 41 (Bretschneider)")
    rnum = math.ceil(math.log10(T_stop * clock) / math.log10(2.0))
    n_pts = int(np.rint(2 * 2 ** rnum / clock + 1))
    f_base = clock / 2.0 ** rnum
    i = np.arange(1, n_pts + 1)
    f = i * f_base
    sig = np.where(f < 1.0 / Tp, float(np.float32(0.07)), float(np.float32(0.09)))
    if icase == 4:
        S = (1.0 / TWOPI ** 4 * G ** 2 / f ** 5 * np.exp(-5.0 / 4.0 * (f * Tp) ** -4)
             * gamma ** np.exp(-((f - 1.0 / Tp) ** 2) / (2.0 * (sig / Tp) ** 2)) * f_base)
    else:
        S = 5.0 / 16.0 / (Tp ** 4 * f ** 5) * Hs * np.exp(-5.0 / 4.0 * (f * Tp) ** -4) * f_base
    Hs_calc = 4.0 * math.sqrt(S.sum())
    amp = np.sqrt(2.0 * S) * Hs / Hs_calc
    return dict(harmo=i - 1, freq=f, ampli=amp, angle=np.zeros(n_pts), phase=TWOPI * park_miller(iseed, n_pts),
                rnum=rnum, clock=clock)


def jonswap_components(Hs: float, Tp: float, gamma: float, n_freq: int, seed: int, rnum: int = 10,
                       clock: float = 32.0, n_dir: int = 1, spread_s: float = 10.0):
    """Seeded JONSWAP components on the clock/2^rnum harmonic grid (SI units).

    Spectrum as in initial_condition.f90:255-283 (sigma 0.07/0.09, Hs rescaled);
    n_freq harmonics starting at the one nearest 0.5/Tp, phases U[0, 2 pi) from
    numpy's PCG64 with ``seed``.  n_dir > 1 spreads each frequency over n_dir
    directions with cos^(2s) weights on [-pi/2, pi/2].
    """
    rng = np.random.default_rng(seed)
    f_base = clock / 2.0 ** rnum
    h0 = max(1, int(round(0.5 / Tp / f_base)))
    harmo = np.arange(h0, h0 + n_freq)
    f = harmo * f_base
    sigma = np.where(f < 1.0 / Tp, 0.07, 0.09)
    S = (1.0 / TWOPI ** 4 * G ** 2 / f ** 5 * np.exp(-1.25 * (f * Tp) ** -4)
         * gamma ** np.exp(-((f - 1.0 / Tp) ** 2) / (2.0 * (sigma / Tp) ** 2)) * f_base)
    Hs_calc = 4.0 * math.sqrt(S.sum())
    amp = np.sqrt(2.0 * S) * Hs / Hs_calc
    if n_dir == 1:
        return dict(harmo=harmo, freq=f, ampli=amp, angle=np.zeros(n_freq),
                    phase=rng.uniform(0.0, TWOPI, n_freq), rnum=rnum, clock=clock)
    th = (np.arange(n_dir) + 0.5) / n_dir * math.pi - math.pi / 2.0
    wgt = np.cos(th) ** (2 * spread_s)
    wgt /= wgt.sum()
    H = np.repeat(harmo, n_dir)
    F = np.repeat(f, n_dir)
    A = (amp[:, None] * np.sqrt(wgt)[None, :]).ravel()
    TH = np.tile(th * 180.0 / math.pi, n_freq)
    PH = rng.uniform(0.0, TWOPI, n_freq * n_dir)
    return dict(harmo=H, freq=F, ampli=A, angle=TH, phase=PH, rnum=rnum, clock=clock)
