"""FP64 CPU oracle for the HOS-NWT time-stepping core.

TEST INFRASTRUCTURE ONLY.  This module restates the reference's algorithm in
numpy/scipy so that the GPU library can be checked against it.  Only the test
suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker / timed CPU baseline.  The product path
(``hos-nwt_amd``) never imports it.

It follows the reference's literal transform sequence ("reference mode"),
including the round trips and the dealias calls the GPU path proves redundant,
so that the GPU path's redundancy elimination is checked by it, not assumed.

Array convention: every 2-D field is a numpy array of shape (n1, n2) indexed
[i1-1, i2-1] (reference i1 = x index, i2 = y index).  Paddle-grid fields have
shape (n3_add, n2).

Pinning of each function (tests/test_oracle_*.py):
  * r2r kinds / normalisations   -> O(N^2) closed-form FFTW definitions
  * S2F/F2S                        -> identity round trips, F2S(a)=sum a_k cos
  * dealias/filter_ext             -> projection identities, product identities
  * phisxy_etaxy derivatives       -> single-mode closed forms
  * hosphis_modes                  -> M=1 closed form W = sum a k tanh k;
                                      convergence to the exact Laplace
                                      solution at order (kA)^(M+1)
  * expon/gtof                     -> closed-form rotation, inverse pair
  * runge4_adapt2                  -> y'=lambda*y RK4 polynomial + Kutta-3
                                      embedded difference (sloshing M=1 case)
  * phiadd_sl separable form       -> brute-force quadruple loop on tiny grids
  * phi_dipoles                    -> Laplace equation, bottom condition and
                                      derivative consistency by finite
                                      differences; dipole far field of
                                      dipole_transfer on the GPU path (linear
                                      wave amplitude)
  * match_surface_wmk              -> its four defining conditions
  * end to end                     -> linear sloshing closed form, energy and
                                      volume conservation
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy import fft as sfft

# ---------------------------------------------------------------------------
# constants (type.f90:38-50)
# ---------------------------------------------------------------------------
G = 9.81
TINY = np.finfo(np.float64).eps
TWOPI = 2.0 * math.pi
L_ADD = 2  # common_vars.f90:47
K_ADD_X_MAX = 700.0  # HOS-NWT.f90:359


# ---------------------------------------------------------------------------
# r2r kinds (FFTW definitions; fourier_FFTW.f90:130-167 selects them)
# ---------------------------------------------------------------------------
def r2r(x: np.ndarray, kind: str, axis: int) -> np.ndarray:
    """Unnormalised FFTW r2r transform along ``axis``.

    REDFT00 = DCT-I, RODFT00 = DST-I, REDFT01 = DCT-III, REDFT10 = DCT-II,
    RODFT01 = DST-III, RODFT10 = DST-II.  scipy's unnormalised types 1-3 use
    exactly FFTW's definitions (factor 2 on the interior sums).
    """
    if kind == "REDFT00":
        return sfft.dct(x, type=1, axis=axis)
    if kind == "RODFT00":
        return sfft.dst(x, type=1, axis=axis)
    if kind == "REDFT01":
        return sfft.dct(x, type=3, axis=axis)
    if kind == "REDFT10":
        return sfft.dct(x, type=2, axis=axis)
    if kind == "RODFT01":
        return sfft.dst(x, type=3, axis=axis)
    if kind == "RODFT10":
        return sfft.dst(x, type=2, axis=axis)
    raise ValueError(kind)


def _s2f_vec(n: int) -> np.ndarray:
    """Per-index S2F factor: 1/(2(n-1)) at ends, doubled inside (fourier_FFTW.f90:178-182)."""
    v = np.full(n, 1.0 / max(1, 2 * (n - 1)))
    v[1:n - 1] *= 2.0
    return v


def _f2s_vec(n: int) -> np.ndarray:
    """Per-index F2S factor: 1 at ends, 1/2 inside (fourier_FFTW.f90:205-210)."""
    v = np.ones(n)
    v[1:n - 1] *= 0.5
    return v


class Fourier:
    """The reference's transform operators on one grid (fourier_FFTW.f90:35-63).

    ``grid`` is (N1, N2); N2 == 1 means a 2-D tank (1-D transforms).
    """

    def __init__(self, n1: int, n2: int):
        self.n1, self.n2 = n1, n2
        self.two_d = n2 != 1
        self.s2f = _s2f_vec(n1)[:, None] * (_s2f_vec(n2)[None, :] if self.two_d else 1.0)
        self.f2s = _f2s_vec(n1)[:, None] * (_f2s_vec(n2)[None, :] if self.two_d else 1.0)

    def _execute(self, inp: np.ndarray, cx: str, cy: str) -> np.ndarray:
        """execute / execute_big (fourier_FFTW.f90:271-302, :621-653)."""
        a = np.array(inp, dtype=np.float64, copy=True)
        n1, n2 = self.n1, self.n2
        if self.two_d and cy == "sin":
            a[:, 0] = 0.0
            a[:, n2 - 1] = 0.0
        if cx == "cos":
            a = r2r(a, "REDFT00", 0)
        else:
            b = np.zeros_like(a)
            b[1:n1 - 1, :] = r2r(a[1:n1 - 1, :], "RODFT00", 0)
            a = b
        if self.two_d:
            if cy == "cos":
                a = r2r(a, "REDFT00", 1)
            else:
                b = np.zeros_like(a)
                b[:, 1:n2 - 1] = r2r(a[:, 1:n2 - 1], "RODFT00", 1)
                a = b
        return a

    def space_2_fourier(self, f, cx="cos", cy="cos"):
        """DCT_FFTW / DCT_FFTW_big (fourier_FFTW.f90:235-247, :558-571)."""
        return self._execute(f, cx, cy) * self.s2f

    def fourier_2_space(self, a, cx="cos", cy="cos"):
        """IDCT_FFTW / IDCT_FFTW_big (fourier_FFTW.f90:251-267, :575-588)."""
        return self._execute(np.asarray(a) * self.f2s, cx, cy)


class FourierAdd:
    """Paddle-grid transforms (n3_add x n2): z DCT/DST-II/III tensor y DCT/DST-I.

    DCT_FFTW_add / IDCT_FFTW_add / execute_add (fourier_FFTW.f90:432-492).
    Forward uses REDFT01/RODFT01 along z, inverse REDFT10/RODFT10.
    """

    def __init__(self, n3_add: int, n2: int):
        self.n3_add, self.n2 = n3_add, n2
        self.two_d = n2 != 1
        ys2f = _s2f_vec(n2)[None, :] if self.two_d else 1.0
        yf2s = _f2s_vec(n2)[None, :] if self.two_d else 1.0
        self.s2f = np.full((n3_add, 1), 1.0 / n3_add) * ys2f
        self.f2s = np.full((n3_add, 1), 0.5) * yf2s
        if not self.two_d:
            self.s2f = np.full((n3_add, 1), 1.0 / n3_add)
            self.f2s = np.full((n3_add, 1), 0.5)

    def _execute(self, inp, cz, cy, fb):
        a = np.array(inp, dtype=np.float64, copy=True)
        n2 = self.n2
        if self.two_d and cy == "sin":
            a[:, 0] = 0.0
            a[:, n2 - 1] = 0.0
        if fb == 1:
            kind = "REDFT01" if cz == "cos" else "RODFT01"
        else:
            kind = "REDFT10" if cz == "cos" else "RODFT10"
        a = r2r(a, kind, 0)
        if self.two_d:
            if cy == "cos":
                a = r2r(a, "REDFT00", 1)
            else:
                b = np.zeros_like(a)
                b[:, 1:n2 - 1] = r2r(a[:, 1:n2 - 1], "RODFT00", 1)
                a = b
        return a

    def space_2_fourier(self, f, cz="cos", cy="cos"):
        return self._execute(f, cz, cy, 1) * self.s2f

    def fourier_2_space(self, a, cz="cos", cy="cos"):
        return self._execute(np.asarray(a) * self.f2s, cz, cy, -1)


def fourier_2_space_y_axis(a: np.ndarray, cy: str, axis: int) -> np.ndarray:
    """IDCT_FFTW_y applied along ``axis`` (fourier_FFTW.f90:362-380, :384-407).

    For a length-n2 line: cos -> sum_j a_j cos(pi j i/(n2-1));
    sin -> sum_j a_j sin(pi j i/(n2-1)) with end modes and end nodes zero.
    """
    a = np.moveaxis(np.asarray(a, dtype=np.float64), axis, -1)
    n2 = a.shape[-1]
    if n2 == 1:  # 2-D tank: the y "transform" is the identity (ky = 0)
        out = a.copy() if cy == "cos" else np.zeros_like(a)
        return np.moveaxis(out, -1, axis)
    v = a * _f2s_vec(n2)
    if cy == "cos":
        out = r2r(v, "REDFT00", -1)
    else:
        out = np.zeros_like(v)
        out[..., 1:n2 - 1] = r2r(v[..., 1:n2 - 1], "RODFT00", -1)
    return np.moveaxis(out, -1, axis)


# ---------------------------------------------------------------------------
# parameters and static tables (HOS-NWT.f90:99-406)
# ---------------------------------------------------------------------------
@dataclass
class Params:
    """Nondimensional run parameters (input.f90:62-133 after HOS-NWT.f90:99-119)."""

    n1: int
    n2: int
    mHOS: int
    xlen: float
    ylen: float = 1.0
    n3: int = 33
    p1: int | None = None  # default: mHOS  (dealiasing factor (p+1)/2)
    p2: int | None = None  # default: mHOS in 3-D, 1 in 2-D
    Nd1: int | None = None  # extended grid; default ((p1+1)*n1)//2
    Nd2: int | None = None
    icase: int = 2
    igeom: int = 1
    d_hinge: float = 0.0
    iramp: int = 1
    Tramp: float = 0.0
    T_stop: float = 1.0e30
    T_mono: float = 1.0
    iabsnb: int = 1
    xabsf: float = 0.8
    coeffabsf: float = 1.0
    i_wmk: int = 1
    # spectral filters (common_vars.f90:43): x, y free-surface modes, z paddle modes
    coeffilt: tuple = (1.0, 1.0, 1.0)
    # wavemaker spectral content: pos_ini (n2 x n_freq complex) and freq (n_freq)
    freq: np.ndarray = field(default_factory=lambda: np.zeros(0))
    pos_ini: np.ndarray = field(default_factory=lambda: np.zeros((1, 0), complex))
    # y derivative of pos_ini (wavemaking.f90:147-149), needed for i_wmk >= 2; None: zero
    dposdy_ini: np.ndarray | None = None
    # dipole wavemaker (igeom = 3): one row per wave component, mu(:,1..4) of dipoles_ini
    # (resol_wmkr.f90:1448-1455), the component frequencies, the dipole abscissa and the
    # number of image pairs on each side (input.f90:137-139, HOS-NWT.f90:111)
    dip_mu: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), complex))
    dip_freq: np.ndarray = field(default_factory=lambda: np.zeros(0))
    x_dip: float = 2.5
    n_side: int = 6

    def __post_init__(self):
        if self.p1 is None:
            self.p1 = self.mHOS
        if self.p2 is None:
            self.p2 = self.mHOS if self.n2 != 1 else 1
        if self.Nd1 is None:
            self.Nd1 = ((self.p1 + 1) * self.n1) // 2
        if self.Nd2 is None:
            self.Nd2 = ((self.p2 + 1) * self.n2) // 2
        if self.n2 == 1:
            self.ylen = 1.0
        self.n3_add = L_ADD * (self.n3 - 1)
        if self.dposdy_ini is None:
            self.dposdy_ini = np.zeros_like(np.asarray(self.pos_ini, complex))
        self.has_wmk = self.icase in (2, 3, 31, 32, 33, 4, 41, 42) and self.igeom in (1, 2)
        self.has_dip = self.icase in (2, 3, 31, 32, 33, 4, 41, 42) and self.igeom == 3
        if self.i_wmk == 3 and min(self.coeffilt) < 1.0:
            # wmk_3rdorder filters its INTENT(IN) stage inputs in place (resol_wmkr.f90:1019-1030):
            # what that does to the caller's arrays is left to the compiler, so it is not restated
            raise ValueError("i_wmk = 3 with coeffilt < 1 is not supported")


class Tables:
    """Wavenumbers, omega, kth, kth_ext and wavemaker tables (HOS-NWT.f90:269-406)."""

    def __init__(self, p: Params):
        self.p = p
        n1, n2, M = p.n1, p.n2, p.mHOS
        pixlen = math.pi / p.xlen
        piylen = math.pi / p.ylen if n2 != 1 else 0.0
        self.kx = np.arange(n1) * pixlen
        self.ky = np.arange(n2) * piylen
        self.k = np.sqrt(self.kx[:, None] ** 2 + self.ky[None, :] ** 2)
        self.omega = np.sqrt(self.k * np.tanh(self.k))
        self.kth = self._kth(self.k, M)
        kxe = np.arange(p.Nd1) * pixlen
        kye = np.arange(p.Nd2) * piylen
        ke = np.sqrt(kxe[:, None] ** 2 + kye[None, :] ** 2)
        self.kth_ext = self._kth(ke, M)
        # kth2_ext: k^j / tanh k for odd j (0 at k = 0), k^j for even j (HOS-NWT.f90:310-327)
        self.kth2_ext = np.zeros(ke.shape + (M,))
        thk = np.tanh(ke)
        for j in range(1, M + 1):
            if j % 2 == 1:
                with np.errstate(divide="ignore", invalid="ignore"):
                    v = ke ** j / thk
                v[0, 0] = 0.0
                self.kth2_ext[..., j - 1] = v
            else:
                self.kth2_ext[..., j - 1] = ke ** j
        # mesh
        self.delx = p.xlen / (n1 - 1)
        self.dely = p.ylen / (n2 - 1) if n2 != 1 else 0.0
        self.cx = np.arange(n1) * self.delx
        self.cy = np.arange(n2) * self.dely
        self.delz = 1.0 / (p.n3 - 1)
        self.cz = np.arange(p.n3) * self.delz - 1.0
        # wavemaker modes (HOS-NWT.f90:330-339)
        xlen_add = 2.0 * L_ADD
        pixlen_add = math.pi / xlen_add
        n3a = p.n3_add
        self.kx_add = (2 * np.arange(1, n3a + 1) - 1) * pixlen_add
        self.k_add = np.sqrt(self.kx_add[:, None] ** 2 + self.ky[None, :] ** 2)
        self.k_add_thk_add = self.k_add * np.tanh(self.k_add * p.xlen)
        # beach (initial_condition.f90:724-794, full-width wavemaker, no back beach)
        self.nu = np.zeros((n1, n2))
        if p.iabsnb != 0:
            x0 = p.xlen * p.xabsf
            front = abs(p.xabsf - 1.0) > TINY
            if front:
                x2 = p.xlen - x0
                sel = self.cx >= x0
                x3 = (self.cx[sel] - x0) / x2
                self.nu[sel, :] = (p.coeffabsf * x3 * x3)[:, None]
        # beach x derivative as the reference builds it, d nu / d x3 = 2 c x3 without the
        # 1/(xlen - x0) of the chain rule (initial_condition.f90:768); no side beach: dnuy = 0
        self.dnux = np.zeros((n1, n2))
        self.dnuy = np.zeros((n1, n2))
        if p.iabsnb != 0 and abs(p.xabsf - 1.0) > TINY:
            x0 = p.xlen * p.xabsf
            sel = self.cx >= x0
            x3 = (self.cx[sel] - x0) / (p.xlen - x0)
            self.dnux[sel, :] = (p.coeffabsf * 2.0 * x3)[:, None]
        # wavemaker vertical shape and its derivative (wavemaking.f90:184-250)
        self.gz = np.array([gzed(z, p.igeom, p.d_hinge) for z in self.cz])
        self.dgz = np.array([dgzeddz(z, p.igeom, p.d_hinge) for z in self.cz])

    @staticmethod
    def _kth(k: np.ndarray, M: int) -> np.ndarray:
        out = np.zeros(k.shape + (M,))
        thk = np.tanh(k)
        for j in range(1, M + 1):
            out[..., j - 1] = k ** j * thk if j % 2 == 1 else k ** j
        return out

    def csh_slice(self, i3: int):
        """csh_add_x(:,:,i3) and k_add_sh_add_x(:,:,i3) (HOS-NWT.f90:364-405)."""
        p = self.p
        k = self.k_add[i3][None, :]  # (1, n2)
        k2 = 2.0 * k
        cx = self.cx[:, None]
        k_add_x = k2 * p.xlen
        expon3 = np.where(k_add_x <= K_ADD_X_MAX, np.exp(-np.minimum(k_add_x, K_ADD_X_MAX)), 0.0)
        expon3p1 = expon3 + 1.0
        k_add_x1 = k * cx
        expon1 = np.where(k_add_x1 <= K_ADD_X_MAX, np.exp(-np.minimum(k_add_x1, K_ADD_X_MAX)) / expon3p1, 0.0)
        k_add_x2 = k2 * (p.xlen - cx)
        expon2 = np.where(k_add_x2 <= K_ADD_X_MAX, np.exp(-np.minimum(k_add_x2, K_ADD_X_MAX)), 0.0)
        expon12 = np.where(k_add_x1 + k_add_x2 <= K_ADD_X_MAX, expon1 * expon2, 0.0)
        csh = expon1 + expon12
        sh = expon1 - expon12
        return csh, sh * k


def gzed(z: float, igeom: int, d_hinge: float) -> float:
    """Paddle shape g(z) (wavemaking.f90:184-215)."""
    delta = min(d_hinge, (1.0 - d_hinge) / 10.0)
    if igeom == 1:
        return 1.0
    if z >= d_hinge - 1.0 + delta:
        return (z + 1.0 - d_hinge) / (1.0 - d_hinge)
    if z <= d_hinge - 1.0 - delta:
        return 0.0
    u = (z - (d_hinge - 1.0)) / delta
    return delta / (2.0 * (1.0 - d_hinge)) * (5.0 / 16.0 + u + 15.0 / 16.0 * (u ** 2 - u ** 4 / 3.0 + u ** 6 / 15.0))


# ---------------------------------------------------------------------------
# wavemaker motion per stage (resol_wmkr.f90:1268-1397, wavemaking.f90:259-304)
# ---------------------------------------------------------------------------
def dgzeddz(z: float, igeom: int, d_hinge: float) -> float:
    """Vertical derivative of the paddle shape (wavemaking.f90:222-250)."""
    delta = min(d_hinge, (1.0 - d_hinge) / 10.0)
    if igeom == 1:
        return 0.0
    if z >= d_hinge - 1.0 + delta:
        return 1.0 / (1.0 - d_hinge)
    if z <= d_hinge - 1.0 - delta:
        return 0.0
    u = (z - (d_hinge - 1.0)) / delta
    return 1.0 / (1.0 - d_hinge) * (0.5 + 15.0 / 16.0 * (u - 2.0 / 3.0 * u ** 3 + u ** 5 / 5.0))


def ramp(t: float, p: Params):
    """alpha, alpha_t, alpha_tt of the time ramp (resol_wmkr.f90:1268-1397)."""
    Tr, Ts = p.Tramp, p.T_stop
    if p.iramp == 11:
        Tsw = 3.0 / math.sqrt(5.0 / 9.81)
        return (t / Tsw, 1.0 / Tsw, 0.0) if t <= Tsw else (1.0, 0.0, 0.0)
    if p.iramp == 33:
        Tsw = -math.log(0.1) / (10.0 * p.T_mono)
        a = 1.0 - math.exp(-Tsw * t)
        return a, (1.0 - a) * Tsw, (1.0 - a) * Tsw * Tsw
    if t <= Tr:
        if p.iramp == 0:
            return 1.0, 0.0, 0.0
        if p.iramp == 1:
            return t / Tr, 1.0 / Tr, 0.0
        if p.iramp == 2:
            s = t / Tr
            return s * s, 2.0 * s / Tr, 2.0 / (Tr * Tr)
        if p.iramp == 21:
            T2 = Tr ** 2
            if t <= Tr / 2.0:
                return 2.0 * t * t / T2, 4.0 * t / T2, 4.0 / T2
            ts = t / Tr - 1.0
            return 1.0 - 2.0 * ts ** 2, -4.0 * ts / Tr, -4.0 / T2
        if p.iramp == 4:
            s = t / Tr
            s2 = s * s
            return s2 * s2, 4.0 * s2 * s / Tr, 12.0 * s2 / (Tr * Tr)
        if p.iramp == 41:
            T2 = Tr ** 2
            rap = 8.0 * t / Tr
            rap2 = t * t / T2
            rap22 = 3.0 * rap2
            return ((rap22 - rap + 6.0) * rap2, (4.0 * rap22 - 3.0 * rap + 12.0) * t / T2,
                    (12.0 * rap22 - 6.0 * rap + 12.0) / T2)
        raise ValueError("Unknown ramp")
    if t >= Ts + Tr:
        return 0.0, 0.0, 0.0
    if t >= Ts:
        ti = Ts + Tr - t
        if p.iramp == 1:
            return ti / Tr, -1.0 / Tr, 0.0
        if p.iramp == 2:
            s = ti / Tr
            return s * s, -2.0 * s / Tr, 2.0 / (Tr * Tr)
        if p.iramp == 21:
            T2 = Tr ** 2
            if ti <= Tr / 2.0:
                return 2.0 * ti * ti / T2, -4.0 * ti / T2, 4.0 / T2
            ts = ti / Tr - 1.0
            return 1.0 - 2.0 * ts ** 2, 4.0 * ts / Tr, -4.0 / T2
        if p.iramp == 4:
            s = ti / Tr
            s2 = s * s
            return s2 * s2, -4.0 * s2 * s / Tr, 12.0 * s2 / (Tr * Tr)
        if p.iramp == 41:
            T2 = Tr ** 2
            rap = 8.0 * ti / Tr
            rap2 = ti * ti / T2
            rap22 = 3.0 * rap2
            return ((rap22 - rap + 6.0) * rap2, -(4.0 * rap22 - 3.0 * rap + 12.0) * ti / T2,
                    (12.0 * rap22 - 6.0 * rap + 12.0) / T2)
        # the reference leaves alpha unchanged for iramp 0 here; iramp 0 has Tr=0
        return 1.0, 0.0, 0.0
    return 1.0, 0.0, 0.0


def pos_1st(t: float, p: Params, tb: Tables):
    """X, X_t, X_tt on the paddle, shape (n3, n2) (wavemaking.f90:259-304)."""
    w = TWOPI * p.freq  # (nf,)
    e = np.exp(1j * w * t)  # (nf,)
    # sum over frequencies of pos_ini(i2,f) e^{i w t} (i w)^d, then times g(z)
    base = p.pos_ini * e[None, :]  # (n2, nf)
    s0 = base.sum(axis=1)
    s1 = (base * (1j * w)[None, :]).sum(axis=1)
    s2 = (base * (1j * w)[None, :] ** 2).sum(axis=1)
    gz = tb.gz[:, None]
    return (gz * s0[None, :]).real, (gz * s1[None, :]).real, (gz * s2[None, :]).real


def dpos_1st(t: float, p: Params, tb: Tables):
    """dX/dy, d2X/dydt, dX/dz, d2X/dzdt on the paddle, shape (n3, n2) (wavemaking.f90:306-406)."""
    w = TWOPI * p.freq
    e = np.exp(1j * w * t)
    by = p.dposdy_ini * e[None, :]
    bz = p.pos_ini * e[None, :]
    sy0, sy1 = by.sum(axis=1), (by * (1j * w)[None, :]).sum(axis=1)
    sz0, sz1 = bz.sum(axis=1), (bz * (1j * w)[None, :]).sum(axis=1)
    gz, dgz = tb.gz[:, None], tb.dgz[:, None]
    return ((gz * sy0[None, :]).real, (gz * sy1[None, :]).real,
            (dgz * sz0[None, :]).real, (dgz * sz1[None, :]).real)


def dipole_transfer(k: float, z_dip: float) -> complex:
    """Transfer function of a dipole at depth z_dip, TF_adim_new 'dipole' (linear_wave.f90:316-327)."""
    # IEEE arithmetic as the reference's: k^2 - w^4 rounds to 0 once tanh k rounds to 1
    # (k > 18.7), and the finite-depth branch then gives an infinite transfer function
    k = np.float64(k)
    k2 = k * k
    w2 = k * np.tanh(k)
    with np.errstate(divide="ignore", invalid="ignore"):
        if k <= 20.0:
            w4 = w2 * w2
            return complex((k2 - w4 + w2) / (2.0 * np.sqrt(w2) * (k2 - w4) * np.cosh(k) * np.sinh(k * (z_dip + 1.0))))
        return complex(np.sqrt(w2) / (2.0 * np.exp(k * z_dip) * k2))


def dipoles_ini(freq, k, ampli, phase, icase: int, x_dip: float, z_dipmono: float = -0.25) -> np.ndarray:
    """mu(:,1..4) of the dipole wavemaker, one row per wave component (resol_wmkr.f90:1404-1455):
    the depth chosen from the component's nondimensional frequency, its transfer function,
    mu_z = a TF e^{i(k x_dip + phase)}, mu_x = -i mu_z tanh(k (z_dip + 1)) and the bottom image's
    depth -2 - z_dip."""
    freq = np.asarray(freq, float)
    z = np.where(freq > 1.0, -0.35, np.where(freq > 0.6, -0.2, np.where(freq > 0.5, -0.4, -0.6)))
    if icase == 2:
        z = np.full_like(freq, z_dipmono)
    mu = np.zeros((len(freq), 4), complex)
    for f in range(len(freq)):
        tf = dipole_transfer(float(k[f]), float(z[f]))
        mu[f, 2] = z[f]
        mu[f, 3] = -2.0 - z[f]
        mu[f, 1] = ampli[f] * tf * np.exp(1j * (k[f] * x_dip + phase[f]))
        mu[f, 0] = -1j * mu[f, 1] * math.tanh(k[f] * (z[f] + 1.0))
    return mu


def dipole_images(p: "Params", x: np.ndarray):
    """delta_x(:, :, -n_side:n_side) of dipoles_ini (resol_wmkr.f90:1461-1481): the main dipole at
    x_dip and its wall images; returns {j: x - x_j}."""
    dx = {0: x - p.x_dip}
    for j in range(1, p.n_side + 1):
        if j % 2 == 0:
            right, left = p.x_dip + p.xlen * j, p.x_dip - p.xlen * j
        else:
            right, left = -p.x_dip + p.xlen * (j + 1), -p.x_dip - p.xlen * (j - 1)
        dx[j] = x - right
        dx[-j] = x - left
    return dx


def phi_dipoles(eta: np.ndarray, t: float, alpha: float, alphat: float, p: "Params", x: np.ndarray):
    """Potential of the dipole wavemaker and its t, x, z derivatives at z = eta, each component's
    dipole with its bottom image and n_side pairs of wall images (+ their bottom images), the wall
    images' x moment alternating in sign (phi_dipoles, resol_wmkr.f90:1492-1647).  beta = 1,
    beta_x = 0 (dipoles_ini); phiy is not formed."""
    dx = dipole_images(p, x)
    phi = np.zeros(eta.shape, complex)
    dphi_dt = np.zeros(eta.shape, complex)
    dphi_dx = np.zeros(eta.shape, complex)
    dphi_dz = np.zeros(eta.shape, complex)
    for f in range(len(p.dip_freq)):
        mu1, mu2, mu3, mu4 = p.dip_mu[f]
        dz_up = eta - mu3.real
        dz_down = eta - mu4.real
        resu = np.zeros(eta.shape, complex)
        resu_dx = np.zeros(eta.shape, complex)
        resu_dz = np.zeros(eta.shape, complex)
        terms = [(0, 1.0, dz_up, 1.0), (0, 1.0, dz_down, -1.0)]
        for j in range(1, p.n_side + 1):
            sg = -1.0 if j % 2 == 1 else 1.0
            terms += [(-j, sg, dz_up, 1.0), (j, sg, dz_up, 1.0), (j, sg, dz_down, -1.0), (-j, sg, dz_down, -1.0)]
        for j, sg, dz, up in terms:
            d = dx[j]
            r2 = 1.0 / (d * d + dz * dz)
            muxor2 = sg * mu1 * r2
            muzor2 = mu2 * r2
            tmp = muxor2 * d + up * muzor2 * dz
            resu = resu + tmp
            resu_dx = resu_dx + muxor2 - tmp * 2.0 * d * r2
            resu_dz = resu_dz + up * muzor2 - tmp * 2.0 * dz * r2
        w = TWOPI * p.dip_freq[f]
        eiwt = np.exp(1j * w * t)
        phi = phi + resu * eiwt
        dphi_dt = dphi_dt + resu * 1j * w * eiwt
        dphi_dx = dphi_dx + resu_dx * eiwt
        dphi_dz = dphi_dz + resu_dz * eiwt
    phi, dphi_dt, dphi_dx, dphi_dz = phi / TWOPI, dphi_dt / TWOPI, dphi_dx / TWOPI, dphi_dz / TWOPI
    return ((phi * alpha).real, (phi * alphat + dphi_dt * alpha).real, (dphi_dx * alpha).real,
            (dphi_dz * alpha).real)


def match_surface_coeffs(f, df, d2f):
    """Coefficients c1..c4 of the C2 continuation polynomial.

    P(xx) = (c1 xx^7 + c2 xx^5 + c3 xx^3 + c4 xx) exp(-6 (xx + 1)) with
    P(-1)=f, P'(-1)=df, P''(-1)=d2f and P(-2/3)=0 (resol_wmkr.f90:1655-1741,
    l_add = 2).  Solved as the 4x4 linear system those conditions define.
    """
    f = np.atleast_1d(np.asarray(f, float))
    A = np.zeros((4, 4))
    x0, a = -1.0, 6.0
    powers = [7, 5, 3, 1]
    E0 = math.exp(-a * (x0 + 1.0))
    for col, n in enumerate(powers):
        # q = x^n ; P = q E ; P' = (q' - a q)E ; P'' = (q'' - 2a q' + a^2 q)E
        q = x0 ** n
        q1 = n * x0 ** (n - 1)
        q2 = n * (n - 1) * x0 ** (n - 2) if n >= 2 else 0.0
        A[0, col] = q * E0
        A[1, col] = (q1 - a * q) * E0
        A[2, col] = (q2 - 2 * a * q1 + a * a * q) * E0
        xb = -2.0 / 3.0
        A[3, col] = xb ** n * math.exp(-a * (xb + 1.0))
    rhs = np.stack([f, np.atleast_1d(df), np.atleast_1d(d2f), np.zeros_like(f)])
    return np.linalg.solve(A, rhs)  # (4, n)


def match_surface_wmk(data: np.ndarray, n3: int, delz: float) -> np.ndarray:
    """Continue a paddle RHS from i3=n3 up to n3_add (resol_wmkr.f90:1655-1741)."""
    d = data.copy()
    n3a = d.shape[0]
    # reference rows n3-3..n3 are python n3-4..n3-1
    fm3, fm2, fm1, f0 = d[n3 - 4], d[n3 - 3], d[n3 - 2], d[n3 - 1]
    df = (fm2 - 4.0 * fm1 + 3.0 * f0) / (2.0 * delz)
    d2f = (-fm3 + 4.0 * fm2 - 5.0 * fm1 + 2.0 * f0) / delz ** 2
    c = match_surface_coeffs(f0, df, d2f)  # (4, n2)
    i3 = np.arange(n3 + 1, n3a + 1)
    xx = (i3 - 1) / (n3 - 1) - float(L_ADD)
    xx2 = xx * xx
    ex = np.exp(-6.0 * (xx + 1.0))
    pxx = (((c[0][None, :] * xx2[:, None] + c[1][None, :]) * xx2[:, None] + c[2][None, :]) * xx2[:, None]
           + c[3][None, :]) * xx[:, None] * ex[:, None]
    d[n3:n3a, :] = pxx
    return d


# ---------------------------------------------------------------------------
# additional potential on the free surface
# ---------------------------------------------------------------------------
def phiadd_sl(a_add, da_add, eta, tb: Tables, need=("phi", "phix", "phiy", "phiz", "phit")):
    """Separable evaluation of phiadd_SL (resol_wmkr.f90:1124-1187).

    phi(i1,i2) = sum_{i3} cos(kx_add(i3)(eta+1)) B(i1,i3,i2) with
    B(i1,i3,.) = F2S_y[ a(i3,.) C(i1,.,i3) ]; identical to the reference's
    quadruple loop in exact arithmetic (pinned by brute force in the tests).
    Pass eta=None for z=0 (phiadd_1st_SWEET, resol_wmkr.f90:485-532).
    """
    p = tb.p
    n1, n2 = p.n1, p.n2
    out = {k: np.zeros((n1, n2)) for k in need}
    etap1 = np.ones((n1, n2)) if eta is None else eta + 1.0
    for i3 in range(p.n3_add):
        csh, ksh = tb.csh_slice(i3)  # (n1, n2) indexed [i1, j2]
        kxa = tb.kx_add[i3]
        c = np.cos(kxa * etap1)
        s = np.sin(kxa * etap1)
        a = a_add[i3][None, :]
        da = da_add[i3][None, :]
        if "phi" in need:
            out["phi"] += c * fourier_2_space_y_axis(a * csh, "cos", 1)
        if "phit" in need:
            out["phit"] += c * fourier_2_space_y_axis(da * csh, "cos", 1)
        if "phix" in need:
            out["phix"] -= c * fourier_2_space_y_axis(a * ksh, "cos", 1)
        if "phiy" in need:
            out["phiy"] -= c * fourier_2_space_y_axis(a * csh * tb.ky[None, :], "sin", 1)
        if "phiz" in need:
            out["phiz"] -= s * fourier_2_space_y_axis(a * csh * kxa, "cos", 1)
        # fields of phiadd_1st_SWEET needed by i_wmk >= 2 (resol_wmkr.f90:500-524)
        if "phizz" in need:
            out["phizz"] -= c * fourier_2_space_y_axis(a * csh * kxa * kxa, "cos", 1)
        if "phixt" in need:
            out["phixt"] -= c * fourier_2_space_y_axis(da * ksh, "cos", 1)
        if "phiyt" in need:
            out["phiyt"] -= c * fourier_2_space_y_axis(da * csh * tb.ky[None, :], "sin", 1)
        if "phizt" in need:
            out["phizt"] -= s * fourier_2_space_y_axis(da * csh * kxa, "cos", 1)
        if "phixx" in need:
            out["phixx"] += c * fourier_2_space_y_axis(a * csh * tb.k_add[i3][None, :] ** 2, "cos", 1)
        if "phixz" in need:
            out["phixz"] += s * fourier_2_space_y_axis(a * ksh * kxa, "cos", 1)
        if "phiyz" in need:
            out["phiyz"] += s * fourier_2_space_y_axis(a * csh * tb.ky[None, :], "sin", 1)
    return out


def phiadd_sl_bruteforce(a_add, da_add, eta, tb: Tables):
    """The reference's literal quadruple loop (resol_wmkr.f90:1149-1175); tiny grids only."""
    p = tb.p
    n1, n2 = p.n1, p.n2
    phi, phit, phix, phiy, phiz = (np.zeros((n1, n2)) for _ in range(5))
    for i2 in range(n2):
        for i1 in range(n1):
            for i3 in range(p.n3_add):
                csh, ksh = tb.csh_slice(i3)
                e = 0.0 if eta is None else eta[i1, i2]
                kx_e = tb.kx_add[i3] * (e + 1.0)
                ce, se = math.cos(kx_e), math.sin(kx_e)
                for j2 in range(n2):
                    cy = math.cos(tb.ky[j2] * tb.cy[i2])
                    sy = math.sin(tb.ky[j2] * tb.cy[i2])
                    phi[i1, i2] += a_add[i3, j2] * ce * csh[i1, j2] * cy
                    phit[i1, i2] += da_add[i3, j2] * ce * csh[i1, j2] * cy
                    phix[i1, i2] -= a_add[i3, j2] * ce * ksh[i1, j2] * cy
                    phiy[i1, i2] -= a_add[i3, j2] * ce * csh[i1, j2] * tb.ky[j2] * sy
                    phiz[i1, i2] -= a_add[i3, j2] * se * csh[i1, j2] * tb.kx_add[i3] * cy
    return dict(phi=phi, phit=phit, phix=phix, phiy=phiy, phiz=phiz)


# ---------------------------------------------------------------------------
# solver state
# ---------------------------------------------------------------------------
# ---------------------------------------------------------------------------
# spectral filters (filtering.f90)
# ---------------------------------------------------------------------------
def filter_nfilt(n: int, coeff: float) -> int:
    """Number of kept modes (filtering.f90:53-55): ABS(INT(n*coeff)), n when coeff >= 0.9999."""
    return n if coeff >= 0.9999 else abs(int(n * coeff))


def filter_weights(n: int, coeff: float, first_dir: bool = True) -> np.ndarray:
    """Weights of modes i = 1..n along one direction (filtering_base, filtering.f90:334-525).

    Returned as an array over 0-based mode index i-1.  The reference's integer
    quotients REAL(i/n) (types -1 and the exponential cut-off) are kept, and so
    is the exponential y branch that only touches the last x mode, which
    ``first_dir=False`` cannot express per direction: callers apply
    :func:`filter_modes` for that case.
    """
    nf = filter_nfilt(n, coeff)
    i = np.arange(1, n + 1)
    w = np.ones(n)
    tiny = TINY
    if coeff >= 0.0:
        w[i > nf] = 0.0
    elif abs(coeff + 1.0) < tiny:
        w = 1.0 - (i // n).astype(float)
    elif abs(coeff + 2.0) < tiny:
        a = TWOPI * i / n
        w = np.sin(a) / a
    elif abs(coeff + 3.0) < tiny:
        w = (1.0 + np.cos(TWOPI * i / n)) / 2.0
    elif abs(coeff + 4.0) < tiny:
        s0 = (1.0 + np.cos(TWOPI * i / n)) / 2.0
        w = s0 ** 4 * (35.0 - 84.0 * s0 + 70.0 * s0 ** 2 - 20.0 * s0 ** 3)
    elif abs(coeff + 5.0) < tiny:
        from scipy.special import erfc
        pp = 10
        s0 = (i - 0.75 * n) / (n + 1.0)
        q = (4.0 / 3.0 * s0) ** 2
        # -log(1-q)/q -> 1 as q -> 0: the mode i = 3n/4 (s0 = 0) takes the limit
        # value 1/2 where the literal expression would be 0/0
        r = np.ones(n)
        nz = q > 0.0
        r[nz] = -np.log(1.0 - q[nz]) / q[nz]
        w = 0.5 * erfc(2.0 * math.sqrt(pp) * s0 * np.sqrt(r))
    else:
        w = np.ones(n)
        sel = i > nf
        w[sel] = np.exp(-2.0 * (TWOPI * (i[sel] // n) - TWOPI * (nf // n)) ** 4)
    return w


def filter_modes(a: np.ndarray, coeff: float, axis: int) -> np.ndarray:
    """filtering_base on a modal (n1, n2) array along x (axis 0) or y (axis 1)."""
    n1, n2 = a.shape
    out = a.copy()
    if axis == 0:
        out *= filter_weights(n1, coeff)[:, None]
        return out
    nf = filter_nfilt(n2, coeff)
    if coeff < 0.0 and not any(abs(coeff + c) < TINY for c in (1.0, 2.0, 3.0, 4.0, 5.0)):
        # exponential cut-off along y touches only the last x mode (filtering.f90:513-517)
        j = np.arange(1, n2 + 1)
        sel = j > nf
        out[n1 - 1, sel] *= np.exp(-2.0 * (TWOPI * (j[sel] // n2) - TWOPI * (nf // n2)) ** 4)
        return out
    out *= filter_weights(n2, coeff)[None, :]
    return out


@dataclass
class State:
    """Driver state copied into runge4_adapt2 each attempt (HOS-NWT.f90:506-515)."""

    eta: np.ndarray
    phis: np.ndarray
    a1st_add: np.ndarray
    a_add_ramp: np.ndarray
    a1st: np.ndarray
    eta1st: np.ndarray
    # second-order wavemaker state (i_wmk >= 2): paddle modes, modal and physical FS parts
    a2nd_add: np.ndarray | None = None
    a2nd: np.ndarray | None = None
    eta2nd: np.ndarray | None = None
    # third-order wavemaker paddle modes (i_wmk = 3)
    a3rd_add: np.ndarray | None = None

    def __post_init__(self):
        if self.a2nd_add is None:
            self.a2nd_add = np.zeros_like(self.a1st_add)
        if self.a2nd is None:
            self.a2nd = np.zeros_like(self.a1st)
        if self.eta2nd is None:
            self.eta2nd = np.zeros_like(self.eta1st)
        if self.a3rd_add is None:
            self.a3rd_add = np.zeros_like(self.a1st_add)

    def copy(self):
        return State(*(x.copy() for x in (self.eta, self.phis, self.a1st_add, self.a_add_ramp,
                                          self.a1st, self.eta1st, self.a2nd_add, self.a2nd, self.eta2nd,
                                          self.a3rd_add)))


class Oracle:
    """Reference-mode restatement of solveHOS2 / runge4_adapt2 and the driver loop."""

    def __init__(self, p: Params, wmk_direct: bool = False):
        self.p = p
        self.tb = Tables(p)
        self.F = Fourier(p.n1, p.n2)
        self.FB = Fourier(p.Nd1, p.Nd2)
        self.FA = FourierAdd(p.n3_add, p.n2)
        self.second = None  # (da2nd, deta2nd, da_add_2) of the last solve_hos2, i_wmk >= 2
        self.third = None  # da_add_3 of the last solve_hos2, i_wmk = 3
        self.wmk_direct = wmk_direct
        # module globals written inside the step (resol_wmkr.f90:1268-1397, wavemaking.f90:277-295)
        self.alpha = self.alphat = self.alphatt = 0.0
        self.pos1st = np.zeros((p.n3, p.n2))
        self.dpos1stdt = np.zeros((p.n3, p.n2))
        self.d2pos1stdt2 = np.zeros((p.n3, p.n2))
        self.time = 0.0
        self.etax = self.etay = None
        self.phi_add = np.zeros((p.n1, p.n2))
        self.deta = np.zeros((p.n1, p.n2))
        self.n_rhs = 0
        self.count = {"big": 0, "base": 0}
        self.endramp = 0  # HOS-NWT.f90:527,546-549

    # ----- extended-grid helpers -----------------------------------------
    def extend(self, a):
        """extend (variables.f90:243-266)."""
        p = self.p
        e = np.zeros((p.Nd1, p.Nd2))
        e[:p.n1, :p.n2] = a
        return e

    def s2f_big(self, f, cx="cos", cy="cos"):
        self.count["big"] += 1
        return self.FB.space_2_fourier(f, cx, cy)

    def f2s_big(self, a, cx="cos", cy="cos"):
        self.count["big"] += 1
        return self.FB.fourier_2_space(a, cx, cy)

    def s2f(self, f, cx="cos", cy="cos"):
        self.count["base"] += 1
        return self.F.space_2_fourier(f, cx, cy)

    def f2s(self, a, cx="cos", cy="cos"):
        self.count["base"] += 1
        return self.F.fourier_2_space(a, cx, cy)

    def _dealias_active(self, order):
        p = self.p
        return ((p.p1 != 1 and (order - 1) % (p.p1 - 1) == 0)
                or (p.p2 != 1 and (order - 1) % (p.p2 - 1) == 0))

    def dealias(self, order, x, cx="cos", cy="cos"):
        """dealias (dealiasing.f90:39-92)."""
        p = self.p
        if not self._dealias_active(order):
            return x
        a = self.s2f_big(x, cx, cy)
        if p.p1 != 1 and (order - 1) % (p.p1 - 1) == 0:
            a[p.n1:p.Nd1, :] = 0.0
        if p.n2 != 1 and p.p2 != 1 and (order - 1) % (p.p2 - 1) == 0:
            a[:, p.n2:p.Nd2] = 0.0
        return self.f2s_big(a, cx, cy)

    def filter_ext(self, x, cx="cos", cy="cos"):
        """filter_ext (dealiasing.f90:96-134)."""
        p = self.p
        a = self.s2f_big(x, cx, cy)
        return self.f2s(a[:p.n1, :p.n2], cx, cy)

    # ----- integrating factor (variablechange.f90:37-130) ----------------
    def expon(self, n1_, n2_, dt):
        c = np.cos(self.tb.omega * dt)
        s = np.sin(self.tb.omega * dt)
        return n1_ * c - n2_ * s, n2_ * c + n1_ * s

    def gtof(self, g1, g2, dt):
        c = np.cos(-self.tb.omega * dt)
        s = np.sin(-self.tb.omega * dt)
        return g1 * c - g2 * s, g2 * c + g1 * s

    # ----- phisxy_etaxy (resol_HOS.f90:924-1078) --------------------------
    def phisxy_etaxy(self, phisrk, etark):
        p, tb = self.p, self.tb
        M = p.mHOS
        as_ = self.s2f(etark)
        etapm = [None] * (M + 1)
        etapm[1] = self.f2s_big(self.extend(as_))
        for j in range(1, M):
            e = etapm[j] / (j + 1) * etapm[1]
            etapm[j + 1] = self.dealias(j + 1, e)
        etax = self.f2s_big(self.extend(-tb.kx[:, None] * as_), "sin", "cos")
        if p.n2 != 1:
            etay = self.f2s_big(self.extend(-tb.ky[None, :] * as_), "cos", "sin")
        else:
            etay = np.zeros_like(etax)
        gradeta2 = self.dealias(2, etax * etax + etay * etay)
        as_ = self.s2f(phisrk)
        phisx = self.f2s_big(self.extend(-tb.kx[:, None] * as_), "sin", "cos")
        if p.n2 != 1:
            phisy = self.f2s_big(self.extend(-tb.ky[None, :] * as_), "cos", "sin")
        else:
            phisy = np.zeros_like(phisx)
        return etapm, etax, etay, gradeta2, phisx, phisy

    # ----- HOSphis_modes (resol_HOS.f90:665-915) --------------------------
    def hosphis_modes(self, phisrk, etapm, gradeta2):
        p, tb = self.p, self.tb
        M = p.mHOS
        kth_ext = tb.kth_ext
        phim = self.s2f(phisrk)
        aHOS = [None] * (M + 2)
        aHOS[1] = self.extend(phim)
        phim = self.f2s(phim)  # result unused in the reference too
        phizm = [None] + [np.zeros((p.Nd1, p.Nd2)) for _ in range(M)]
        for iHOS in range(1, M + 1):
            phimext = np.zeros((p.Nd1, p.Nd2))
            for j in range(iHOS, 1, -1):
                jm1 = j - 1
                j2 = iHOS - jm1
                phizi = self.f2s_big(aHOS[j2] * kth_ext[:, :, j - 1])
                phimext = phimext - phizi * etapm[j]
                phizm[iHOS] = phizm[iHOS] + phizi * etapm[jm1]
            phizi = self.f2s_big(aHOS[iHOS] * kth_ext[:, :, 0])
            phimext = phimext - phizi * etapm[1]
            phizm[iHOS] = phizm[iHOS] + phizi
            if iHOS > 1:
                phizm[iHOS] = self.dealias(iHOS, phizm[iHOS])
            if iHOS < M:
                phimext = self.dealias(iHOS + 1, phimext)
                aHOS[iHOS + 1] = self.s2f_big(phimext)
        phizMm2 = np.zeros((p.Nd1, p.Nd2))
        phizm2 = [None] + [np.zeros((p.Nd1, p.Nd2)) for _ in range(M)]
        if M >= 2:
            for j in range(1, M - 1):
                phizMm2 = phizMm2 + phizm[j]
                phizm2[j] = phizm[j].copy()
            phizMm1 = phizMm2 + phizm[M - 1]
            phiz = self.filter_ext(phizMm1.copy()) + self.filter_ext(phizm[M].copy())
        else:
            phizMm1 = np.zeros((p.Nd1, p.Nd2))
            phiz = self.filter_ext(phizm[1].copy())
        # phiz2Mm2, read only by the SWENSE output (resol_HOS.f90:807-823; mHOSm3div2,
        # mHOSdiv2m1 of common_vars.f90:49); phizm2(j+1) accumulates in place as there
        phiz2Mm2 = np.zeros((p.Nd1, p.Nd2))
        if M >= 4:
            m3d2 = max(1, (M - 3) // 2)
            d2m1 = max(1, M // 2 - 1)
            for j in range(1, m3d2 + 1):
                phiz2Mm2 = phiz2Mm2 + 2.0 * phizm2[j] * phizm2[M - 2 - j]
                phizm2[j + 1] = phizm2[j + 1] + phizm2[j]
            phiz2Mm2 = phiz2Mm2 + phizm2[d2m1] ** 2
        W1 = self.filter_ext(phizm[1].copy())
        geta2phiz = np.zeros((p.Nd1, p.Nd2))
        for j in range(1, M - 2):
            geta2phiz = geta2phiz + self.dealias(j + 2, gradeta2 * phizm[j])
        if M >= 3:
            geta2phiz = geta2phiz + self.dealias(M, gradeta2 * phizm[max(M - 2, 1)])
        geta2phiz2 = np.zeros((p.Nd1, p.Nd2))
        temp1 = np.zeros((p.Nd1, p.Nd2))
        for j in range(1, (M - 2) // 2 + 1):
            temp2 = self.dealias(2 * j, phizm[j] * phizm[j])
            temp1 = temp1 + temp2
            geta2phiz2 = geta2phiz2 + self.dealias(2 * j + 2, gradeta2 * temp2)
        for j in range(2, M - 2):
            for q in range(1, min(M - 2 - j, j - 1) + 1):
                temp2 = self.dealias(j + q, 2.0 * phizm[j] * phizm[q])
                temp1 = temp1 + temp2
                geta2phiz2 = geta2phiz2 + self.dealias(j + q + 2, gradeta2 * temp2)
        if M >= 2:
            temp2 = self.dealias(2 * (M // 2), phizm[M // 2] * phizm[M // 2])
            temp1 = temp1 + temp2
        for j in range(2, M):
            for q in range(max(0, min(M - 2 - j, j - 1)) + 1, min(M - j, j - 1) + 1):
                temp2 = self.dealias(j + q, 2.0 * phizm[j] * phizm[q])
                temp1 = temp1 + temp2
        phiz2 = temp1
        self.last_aHOS = aHOS  # modal potentials of orders 1..M on the extended grid
        return dict(phiz=phiz, W1=W1, phizMm1=phizMm1, geta2phiz=geta2phiz,
                    geta2phiz2=geta2phiz2, phiz2=phiz2, phizMm2=phizMm2, phiz2Mm2=phiz2Mm2)

    # ----- wavemaker first order (resol_wmkr.f90:109-210) ----------------
    SECOND_FIELDS = ("phi", "phit", "phix", "phiy", "phiz", "phizz", "phixt", "phiyt", "phizt", "phixx",
                     "phixz", "phiyz")

    def solve_sweet_1(self, a1st, eta1st, a_add_1):
        p, tb = self.p, self.tb
        self.alpha, self.alphat, self.alphatt = ramp(self.time, p)
        self.pos1st, self.dpos1stdt, self.d2pos1stdt2 = pos_1st(self.time, p, tb)
        al, alt, altt = self.alpha, self.alphat, self.alphatt
        n3, n3a = p.n3, p.n3_add
        d1 = np.zeros((n3a, p.n2))
        dr = np.zeros((n3a, p.n2))
        d1[:n3] = -al * self.d2pos1stdt2 - 2.0 * alt * self.dpos1stdt - altt * self.pos1st
        dr[:n3] = -alt * self.dpos1stdt - altt * self.pos1st
        dr = match_surface_wmk(dr, n3, tb.delz)
        d1 = match_surface_wmk(d1, n3, tb.delz)
        d1 = self.FA.space_2_fourier(d1)
        dr = self.FA.space_2_fourier(dr)
        if p.coeffilt[2] < 1.0:
            d1 = filter_modes(d1, p.coeffilt[2], 0)
            dr = filter_modes(dr, p.coeffilt[2], 0)
        d1 = d1 / tb.k_add_thk_add
        dr = dr / tb.k_add_thk_add
        second = p.i_wmk >= 2
        # phiadd_1st_SWEET at z=0 (resol_wmkr.f90:446-574); the RK path needs phit and phiz,
        # the second-order wavemaker also its derivatives
        if self.wmk_direct and not second:
            sl = phiadd_sl_bruteforce(a_add_1, d1, None, tb)
        else:
            sl = phiadd_sl(a_add_1, d1, None, tb, need=self.SECOND_FIELDS if second else ("phit", "phiz"))
        # phidphiz (resol_wmkr.f90:583-636)
        phiz1st = self.f2s(a1st * tb.kth[:, :, 0]) + sl["phiz"]
        da1 = -eta1st - sl["phit"] - p.iabsnb * tb.nu * phiz1st
        deta1 = phiz1st
        da1 = self.s2f(da1)
        if second:
            self.first, da1 = self.d1st_for_2nd(a1st, da1, sl, a_add_1, d1, phiz1st)
        return da1, deta1, d1, dr

    # ----- first-order fields for the second-order wavemaker (resol_wmkr.f90:645-793) ----
    def d1st_for_2nd(self, a1, da1, sl, a_add_1, d1, phiz1st):
        """Returns (fields, da1): first-order potential derivatives on the paddle (n3 x n2)
        and on the free surface (n1 x n2), and da1 after the optional x/y filters."""
        p, tb = self.p, self.tb
        a1 = a1.copy()
        if p.coeffilt[0] < 1.0:
            a1 = filter_modes(a1, p.coeffilt[0], 0)
            da1 = filter_modes(da1, p.coeffilt[0], 0)
        if p.n2 != 1 and p.coeffilt[1] < 1.0:
            a1 = filter_modes(a1, p.coeffilt[1], 1)
            da1 = filter_modes(da1, p.coeffilt[1], 1)
        f = self.paddle_fields(a1, da1, a_add_1, d1)
        kyb = tb.ky[None, :]
        k = tb.k
        # free-surface first-order fields (:711-777)
        kx = tb.kx[:, None]
        kth1 = tb.kth[:, :, 0]
        f["phix1st"] = self.f2s(-a1 * kx, "sin", "cos") + sl["phix"]
        f["phiy1st"] = self.f2s(-a1 * kyb, "cos", "sin") + sl["phiy"]
        f["phizt1st"] = self.f2s(da1 * kth1) + sl["phizt"]
        f["phixx1st"] = self.f2s(-a1 * kx * kx) + sl["phixx"]
        f["phizz1st"] = self.f2s(a1 * k * k) + sl["phizz"]
        phixz1st = self.f2s(-a1 * kx * kth1, "sin", "cos") + sl["phixz"]
        phiyz1st = self.f2s(-a1 * kyb * kth1, "cos", "sin") + sl["phiyz"]
        f["etax1st"] = (self.f2s(da1 * kx, "sin", "cos") - sl["phixt"]
                        - p.iabsnb * (tb.dnux * phiz1st + tb.nu * phixz1st))
        f["etay1st"] = (self.f2s(da1 * kyb, "cos", "sin") - sl["phiyt"]
                        - p.iabsnb * (tb.dnuy * phiz1st + tb.nu * phiyz1st))
        f["phiz1st"] = phiz1st
        return f, da1

    def paddle_fields(self, a_fs, da_fs, a_add, d_add):
        """phi_xx, phi_y, phi_z and their time derivatives on the paddle (n3 x n2) of one order
        of the potential: the additional part (phiadd_1st_SWEET :535-566, wmk_3rdorder
        :982-1012) plus the free-surface modes at x = 0 (d1st_for_2nd :686-709, wmk_3rdorder
        :1045-1073; tables HOS-NWT.f90:342-355), y-transformed.  Keys axx, axxt, ay, ayt, az, azt."""
        p, tb = self.p, self.tb
        n3 = p.n3
        FA = self.FA
        ka2 = tb.k_add ** 2
        kyb = tb.ky[None, :]
        kxa = tb.kx_add[:, None]
        add = dict(
            axx=FA.fourier_2_space(a_add * ka2, "cos", "cos")[:n3],
            axxt=FA.fourier_2_space(d_add * ka2, "cos", "cos")[:n3],
            ay=FA.fourier_2_space(-a_add * kyb, "cos", "sin")[:n3],
            ayt=FA.fourier_2_space(-d_add * kyb, "cos", "sin")[:n3],
            az=FA.fourier_2_space(-a_add * kxa, "sin", "cos")[:n3],
            azt=FA.fourier_2_space(-d_add * kxa, "sin", "cos")[:n3],
        )
        k = tb.k
        kx2 = (tb.kx ** 2)[:, None]
        sums = {key: np.zeros((n3, p.n2)) for key in ("axx", "axxt", "ay", "ayt", "az", "azt")}
        for i3 in range(n3):
            cxa = i3 * tb.delz  # cx_add (HOS-NWT.f90:242)
            ch = np.cosh(k * cxa) / np.cosh(k)
            sh = k * np.sinh(k * cxa) / np.cosh(k)
            sums["axx"][i3] = -(a_fs * ch * kx2).sum(axis=0)
            sums["axxt"][i3] = -(da_fs * ch * kx2).sum(axis=0)
            sums["ay"][i3] = -(a_fs * ch * kyb).sum(axis=0)
            sums["ayt"][i3] = -(da_fs * ch * kyb).sum(axis=0)
            sums["az"][i3] = (a_fs * sh).sum(axis=0)
            sums["azt"][i3] = (da_fs * sh).sum(axis=0)
        f = {}
        for key, cy in (("axx", "cos"), ("axxt", "cos"), ("ay", "sin"), ("ayt", "sin"), ("az", "cos"),
                        ("azt", "cos")):
            f[key] = (fourier_2_space_y_axis(sums[key], cy, 1) if p.n2 != 1 else sums[key]) + add[key]
        return f

    # ----- wavemaker third order (resol_wmkr.f90:338-437, :942-1115) --------------
    def wmk_3rdorder(self, a2nd, da2nd, a_add_1, d1, a_add_2, d2):
        """First-order third derivatives (additional part only: the free-surface part is a sine
        series in x, zero on the paddle) and second-order paddle fields (wmk_3rdorder
        :982-1098).  Keys xxx, xxxt, xy, xyt, xz, xzt and the paddle_fields keys of order 2."""
        p, tb = self.p, self.tb
        n3 = p.n3
        FA = self.FA
        kt = tb.k_add_thk_add
        k2 = tb.k_add ** 2
        kyb = tb.ky[None, :]
        kxa = tb.kx_add[:, None]
        first = dict(
            xxx=FA.fourier_2_space(-a_add_1 * k2 * kt, "cos", "cos")[:n3],
            xxxt=FA.fourier_2_space(-d1 * k2 * kt, "cos", "cos")[:n3],
            xy=FA.fourier_2_space(a_add_1 * kyb * kt, "cos", "sin")[:n3],
            xyt=FA.fourier_2_space(d1 * kyb * kt, "cos", "sin")[:n3],
            xz=FA.fourier_2_space(a_add_1 * kxa * kt, "sin", "cos")[:n3],
            xzt=FA.fourier_2_space(d1 * kxa * kt, "sin", "cos")[:n3],
        )
        second = self.paddle_fields(a2nd, da2nd, a_add_2, d2)
        return first, second

    def solve_sweet_3(self, a2nd, da2nd, a_add_1, d1, a_add_2, d2):
        """Third-order paddle condition and modes (solveHOS_SWEET_3, resol_wmkr.f90:338-437).
        Returns da3rd_add (n3_add x n2 modal).  phiadd_3rd (:872-933) is not formed: the
        right-hand side reads none of its fields."""
        p, tb = self.p, self.tb
        n3, n3a = p.n3, p.n3_add
        f1, f2 = self.wmk_3rdorder(a2nd, da2nd, a_add_1, d1, a_add_2, d2)
        al, alt = self.alpha, self.alphat
        X, Xt = self.pos1st, self.dpos1stdt
        Xy, Xyt, Xz, Xzt = dpos_1st(self.time, p, tb)
        t_al = -al * (-Xt * f2["axx"] - X * f2["axxt"] + Xyt * f2["ay"] + Xy * f2["ayt"]
                      + Xzt * f2["az"] + Xz * f2["azt"])
        t_alt = -alt * (-X * f2["axx"] + Xy * f2["ay"] + Xz * f2["az"])
        t_al2 = -al ** 2 * (-(X ** 2) / 2.0 * f1["xxxt"] - Xt * X * f1["xxx"] + Xt * Xy * f1["xy"]
                            + X * Xy * f1["xyt"] + X * Xyt * f1["xy"] + Xt * Xz * f1["xz"]
                            + X * Xz * f1["xzt"] + X * Xzt * f1["xz"])
        t_alalt = -al * alt * (-(X ** 2) * f1["xxx"] + 2.0 * X * Xy * f1["xy"] + 2.0 * X * Xz * f1["xz"])
        d3 = np.zeros((n3a, p.n2))
        d3[:n3] = t_al + t_alt + t_al2 + t_alalt
        d3 = match_surface_wmk(d3, n3, tb.delz)
        d3 = self.FA.space_2_fourier(d3)
        if p.coeffilt[2] < 1.0:
            d3 = filter_modes(d3, p.coeffilt[2], 0)
        return d3 / tb.k_add_thk_add

    # ----- wavemaker second order (resol_wmkr.f90:219-329) -----------------
    def solve_sweet_2(self, a2nd, eta2nd, eta1st, a2nd_add):
        """Returns (da2nd modal, deta2nd physical, da2nd_add modal, phiadd_2nd fields)."""
        p, tb = self.p, self.tb
        f = self.first
        n3, n3a = p.n3, p.n3_add
        al, alt = self.alpha, self.alphat
        dy, dyt, dz, dzt = dpos_1st(self.time, p, tb)
        pos, post = self.pos1st, self.dpos1stdt
        noramp = -pos * f["axx"] + dy * f["ay"] + dz * f["az"]
        d2 = np.zeros((n3a, p.n2))
        d2[:n3] = (-al * (-pos * f["axxt"] - post * f["axx"] + dy * f["ayt"] + dyt * f["ay"]
                          + dz * f["azt"] + dzt * f["az"]) - alt * noramp)
        d2 = match_surface_wmk(d2, n3, tb.delz)
        d2 = self.FA.space_2_fourier(d2)
        if p.coeffilt[2] < 1.0:
            d2 = filter_modes(d2, p.coeffilt[2], 0)
        d2 = d2 / tb.k_add_thk_add
        # phiadd_2nd (resol_wmkr.f90:802-864): the z = 0 sums of phiadd_SL
        sl2 = phiadd_sl(a2nd_add, d2, None, tb, need=("phi", "phix", "phiy", "phit", "phiz"))
        # phidphiz for the second order, then the free-surface conditions (:304-318)
        phiz2nd = self.f2s(a2nd * tb.kth[:, :, 0]) + sl2["phiz"]
        deta2 = phiz2nd - f["phix1st"] * f["etax1st"] - f["phiy1st"] * f["etay1st"] + eta1st * f["phizz1st"]
        da2 = (-eta2nd - sl2["phit"] - 0.5 * (f["phix1st"] ** 2 + f["phiy1st"] ** 2 + f["phiz1st"] ** 2)
               - eta1st * f["phizt1st"] - p.iabsnb * tb.nu * deta2)
        return self.s2f(da2), deta2, d2, sl2

    # ----- solveHOS2 (resol_HOS.f90:270-656) ------------------------------
    def solve_hos2(self, a1st, eta1st, G2, G1, a_add_1, delta_t, a2nd=None, eta2nd=None, a_add_2=None,
                   a_add_3=None):
        """One RHS.  Returns (dG1, dG2, da1st, deta1st, da_add_1, da_add_ramp); with i_wmk >= 2
        the second-order derivatives (da2nd, deta2nd, da_add_2) are left in ``self.second``,
        and with i_wmk = 3 da_add_3 in ``self.third``."""
        p, tb = self.p, self.tb
        self.second = None
        self.third = None
        self.n_rhs += 1
        etark, phisrk = self.gtof(G1, G2, delta_t)
        om = tb.omega.copy()
        om[0, 0] = 1.0
        ph = phisrk / om
        ph[0, 0] = phisrk[0, 0]
        phisrk = ph
        NL1 = self.f2s(etark)
        etark = etark.copy()
        etark[0, 0] = self._volume_mode(NL1[0, 0])
        self.last_eta_modes = etark.copy()  # modal eta of this evaluation (after the volume fix)
        self.last_phis_modes = ph.copy()    # modal phis / omega-scaled potential of order 1
        etark = self.f2s(etark)
        phisrk = self.f2s(phisrk)
        etapm, etax, etay, gradeta2, phisx, phisy = self.phisxy_etaxy(phisrk, etark)
        self.etax, self.etay = etax, etay
        h = self.hosphis_modes(phisrk, etapm, gradeta2)
        # module globals of this evaluation read by the SWENSE output (velocities.f90:501-784)
        self.mod = dict(etapm=etapm, etax=etax, etay=etay, phisx=phisx, phisy=phisy, h=h)
        detarkext = -phisx * etax - phisy * etay
        dphisrkext = -phisx * phisx - phisy * phisy
        detarkext = self.dealias(2, detarkext)
        dphisrkext = self.dealias(2, dphisrkext)
        detarkext = detarkext + h["geta2phiz"]
        dphisrkext = dphisrkext + h["geta2phiz2"] + h["phiz2"]
        detark = self.filter_ext(detarkext)
        dphisrk = self.filter_ext(dphisrkext)
        n1, n2, n3a = p.n1, p.n2, p.n3_add
        if p.has_wmk:
            da1, deta1, d1, dr = self.solve_sweet_1(a1st, eta1st, a_add_1)
            a_tot, d_tot = a_add_1, d1
            if p.i_wmk >= 2:
                z1 = np.zeros((n1, n2))
                a2 = z1 if a2nd is None else a2nd
                e2 = z1 if eta2nd is None else eta2nd
                aa2 = np.zeros((n3a, n2)) if a_add_2 is None else a_add_2
                da2, deta2, d2, _ = self.solve_sweet_2(a2, e2, eta1st, aa2)
                self.second = (da2, deta2, d2)
                # total additional potential (resol_HOS.f90:498-502)
                a_tot, d_tot = a_add_1 + aa2, d1 + d2
                if p.i_wmk >= 3:
                    aa3 = np.zeros((n3a, n2)) if a_add_3 is None else a_add_3
                    d3 = self.solve_sweet_3(a2, da2, a_add_1, d1, aa2, d2)
                    self.third = d3
                    a_tot, d_tot = a_tot + aa3, d_tot + d3
            if self.wmk_direct:
                sl = phiadd_sl_bruteforce(a_tot, d_tot, etark, tb)
            else:
                sl = phiadd_sl(a_tot, d_tot, etark, tb)
            phi_add, phix_add, phiy_add = sl["phi"], sl["phix"], sl["phiy"]
            phiz_add, phit_add = sl["phiz"], sl["phit"]
        elif p.has_dip:
            # ramp + phi_dipoles at z = eta (resol_HOS.f90:494-496); phiy_add keeps its zero
            da1 = np.zeros((n1, n2))
            deta1 = np.zeros((n1, n2))
            d1 = np.zeros((n3a, n2))
            dr = np.zeros((n3a, n2))
            self.alpha, self.alphat, self.alphatt = ramp(self.time, p)
            phi_add, phit_add, phix_add, phiz_add = phi_dipoles(etark, self.time, self.alpha, self.alphat, p,
                                                                tb.cx[:, None] + 0.0 * etark)
            phiy_add = np.zeros((n1, n2))
        else:
            da1 = np.zeros((n1, n2))
            deta1 = np.zeros((n1, n2))
            d1 = np.zeros((n3a, n2))
            dr = np.zeros((n3a, n2))
            phi_add = phix_add = phiy_add = phiz_add = phit_add = np.zeros((n1, n2))
        self.phi_add = phi_add
        # phiz_etax_ext / phiz_etay_ext are filtered only for PRESENT(c) (resol_HOS.f90:522-530)
        self.filter_ext(etax * h["phizMm1"], "sin", "cos")
        self.filter_ext(etay * h["phizMm1"], "cos", "sin")
        phix_add = self.s2f(phix_add)
        phiy_add = self.s2f(phiy_add)
        phiz_add = self.s2f(phiz_add)
        phix_add_ext = self.extend(phix_add)
        phiy_add_ext = self.extend(phiy_add)
        phiz_add_ext = self.extend(phiz_add)
        phix_add = self.f2s(phix_add)
        phiy_add = self.f2s(phiy_add)
        phiz_add = self.f2s(phiz_add)
        self.mod.update(phix_add=phix_add, phiy_add=phiy_add, phiz_add=phiz_add, phit_add=phit_add)
        phix_add_ext = self.f2s_big(phix_add_ext)
        phiy_add_ext = self.f2s_big(phiy_add_ext)
        phiz_add_ext = self.f2s_big(phiz_add_ext)
        gg_ext = etax * phix_add_ext + etay * phiy_add_ext
        gp_ext = (phisx * phix_add_ext + phisy * phiy_add_ext
                  + 0.5 * (phix_add_ext * phix_add_ext + phiy_add_ext * phiy_add_ext
                           + phiz_add_ext * phiz_add_ext))
        gradeta_gradphi = self.filter_ext(gg_ext)
        gradphis_gradphi = self.filter_ext(gp_ext)
        if p.mHOS == 1:
            dphin = h["phiz"] + phiz_add
            detark = dphin - h["W1"]
            dphisrk = -phit_add - tb.nu * dphin
        else:
            dphin = detark + h["phiz"] + phiz_add - gradeta_gradphi
            detark = dphin - h["W1"]
            dphisrk = 0.5 * dphisrk - phit_add - tb.nu * dphin - gradphis_gradphi
        if p.coeffilt[0] < 1.0:
            detark = self.filtering(detark, p.coeffilt[0], 0)
            dphisrk = self.filtering(dphisrk, p.coeffilt[0], 0)
        if p.n2 != 1 and p.coeffilt[1] < 1.0:
            detark = self.filtering(detark, p.coeffilt[1], 1)
            dphisrk = self.filtering(dphisrk, p.coeffilt[1], 1)
        NL1 = self.s2f(detark)
        NL2 = self.s2f(dphisrk)
        dphisrk_1 = NL2[0, 0]
        NL2 = NL2 * tb.omega
        dG1, dG2 = self.expon(NL1, NL2, delta_t)
        dG2[0, 0] = dphisrk_1
        return dG1, dG2, da1, deta1, d1, dr

    def _volume_mode(self, nl11):
        """Mean-volume mode from the paddle-corner value (resol_HOS.f90:358-364)."""
        p = self.p
        x = self.alpha * self.pos1st[p.n3 - 1, 0] / p.xlen
        if p.igeom == 2:
            return 0.5 * self.alpha * (1 - p.d_hinge) * self.pos1st[p.n3 - 1, 0] / p.xlen + nl11 * x
        if p.igeom == 1:
            return x + nl11 * x
        return 0.0

    # ----- filtering_x / filtering_y (filtering.f90:37-134) -----------------
    def filtering(self, f, coeff, axis):
        """Filter a physical base-grid field: S2F, filtering_base, F2S."""
        return self.f2s(filter_modes(self.s2f(f), coeff, axis))

    # ----- check_slope (initial_condition.f90:687-716) ------------------------
    def check_slope(self, st: State, etax_ext, etay_ext, threshold=8.0, coeff=0.7):
        """Filter eta and phis when a slope of the extended-grid eta exceeds the threshold.

        Returns (state, filtered_x, filtered_y); etax_ext / etay_ext are the
        extended-grid slopes of the last right-hand side.
        """
        sx = float(np.max(np.abs(etax_ext))) if etax_ext is not None else 0.0
        sy = float(np.max(np.abs(etay_ext))) if etay_ext is not None else 0.0
        st = st.copy()
        fx = sx > threshold
        fy = sy > threshold and self.p.n2 != 1
        if fx:
            st.phis = self.filtering(st.phis, coeff, 0)
            st.eta = self.filtering(st.eta, coeff, 0)
        if fy:
            st.phis = self.filtering(st.phis, coeff, 1)
            st.eta = self.filtering(st.eta, coeff, 1)
        return st, fx, fy

    # ----- runge4_adapt2 (runge_kutta.f90:81-360) ----------------------------
    A = np.array([[0, 0, 0, 0, 0], [0.5, 0, 0, 0, 0], [0, 0.5, 0, 0, 0], [0, 0, 1.0, 0, 0],
                  [-1.0, 2.0, 0, 0, 0]])
    B = np.array([1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 0.0])
    C = np.array([0.0, 0.5, 0.5, 1.0, 1.0])
    E = np.array([1.0 / 6.0, 2.0 / 3.0, 0.0, 0.0, 1.0 / 6.0])
    RK_P = 4

    def runge4_adapt2(self, t_o, st: State, h):
        """One RK4(3) attempt; returns (new State, err, deta) without committing."""
        p, tb = self.p, self.tb
        G1 = self.s2f(st.eta)
        G2 = self.s2f(st.phis)
        phisrk0 = G2[0, 0]
        G2 = G2 * tb.omega
        s = 5
        kG1, kG2, ka1, ke1, kad1, kadr = ([None] * s for _ in range(6))
        # second- and third-order wavemaker state (zero below i_wmk = 2, 3)
        ka2, ke2, kad2, kad3 = ([None] * s for _ in range(4))

        def second_out():
            if self.second is None:
                return np.zeros_like(st.a2nd), np.zeros_like(st.eta2nd), np.zeros_like(st.a2nd_add)
            return self.second

        def third_out():
            return np.zeros_like(st.a3rd_add) if self.third is None else self.third

        self.time = t_o
        out = self.solve_hos2(st.a1st, st.eta1st, G2, G1, st.a1st_add, h * self.C[0], st.a2nd, st.eta2nd,
                              st.a2nd_add, st.a3rd_add)
        kG1[0], kG2[0], ka1[0], ke1[0], kad1[0], kadr[0] = out
        ka2[0], ke2[0], kad2[0] = second_out()
        kad3[0] = third_out()
        kG2[0][0, 0] = 0.0
        for jl in range(1, s):
            def comb(arr_list, base):
                acc = np.zeros_like(base)
                for ns in range(jl):
                    acc = acc + arr_list[ns] * self.A[jl, ns]
                return base + h * acc
            G1m = comb(kG1, G1)
            G2m = comb(kG2, G2)
            aad = comb(kad1, st.a1st_add)
            a1m = comb(ka1, st.a1st)
            e1m = comb(ke1, st.eta1st)
            a2m = comb(ka2, st.a2nd)
            e2m = comb(ke2, st.eta2nd)
            aad2 = comb(kad2, st.a2nd_add)
            aad3 = comb(kad3, st.a3rd_add)
            self.time = t_o + h * self.C[jl]
            out = self.solve_hos2(a1m, e1m, G2m, G1m, aad, h * self.C[jl], a2m, e2m, aad2, aad3)
            kG1[jl], kG2[jl], ka1[jl], ke1[jl], kad1[jl], kadr[jl] = out
            ka2[jl], ke2[jl], kad2[jl] = second_out()
            kad3[jl] = third_out()
            kG2[jl][0, 0] = 0.0

        def bsum(arr_list, w):
            acc = np.zeros_like(arr_list[0])
            for j in range(s):
                acc = acc + arr_list[j] * w[j]
            return acc
        dG1m, dG2m = bsum(kG1, self.B), bsum(kG2, self.B)
        dad1, dadr = bsum(kad1, self.B), bsum(kadr, self.B)
        da1, de1 = bsum(ka1, self.B), bsum(ke1, self.B)
        er1, er2, era = bsum(kG1, self.E), bsum(kG2, self.E), bsum(kad1, self.E)
        dad2, da2, de2 = bsum(kad2, self.B), bsum(ka2, self.B), bsum(ke2, self.B)
        era2 = bsum(kad2, self.E)
        dad3, era3 = bsum(kad3, self.B), bsum(kad3, self.E)
        G1 = G1 + h * dG1m
        G2 = G2 + h * dG2m
        a_add_1 = st.a1st_add + h * dad1
        a_add_r = st.a_add_ramp + h * dadr
        a1st = st.a1st + h * da1
        eta1st = st.eta1st + h * de1
        a_add_2 = st.a2nd_add + h * dad2
        a2nd = st.a2nd + h * da2
        eta2nd = st.eta2nd + h * de2
        a_add_3 = st.a3rd_add + h * dad3
        er1 = np.abs((er1 - dG1m) * h) / (np.max(np.abs(G1)) + TINY)
        er2 = np.abs((er2 - dG2m) * h) / (np.max(np.abs(G2)) + TINY)
        era = np.abs((era - dad1) * h) / (np.max(np.abs(a_add_1)) + TINY)
        era2 = np.abs((era2 - dad2) * h) / (np.max(np.abs(a_add_2)) + TINY)
        era3 = np.abs((era3 - dad3) * h) / (np.max(np.abs(a_add_3)) + TINY)
        # a_add_3 is identically zero below a third-order wavemaker, and so is a_add_2 below
        # a second-order one: their error terms are 0/(0+tiny) = 0 (runge_kutta.f90:273-279)
        err = max(er1.max(), er2.max(), era.max(), era2.max(), era3.max())
        etark, phisrk = self.gtof(G1, G2, h)
        deta, _ = self.gtof(kG1[s - 1], kG2[s - 1], h)
        om = tb.omega.copy()
        om[0, 0] = 1.0
        ph = phisrk / om
        ph[0, 0] = phisrk0
        d = deta + ph * tb.omega ** 2
        d[0, 0] = deta[0, 0]
        modesFS = self.f2s(etark)
        etark = etark.copy()
        etark[0, 0] = self._volume_mode(modesFS[0, 0])
        self.modesFS = etark.copy()
        new = State(self.f2s(etark), self.f2s(ph), a_add_1, a_add_r, a1st, eta1st, a_add_2, a2nd, eta2nd,
                    a_add_3)
        deta = self.f2s(d)
        return new, err, deta

    # ----- initial state (initial_condition.f90:314-482) -------------------------
    def initialize(self, eta0=None, phis0=None):
        p, tb = self.p, self.tb
        n1, n2 = p.n1, p.n2
        eta = np.zeros((n1, n2)) if eta0 is None else np.array(eta0, float)
        phis = np.zeros((n1, n2)) if phis0 is None else np.array(phis0, float)
        z = np.zeros((n1, n2))
        za = np.zeros((p.n3_add, n2))
        self.time = 0.0
        if p.has_wmk:
            self.alpha, self.alphat, self.alphatt = ramp(0.0, p)
            self.pos1st, self.dpos1stdt, self.d2pos1stdt2 = pos_1st(0.0, p, tb)
            a = np.zeros((p.n3_add, n2))
            a[:p.n3] = -self.alpha * self.dpos1stdt - self.alphat * self.pos1st
            a = match_surface_wmk(a, p.n3, tb.delz)
            a = self.FA.space_2_fourier(a) / tb.k_add_thk_add
            da1, deta1, d1, dr = self.solve_sweet_1(z, z, a)
            sl = phiadd_sl(a, d1, eta, tb, need=("phi",))
            self.phi_add = sl["phi"]
            return State(eta, phis, a, a.copy(), z.copy(), z.copy())
        if p.has_dip:
            # solveHOS at t = 0 (initial_condition.f90:467-478): the dipole potential at rest
            self.alpha, self.alphat, self.alphatt = ramp(0.0, p)
            self.phi_add = phi_dipoles(eta, 0.0, self.alpha, self.alphat, p, tb.cx[:, None] + 0.0 * eta)[0]
        return State(eta, phis, za, za.copy(), z.copy(), z.copy())

    # ----- SWENSE output (velocities.f90:40-784, runge_kutta.f90:327-333) ------------
    VEL_KINDS = (("sin", "cos"), ("cos", "sin"), ("cos", "cos"), ("cos", "cos"))  # x, y, z, t

    def init_velocity_comp(self, t, h, G1, G2, eta_l, phis_l, a1st, eta1st, a_add, a2nd=None, eta2nd=None):
        """init_velocity_comp (velocities.f90:501-784) for a wavemaker tank (igeom 1, 2): the
        right-hand side re-evaluated at the state (G1, G2) of an attempt of step h ending at t,
        the full time derivatives, and the free-surface velocities and phi_t at z = eta
        (phimx, phimy, phimz, phimt; n1 x n2, physical).  Also sets modesFSt, modesadd and
        modesaddt.  a_add: the attempt's paddle modes (a_add_1, a_add_2, a_add_3; None = 0)."""
        p, tb = self.p, self.tb
        if not p.has_wmk:
            raise ValueError("the SWENSE output is restated for wavemaker tanks (igeom 1, 2)")
        self.time = t
        a1, a2, a3 = (np.zeros((p.n3_add, p.n2)) if a is None else a for a in a_add)
        z = np.zeros((p.n1, p.n2))
        a2nd = z if a2nd is None else a2nd
        eta2nd = z if eta2nd is None else eta2nd
        dG1, dG2, da1, deta1, d1, dr = self.solve_hos2(a1st, eta1st, G2, G1, a1, h,
                                                       a2nd=a2nd if p.i_wmk >= 2 else None,
                                                       eta2nd=eta2nd if p.i_wmk >= 2 else None,
                                                       a_add_2=a2 if p.i_wmk >= 2 else None,
                                                       a_add_3=a3 if p.i_wmk >= 3 else None)
        d2 = self.second[2] if self.second is not None else np.zeros_like(d1)
        d3 = self.third if self.third is not None else np.zeros_like(d1)
        m = self.mod
        hm = m["h"]
        dG2 = dG2.copy()
        dphis3_1 = dG2[0, 0]
        dG2[0, 0] = 0.0
        eta3 = self.s2f(eta_l)
        phis3 = self.s2f(phis_l)
        deta3, dphis3 = self.gtof(dG1, dG2, h)
        om = tb.omega
        d3p = dphis3.copy()
        de3 = deta3.copy()
        mask = np.ones_like(om, bool)
        mask[0, 0] = False
        d3p[mask] = dphis3[mask] / om[mask] - eta3[mask]
        de3[mask] = deta3[mask] + phis3[mask] * om[mask] ** 2
        d3p[0, 0] = dphis3_1 - eta3[0, 0]
        self.modesFSt = de3.copy()
        dphis3 = self.f2s(d3p)
        self.modesadd = a1 + a2 + a3
        self.modesaddt = d1 + d2 + d3
        etax, etay, phisx, phisy = m["etax"], m["etay"], m["phisx"], m["phisy"]
        etax2 = self.dealias(2, etax * etax)
        etay2 = self.dealias(2, etay * etay)
        phiz_etax = self.dealias(2, etax * hm["phizMm1"], "sin", "cos")
        phiz_etay = self.dealias(2, etay * hm["phizMm1"], "cos", "sin")
        pm2_etax = self.dealias(2, etax * hm["phizMm2"], "sin", "cos")
        pm2_etay = self.dealias(2, etay * hm["phizMm2"], "cos", "sin")
        del etax2, etay2  # they feed only the pressure check of the reference
        pxa, pya, pza = (self.s2f(m[k]) for k in ("phix_add", "phiy_add", "phiz_add"))
        phimz = self.s2f(hm["phiz"])
        pxa_e, pya_e, pza_e, pz_e = (self.f2s_big(self.extend(a)) for a in (pxa, pya, pza, phimz))
        phimz = self.f2s(phimz)
        phimx_ext = phisx - etax * hm["phizMm1"]
        phimy_ext = phisy - etay * hm["phizMm1"]
        detaphiz = (hm["geta2phiz2"] + hm["phiz2"] - phisx * pm2_etax - phisy * pm2_etay - pxa_e * phiz_etax
                    - pya_e * phiz_etay + pza_e * pz_e)
        phimx_ext = self.dealias(2, phimx_ext, "sin", "cos")
        phimy_ext = self.dealias(2, phimy_ext, "cos", "sin")
        phimt = dphis3 - self.filter_ext(detaphiz)
        phimx = self.filter_ext(phimx_ext, "sin", "cos")
        phimy = self.filter_ext(phimy_ext, "cos", "sin")
        return phimx, phimy, phimz, phimt

    def _vel_level(self, aH, vit, etapm_a, etapm_b, iH_dealias_shift):
        """One HOS re-expansion of the four velocity fields (the order loop shared by
        HOSvel_SWEET and HOSvel2_SWEET): aH[f][j] modal potentials by order, vit[f] physical
        extended fields; etapm_a subtracted (the level left), etapm_b added (the level reached,
        HOSvel2_SWEET only)."""
        p, tb = self.p, self.tb
        M = p.mHOS
        for iH in range(1, M + 1):
            pa = [np.zeros((p.Nd1, p.Nd2)) for _ in range(4)]
            pb = [np.zeros((p.Nd1, p.Nd2)) for _ in range(4)]
            for j in list(range(iH, 1, -1)) + [None]:
                if j is None:
                    jj, j2 = 1, iH
                else:
                    jj, j2 = j, iH - (j - 1)
                for f, (cx, cy) in enumerate(self.VEL_KINDS):
                    mult = tb.kth2_ext[:, :, jj - 1] if f == 2 else tb.kth_ext[:, :, jj - 1]
                    phizi = self.f2s_big(aH[f][j2] * mult, cx, cy)
                    pa[f] = pa[f] - phizi * etapm_a[jj]
                    if etapm_b is not None:
                        pb[f] = pb[f] + phizi * etapm_b[jj]
            for f, (cx, cy) in enumerate(self.VEL_KINDS):
                vit[f] = vit[f] + pa[f] + (pb[f] if etapm_b is not None else 0.0)
                vit[f] = self.dealias(iH + iH_dealias_shift, vit[f], cx, cy)
                if iH < M:
                    aH[f][iH + 1] = self.s2f_big(self.dealias(iH + 1, pa[f], cx, cy), cx, cy)
        return vit

    def hosvel2_sweet(self, meta2, t, h, G1, G2, eta_l, phis_l, a1st, eta1st, a_add, a2nd=None, eta2nd=None):
        """HOSvel2_SWEET (velocities.f90:223-492) + HOSvel_SWEET (:40-216): the velocity and
        phi_t modes at z = 0, moved from z = eta in meta2 levels (eta (meta2 - l) / meta2).
        Returns (modesspecx, modesspecy, modesspecz, modesspect), n1 x n2 modal; the coeffilt
        filters of the reference's last step are not restated (coeffilt = 1)."""
        p = self.p
        M = p.mHOS
        if min(p.coeffilt) < 1.0:
            raise ValueError("the SWENSE output is restated without filters")
        phim = self.init_velocity_comp(t, h, G1, G2, eta_l, phis_l, a1st, eta1st, a_add, a2nd, eta2nd)
        etapm2 = [None] + [self.mod["etapm"][j].copy() for j in range(1, M + 1)]
        aH = [[None] * (M + 2) for _ in range(4)]
        vit = []
        for f, (cx, cy) in enumerate(self.VEL_KINDS):
            a = self.extend(self.s2f(phim[f], cx, cy))
            aH[f][1] = a
            vit.append(self.f2s_big(a, cx, cy))
        for ii1 in range(1, meta2):
            etapm = [None] + [etapm2[j].copy() for j in range(1, M + 1)]
            eta2 = eta_l * (meta2 - 1 - (ii1 - 1)) / meta2
            as2 = self.s2f(eta2)
            etapm2 = [None] * (M + 1)
            etapm2[1] = self.f2s_big(self.extend(as2))
            for j in range(1, M):
                etapm2[j + 1] = self.dealias(j, etapm2[j] / (j + 1) * etapm2[1])
            vit = self._vel_level(aH, vit, etapm, etapm2, 0)
            for f, (cx, cy) in enumerate(self.VEL_KINDS):
                aH[f][1] = self.s2f_big(vit[f], cx, cy)
        # HOSvel_SWEET: from the last level to z = 0
        for f, (cx, cy) in enumerate(self.VEL_KINDS):
            aH[f][1] = self.s2f_big(vit[f], cx, cy)
        vit = self._vel_level(aH, vit, etapm2, None, 1)
        out = []
        for f, (cx, cy) in enumerate(self.VEL_KINDS):
            out.append(self.s2f(self.filter_ext(vit[f], cx, cy), cx, cy))
        return tuple(out)

    # ----- volume / energy (vol_energy.f90:36-119) ----------------------------
    def volume(self, eta):
        p, tb = self.p, self.tb
        w2 = np.ones(p.n2)
        if p.n2 > 1:
            w2[0] = w2[-1] = 0.5
        sx = (eta * w2[None, :]).sum(axis=1) if p.n2 > 1 else eta[:, 0]
        w1 = np.ones(p.n1)
        w1[0] = w1[-1] = 0.5
        v = (sx * w1).sum() * tb.delx / p.xlen
        if p.n2 != 1:
            v = v / p.ylen * tb.dely
        return v

    def energy(self, eta, phis, deta):
        p, tb = self.p, self.tb
        def integ(f):
            w2 = np.ones(p.n2)
            if p.n2 > 1:
                w2[0] = w2[-1] = 0.5
            sx = (f * w2[None, :]).sum(axis=1) if p.n2 > 1 else f[:, 0]
            w1 = np.ones(p.n1)
            w1[0] = w1[-1] = 0.5
            v = (sx * w1).sum() * tb.delx / 2.0
            return v * tb.dely if p.n2 != 1 else v
        ep = integ(eta * eta)
        ek = integ(phis * deta)
        return ep, ek, ep + ek


# ---------------------------------------------------------------------------
# adaptive controller (HOS-NWT.f90:498-563)
# ---------------------------------------------------------------------------
def dt_stab(p: Params) -> float:
    """RK4 stability bound dt_rk4 (HOS-NWT.f90:132-176)."""
    c = math.sqrt(8.0 / math.pi ** 2)
    d1 = c * p.xlen / (p.n1 - 1)
    d2 = c * p.ylen / (p.n2 - 1) if p.n2 != 1 else d1 + 1.0
    d3 = c * (2.0 * L_ADD) / (L_ADD * (p.n3 - 1))
    if p.icase == 1 or p.igeom == 3:
        return min(d1, d2)
    if p.n2 == 1:
        return min(d1, d3)
    return min(d1, d2, d3)


def advance(orc: Oracle, st: State, t_cur: float, t_next: float, h_rk: float, toler: float,
            stab: float, delt_out: float, log=None, replay=None):
    """Inner loop DO WHILE(time_cur < time_next) with accept/reject (HOS-NWT.f90:498-571).

    ``replay``: optional list of (h_loc, accept) forcing the decisions (parity replay).
    Returns (state, t_cur, h_rk, n_acc, n_rej, deta).
    """
    n_acc = n_rej = 0
    deta = None
    k = 0
    p = orc.p
    while t_cur < t_next:
        h_loc = h_rk
        if t_next - t_cur < h_rk:
            h_loc = t_next - t_cur
        new, err, d = orc.runge4_adapt2(t_cur, st, h_loc)
        if replay is not None:
            accept = replay[k][1]
        else:
            accept = not (err > toler and h_rk > 0.001 * 0.95 * stab)
        if log is not None:
            log.append((h_loc, accept, err))
        k += 1
        if not accept:
            dt_correc = 0.9 * (err / toler) ** (-(1.0 / orc.RK_P))
            n_rej += 1
            if abs(p.Tramp - (t_cur + h_loc)) <= TINY and p.Tramp > t_cur:
                orc.endramp = 0
        else:
            st = new
            deta = d
            orc.deta = d
            if (p.iramp == 1 and p.Tramp - (t_cur + h_loc) <= TINY and p.Tramp > t_cur
                    and orc.endramp != 1):
                st.a1st_add = st.a1st_add - st.a_add_ramp
                orc.endramp = 1
            t_cur = t_cur + h_loc
            dt_correc = 2.0 if abs(err) < TINY else 0.9 * (err / toler) ** (-(1.0 / orc.RK_P))
            n_acc += 1
        h_rk = h_rk * dt_correc
        h_rk = min(h_rk, 0.95 * stab, delt_out)
        h_rk = max(h_rk, 0.0001 * 0.95 * stab, 0.0001 * delt_out)
    return st, t_cur, h_rk, n_acc, n_rej, deta
