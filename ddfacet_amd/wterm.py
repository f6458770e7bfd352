"""W-projection convolution functions and the prolate-spheroidal taper of one facet.

Host-side builder with the interface of ``ModCF.ClassWTermModified``
(DDFacet/Imager/ModCF.py:297-597): attributes ``Wplanes``, ``WplanesConj`` (lists of complex64
``[S*OverS, S*OverS]`` planes in the reorganised ``[OverS][OverS][S][S]`` layout), ``Sups``,
``ifzfCF`` (the taper on the padded facet grid), ``Cu``, ``Cv``, ``RefWave``, ``wmax``, ``OverS``.

Steps, each following the reference:
  * ``eval_sphe``      Schwab rational approximation (DDFacet/ToolsDir/ModTaper.py:76-111)
  * ``sphe2d``         radial evaluation on an N x N grid over [-1, 1]^2 (ModTaper.py:38-47)
  * ``_fft2/_ifft2``   centred, 1/N^2-normalised transforms (ModCF.py:52-65)
  * ``zero_pad``       centring rule for odd / even outputs (ModCF.py:68-84)
  * ``SpheMachine``    cropped-spectrum taper and its padded inverse (ModCF.py:87-172)
  * ``give_dn``        order-5 least-squares fit of n(l, m) - n0 (ModCF.py:267-295,
                        ModFitPoly2D.py:30-58)
  * ``reorg``          [OverS][OverS][S][S] reorganisation (ModCF.py:364-372)
  * ``ClassWTermModified.init_w``   w-plane construction (ModCF.py:374-597)
The optional ``sups`` argument replaces the SupMax estimate (ModCF.py:389-422) with explicit
per-plane supports (forced odd, as ModCF.py:477-478 does).
"""
import itertools

import numpy as np
import scipy.fftpack
from scipy.interpolate import interp1d

C_LIGHT = 299792458.0

_P = np.array([[8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1],
               [4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2]])
_Q = np.array([[1.0000000e0, 8.212018e-1, 2.078043e-1],
               [1.0000000e0, 9.599102e-1, 2.918724e-1]])


def eval_sphe(nu):
    """Prolate spheroidal approximation of Schwab (1984) on |nu| <= 1, zero outside."""
    nu = np.abs(np.asarray(nu, dtype=np.float64))
    out = np.zeros_like(nu)
    for part, lo, hi, end in ((0, 0.0, 0.75, 0.75), (1, 0.75, 1.0, 1.0)):
        sel = (nu >= lo) & ((nu < hi) if part == 0 else (nu <= hi))
        x = nu[sel]
        nusq = x * x
        d = nusq - end * end
        top = np.full_like(x, _P[part][0])
        dp = d.copy()
        for k in range(1, 5):
            top += _P[part][k] * dp
            dp = dp * d
        bot = np.full_like(x, _Q[part][0])
        dp = d.copy()
        for k in range(1, 3):
            bot += _Q[part][k] * dp
            dp = dp * d
        out[sel] = (1.0 - nusq) * (top / bot)
    return out


def sphe2d(npix):
    x, y = np.mgrid[-1:1:1j * npix, -1:1:1j * npix]
    return eval_sphe(np.sqrt(x ** 2 + y ** 2))


def _fft2(a):
    return scipy.fftpack.fftshift(scipy.fftpack.fft2(scipy.fftpack.ifftshift(a))) / np.float64(a.size)


def _ifft2(a):
    return scipy.fftpack.fftshift(scipy.fftpack.ifft2(scipy.fftpack.ifftshift(a))) * np.float64(a.size)


def zero_pad(a, outshape):
    nx = a.shape[0]
    b = np.zeros((outshape, outshape), dtype=a.dtype)
    off = (outshape - nx) // 2 + (1 if outshape % 2 == 0 else 0)
    b[off:off + nx, off:off + nx] = a
    return b


class SpheMachine:
    def __init__(self, support=11, support_calc=111):
        self.support = support
        xc = support_calc // 2
        cf = np.complex128(sphe2d(support_calc))
        fcf = _fft2(cf)
        h = support // 2
        self.small_fcf = fcf[xc - h:xc + h + 1, xc - h:xc + h + 1].copy()
        self.small_cf = cf

    def make_sphe(self, npix):
        """Taper on an npix grid: inverse of the zero-padded cropped spectrum, negatives -> 1e-10."""
        ifz = _ifft2(zero_pad(self.small_fcf, npix))
        ifz[ifz < 0] = 1e-10
        return self.small_cf, self.small_fcf, ifz


def polyfit2d(x, y, z, order):
    g = np.zeros((x.size, (order + 1) ** 2))
    for k, (i, j) in enumerate(itertools.product(range(order + 1), range(order + 1))):
        g[:, k] = x ** i * y ** j
    m, _, _, _ = np.linalg.lstsq(g, z, rcond=-1)
    return m


def polyval2d(x, y, m):
    order = int(np.sqrt(len(m))) - 1
    z = np.zeros_like(x)
    for a, (i, j) in zip(m, itertools.product(range(order + 1), range(order + 1))):
        z += a * x ** i * y ** j
    return z


def give_dn(l0, m0, rad=1.0, order=4):
    """Fit n(l0+dl, m0+dm) - n(l0, m0) by a 2-D polynomial; return (Cl, Cm, coefs with the
    linear terms zeroed)."""
    npt = 100
    l, m = np.mgrid[l0 - rad:l0 + rad:npt * 1j, m0 - rad:m0 + rad:npt * 1j]
    dl = (l - l0).ravel()
    dm = (m - m0).ravel()
    if np.max(np.abs((dl + l0) ** 2 + (dm + m0) ** 2)) > 1.0:
        return 0.0, 0.0, np.zeros((order + 1) ** 2, np.float32)
    y = np.sqrt(1 - (dl + l0) ** 2 - (dm + m0) ** 2) - np.sqrt(1 - l0 ** 2 - m0 ** 2)
    coef = polyfit2d(dl, dm, y, order).reshape((order + 1, order + 1))
    cl, cm = coef[0, 1], coef[1, 0]
    coef[0, 1] = 0
    coef[1, 0] = 0
    return cl, cm, coef.ravel()


def reorg(a, overs):
    s = a.shape[0] // overs
    b = np.zeros((overs, overs, s, s), dtype=a.dtype)
    for i in range(overs):
        for j in range(overs):
            b[i, j] = a[i::overs, j::overs]
    return b.reshape(a.shape)


class ClassWTermModified:
    """CF tables of one facet (see module docstring).  ``Cell`` in arcsec, ``Npix`` the padded
    facet size, ``lmShift`` the facet centre (l0, m0) in radians."""

    def __init__(self, Cell=10, Sup=15, Nw=11, wmax=30000, Npix=101, Freqs=np.array([100.e6]),
                 OverS=11, lmShift=None, sups=None, IDFacet=0):
        self.Nw = max(1, int(Nw))
        self.Cell = Cell
        self.Sup = Sup
        self.Npix = Npix
        self.Freqs = np.asarray(Freqs, dtype=np.float64)
        self.OverS = OverS
        self.lmShift = lmShift
        self.IDFacet = IDFacet
        self.RefWave = C_LIGHT / self.Freqs[-1]
        self.wmax = wmax
        self.SpheM = SpheMachine(support=Sup)
        self.CF, self.fCF, self.ifzfCF = self.SpheM.make_sphe(Npix)
        self._init_w(sups)
        self.Sphe = np.float32(self.ifzfCF.real)
        self.InvSphe = np.float32(1.0 / np.float64(self.ifzfCF.real))

    def _sup_max(self, lrad):
        supmax = 501
        _, _, sphe_w = self.SpheM.make_sphe(supmax)
        l, m = np.mgrid[-lrad * np.sqrt(2.):np.sqrt(2.) * lrad:supmax * 1j,
                        -lrad * np.sqrt(2.):np.sqrt(2.) * lrad:supmax * 1j]
        n_1 = np.sqrt(1. - l ** 2 - m ** 2) - 1
        w = np.exp(-2. * 1j * np.pi * (self.wmax / self.RefWave) * n_1) * sphe_w
        fw = _fft2(w)
        fw1d = np.abs(fw[(supmax - 1) // 2, :])
        fw1d /= np.max(fw1d)
        fw1d = fw1d[(supmax - 1) // 2:]
        ind = np.argsort(fw1d)
        try:
            return np.int64(interp1d(fw1d[ind], np.arange(fw1d.shape[0])[ind])(np.array([1. / 1000]))[0])
        except Exception:
            return self.Sup

    def _init_w(self, sups):
        nw, overs = self.Nw, self.OverS
        lrad = (self.Npix / 2.) * self.Cell / 3600. * np.pi / 180.
        if sups is None:
            supmax = self._sup_max(lrad)
            sups = np.int64(np.linspace(self.Sup, np.max([supmax, self.Sup]), nw))
        sups = np.array(sups, dtype=np.int64).copy()
        if sups.size != nw:
            raise ValueError("need one support per w-plane")
        wlist = np.linspace(0, self.wmax, nw)
        l0, m0 = (0., 0.) if self.lmShift is None else self.lmShift
        radmax = (1. - np.sqrt(l0 ** 2 + m0 ** 2)) / np.sqrt(2)
        rad = max(min(3 * lrad, 0.99 * radmax), lrad)
        self.Cv, self.Cu, coef = give_dn(l0, m0, rad=rad, order=5)
        self.Wplanes, self.WplanesConj = [], []
        for i in range(nw):
            if not sups[i] % 2:
                sups[i] += 1
            _, _, this_sphe = self.SpheM.make_sphe(int(sups[i]))
            if nw <= 1:
                w = np.abs(this_sphe)
            else:
                dx = 2 * lrad / sups[i]
                l, m = np.mgrid[-lrad + dx / 2:lrad - dx / 2:sups[i] * 1j,
                                -lrad + dx / 2:lrad - dx / 2:sups[i] * 1j]
                n_1 = polyval2d(l, m, coef)
                w = np.exp(-2. * 1j * np.pi * (wlist[i] / self.RefWave) * n_1) * np.abs(this_sphe)
            zw = zero_pad(w, w.shape[0] * overs)
            fzw = reorg(np.complex64(_fft2(zw)), overs)
            fzwc = reorg(np.complex64(_fft2(np.conj(zw))), overs)
            self.Wplanes.append(np.require(fzw, requirements=["A", "C"]))
            self.WplanesConj.append(np.require(fzwc, requirements=["A", "C"]))
        self.Sups = sups
        self.wmap = wlist

    def winfos(self):
        return np.array([self.RefWave, self.wmax, len(self.Wplanes), self.OverS], dtype=np.float64)
