"""W-projection CF tables of many facets built on the GPU (ddf_cf_build).

Replaces the per-facet ``ModCF.ClassWTermModified`` construction (DDFacet/Imager/ModCF.py:374-597,
restated on the host in ``wterm.py``) for a whole facet set: the host computes what ModCF
computes once per facet (the order-5 fit of n(l, m) - n0, ``wterm.give_dn``) or once per support
(the taper ``|SpheMachine.make_sphe(S)|``), the device builds every w-plane and its conjugate
(screen, zero pad, centred FP64 FFT, complex64, [OverS][OverS][S][S]) in one call.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import check, lib
from .device import require_gpu, stream_handle
from .wterm import C_LIGHT, SpheMachine, give_dn


class DeviceCFSet:
    """CF tables of `facets_lm` (list of (l0, m0) in radians) on a padded facet grid of `npix`
    pixels of `cell_arcsec`.  ``data`` is complex64 [nfacets, stride] on the device; facet f's plane
    k (put order) starts at ``data[f, offsets[k]]`` and has (sups[k % nw] * overs)^2 entries."""

    def __init__(self, facets_lm, npix, cell_arcsec, freqs, nw, sups, overs, wmax, sup0=None, device=None,
                 stream=None):
        dev = device or require_gpu()
        nw = max(1, int(nw))
        sups = np.array(sups, dtype=np.int64).copy()
        if sups.size != nw:
            raise ValueError("need one support per w-plane")
        sups[sups % 2 == 0] += 1
        self.nw, self.overs, self.sups = nw, int(overs), sups.astype(np.int32)
        freqs = np.asarray(freqs, np.float64)
        self.ref_wave = C_LIGHT / freqs[-1]
        self.wmax = float(wmax)
        lrad = (npix / 2.) * cell_arcsec / 3600. * np.pi / 180.
        self.lrad = lrad
        wlist = np.linspace(0, wmax, nw)
        coefs = np.zeros((len(facets_lm), 36), np.float64)
        self.cu = np.zeros(len(facets_lm))
        self.cv = np.zeros(len(facets_lm))
        for f, (l0, m0) in enumerate(facets_lm):
            radmax = (1. - np.sqrt(l0 ** 2 + m0 ** 2)) / np.sqrt(2)
            rad = max(min(3 * lrad, 0.99 * radmax), lrad)
            cv, cu, c = give_dn(l0, m0, rad=rad, order=5)
            coefs[f] = np.asarray(c, np.float64)
            self.cu[f], self.cv[f] = cu, cv
        sphe = SpheMachine(support=int(sup0 or sups[0]))
        tapers = np.concatenate([np.abs(sphe.make_sphe(int(s))[2]).astype(np.float64).ravel() for s in sups])
        stride = int(2 * np.sum((sups.astype(np.int64) * overs) ** 2))
        self.data = torch.empty((len(facets_lm), stride), dtype=torch.complex64, device=dev)
        offs = np.zeros(2 * nw, np.int64)
        keep = (self.sups, wlist, coefs, tapers)
        spec = abi.ddf_cf_spec(nw, self.overs, len(facets_lm), 0, self.sups.ctypes.data, wlist.ctypes.data,
                               coefs.ctypes.data, tapers.ctypes.data, float(self.ref_wave), float(lrad))
        check(lib().ddf_cf_build(C.byref(spec), C.c_void_p(self.data.data_ptr()),
                                 offs.ctypes.data_as(C.POINTER(C.c_int64)), C.c_void_p(stream_handle(stream))))
        del keep
        self.offsets = offs
        self.wlist = wlist

    def plane(self, f, k):
        """Plane k (put order: Wplanes then WplanesConj) of facet f, [S*OverS, S*OverS]."""
        s = int(self.sups[k % self.nw]) * self.overs
        o = int(self.offsets[k])
        return self.data[f, o:o + s * s].view(s, s)
