"""Facet geometry and the image side of the major cycle on the GPU.

Geometry: the regular nfx x nfx tessellation of an m x m main image (ClassFacetMachine.
setFacetsLocs): facet centres ``pixCentral``, facet (l0, m0), and the spatial weights ``SW`` --
the facet polygon mask convolved with a Gaussian of ``MixingWidth`` pixels, normalised to 1 and
clipped below 1e-3, built the way ClassFacetMachine._initcf_worker builds it
(ClassFacetMachine.py:785-818, ModFFTW.py:654-689, 760-847).  Main-image pixel (X, Y) maps to (l, m) = ((X - m//2) cell, (Y - m//2) cell), the
orientation fixed by ClassDDEGridMachine.testGrid (ClassDDEGridMachine.py:73-83, 229-271) and
the facet flips of FacetsToIm_Channel / GiveModelTessel.

Device functions (ddf_fft2c, ddf_facet_norm, ddf_stitch, ddf_model_to_grid) are wrapped for torch
tensors; see include/ddfgpu.h for their reference counterparts.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import check, lib
from .device import stream_handle


class FacetGeometry:
    """The regular nfx x nfx tessellation of ClassFacetMachine.setFacetsLocs
    (ClassFacetMachine.py:414-485): the m x m main image is nfx facets of facet_npix pixels a
    side (m = nfx * facet_npix), facet centres one facet apart, each facet a square polygon of
    half-width facet_npix / 2 pixels around its centre, gridded on a padded_npix grid."""

    def __init__(self, m, nfx, facet_npix, padded_npix, cell_rad, mixing_width=10.0):
        self.m = int(m)
        self.nfx = int(nfx)
        self.nf = self.nfx * self.nfx
        self.facet_npix = int(facet_npix)
        if self.m != self.nfx * self.facet_npix:
            raise ValueError("the main image must be nfx facets wide")
        self.n = int(padded_npix)
        self.cell = float(cell_rad)
        centres = [i * self.facet_npix + self.facet_npix // 2 for i in range(self.nfx)]
        self.pix = [(centres[ix], centres[iy]) for iy in range(self.nfx) for ix in range(self.nfx)]
        c = self.m // 2
        self.lm = [((x - c) * self.cell, (y - c) * self.cell) for x, y in self.pix]
        self.mixing_width = mixing_width
        self._sw = None

    def spatial_weights(self):
        """SW of every facet, float32 [nf, n, n], axes (l, m) (cf_dict["SW"],
        ClassFacetMachine._initcf_worker, ClassFacetMachine.py:785-818):

        * the facet polygon (and the image's outer square) tested at the points of
          mgrid[l0 - R : l0 + R : n j] with R = n cell / 2, i.e. n points spaced n / (n - 1)
          pixels apart (strict containment; the regular polygons' edges fall between samples);
        * convolved with the unit-peak Gaussian of ModFFTW.GiveGauss(n, 1, (w, w, 0)), sampled on
          the same n / (n - 1) spacing, as ModFFTW._convolveSingleGaussianFFTW does it: both
          zero-padded to EstimateNpix(n) (no padding when n is already a fast size of at least
          513), multiplied in the Fourier domain by |rfft2(PSF)|, i.e. a circular convolution;
        * divided by its maximum, values below 1e-3 set to 0.
        """
        if self._sw is not None:
            return self._sw
        from .configs import estimate_npix
        n, m = self.n, self.m
        s = n / (n - 1.0)
        off = (np.arange(n) - (n - 1) / 2.0) * s              # sample offsets from the facet centre (px)
        half = self.facet_npix / 2.0
        u = -n / 2.0 + np.arange(n) * s                       # GiveGauss grid, uvscale = n / 2
        g1 = np.exp(-u * u / (2.0 * self.mixing_width ** 2))
        npad = max(n, estimate_npix(n)[1])
        pe = max(int(np.ceil((estimate_npix(n)[1] - n) / 2.0) * 2), 0) // 2
        psf = np.zeros((npad, npad))
        psf[pe:pe + n, pe:pe + n] = np.outer(g1, g1)
        fpsf = np.abs(np.fft.rfft2(np.fft.ifftshift(psf)))
        out = np.zeros((self.nf, n, n), np.float32)
        for i, (xc, yc) in enumerate(self.pix):
            inx = (np.abs(off) < half) & (np.abs(xc - m // 2 + off) < m / 2.0)
            iny = (np.abs(off) < half) & (np.abs(yc - m // 2 + off) < m / 2.0)
            mask = np.zeros((npad, npad))
            mask[pe:pe + n, pe:pe + n] = np.outer(inx, iny)
            a = np.fft.fftshift(np.fft.irfft2(np.fft.rfft2(np.fft.ifftshift(mask)) * fpsf, s=mask.shape))
            sw = a[pe:pe + n, pe:pe + n]
            sw = sw / sw.max()
            sw[sw < 1e-3] = 0.0
            out[i] = sw
        self._sw = out
        return out


def _facets_struct(geom, sphe, inv_sphe, sw, dirty=None, model=None, sumwt=None, sum_jones=None):
    n = geom.nf
    arr = (abi.ddf_image_facet * n)()
    per = lambda t, i: 0 if t is None else t[i].data_ptr()  # noqa: E731
    for i in range(n):
        xc, yc = geom.pix[i]
        arr[i].xc, arr[i].yc = xc, yc
        arr[i].sphe = per(sphe, i)
        arr[i].inv_sphe = per(inv_sphe, i)
        arr[i].sw = per(sw, i)
        arr[i].dirty = per(dirty, i)
        arr[i].model = per(model, i)
        arr[i].sumwt = per(sumwt, i)
        arr[i].sum_jones = per(sum_jones, i)
    return arr


def fft2c(planes, inverse, scale=1.0, stream=None):
    """In-place centred FFT of a contiguous complex64 tensor [..., n, n]."""
    assert planes.dtype == torch.complex64 and planes.is_contiguous()
    n = planes.shape[-1]
    check(lib().ddf_fft2c(C.c_void_p(planes.data_ptr()), n, planes.numel() // (n * n), int(inverse), float(scale),
                          C.c_void_p(stream_handle(stream))))


def grid_to_image(grids, overs, stream=None):
    """GridToIm of every facet grid, in place (grids complex64 [nf, nband, npol, n, n])."""
    fft2c(grids, True, float(overs) ** 2, stream)


def facet_norm(geom, sw, out, stream=None):
    arr = _facets_struct(geom, sw, None, sw)
    check(lib().ddf_facet_norm(arr, geom.nf, geom.n, geom.m, C.c_void_p(out.data_ptr()),
                               C.c_void_p(stream_handle(stream))))


def stitch(geom, sphe, inv_sphe, sw, dirty, sumwt, sum_jones, norm, image, stream=None):
    nband, npol = int(image.shape[0]), int(image.shape[1])
    arr = _facets_struct(geom, sphe, inv_sphe, sw, dirty=dirty, sumwt=sumwt, sum_jones=sum_jones)
    check(lib().ddf_stitch(arr, geom.nf, geom.n, nband, npol, geom.m, C.c_void_p(norm.data_ptr()),
                           C.c_void_p(image.data_ptr()), C.c_void_p(stream_handle(stream))))


def model_to_grid(geom, sphe, sw, image, norm, models, overs, stream=None):
    nband, npol = int(image.shape[0]), int(image.shape[1])
    arr = _facets_struct(geom, sphe, None, sw, model=models)
    check(lib().ddf_model_to_grid(arr, geom.nf, geom.n, nband, npol, geom.m, C.c_void_p(image.data_ptr()),
                                  C.c_void_p(norm.data_ptr()), int(overs), C.c_void_p(stream_handle(stream))))
