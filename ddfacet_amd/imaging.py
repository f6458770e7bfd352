"""Facet geometry and the image side of the major cycle on the GPU.

Geometry: a regular nfx x nfx tessellation of an m x m main image (the layout
ClassFacetMachineTessel produces for a square grid of facets): facet centres ``pixCentral``,
facet (l0, m0), and the spatial weights ``SW`` -- the facet mask smoothed by a Gaussian of
sigma ``MixingWidth`` pixels, normalised to 1 and clipped below 1e-3 (ClassFacetMachine.py:
766-820).  Main-image pixel (X, Y) maps to (l, m) = ((X - m//2) cell, (Y - m//2) cell), the
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
    def __init__(self, m, nfx, facet_npix, padded_npix, cell_rad, mixing_width=10.0):
        self.m = int(m)
        self.nfx = int(nfx)
        self.nf = self.nfx * self.nfx
        self.facet_npix = int(facet_npix)
        self.n = int(padded_npix)
        self.cell = float(cell_rad)
        step = self.m / self.nfx
        centres = [int(round(step * (i + 0.5))) for i in range(self.nfx)]
        self.pix = [(centres[ix], centres[iy]) for iy in range(self.nfx) for ix in range(self.nfx)]
        c = self.m // 2
        self.lm = [((x - c) * self.cell, (y - c) * self.cell) for x, y in self.pix]
        self.mixing_width = mixing_width
        self._sw = None

    def spatial_weights(self):
        """SW of every facet, float32 [nf, n, n] (image orientation, cf_dict["SW"])."""
        if self._sw is not None:
            return self._sw
        from scipy.ndimage import gaussian_filter
        n, m = self.n, self.m
        out = np.zeros((self.nf, n, n), np.float32)
        half = self.m / self.nfx / 2.0
        a = np.arange(n)
        for i, (xc, yc) in enumerate(self.pix):
            X = xc - n // 2 + a
            Y = yc - n // 2 + a
            inx = (np.abs(X - xc) <= half) & (X >= 0) & (X < m)
            iny = (np.abs(Y - yc) <= half) & (Y >= 0) & (Y < m)
            mask = (inx[:, None] & iny[None, :]).astype(np.float64)
            sw = gaussian_filter(mask, self.mixing_width, mode="constant")
            sw /= sw.max()
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
