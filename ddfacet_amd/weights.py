"""Imaging weights on the device: the uniform / Briggs weighting of ClassVisServer for one chunk.

Mirrors ``ClassVisServer._CalcWeights_serial`` (DDFacet/Data/ClassVisServer.py:1068-1215) for the
arrays of one chunk: the weight-grid geometry (:1122-1133), the uv-grid index of every visibility
(``_uv_to_index``, :983-1004), the scatter-add of the weights onto the grid
(``_accumulateWeights_handler`` through ``pyAccumulateWeightsOntoGrid``, :1006-1019, here
``ddf_weights_accumulate`` on the int32 index), the robust
transform of the grid (:1188-1197) and the renormalisation (``_finalizeWeights_handler``,
:1028-1040).  Every step is a kernel of libddfgpu.so; torch provides the device buffers.
"""
import ctypes as C
import math

import numpy as np
import torch

from ._lib import check, lib
from .device import require_gpu, stream_handle, to_device

LIGHT_SPEED = 299792458.0  # ClassVisServer._cc


def grid_geometry(uvmax_wavelengths, cell):
    """(xymax, npixx, npix) of the weight grid (ClassVisServer.py:1122-1133): U from -xymax to
    xymax, V from 0 to xymax (negative-v points are mirrored)."""
    xymax = int(math.floor(uvmax_wavelengths / cell)) + 1
    npixx = xymax * 2 + 1
    npixy = xymax + 1
    return xymax, npixx, npixx * npixy


def uvmax_wavelengths(uvw, freqs, rowflags=None):
    """max |u|, |v| of the unflagged rows in wavelengths at the highest frequency
    (ClassVisServer.py:951, 1112)."""
    uv = np.abs(np.asarray(uvw)[:, :2])
    if rowflags is not None:
        uv = uv[~np.asarray(rowflags)]
    return float(uv.max() * np.max(freqs) / LIGHT_SPEED)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def imaging_weights(uvw, weights, freqs, cell, uvmax, weighting="briggs", robust=0.0, chan_band=None, nbands=1,
                    stream=None, return_grid=False):
    """Imaging weights of one chunk.

    uvw [nrow, 3] float64, weights [nrow, nchan] float32 (flagged visibilities already at 0, as
    ``_loadWeights_handler`` leaves them), freqs [nchan] float64; cell: the weight-grid cell in
    wavelengths (1 / (Super FOV)); uvmax: the chunk's uvmax in wavelengths; chan_band [nchan]: the
    band of each channel when nbands > 1.  Returns the float32 weights on the device (natural: the
    weights unchanged) and, with ``return_grid``, the float64 weight grid [nbands, npix].
    """
    if weighting not in ("natural", "uniform", "briggs", "robust"):
        raise ValueError("unknown weighting %r" % (weighting,))
    dev = require_gpu()
    st = stream_handle(stream)
    w = to_device(np.asarray(weights, np.float32) if not isinstance(weights, torch.Tensor) else weights, dev).clone()
    if w.dtype != torch.float32 or w.dim() != 2:
        raise TypeError("weights must be float32 [nrow, nchan]")
    nrow, nchan = w.shape
    if weighting == "natural":
        return (w, None) if return_grid else w
    u = to_device(np.ascontiguousarray(uvw, np.float64) if not isinstance(uvw, torch.Tensor) else uvw, dev)
    f = to_device(np.ascontiguousarray(freqs, np.float64), dev)
    if u.dtype != torch.float64 or u.shape != (nrow, 3):
        raise TypeError("uvw must be float64 [nrow, 3]")
    cb = None
    if nbands > 1:
        if chan_band is None:
            raise ValueError("chan_band is required when nbands > 1")
        cb = to_device(np.ascontiguousarray(chan_band, np.int32), dev)
    xymax, npixx, npix = grid_geometry(uvmax, cell)
    index = torch.empty((nrow, nchan), dtype=torch.int32, device=dev)
    L = lib()
    check(L.ddf_weights_index(_ptr(u), _ptr(w), nrow, nchan, _ptr(f), _ptr(cb) if cb is not None else None,
                              float(cell), npix, npixx, nbands, xymax, _ptr(index), C.c_void_p(st)))
    grid = torch.zeros((nbands, npix), dtype=torch.float64, device=dev)
    check(L.ddf_weights_accumulate(_ptr(grid), grid.numel(), _ptr(w), _ptr(index), nrow, nchan, C.c_void_p(st)))
    if weighting in ("briggs", "robust"):
        work = torch.empty(int(L.ddf_weights_briggs_work_size(nbands)), dtype=torch.float64, device=dev)
        check(L.ddf_weights_briggs(_ptr(grid), nbands, npix, float(robust), _ptr(work), work.numel(),
                                   C.c_void_p(st)))
    check(L.ddf_weights_finalize(_ptr(w), _ptr(index), w.numel(), _ptr(grid), grid.numel(), C.c_void_p(st)))
    return (w, grid) if return_grid else w
