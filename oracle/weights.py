"""TEST INFRASTRUCTURE -- the CPU oracle of the imaging weights (numpy restatement).

Follows DDFacet/Data/ClassVisServer.py: ``_uv_to_index`` (:983-1004), the scatter-add of
``pyAccumulateWeightsOntoGridNoSem`` (DDFacet/Gridder/GridderSmearPols.cc:55-74), the robust
transform of the weight grid (:1188-1197) and the renormalisation of ``_finalizeWeights_handler``
(:1035-1040).  Only tests/ and bench.py's CPU baseline may use it, as the checker.
"""
import numpy as np

LIGHT_SPEED = 299792458


def uv_to_index(uvw, weights, freqs, cell, npix, npixx, nbands, xymax, chan_band=None):
    uv = np.array(uvw[:, :2], np.float64)
    uv[uv[:, 1] < 0] *= -1
    uv = uv[..., np.newaxis] * np.asarray(freqs, np.float64)[np.newaxis, np.newaxis, :] / LIGHT_SPEED
    uv = np.floor(uv / cell).astype(int)
    x = uv[:, 0, :]
    y = uv[:, 1, :]
    x += xymax
    index = np.zeros((uv.shape[0], len(freqs)), np.int32)
    index[...] = y * npixx + x
    if nbands > 1:
        index += np.asarray(chan_band)[np.newaxis, :] * npix
    index[weights == 0] = 0
    return index


def accumulate(grid, weights, index):
    flat = grid.reshape(-1)
    w = np.asarray(weights, np.float32).ravel()
    idx = np.asarray(index).astype(np.int64).ravel()
    for i in range(w.size):
        if w[i] != 0:
            flat[idx[i]] += w[i]
    return grid


def accumulate_fast(grid, weights, index):
    """accumulate() with np.add.at (the same sums in the same order: numpy applies them in index
    order of the input)."""
    flat = grid.reshape(-1)
    w = np.asarray(weights, np.float32).ravel()
    idx = np.asarray(index).astype(np.int64).ravel()
    nz = w != 0
    np.add.at(flat, idx[nz], w[nz].astype(np.float64))
    return grid


def briggs(grid, robust):
    numeratorSqrt = 5.0 * 10 ** (-robust)
    for band in range(grid.shape[0]):
        grid1 = grid[band, :]
        avgW = (grid1 ** 2).sum() / grid1.sum()
        sSq = numeratorSqrt ** 2 / avgW
        grid1[...] = 1 + grid1 * sSq
    return grid


def finalize(weights, index, grid):
    g = grid.reshape(-1)
    index = np.array(index)
    index[index >= len(g)] = 0
    w = np.array(weights, np.float32)
    w /= g[index]
    return w


def imaging_weights(uvw, weights, freqs, cell, xymax, npixx, npix, weighting="briggs", robust=0.0, nbands=1,
                    chan_band=None, fast=True):
    w = np.array(weights, np.float32)
    if weighting == "natural":
        return w, None
    index = uv_to_index(uvw, w, freqs, cell, npix, npixx, nbands, xymax, chan_band)
    grid = np.zeros((nbands, npix), np.float64)
    (accumulate_fast if fast else accumulate)(grid, w, index)
    if weighting in ("briggs", "robust"):
        briggs(grid, robust)
    return finalize(w, index, grid), grid
