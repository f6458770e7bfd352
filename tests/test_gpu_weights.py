"""Imaging weights on the device (ddfacet_amd.weights: ddf_weights_index, ddf_accumulate_weights,
ddf_weights_briggs, ddf_weights_finalize) against the oracle (oracle/weights.py, pinned by
tests/test_weights_oracle.py).  Indices bit-exact; the weight grid to 1e-12 relative (the device
scatter-add sums a cell in another order); final float32 weights to 2 ulp (the quotient of two
FP64 values whose last bits may differ, rounded to FP32); zero weights give the reference's
0 / grid[0]."""
import numpy as np
import pytest
import torch

from oracle import weights as ow

pytestmark = pytest.mark.gpu
C0 = 299792458.0


def _chunk(seed, nrow=3000, nchan=16, nbands=1):
    rng = np.random.default_rng(seed)
    uvw = rng.uniform(-2000, 2000, (nrow, 3))
    freqs = np.linspace(0.9e9, 1.67e9, nchan)
    w = rng.uniform(0.5, 1.5, (nrow, nchan)).astype(np.float32)
    w[rng.random((nrow, nchan)) < 0.1] = 0
    uvmax = float(np.abs(uvw[:, :2]).max() * freqs.max() / C0)
    cell = 97.0
    chan_band = np.minimum(np.arange(nchan) * nbands // nchan, nbands - 1).astype(np.int32)
    return uvw, w, freqs, cell, uvmax, chan_band


@pytest.mark.parametrize("weighting,robust,nbands", [("uniform", 0.0, 1), ("briggs", 0.0, 1), ("briggs", -1.5, 2),
                                                    ("briggs", 2.0, 1), ("natural", 0.0, 1)])
def test_weights_match_oracle(weighting, robust, nbands):
    from ddfacet_amd import weights as dw
    print("case", weighting, robust, nbands, flush=True)
    uvw, w, freqs, cell, uvmax, cb = _chunk(7 + nbands, nbands=nbands)
    xymax, npixx, npix = dw.grid_geometry(uvmax, cell)
    got, grid = dw.imaging_weights(uvw, w, freqs, cell, uvmax, weighting=weighting, robust=robust, chan_band=cb,
                                   nbands=nbands, return_grid=True)
    torch.cuda.synchronize()
    want, wgrid = ow.imaging_weights(uvw, w, freqs, cell, xymax, npixx, npix, weighting=weighting, robust=robust,
                                     nbands=nbands, chan_band=cb)
    got = got.cpu().numpy()
    if weighting == "natural":
        assert np.array_equal(got, want)
        return
    np.testing.assert_allclose(grid.cpu().numpy(), wgrid, rtol=1e-12, atol=0)
    nz = w != 0
    np.testing.assert_allclose(got[nz], want[nz], rtol=2.4e-7, atol=0)
    np.testing.assert_array_equal(got[~nz], want[~nz])


def test_weights_index_is_bit_exact():
    import ctypes as C
    from ddfacet_amd import weights as dw
    from ddfacet_amd._lib import check, lib
    uvw, w, freqs, cell, uvmax, cb = _chunk(21, nbands=4)
    xymax, npixx, npix = dw.grid_geometry(uvmax, cell)
    dev = torch.device("cuda")
    tu = torch.from_numpy(uvw).to(dev)
    tw = torch.from_numpy(w).to(dev)
    tf = torch.from_numpy(freqs).to(dev)
    tc = torch.from_numpy(cb).to(dev)
    idx = torch.empty(w.shape, dtype=torch.int32, device=dev)
    p = lambda t: C.c_void_p(t.data_ptr())
    check(lib().ddf_weights_index(p(tu), p(tw), w.shape[0], w.shape[1], p(tf), p(tc), cell, npix, npixx, 4, xymax,
                                  p(idx), C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    want = ow.uv_to_index(uvw, w, freqs, cell, npix, npixx, 4, xymax, chan_band=cb)
    assert np.array_equal(idx.cpu().numpy(), want)


def test_scatter_add_long_runs_of_one_cell():
    """Rows in baseline-time order put runs of consecutive rows in one weight-grid cell: here
    every uvw row is repeated 8-100 times (with sub-cell jitter), so the run-summed scatter-add
    sees runs of 3 to 100 lanes, runs cut by zero weights, runs ending at lane 63 and runs
    crossing the 64-row tile edge.  Grid and final weights against the oracle."""
    from ddfacet_amd import weights as dw
    rng = np.random.default_rng(31)
    base = rng.uniform(-2000, 2000, (70, 3))
    reps = rng.integers(8, 101, base.shape[0])
    uvw = np.repeat(base, reps, axis=0) + rng.normal(0, 1e-3, (int(reps.sum()), 3))
    nrow, nchan = uvw.shape[0], 8
    freqs = np.linspace(0.9e9, 1.67e9, nchan)
    w = rng.uniform(0.5, 1.5, (nrow, nchan)).astype(np.float32)
    w[rng.random((nrow, nchan)) < 0.15] = 0
    uvmax = float(np.abs(uvw[:, :2]).max() * freqs.max() / C0)
    cell = 97.0
    xymax, npixx, npix = dw.grid_geometry(uvmax, cell)
    idx = ow.uv_to_index(uvw, np.ones_like(w), freqs, cell, npix, npixx, 1, xymax)
    same = idx[1:] == idx[:-1]                        # [row, chan]: row r + 1 continues row r's cell
    runs3 = sum(int(np.sum(np.convolve(same[:, c].astype(int), [1, 1], "valid") == 2)) for c in range(nchan))
    cross = int(np.sum(same[63::64]))                 # row 64 k + 63 -> 64 (k + 1) in one cell
    assert runs3 > 1000 and cross > 10, (runs3, cross)
    for weighting in ("uniform", "briggs"):
        got, grid = dw.imaging_weights(uvw, w, freqs, cell, uvmax, weighting=weighting, robust=0.5, return_grid=True)
        torch.cuda.synchronize()
        want, wgrid = ow.imaging_weights(uvw, w, freqs, cell, xymax, npixx, npix, weighting=weighting, robust=0.5)
        np.testing.assert_allclose(grid.cpu().numpy(), wgrid, rtol=1e-12, atol=0)
        got = got.cpu().numpy()
        nz = w != 0
        np.testing.assert_allclose(got[nz], want[nz], rtol=2.4e-7, atol=0)
        np.testing.assert_array_equal(got[~nz], want[~nz])
