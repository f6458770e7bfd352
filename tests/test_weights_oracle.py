"""The imaging-weight oracle (oracle/weights.py) pinned by hand-worked cases: uv-grid indices
(mirroring of negative v, floor of negative u, zero weights, band offsets), the scatter-add, the
uniform renormalisation (every occupied cell's weights sum to 1), the robust transform on a
two-cell grid worked by hand and in its limits (robust -> +inf: natural, -> -inf: uniform up to a
constant)."""
import numpy as np

from oracle import weights as ow

C0 = 299792458.0


def _geom(xymax):
    npixx = 2 * xymax + 1
    return npixx, npixx * (xymax + 1)


def test_uv_to_index_hand_worked():
    # cell 10 wavelengths at 1 m wavelength (freq C0): u = 25 -> floor(2.5) = 2, u = -25 -> -3
    xymax = 5
    npixx, npix = _geom(xymax)
    uvw = np.array([[25.0, 35.0, 0.0],     # x = 2 + 5, y = 3 -> 3 * 11 + 7 = 40
                    [-25.0, 35.0, 0.0],    # x = -3 + 5, y = 3 -> 35
                    [25.0, -35.0, 0.0],    # mirrored to (-25, 35) -> 35
                    [5.0, 5.0, 0.0]])      # weight 0 -> 0
    w = np.array([[1.0], [1.0], [1.0], [0.0]], np.float32)
    idx = ow.uv_to_index(uvw, w, np.array([C0]), 10.0, npix, npixx, 1, xymax)
    assert idx.dtype == np.int32
    assert idx[:, 0].tolist() == [40, 35, 35, 0]


def test_uv_to_index_two_bands():
    # channel 1 at twice the frequency doubles u, v in wavelengths and is offset by one band grid
    xymax = 5
    npixx, npix = _geom(xymax)
    uvw = np.array([[15.0, 22.0, 0.0]])
    w = np.ones((1, 2), np.float32)
    idx = ow.uv_to_index(uvw, w, np.array([C0, 2 * C0]), 10.0, npix, npixx, 2, xymax, chan_band=np.array([0, 1]))
    # ch 0: x = 1 + 5, y = 2 -> 28; ch 1: x = 3 + 5, y = 4 -> 52, + 66 (band 1)
    assert idx[0].tolist() == [28, 118]


def test_accumulate_sums_weights_per_cell_in_input_order():
    grid = np.zeros((1, 6))
    w = np.array([1.5, 0.0, 2.0, 0.25, 3.0], np.float32)
    idx = np.array([2, 2, 5, 2, 0], np.int32)
    ow.accumulate(grid, w, idx)
    assert grid[0].tolist() == [3.0, 0, 1.75, 0, 0, 2.0]
    rng = np.random.default_rng(3)
    w = rng.uniform(0, 2, 5000).astype(np.float32)
    w[rng.random(5000) < 0.2] = 0
    idx = rng.integers(0, 40, 5000).astype(np.int32)
    a, b = np.zeros((1, 40)), np.zeros((1, 40))
    ow.accumulate(a, w, idx)
    ow.accumulate_fast(b, w, idx)
    assert np.array_equal(a, b)


def _random_chunk(seed, nrow=400, nchan=8):
    rng = np.random.default_rng(seed)
    uvw = rng.uniform(-800, 800, (nrow, 3))
    freqs = np.linspace(1.2e9, 1.5e9, nchan)
    w = rng.uniform(0.5, 1.5, (nrow, nchan)).astype(np.float32)
    w[rng.random((nrow, nchan)) < 0.1] = 0
    uvmax = np.abs(uvw[:, :2]).max() * freqs.max() / C0
    cell = 150.0
    xymax = int(np.floor(uvmax / cell)) + 1
    npixx, npix = _geom(xymax)
    return uvw, w, freqs, cell, xymax, npixx, npix


def test_uniform_weights_sum_to_one_per_cell():
    uvw, w, freqs, cell, xymax, npixx, npix = _random_chunk(5)
    out, grid = ow.imaging_weights(uvw, w, freqs, cell, xymax, npixx, npix, weighting="uniform")
    idx = ow.uv_to_index(uvw, w, freqs, cell, npix, npixx, 1, xymax)
    nz = w != 0
    sums = np.zeros(npix)
    np.add.at(sums, idx[nz], out[nz].astype(np.float64))
    occupied = grid[0] > 0
    np.testing.assert_allclose(sums[occupied], 1.0, rtol=1e-6)
    # zero weights point at cell 0 and are divided by it, as in the reference: 0 / 0 = NaN when
    # cell 0 is empty (those visibilities are flagged, so the gridders never read them)
    assert grid[0, 0] == 0
    assert np.all(np.isnan(out[~nz]))


def test_briggs_two_cells_hand_worked():
    # cell A holds weight 1, cell B weights 1 and 2: avgW = (1 + 9) / 4 = 2.5, robust 0 gives
    # sSq = 25 / 2.5 = 10, so the grid becomes A: 11, B: 31
    grid = np.zeros((1, 4))
    w = np.array([1.0, 1.0, 2.0], np.float32)
    idx = np.array([1, 3, 3], np.int32)
    ow.accumulate(grid, w, idx)
    ow.briggs(grid, 0.0)
    assert grid[0].tolist() == [1.0, 11.0, 1.0, 31.0]
    out = ow.finalize(w, idx, grid)
    np.testing.assert_allclose(out, np.array([1 / 11, 1 / 31, 2 / 31], np.float32), rtol=1e-7)


def test_briggs_limits():
    uvw, w, freqs, cell, xymax, npixx, npix = _random_chunk(6)
    nat, _ = ow.imaging_weights(uvw, w, freqs, cell, xymax, npixx, npix, weighting="natural")
    uni, _ = ow.imaging_weights(uvw, w, freqs, cell, xymax, npixx, npix, weighting="uniform")
    hi, _ = ow.imaging_weights(uvw, w, freqs, cell, xymax, npixx, npix, weighting="briggs", robust=8.0)
    lo, _ = ow.imaging_weights(uvw, w, freqs, cell, xymax, npixx, npix, weighting="briggs", robust=-8.0)
    np.testing.assert_allclose(hi, nat, rtol=1e-6)
    nz = w != 0
    ratio = lo[nz].astype(np.float64) / uni[nz]
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-6)


def test_finalize_out_of_range_index_reads_cell_zero():
    grid = np.array([[4.0, 2.0, 8.0]])
    out = ow.finalize(np.array([1.0, 1.0, 1.0], np.float32), np.array([1, 3, 7], np.int32), grid)
    assert out.tolist() == [0.5, 0.25, 0.25]
