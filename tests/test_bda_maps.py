"""BDA block maps (ClassSmearMapping.py:83-135 format, :496-625 mode-1 builder, :348-491 mode-2
builder) against hand-worked cases.

The reference builder is Python that cannot be imported here (it needs the process pool and its
native helpers), so the restatement in ddfacet_amd/bda.py is pinned by cases whose blocks follow
from the builder's rules by hand:

* rows of one baseline on a circle of radius R (constant |uvw|) with equal chord steps s: a
  block closes at the first row where the running uv path length exceeds duv_cut, and at the
  last row; the channel blocks come from the smallest per-row channel count of the block,
  nch = duv_cut nu0 / (|uvw| dfreq), spread over the band by linspace (integer truncation),
  and are cut again wherever chan_map changes;
* baselines are concatenated in the lexicographic order of the string "a0:a1" (so "10:11"
  precedes "2:10");
* the two header words hold the block count (low, high), then one size per block, then the
  records [chStart, chEnd, rows...].
"""
import numpy as np

from ddfacet_amd import bda
from oracle import bda as obda

C = 3e8  # the builder's speed of light (ClassSmearMapping.py:508)


def _circle(nrows, R=1000.0, dtheta=0.01, a0=0, a1=1, row0=0):
    th = np.arange(nrows) * dtheta
    uvw = np.stack([R * np.cos(th), R * np.sin(th), np.zeros(nrows)], 1)
    return uvw, np.full(nrows, a0, np.int32), np.full(nrows, a1, np.int32)


def _freqs(nu0=1e9, dfreq=1e7, nchan=8):
    return nu0 - dfreq * (nchan - 1 - np.arange(nchan))


def _blocks(m):
    return [(c0, c1, list(map(int, rows))) for c0, c1, rows in bda.iter_blocks(m)]


def test_pack_header_sizes_records_and_key_order():
    per = {"2:3": ([3], [0, 4, 7]), "10:11": ([4, 3], [0, 2, 1, 2, 2, 4, 5]), "2:10": ([2], [1, 3])}
    m = bda.pack(per)
    # keys sorted as strings: "10:11" < "2:10" < "2:3"
    want = [4, 0, 4, 3, 2, 3, 0, 2, 1, 2, 2, 4, 5, 1, 3, 0, 4, 7]
    assert m.dtype == np.int32 and m.tolist() == want
    assert bda.nblocks_of(m) == 4
    assert bda.block_starts(m).tolist() == [6, 10, 13, 15]
    assert _blocks(m) == [(0, 2, [1, 2]), (2, 4, [5]), (1, 3, []), (0, 4, [7])]


def _mode1_one_baseline(nrows, chan_map=None, R=1000.0, dtheta=0.01, nch=2.5, cut_steps=2.5):
    """duv_cut = cut_steps chord steps; nch channels per block at radius R."""
    uvw, a0, a1 = _circle(nrows, R, dtheta)
    s = 2 * R * np.sin(dtheta / 2)
    duv_cut = cut_steps * s
    nu0 = 1e9
    dfreq = duv_cut * nu0 / (R * nch)
    freqs = _freqs(nu0, dfreq)
    lrad = 0.01
    dphi = duv_cut * np.pi * lrad * nu0 / C
    chan_map = np.zeros(8, np.int32) if chan_map is None else np.asarray(chan_map, np.int32)
    sizes, flat = obda._blocks_bl_old(np.arange(nrows), uvw, dfreq, freqs, dphi, lrad, chan_map)
    return bda.pack({"0:1": (sizes, flat)})


def test_mode1_rows_close_when_path_exceeds_cut_and_at_last_row():
    m = _mode1_one_baseline(7)
    # path after rows 0, 1, 2 = s, 2s, 3s > 2.5 s: blocks {0,1,2}, {3,4,5}, then {6} (last row);
    # nch = 2.5 -> edges arange(0, 8, 2.5) + [8] -> linspace(0, 8, 5) -> 0, 2, 4, 6, 8
    want = []
    for rows in ([0, 1, 2], [3, 4, 5], [6]):
        for c0 in (0, 2, 4, 6):
            want.append((c0, c0 + 2, rows))
    assert _blocks(m) == want


def test_mode1_channel_blocks_cut_at_band_edges():
    m = _mode1_one_baseline(3, chan_map=[0, 0, 0, 1, 1, 1, 1, 1])
    # one row block {0,1,2} (last row closes it); channel edges 0,2,4,6,8 plus the band edge at 3
    want = [(0, 2, [0, 1, 2]), (2, 3, [0, 1, 2]), (3, 4, [0, 1, 2]), (4, 6, [0, 1, 2]), (6, 8, [0, 1, 2])]
    assert _blocks(m) == want


def test_mode1_fewer_than_one_channel_per_block_gives_single_channels():
    m = _mode1_one_baseline(2, nch=0.4, cut_steps=10.0)
    assert _blocks(m) == [(c, c + 1, [0, 1]) for c in range(8)]


def test_mode1_wide_channel_blocks_cover_the_band_once():
    m = _mode1_one_baseline(4, nch=100.0, cut_steps=1.5)
    # path s, 2s > 1.5 s after row 1: blocks {0,1}, {2,3}; nch >= 8: one block [0, 8)
    assert _blocks(m) == [(0, 8, [0, 1]), (0, 8, [2, 3])]


def test_mode1_single_row_baseline():
    m = _mode1_one_baseline(1)
    assert _blocks(m) == [(0, 2, [0]), (2, 4, [0]), (4, 6, [0]), (6, 8, [0])]


def test_mode1_driver_decorr_and_fov_and_baseline_order():
    """bda_map_mode1: dPhi = sqrt(6 decorr), lrad = cell npix sqrt(2)/2, autocorrelations skipped,
    baselines in string-key order; the same blocks as the direct builder call."""
    R, dtheta = 1000.0, 0.01
    s = 2 * R * np.sin(dtheta / 2)
    duv_cut = 2.5 * s
    nu0 = 1e9
    cell, npix = 1e-5, 1000
    lrad = cell * npix * np.sqrt(2) / 2
    dphi = duv_cut * np.pi * lrad * nu0 / C
    decorr = dphi ** 2 / 6
    dfreq = duv_cut * nu0 / (R * 2.5)
    freqs = _freqs(nu0, dfreq)
    u1, a01, a11 = _circle(7, R, dtheta, a0=2, a1=10)
    u2, a02, a12 = _circle(4, R, dtheta, a0=10, a1=11)
    u3, a03, a13 = _circle(2, R, dtheta, a0=3, a1=3)  # autocorrelation rows: no blocks
    uvw = np.concatenate([u1, u2, u3])
    a0 = np.concatenate([a01, a02, a03])
    a1 = np.concatenate([a11, a12, a13])
    m = obda.bda_map_mode1(uvw, a0, a1, freqs, dfreq, decorr, cell, npix, np.zeros(8, np.int32))
    got = _blocks(m)
    # "10:11" (rows 7-10) first: blocks {7,8,9}, {10}; then "2:10" (rows 0-6) as in the 7-row case
    want = []
    for rows in ([7, 8, 9], [10], [0, 1, 2], [3, 4, 5], [6]):
        for c0 in (0, 2, 4, 6):
            want.append((c0, c0 + 2, rows))
    assert got == want


def test_mode1_row_step_is_the_distance_to_the_next_row():
    """Row i adds |uvw[i+1] - uvw[i]| (the last row repeats the previous step): with positions
    0, 10, 11, 31, 32, 33 along u and a cut of 15, the path after rows 0, 1, 2 is 10, 11, 31, so
    the first block is {0, 1, 2} (a step measured from the previous row would give {0, 1, 2, 3})."""
    x = np.array([0.0, 10.0, 11.0, 31.0, 32.0, 33.0])
    uvw = np.stack([1e6 + x, np.zeros(6), np.zeros(6)], 1)
    nu0, dfreq, lrad = 1e9, 1e7, 0.01
    duv_cut = 15.0
    dphi = duv_cut * np.pi * lrad * nu0 / C
    sizes, flat = obda._blocks_bl_old(np.arange(6), uvw, dfreq, _freqs(nu0, dfreq), dphi, lrad, np.zeros(8, np.int32))
    m = bda.pack({"0:1": (sizes, flat)})
    # nch = 15 * 1e9 / (1e6 * 1e7) < 1 -> single channels
    want = [(c, c + 1, [0, 1, 2]) for c in range(8)] + [(c, c + 1, [3, 4, 5]) for c in range(8)]
    assert _blocks(m) == want


# ---- mode 2 (GiveBlocksRowsListBL, ClassSmearMapping.py:348-491) ------------------------------------
# With nu_max = C / (2 pi) the phase factor 2 pi nu_max / C is 1, so a row's phase step is
# |du,dv| lmax + |dw| nmax and its channel span is nu_max dPhi / (ext dfreq) with
# ext = |u,v| lmax + |w| nmax.
NU1 = C / (2 * np.pi)


def _mode2(uvw, dphi, lmax, nchan=8, dfreq=None, chan_map=None):
    freqs = np.linspace(0.5 * NU1, NU1, nchan)
    dfreq = NU1 / 100.0 if dfreq is None else dfreq
    chan_map = np.zeros(nchan, np.int32) if chan_map is None else np.asarray(chan_map, np.int32)
    ind = np.arange(uvw.shape[0])
    sizes, flat = obda._blocks_bl_mode2(ind, uvw, dfreq, freqs, dphi, lmax, chan_map)
    return _blocks(bda.pack({"0:1": (sizes, flat)}))


def _line_u(n, u0=1e4, step=100.0):
    return np.stack([u0 + step * np.arange(n), np.zeros(n), np.zeros(n)], 1)


def test_mode2_time_blocks_from_accumulated_phase():
    # step 100 along u, lmax 0.01: phase step 1 per row (the last row repeats it); dPhi 2.6:
    # cumsum 1..7 / 2.6 = 0.38, 0.77, 1.15, 1.54, 1.92, 2.31, 2.69 -> blocks 0,0,1,1,1,2,2.  Spans
    # 260 / (100 + i) lie in (2.4, 2.6]: ceil(8 / span) = 4 channel blocks of ceil(8 / 4) = 2
    # channels in every time block.
    got = _mode2(_line_u(7), dphi=2.6, lmax=0.01)
    want = []
    for rows in ([0, 1], [2, 3, 4], [5, 6]):
        for c0 in (0, 2, 4, 6):
            want.append((c0, c0 + 2, rows))
    assert got == want


def test_mode2_channel_size_and_band_cuts():
    # one row block (small steps); span 290 / 100 = 2.9 for every row: ceil(8 / 2.9) = 3 blocks of
    # ceil(8 / 3) = 3 channels -> cuts 0, 3, 6, 8, plus the band edge at 4
    uvw = _line_u(3, step=1e-3)
    got = _mode2(uvw, dphi=2.9, lmax=0.01, chan_map=[0, 0, 0, 0, 1, 1, 1, 1])
    assert got == [(0, 3, [0, 1, 2]), (3, 4, [0, 1, 2]), (4, 6, [0, 1, 2]), (6, 8, [0, 1, 2])]


def test_mode2_spans_below_one_channel_give_single_channels():
    got = _mode2(_line_u(2, step=1e-3), dphi=0.5, lmax=0.01)
    assert got == [(c, c + 1, [0, 1]) for c in range(8)]


def test_mode2_w_steps_weighted_by_n_max():
    # lmax 0.6 -> nmax = 1 - sqrt(1 - 0.36) = 0.2; steps of 5 in w only: phase step 1 per row;
    # dPhi 1.6: cumsum 1, 2, 3, 4 / 1.6 = 0.625, 1.25, 1.875, 2.5 -> blocks 0, 1, 1, 2 (a step
    # counted with lmax instead of nmax would give 3 per row and four blocks).  ext = |w| nmax with
    # w = 1000 + 5 i -> span ~ 100 * 1.6 / 200 < 1: single channels
    n = 4
    uvw = np.stack([np.zeros(n), np.zeros(n), 1000.0 + 5.0 * np.arange(n)], 1)
    got = _mode2(uvw, dphi=1.6, lmax=0.6, nchan=2)
    assert got == [(0, 1, [0]), (1, 2, [0]), (0, 1, [1, 2]), (1, 2, [1, 2]), (0, 1, [3]), (1, 2, [3])]


def test_mode2_single_row_baseline():
    # no step to repeat: one time block; span 260 / 100 = 2.6 -> 4 blocks of 2 channels
    got = _mode2(_line_u(1), dphi=2.6, lmax=0.01)
    assert got == [(c, c + 2, [0]) for c in (0, 2, 4, 6)]


def test_mode2_driver_decorr_fov_and_baseline_order():
    """bda_map_mode2: dPhi = sqrt(6 decorr), lmax = cell npix sqrt(2)/2, autocorrelations skipped,
    baselines in string-key order; the same blocks as the direct builder call."""
    dphi, lmax = 2.6, 0.01
    decorr = dphi ** 2 / 6
    npix = 1000
    cell = lmax / (npix * np.sqrt(2) / 2)
    u1 = _line_u(7)
    u2 = _line_u(1)
    uvw = np.concatenate([u1, u2, _line_u(2)])
    a0 = np.array([2] * 7 + [10] + [3, 3], np.int32)
    a1 = np.array([10] * 7 + [11] + [3, 3], np.int32)
    freqs = np.linspace(0.5 * NU1, NU1, 8)
    m = obda.bda_map_mode2(uvw, a0, a1, freqs, NU1 / 100.0, decorr, cell, npix, np.zeros(8, np.int32))
    got = _blocks(m)
    # "10:11" (row 7, span 2.6) first, then "2:10" (rows 0-6) as in the 7-row case
    want = [(c, c + 2, [7]) for c in (0, 2, 4, 6)]
    for rows in ([0, 1], [2, 3, 4], [5, 6]):
        for c0 in (0, 2, 4, 6):
            want.append((c0, c0 + 2, rows))
    assert got == want
