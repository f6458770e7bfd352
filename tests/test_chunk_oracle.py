"""The chunk reader's row sort and flag selections (oracle/chunk.py) against hand-worked cases:
rows ordered by antenna pair, then time, then row (ClassMS.py:575-586); antenna flagged fractions
counted on the equalised flags before the uv-range cut, and the 90 % rule; uv range on
u^2 + v^2 against km^2 1e6; time range; named antennas (ClassMS.py:1155-1233)."""
import numpy as np

from oracle import chunk


def test_sort_index_pairs_then_time_then_row():
    a0 = [1, 0, 0, 1, 0, 0]
    a1 = [2, 1, 1, 2, 2, 1]
    t = [5.0, 7.0, 3.0, 1.0, 9.0, 3.0]
    # (0,1): rows 2 (t3), 5 (t3, later row), 1 (t7); (0,2): row 4; (1,2): rows 3 (t1), 0 (t5)
    assert chunk.sort_index(a0, a1, t).tolist() == [2, 5, 1, 4, 3, 0]


def test_flag_fractions_use_the_flags_before_the_uv_cut():
    nchan = 4
    a0 = np.array([0, 0, 1, 2])
    a1 = np.array([1, 2, 2, 2])
    uvw = np.array([[100.0, 0, 0], [3000.0, 0, 0], [100.0, 0, 0], [0.0, 0, 0]])
    flags = np.zeros((4, nchan, 2), bool)
    flags[2, :, 1] = True            # row 2 (1-2): every channel flagged on one correlation
    flags[0, 0, 0] = True            # row 0 (0-1): one channel
    frac = chunk.update_flags(flags, uvw, None, a0, a1, 3, uv_range_km=(0.05, 1.0))
    # flags1 before the cut: row 0 one cell, row 2 four cells, row 3 autocorrelation four cells
    # ant 0: rows 0, 1 -> 1 / 8; ant 1: rows 0, 2 -> 5 / 8; ant 2: rows 1, 2, 3 -> 8 / 12
    np.testing.assert_allclose(frac, [1 / 8, 5 / 8, 8 / 12])
    assert flags[1].all()            # |u| = 3 km outside [0.05, 1] km
    assert flags[2].all() and flags[3].all() and not flags[0, 1:].any() and flags[0, 0].all()


def test_ninety_percent_rule_and_named_antennas():
    nchan = 10
    a0 = np.array([0, 0, 1])
    a1 = np.array([1, 2, 2])
    uvw = np.full((3, 3), 100.0)
    flags = np.zeros((3, nchan, 1), bool)
    flags[0, :, 0] = True
    flags[1, :9, 0] = True
    # ant 0: rows 0, 1 -> 19 / 20 = 0.95 > 0.9 -> all its rows flagged; ant 1: 10 / 20; ant 2: 9 / 20
    frac = chunk.update_flags(flags, uvw, None, a0, a1, 3)
    np.testing.assert_allclose(frac, [0.95, 0.5, 0.45])
    assert flags[0].all() and flags[1].all() and not flags[2].any()
    flags = np.zeros((3, nchan, 1), bool)
    chunk.update_flags(flags, uvw, None, a0, a1, 3, named_antennas=[2])
    assert not flags[0].any() and flags[1].all() and flags[2].all()


def test_time_range_and_nan_values():
    a0 = np.array([0, 0, 0])
    a1 = np.array([1, 1, 1])
    uvw = np.full((3, 3), 10.0)
    flags = np.zeros((3, 2, 2), bool)
    data = np.ones((3, 2, 2), np.complex64)
    data[2, 1, 0] = np.nan
    chunk.update_flags(flags, uvw, data, a0, a1, 2, times_unix=np.array([10.0, 20.0, 30.0]), time_range=(15.0, 30.0))
    assert flags[0].all() and not flags[1].any()
    # the 1e9 fill comes before the equalisation: the NaN cell only
    assert flags[2, 1].all() and not flags[2, 0].any() and data[2, 1, 0] == 1e9 and data[2, 1, 1] == 1
