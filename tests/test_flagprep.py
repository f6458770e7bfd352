"""Flag preparation (ClassMS.py:1155-1167) against hand-worked cases."""
import numpy as np

from oracle.flagprep import prepare_flags


def test_autocorrelations_nan_and_equalisation():
    a0 = np.array([0, 1, 2])
    a1 = np.array([1, 1, 0])
    flags = np.zeros((3, 2, 4), bool)
    flags[0, 1, 2] = True                  # one correlation flagged -> the whole (row 0, chan 1)
    data = np.ones((3, 2, 4), np.complex64)
    data[2, 0, 3] = np.nan                 # NaN -> flagged, then equalised over (row 2, chan 0)
    prepare_flags(flags, a0, a1, data)
    want = np.zeros((3, 2), bool)
    want[0, 1] = True
    want[1, :] = True                      # autocorrelation row
    want[2, 0] = True
    assert np.array_equal(flags, np.repeat(want[:, :, None], 4, axis=2))
    # 1e9 goes where flags stood before the equalisation (the reference order): the
    # autocorrelation row, the NaN and the originally flagged correlation
    big = np.zeros((3, 2, 4), bool)
    big[1] = True
    big[2, 0, 3] = True
    big[0, 1, 2] = True
    assert np.all(data[big] == 1e9) and np.all(data[~big] == 1)


def test_without_data_only_flags_change():
    flags = np.zeros((2, 1, 2), bool)
    flags[1, 0, 0] = True
    out = prepare_flags(flags, np.array([0, 0]), np.array([1, 2]))
    assert out is flags
    assert flags.tolist() == [[[False, False]], [[True, True]]]
