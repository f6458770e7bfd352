"""TEST INFRASTRUCTURE: numpy restatement of the chunk reader's row sort and flag selections, the
checker of ddfacet_amd/chunk.py (ddf_sort_baseline_time, ddf_flag_selection).

sort_index    ClassMS.ReadData (DDFacet/Data/ClassMS.py:575-586): sorted(zip(A0, A1, TIME, rows))
update_flags  ClassMS.UpdateFlags (ClassMS.py:1155-1233): autocorrelations, NaNs (values set to
              1e9), equalisation over correlations, then the uv-range cut, the time-range cut
              (on unix times), and every row of an antenna whose flagged fraction, counted on the
              equalised flags before the cuts, exceeds the threshold, or of a named antenna.
"""
import numpy as np


def sort_index(a0, a1, times):
    return np.array([s[3] for s in sorted(zip(np.asarray(a0).tolist(), np.asarray(a1).tolist(),
                                                np.asarray(times).tolist(), range(len(a0))))], np.int64)


def update_flags(flags, uvw, data, a0, a1, na, times_unix=None, uv_range_km=None, time_range=None,
                 named_antennas=(), threshold=0.9):
    """flags bool [nrow, nchan, ncorr] (in place); returns the per-antenna flagged fractions."""
    a0 = np.asarray(a0)
    a1 = np.asarray(a1)
    flags[a0 == a1] = True
    if data is not None:
        flags[np.isnan(data)] = True
        data[flags] = 1e9
    flags1 = flags.any(axis=2)
    flags[flags1] = True
    ants = set()
    if uv_range_km:
        d0, d1 = uv_range_km
        d0 = d0 ** 2 * 1e6
        d1 = d1 ** 2 * 1e6
        duv = (uvw[:, :2] ** 2).sum(1)
        flags[(duv < d0) | (duv > d1), :, :] = True
    if time_range is not None:
        st0, st1 = time_range
        sel = np.logical_or(times_unix < st0, times_unix > st1)
        flags[sel, :, :] = True
    antenna_rows = [(a0 == A) | (a1 == A) for A in range(na)]
    frac = [flags1[rows].sum() / float(flags1[rows].size or 1) for rows in antenna_rows]
    ants.update(A for A, f in enumerate(frac) if f > threshold)
    ants.update(int(A) for A in named_antennas)
    for A in ants:
        flags[antenna_rows[A], :, :] = True
    return np.array(frac)
