"""TEST INFRASTRUCTURE: numpy restatement of the chunk reader's flag preparation
(DDFacet/Data/ClassMS.py:1155-1167), the checker of ddf_prepare_flags and the CPU baseline of
its bench workload.

The gridder tests the flag of correlation 0 only (gridder.h:229); that is valid because the chunk
reader flags autocorrelations, flags NaN visibilities (and overwrites every flagged value with
1e9), and then flags all correlations of a (row, channel) when any of them is flagged.
"""
import numpy as np


def prepare_flags(flags, a0, a1, data=None):
    """flags: bool [nrow, nchan, ncorr], updated in place and returned; data: complex
    [nrow, nchan, ncorr] or None (NaNs flagged, then the values flagged so far set to 1e9, in
    place, before the equalisation over correlations, in the reference's order)."""
    flags[np.asarray(a0) == np.asarray(a1)] = True
    if data is not None:
        flags[np.isnan(data)] = True
        data[flags] = 1e9
    flags[flags.any(axis=2)] = True
    return flags
