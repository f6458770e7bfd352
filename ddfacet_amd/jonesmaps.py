"""Row -> Jones-time and channel -> Jones-channel maps (the integer inputs of the Jones server).

``time_mapping`` restates ``ClassJones.GiveTimeMapping`` (DDFacet/Data/ClassJones.py:415-439): every
solution interval [t0, t1) assigns its index to the rows whose time falls in it, intervals taken
in order, so where intervals overlap the later one wins, and rows outside every interval keep 0.
``vis_to_jones_chan`` restates ``ClassJones.GiveVisToJonesChanMapping`` (ClassJones.py:1004-1012):
each visibility channel maps to the Jones channel whose frequency-domain centre is nearest (the
first such channel on a tie).  Both run once per chunk on the host.
"""
import numpy as np


def time_mapping(t0, t1, times):
    """int32 [nrow]: Jones time index of each row."""
    times = np.asarray(times, np.float64)
    ind = np.zeros(times.size, np.int32)
    for it, (a, b) in enumerate(zip(np.asarray(t0, np.float64), np.asarray(t1, np.float64))):
        ind[(times >= a) & (times < b)] = it
    return ind


def vis_to_jones_chan(chan_freqs, freq_domains):
    """int [nchan]: Jones channel of each visibility channel; freq_domains [nchan_jones, 2]."""
    dom = np.asarray(freq_domains, np.float64)
    centre = (dom[:, 0] + dom[:, 1]) / 2.0
    d = np.abs(np.asarray(chan_freqs, np.float64).reshape(-1, 1) - centre.reshape(1, -1))
    return np.argmin(d, axis=1)
