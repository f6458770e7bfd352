"""Jones time / channel maps (ClassJones.py:415-439, 1004-1012) against hand-worked cases."""
import numpy as np

from ddfacet_amd import jonesmaps


def test_time_mapping_intervals_half_open_and_unsorted_times():
    t0 = [0.0, 10.0, 20.0]
    t1 = [10.0, 20.0, 30.0]
    times = [25.0, 0.0, 9.999, 10.0, 19.0, 20.0, 29.0]
    assert jonesmaps.time_mapping(t0, t1, times).tolist() == [2, 0, 0, 1, 1, 2, 2]


def test_time_mapping_later_interval_wins_and_outside_rows_keep_zero():
    # intervals 0: [0, 20), 1: [10, 30): times in [10, 20) take 1; 30 and -1 are outside: 0
    t0, t1 = [0.0, 10.0], [20.0, 30.0]
    assert jonesmaps.time_mapping(t0, t1, [5.0, 15.0, 25.0, 30.0, -1.0]).tolist() == [0, 1, 1, 0, 0]
    assert jonesmaps.time_mapping(t0, t1, []).dtype == np.int32


def test_vis_to_jones_chan_nearest_centre_first_on_tie():
    # domains centred at 1.5, 3.5, 5.5
    dom = [[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]]
    # 2.5 is equidistant from 1.5 and 3.5: the first (0)
    freqs = [0.0, 1.6, 2.5, 2.6, 4.4, 4.6, 9.0]
    assert jonesmaps.vis_to_jones_chan(freqs, dom).tolist() == [0, 0, 0, 1, 1, 2, 2]


def test_eight_channel_domains_give_integer_division():
    freqs = 0.9e9 + 12.2e6 * np.arange(64)
    dom = np.stack([freqs[0::8] - 6.1e6, freqs[7::8] + 6.1e6], 1)
    assert np.array_equal(jonesmaps.vis_to_jones_chan(freqs, dom), np.arange(64) // 8)
