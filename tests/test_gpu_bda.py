"""BDA block maps built on the device (bda.bda_map_device, ddf_bda_build_*) against the host
builders (oracle/bda.py: bda_map_mode1 / bda_map_mode2, pinned by the hand-worked cases of
tests/test_bda_maps.py): the int32 maps must be identical word for word.

Cases: arrays of 12 and 27 antennas (string key order differs from numeric order), rows in
random order with autocorrelations mixed in, baselines of one row, bands that cut channel blocks,
decorrelation settings from one channel per block to the whole band, rows at uvw = 0, only
autocorrelations (an empty map), and the 81-facet MeerKAT-like chunk (1.2 M rows, 5.5 M blocks).
"""
import numpy as np
import pytest

from ddfacet_amd import bda
from oracle import bda as obda

pytestmark = pytest.mark.gpu


def _random_chunk(seed, na=12, ntime=40, nchan=16, nband=3, shuffle=True, zero_rows=0, radius=3000.0, dt=30.0,
                  band=(0.9e9, 1.67e9)):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-radius, radius, (na, 3))
    pos[:, 2] *= 0.1
    a0, a1, t = [], [], []
    for ti in range(ntime):
        for p in range(na):
            for q in range(p, na):
                if rng.random() < 0.9 or p == q:
                    a0.append(p)
                    a1.append(q)
                    t.append(ti)
    a0, a1, t = np.array(a0, np.int32), np.array(a1, np.int32), np.array(t)
    # a few rows with the antennas swapped: (q, p) is a baseline of its own
    sw = rng.random(a0.size) < 0.05
    a0[sw], a1[sw] = a1[sw].copy(), a0[sw].copy()
    ha = t * (2 * np.pi / 86400 * dt)
    b = pos[a1] - pos[a0]
    uvw = np.stack([b[:, 0] * np.cos(ha) - b[:, 1] * np.sin(ha), b[:, 0] * np.sin(ha) + b[:, 1] * np.cos(ha),
                    b[:, 2] + 0.3 * b[:, 0] * np.sin(ha)], 1)
    if zero_rows:
        uvw[rng.choice(uvw.shape[0], zero_rows, replace=False)] = 0.0
    if shuffle:
        perm = rng.permutation(a0.size)
        uvw, a0, a1 = uvw[perm], a0[perm], a1[perm]
    # a single-row baseline: drop all but one row of one pair
    p, q = 1, 2
    sel = np.nonzero((a0 == p) & (a1 == q))[0]
    keep = np.ones(a0.size, bool)
    keep[sel[1:]] = False
    uvw, a0, a1 = np.ascontiguousarray(uvw[keep]), a0[keep], a1[keep]
    freqs = np.linspace(band[0], band[1], nchan)
    dfreq = freqs[1] - freqs[0]
    chan_map = (np.arange(nchan) * nband // nchan).astype(np.int32)
    return uvw, a0, a1, freqs, dfreq, chan_map


CASES = [  # seed, antennas, decorrelation, facet px, cell ("), array radius (m), time step (s), band (Hz)
    (1, 12, 0.02, 455, 1.5, 3000.0, 30.0, (0.9e9, 1.67e9)),
    (2, 12, 0.2, 1500, 4.0, 1000.0, 2.0, (1.2e9, 1.3e9)),
    (3, 27, 0.002, 800, 2.0, 500.0, 1.0, (1.2e9, 1.25e9)),
    (4, 12, 0.6, 3000, 30.0, 300.0, 4.0, (1.3e9, 1.4e9)),
    (5, 27, 0.02, 50, 1.0, 3000.0, 8.0, (0.9e9, 1.67e9)),
    (6, 12, 0.05, 455, 1.5, 200.0, 10.0, (1.0e9, 1.1e9)),
]


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("case", CASES)
def test_device_map_equals_host_builder(mode, case):
    seed, na, decorr, npix, cell_as, radius, dt, band = case
    print("case", mode, case, flush=True)
    uvw, a0, a1, freqs, dfreq, chan_map = _random_chunk(seed, na=na, zero_rows=3 if mode == 2 else 0, radius=radius,
                                                        dt=dt, band=band)
    cell = cell_as * np.pi / 180 / 3600
    host = (obda.bda_map_mode1 if mode == 1 else obda.bda_map_mode2)(uvw, a0, a1, freqs, dfreq, decorr, cell, npix,
                                                                    chan_map)
    dev = bda.bda_map_device(uvw, a0, a1, freqs, dfreq, decorr, cell, npix, chan_map, mode=mode).cpu().numpy()
    print("blocks %d, words %d" % (bda.nblocks_of(host), host.size), flush=True)
    assert bda.nblocks_of(host) > 0
    np.testing.assert_array_equal(dev, host)


def test_only_autocorrelations_give_an_empty_map():
    uvw, a0, a1, freqs, dfreq, chan_map = _random_chunk(6)
    auto = a0 == a1
    uvw, a0, a1 = np.ascontiguousarray(uvw[auto]), a0[auto], a1[auto]
    for mode in (1, 2):
        dev = bda.bda_map_device(uvw, a0, a1, freqs, dfreq, 0.02, 1e-5, 500, chan_map, mode=mode).cpu().numpy()
        assert dev.tolist() == [0, 0]


def test_device_map_of_the_81_facet_chunk():
    """The 81-facet MeerKAT-like chunk's grid map (mode 1, GridDecorr 0.02, facet field of view):
    1.2 M rows of 2016 baselines, 64 channels."""
    import time

    import torch

    from ddfacet_amd import configs, synth
    freqs = np.linspace(0.9e9, 1.67e9, 64)
    obs = synth.make_observation("meerkat", 64, 600, 8.0, freqs, -30.0, -30.7, 2000, sort_by_baseline=True)
    cell = 1.5 * synth.ARCSEC
    nopt, _ = configs.estimate_npix(4096 / 9, 1.7)
    chan_map = np.zeros(64, np.int32)
    t0 = time.perf_counter()
    host = obda.bda_map_mode1(obs.uvw, obs.a0, obs.a1, freqs, obs.dfreq, 0.02, cell, nopt, chan_map)
    t_host = time.perf_counter() - t0
    t_uvw = torch.from_numpy(np.ascontiguousarray(obs.uvw)).cuda()
    t_a0 = torch.from_numpy(np.int32(obs.a0)).cuda()
    t_a1 = torch.from_numpy(np.int32(obs.a1)).cuda()
    t_cm = torch.from_numpy(chan_map).cuda()
    bda.bda_map_device(t_uvw, t_a0, t_a1, freqs, obs.dfreq, 0.02, cell, nopt, t_cm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev = bda.bda_map_device(t_uvw, t_a0, t_a1, freqs, obs.dfreq, 0.02, cell, nopt, t_cm)
    t_dev = time.perf_counter() - t0
    print("81-facet chunk map: %d blocks, %d words; host %.2f s, device %.1f ms"
          % (bda.nblocks_of(host), host.size, t_host, t_dev * 1e3), flush=True)
    np.testing.assert_array_equal(dev.cpu().numpy(), host)
