"""The benchmark configurations as concrete problems.

cfg1  plain single facet: 27 antennas in a 20 km disk, 351 baselines, 100 x 10 s, 16 channels
      1.2-1.6 GHz, 512 x 512 grid at 2", Nw = 8, S = 7, OverS = 11, identity Jones, one BDA
      block per (row, channel), 10 point sources + noise, 1% flags.
cfg3  precision study: the cfg1 array with 256 channels over 0.9-1.67 GHz, 3 x 3 facets of 297 px
      at 2" padded to 525 (EstimateNpix of a 301 px facet with Padding 1.7, the cfg1 grid size;
      Nw = 8, S = 7, OverS = 11), 20 point sources at pixel centres 4-12 px from the internal
      facet boundaries, visibilities from the FP64 DFT, no noise, no flags.
cfg4  uneven batched workload: the cfg1 array (27 antennas, 20 km disk, 100 x 10 s, 16 channels
      1.2-1.6 GHz) with baseline-dependent averaging k_t = k_f in {1, 2, 4, 8, 16} by baseline-length
      quintile (longest -> 1, shortest -> 16; blocks cut at the 4-channel band edges, so the
      frequency factor is at most 4), NBand = 4 sub-band grids, 121 facets (11 x 11) of 567 px
      padded to 975 at 2" on a 6237 px image, Nw = 8, S = 7, OverS = 11, no Jones, 10 point
      sources + noise (seeds 4000 + k).
cfg5  differentiable forward model: the cfg1 array at 300 x 10 s, 32 channels 1.2-1.6 GHz in 4
      bands, 25 facets (5 x 5) of 405 px padded to 693 at 2" (image 2025 px), Nw = 8, S = 7,
      OverS = 11, one BDA block per visibility, 100 point sources at pixel centres (log-uniform
      0.01-1 Jy) + complex noise sigma 0.05 (seeds 5000 + k).
cfg2  81-facet MeerKAT-like: 64 antennas (dense 1 km core + 3 arms to 8 km), 2016 baselines,
      600 x 8 s, 64 channels 0.9-1.67 GHz, 4096 x 4096 image at 1.5" in 9 x 9 facets of 455 px
      padded to 735 (EstimateNpix with Padding 1.7), Nw = 32 with supports odd(linspace(7, 15)),
      OverS = 11, per-visibility DD Jones [8 x 81 x 64 x 8] (10-min x 8-channel bins, nearest
      facet direction), BDA maps of the default builder (GridDecorr = DegridDecorr = 0.02, facet
      FoV), 200 power-law point sources + noise, 1% flags.
Seeds: 1000 * cfg + k (k = 0 array, 1 sky, 2 noise, 3 flags, 4 Jones).
"""
import time

import numpy as np

from . import bda as bdamod
from . import jonesmaps
from . import synth
from .problems import Facet, Problem, chan_equidistant, make_cf_facet

C_LIGHT = 299792458.0

# odd fast FFT sizes (factors 3, 5, 7, optionally one 11 or 13), ModToolBox.py:273-293
def _fast_odd_sizes(nlim=200000):
    sizes = np.array([1])
    for base in (3, 5, 7):
        sizes = (sizes[None, :] * base ** np.arange(15)[:, None]).ravel()
    sizes = (sizes[None, :] * np.array([1, 11, 13])[:, None]).ravel()
    return np.array(sorted(sizes[(sizes < nlim) & (sizes > 64)]))


_ODD = _fast_odd_sizes()


def estimate_npix(npix, padding=1.0, min_size_fft=513):
    """EstimateNpix (DDFacet/ToolsDir/ModToolBox.py:64-88): closest odd fast size of the facet,
    then of the padded facet (at least min_size_fft)."""
    n = int(round(npix))
    n = int(_ODD[np.argmin(np.abs(n - _ODD))])
    nopt = n
    n = n * padding
    if n < min_size_fft:
        n = min_size_fft
    n = int(round(n))
    n = int(_ODD[np.argmin(np.abs(n - _ODD))])
    return nopt, n


def cfg1(noise=0.01, log=print):
    t0 = time.time()
    nchan = 16
    freqs = np.linspace(1.2e9, 1.6e9, nchan)
    obs = synth.make_observation("disk", 27, 100, 10.0, freqs, 45.0, -30.7, 1000, diameter_m=20000.0)
    rng = np.random.default_rng(1001)
    cell_as = 2.0
    cell = cell_as * synth.ARCSEC
    npix = 512
    sky = synth.sky_loguniform(10, 0.45 * npix * cell, rng)
    R = obs.nrow
    vis = synth.predict_point_sources(obs.uvw, freqs, sky).numpy()
    nr = np.random.default_rng(1002)
    vis += (noise / np.sqrt(2)) * (nr.standard_normal(vis.shape) + 1j * nr.standard_normal(vis.shape)).astype(np.complex64)
    flags = synth.random_flags(R, nchan, 4, 0.01, 1003)
    weights = np.ones((R, nchan), np.float32)
    wmax = float(np.abs(obs.uvw[:, 2]).max())
    f = make_cf_facet(0.0, 0.0, npix, cell_as, freqs, 8, [7] * 8, 11, wmax)
    bmap = bdamod.bda_map_per_visibility(obs.a0, obs.a1, nchan)
    chan_map = np.zeros(nchan, np.int32)
    log("cfg1 built in %.1fs: %d rows x %d chan" % (time.time() - t0, R, nchan))
    P = Problem(uvw=obs.uvw, vis=vis, flags=flags, weights=weights, freqs=freqs, a0=obs.a0, a1=obs.a1,
                chan_map=chan_map, chan_map_degrid=chan_map.copy(), bda_grid=bmap, bda_degrid=bmap,
                facets=[f], npix=npix, nband=1, npol=1, incr=np.array([-cell, cell]),
                corr_codes=[9, 10, 11, 12], stokes_codes=[1])
    P.extra.update(sky=sky, cell_arcsec=cell_as, image_npix=npix, name="cfg1")
    return P


def cfg2(nfx=9, image_npix=4096, ntime=600, nchan=64, nant=64, sort_by_baseline=True,
         with_jones=True, with_vis=True, device="cpu", log=print, time_frac=1.0, bda_builder=None):
    """The 81-facet MeerKAT-like workload.  `time_frac` < 1 keeps the first fraction of the
    timesteps (for bounded CPU samples); everything else is unchanged.  `bda_builder(uvw, a0, a1,
    freqs, dfreq, decorr, cell, facet_npix, chan_map)` returns the int32 BDA map (default: the
    mode-1 map built on the GPU, bda.bda_map_device)."""
    t0 = time.time()
    freqs = np.linspace(0.9e9, 1.67e9, nchan)
    nt = max(1, int(round(ntime * time_frac)))
    obs = synth.make_observation("meerkat", nant, nt, 8.0, freqs, -30.0, -30.7, 2000,
                                 sort_by_baseline=sort_by_baseline)
    # time steps of the full observation, so a sample is a prefix of the full run
    if nt != ntime:
        full = synth.make_observation("meerkat", nant, ntime, 8.0, freqs, -30.0, -30.7, 2000)
        obs = synth.make_observation("meerkat", nant, ntime, 8.0, freqs, -30.0, -30.7, 2000,
                                     sort_by_baseline=sort_by_baseline)
        keep = obs.tidx < nt
        obs.uvw, obs.a0, obs.a1, obs.tidx = (np.ascontiguousarray(obs.uvw[keep]), obs.a0[keep].copy(),
                                             obs.a1[keep].copy(), obs.tidx[keep].copy())
        del full
    R = obs.nrow
    cell_as = 1.5
    cell = cell_as * synth.ARCSEC
    facet_px = image_npix / nfx
    nopt, npad = estimate_npix(facet_px, 1.7)
    log("cfg2: %d rows x %d chan, facets %d -> %d padded" % (R, nchan, nopt, npad))
    cen = synth.facet_centres(nfx, nopt, cell)
    wmax = float(np.abs(obs.uvw[:, 2]).max())
    sups = [s if s % 2 else s + 1 for s in np.int64(np.linspace(7, 15, 32))]
    facets = []
    for i in range(nfx * nfx):
        f = make_cf_facet(float(cen[i, 0]), float(cen[i, 1]), npad, cell_as, freqs, 32, sups, 11, wmax)
        f.kms_dir = i
        facets.append(f)
    log("cfg2: CF tables in %.1fs" % (time.time() - t0))
    chan_map = np.zeros(nchan, np.int32)
    if bda_builder is None:
        bmap = bdamod.bda_map_device(obs.uvw, obs.a0, obs.a1, freqs, obs.dfreq, 0.02, cell, nopt, chan_map).cpu().numpy()
    else:
        bmap = bda_builder(obs.uvw, obs.a0, obs.a1, freqs, obs.dfreq, 0.02, cell, nopt, chan_map)
    log("cfg2: BDA map (%d blocks, %d words) in %.1fs" % (bdamod.nblocks_of(bmap), bmap.size, time.time() - t0))
    jones = None
    if with_jones:
        # 10-min solution intervals and 8-channel solution domains, mapped as the Jones server's
        # inputs are (ClassJones.GiveTimeMapping / GiveVisToJonesChanMapping)
        ntj = 8
        tsec = obs.tidx * 8.0
        tmap = jonesmaps.time_mapping(600.0 * np.arange(ntj), 600.0 * np.arange(1, ntj + 1), tsec)
        dfc = (freqs[-1] - freqs[0]) / (nchan - 1)
        doms = np.stack([freqs[0::8] - dfc / 2, freqs[7::8] + dfc / 2], 1)
        cmap = jonesmaps.vis_to_jones_chan(freqs, doms).astype(np.int32)
        jones = dict(mode=1, interp=0, apply_amp=1, apply_phase=1, scale=0, calib_error=0.0,
                     kms=synth.random_jones(ntj, nfx * nfx, nant, nchan // 8, 2004),
                     kms_time=tmap, kms_chan=cmap)
    rng = np.random.default_rng(2001)
    sky = synth.sky_powerlaw(200, 0.5 * image_npix * cell, rng)
    vis = None
    if with_vis:
        vis = cfg2_visibilities(obs.uvw, freqs, sky, device)
        log("cfg2: visibilities in %.1fs" % (time.time() - t0))
    flags = synth.random_flags(R, nchan, 4, 0.01, 2003)
    weights = np.ones((R, nchan), np.float32)
    P = Problem(uvw=obs.uvw, vis=vis, flags=flags, weights=weights, freqs=freqs, a0=obs.a0, a1=obs.a1,
                chan_map=chan_map, chan_map_degrid=chan_map.copy(), bda_grid=bmap, bda_degrid=bmap,
                facets=facets, npix=npad, nband=1, npol=1, incr=np.array([-cell, cell]),
                corr_codes=[9, 10, 11, 12], stokes_codes=[1], jones=jones)
    P.extra.update(sky=sky, cell_arcsec=cell_as, image_npix=image_npix, facet_npix=nopt, name="cfg2",
                   nfx=nfx, equidistant=chan_equidistant(freqs))
    return P


def cfg2_visibilities(uvw, freqs, sky, device="cpu"):
    """The 81-facet workload's visibilities: FP32 point-source predict plus complex Gaussian
    noise sigma 0.01 (seed 2002), on `device` (a torch tensor there, numpy on the CPU)."""
    import torch
    v = synth.predict_point_sources(uvw, freqs, sky, device=device, chunk_rows=4096)
    g = torch.Generator(device=device).manual_seed(2002)
    v += torch.randn(v.shape, generator=g, device=device, dtype=torch.complex64) * 0.01
    return v if device != "cpu" else v.numpy()


def small_cycle(seed=5, nfx=3, m=243, n=147, cell_arcsec=20.0, nant=7, ntime=10, nchan=6, nsrc=6,
                jones=False, nw=4, sups=(7, 7, 9, 9), overs=11, flag_frac=0.0, bda="pervis", nband=1):
    """A small faceted major cycle: nfx x nfx facets on an m x m image (padded facets n px),
    point sources at pixel centres, visibilities from the FP64 DFT.  Returns (Problem, geometry)."""
    from .imaging import FacetGeometry
    cell = cell_arcsec * synth.ARCSEC
    geom = FacetGeometry(m, nfx, m // nfx, n, cell, mixing_width=4.0)
    freqs = np.linspace(1.2e9, 1.5e9, nchan)
    umax = 0.35 / cell
    diam = umax * C_LIGHT / freqs.max()
    obs = synth.make_observation("disk", nant, ntime, 900.0, freqs, 45.0, -30.7, seed, diameter_m=diam)
    rng = np.random.default_rng(seed + 1)
    X = rng.integers(m // 8, m - m // 8, nsrc)
    Y = rng.integers(m // 8, m - m // 8, nsrc)
    flux = rng.uniform(0.5, 2.0, nsrc)
    l = (X - m // 2) * cell
    mm = (Y - m // 2) * cell
    sky = synth.Sky(flux, l, mm)
    vis = synth.predict_point_sources(obs.uvw, freqs, sky).numpy()
    R = obs.nrow
    flags = synth.random_flags(R, nchan, 4, flag_frac, seed + 3)
    weights = np.ones((R, nchan), np.float32)
    wmax = float(np.abs(obs.uvw[:, 2]).max()) + 1e-3
    facets = []
    for i in range(geom.nf):
        l0, m0 = geom.lm[i]
        f = make_cf_facet(l0, m0, n, cell_arcsec, freqs, nw, list(sups), overs, wmax)
        f.kms_dir = i
        facets.append(f)
    chan_map = (np.arange(nchan) * nband // nchan).astype(np.int32)
    if bda == "pervis":
        bmap = bdamod.bda_map_per_visibility(obs.a0, obs.a1, nchan)
    else:
        bmap = bdamod.bda_map_factors(obs.a0, obs.a1, nchan, lambda key: (2, 2), chan_map=chan_map)
    jd = None
    if jones:
        jd = dict(mode=1, interp=0, apply_amp=1, apply_phase=1, scale=0, calib_error=0.0,
                  kms=synth.random_jones(2, geom.nf, nant, 2, seed + 4, 0.05, 0.2),
                  kms_time=np.minimum((obs.tidx * 2) // ntime, 1).astype(np.int32),
                  kms_chan=np.minimum((np.arange(nchan) * 2) // nchan, 1).astype(np.int32))
    P = Problem(uvw=obs.uvw, vis=vis, flags=flags, weights=weights, freqs=freqs, a0=obs.a0, a1=obs.a1,
                chan_map=chan_map, chan_map_degrid=chan_map.copy(), bda_grid=bmap, bda_degrid=bmap,
                facets=facets, npix=n, nband=nband, npol=1, incr=np.array([-cell, cell]),
                corr_codes=[9, 10, 11, 12], stokes_codes=[1], jones=jd)
    P.extra.update(sky=sky, src_pix=(X, Y), cell=cell, image_npix=m, name="small_cycle")
    return P, geom


def cfg3(nsrc=20, nfx=3, facet=297, n=525, nchan=256, cell_arcsec=2.0, chan_block=1, log=print):
    """The precision-study workload (see the module docstring).  Returns (Problem, geometry).
    chan_block > 1 replaces the per-visibility blocks by blocks of one row and chan_block
    channels (the gridders then walk chan_block channels per block)."""
    from .imaging import FacetGeometry
    t0 = time.time()
    freqs = np.linspace(0.9e9, 1.67e9, nchan)
    obs = synth.make_observation("disk", 27, 100, 10.0, freqs, 45.0, -30.7, 3000, diameter_m=20000.0)
    cell = cell_arcsec * synth.ARCSEC
    m = nfx * facet
    geom = FacetGeometry(m, nfx, facet, n, cell, mixing_width=4.0)
    rng = np.random.default_rng(3001)
    # sources 4-12 px from an internal facet boundary (or from the image edge with one facet)
    bounds = [facet * k for k in range(1, nfx)] or [0, m]
    b = np.array(bounds)[rng.integers(0, len(bounds), nsrc)]
    side = np.where(b == 0, 1, np.where(b == m, -1, rng.choice([-1, 1], nsrc)))
    across = b + side * rng.integers(4, 13, nsrc) - (side < 0)
    along = rng.integers(12, m - 12, nsrc)
    vertical = rng.integers(0, 2, nsrc).astype(bool)
    X = np.where(vertical, across, along)
    Y = np.where(vertical, along, across)
    flux = rng.uniform(0.5, 2.0, nsrc)
    sky = synth.Sky(flux, (X - m // 2) * cell, (Y - m // 2) * cell)
    vis = synth.predict_point_sources(obs.uvw, freqs, sky).numpy()
    R = obs.nrow
    flags = np.zeros((R, nchan, 4), np.uint8)
    weights = np.ones((R, nchan), np.float32)
    wmax = float(np.abs(obs.uvw[:, 2]).max()) + 1e-3
    facets = []
    for i in range(geom.nf):
        l0, m0 = geom.lm[i]
        f = make_cf_facet(l0, m0, n, cell_arcsec, freqs, 8, [7] * 8, 11, wmax)
        f.kms_dir = i
        facets.append(f)
    if chan_block > 1:
        bmap = bdamod.bda_map_factors(obs.a0, obs.a1, nchan, lambda key: (1, chan_block))
    else:
        bmap = bdamod.bda_map_per_visibility(obs.a0, obs.a1, nchan)
    chan_map = np.zeros(nchan, np.int32)
    P = Problem(uvw=obs.uvw, vis=vis, flags=flags, weights=weights, freqs=freqs, a0=obs.a0, a1=obs.a1,
                chan_map=chan_map, chan_map_degrid=chan_map.copy(), bda_grid=bmap, bda_degrid=bmap,
                facets=facets, npix=n, nband=1, npol=1, incr=np.array([-cell, cell]),
                corr_codes=[9, 10, 11, 12], stokes_codes=[1])
    P.extra.update(sky=sky, src_pix=(X, Y), cell=cell, image_npix=m, name="cfg3",
                   equidistant=chan_equidistant(freqs))
    log("cfg3 built in %.1fs: %d rows x %d chan, %d facets" % (time.time() - t0, R, nchan, geom.nf))
    return P, geom


def cfg4(nfx=11, image_npix=6144, cell_arcsec=2.0, nband=4, noise=0.01, log=print):
    """The uneven batched workload (see the module docstring)."""
    t0 = time.time()
    nchan = 16
    freqs = np.linspace(1.2e9, 1.6e9, nchan)
    obs = synth.make_observation("disk", 27, 100, 10.0, freqs, 45.0, -30.7, 4000, diameter_m=20000.0,
                                 sort_by_baseline=True)
    cell = cell_arcsec * synth.ARCSEC
    nopt, npad = estimate_npix(image_npix / nfx, 1.7)
    # averaging factor per baseline by length quintile (longest -> 1, shortest -> 16)
    keys = ["%d:%d" % k for k in zip(obs.a0.tolist(), obs.a1.tolist())]
    bl_len = {}
    for k, q in zip(keys, np.hypot(obs.uvw[:, 0], obs.uvw[:, 1])):
        bl_len.setdefault(k, []).append(q)
    lens = {k: float(np.mean(v)) for k, v in bl_len.items()}
    edges = np.quantile(np.array(list(lens.values())), [0.2, 0.4, 0.6, 0.8])
    factors = (16, 8, 4, 2, 1)
    fac = {k: factors[int(np.searchsorted(edges, L, side="right"))] for k, L in lens.items()}
    chan_map = (np.arange(nchan) * nband // nchan).astype(np.int32)
    bmap = bdamod.bda_map_factors(obs.a0, obs.a1, nchan, lambda key: (fac[key], fac[key]), chan_map=chan_map)
    wmax = float(np.abs(obs.uvw[:, 2]).max()) + 1e-3
    cen = synth.facet_centres(nfx, nopt, cell)
    facets = []
    for i in range(nfx * nfx):
        f = make_cf_facet(float(cen[i, 0]), float(cen[i, 1]), npad, cell_arcsec, freqs, 8, [7] * 8, 11, wmax)
        f.kms_dir = i
        facets.append(f)
    rng = np.random.default_rng(4001)
    sky = synth.sky_loguniform(10, 0.45 * nfx * nopt * cell, rng)
    vis = synth.predict_point_sources(obs.uvw, freqs, sky).numpy()
    nr = np.random.default_rng(4002)
    vis += (noise / np.sqrt(2)) * (nr.standard_normal(vis.shape) + 1j * nr.standard_normal(vis.shape)).astype(np.complex64)
    R = obs.nrow
    flags = synth.random_flags(R, nchan, 4, 0.01, 4003)
    weights = np.ones((R, nchan), np.float32)
    P = Problem(uvw=obs.uvw, vis=vis, flags=flags, weights=weights, freqs=freqs, a0=obs.a0, a1=obs.a1,
                chan_map=chan_map, chan_map_degrid=chan_map.copy(), bda_grid=bmap, bda_degrid=bmap,
                facets=facets, npix=npad, nband=nband, npol=1, incr=np.array([-cell, cell]),
                corr_codes=[9, 10, 11, 12], stokes_codes=[1])
    counts = np.bincount(np.array([fac[k] for k in keys]), minlength=17)[[1, 2, 4, 8, 16]]
    P.extra.update(sky=sky, cell_arcsec=cell_arcsec, image_npix=nfx * nopt, facet_npix=nopt, name="cfg4",
                   nfx=nfx, equidistant=chan_equidistant(freqs), rows_per_factor=counts.tolist())
    log("cfg4: %d rows x %d chan, %d facets %d -> %d px, %d BDA blocks, built in %.1fs"
        % (R, nchan, nfx * nfx, nopt, npad, bdamod.nblocks_of(bmap), time.time() - t0))
    return P


def cfg5(nfx=5, image_npix=2048, cell_arcsec=2.0, nband=4, ntime=300, nchan=32, nsrc=100, noise=0.05,
         device="cpu", log=print):
    """The differentiable forward-model workload (see the module docstring).  Returns
    (Problem, geometry); the visibilities live on `device`."""
    from .imaging import FacetGeometry
    t0 = time.time()
    freqs = np.linspace(1.2e9, 1.6e9, nchan)
    obs = synth.make_observation("disk", 27, ntime, 10.0, freqs, 45.0, -30.7, 5000, diameter_m=20000.0)
    cell = cell_arcsec * synth.ARCSEC
    nopt, npad = estimate_npix(image_npix / nfx, 1.7)
    m = nfx * nopt
    geom = FacetGeometry(m, nfx, nopt, npad, cell, mixing_width=10.0)
    rng = np.random.default_rng(5001)
    X = rng.integers(m // 8, m - m // 8, nsrc)
    Y = rng.integers(m // 8, m - m // 8, nsrc)
    flux = np.exp(rng.uniform(np.log(0.01), np.log(1.0), nsrc))
    sky = synth.Sky(flux, (X - m // 2) * cell, (Y - m // 2) * cell)
    import torch
    v = synth.predict_point_sources(obs.uvw, freqs, sky, device=device, chunk_rows=8192)
    g = torch.Generator(device=device).manual_seed(5002)
    v += torch.randn(v.shape, generator=g, device=device, dtype=torch.complex64) * noise
    vis = v if device != "cpu" else v.numpy()
    R = obs.nrow
    flags = np.zeros((R, nchan, 4), np.uint8)
    weights = np.ones((R, nchan), np.float32)
    wmax = float(np.abs(obs.uvw[:, 2]).max()) + 1e-3
    facets = []
    for i in range(geom.nf):
        l0, m0 = geom.lm[i]
        f = make_cf_facet(l0, m0, npad, cell_arcsec, freqs, 8, [7] * 8, 11, wmax)
        f.kms_dir = i
        facets.append(f)
    chan_map = (np.arange(nchan) * nband // nchan).astype(np.int32)
    bmap = bdamod.bda_map_per_visibility(obs.a0, obs.a1, nchan)
    P = Problem(uvw=obs.uvw, vis=vis, flags=flags, weights=weights, freqs=freqs, a0=obs.a0, a1=obs.a1,
                chan_map=chan_map, chan_map_degrid=chan_map.copy(), bda_grid=bmap, bda_degrid=bmap,
                facets=facets, npix=npad, nband=nband, npol=1, incr=np.array([-cell, cell]),
                corr_codes=[9, 10, 11, 12], stokes_codes=[1])
    P.extra.update(sky=sky, src_pix=(X, Y), cell=cell, image_npix=m, name="cfg5",
                   equidistant=chan_equidistant(freqs))
    log("cfg5: %d rows x %d chan, %d facets %d -> %d px, image %d, built in %.1fs"
        % (R, nchan, geom.nf, nopt, npad, m, time.time() - t0))
    return P, geom
