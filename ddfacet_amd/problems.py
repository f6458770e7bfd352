"""Gridding problems assembled as the reference call sites assemble them.

A ``Problem`` holds one observation chunk (uvw, vis, flags, weights, channel maps, BDA maps,
Jones tables) plus per-facet geometry and CF tables, and produces the positional argument lists
of pyGridderWPol32|64 / pyDeGridderWPol exactly as ClassDDEGridMachine.put / .get build them
(DDFacet/Imager/ClassDDEGridMachine.py:681-818, 922-1108; Jones list :571-679, 759-776;
FacetInfos :748-749; ChanEquidistant :711-716).
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import bda as bdamod
from . import synth
from .wterm import ClassWTermModified

C_LIGHT = 299792458.0


def chan_equidistant(freqs):
    if freqs.size > 1:
        df = freqs[1:] - freqs[:-1]
        return int(np.max(np.abs(df - np.mean(df))) < 1.)
    return 0


@dataclass
class Facet:
    l0: float
    m0: float
    cu: float
    cv: float
    wplanes: list
    wplanes_conj: list
    winfos: np.ndarray
    sphe: Optional[np.ndarray] = None
    kms_dir: int = 0
    beam_dir: int = 0


@dataclass
class Problem:
    uvw: np.ndarray
    vis: np.ndarray
    flags: np.ndarray
    weights: np.ndarray
    freqs: np.ndarray
    a0: np.ndarray
    a1: np.ndarray
    chan_map: np.ndarray
    chan_map_degrid: np.ndarray
    bda_grid: np.ndarray
    bda_degrid: np.ndarray
    facets: List[Facet]
    npix: int
    nband: int
    npol: int
    incr: np.ndarray
    corr_codes: list
    stokes_codes: list
    jones: Optional[dict] = None       # kms/beam tables, maps, mode, interp, coefs, flags
    decorr: Optional[tuple] = None     # (uvw_dt, DT, Dnu, tsmear, fsmear) ; l,m from facet
    sparsification: Optional[np.ndarray] = None
    extra: dict = field(default_factory=dict)

    # ---- argument lists ------------------------------------------------------------------
    def _ljones(self, facet, nband, sum_jones, sum_jones_chan, grid_side):
        j = self.jones
        if j is None:
            return []
        empty4 = np.array([], np.complex64).reshape((0, 0, 0, 0))
        kms = j.get("kms", empty4)
        beam = j.get("beam", empty4)
        mode = j.get("mode", 1)
        if grid_side is False and j.get("degrid_mode") is not None:
            mode = j["degrid_mode"]
        if sum_jones is None:
            sum_jones = np.zeros((2, nband), np.float64)
        if sum_jones_chan is None:
            sum_jones_chan = np.zeros((2, self.freqs.size), np.float64)
        return [kms,
                j.get("kms_time", np.array([], np.int32)),
                beam,
                j.get("beam_time", np.array([], np.int32)),
                self.a0, self.a1,
                np.array([facet.kms_dir], np.int32),
                np.float32(j.get("coefs", np.array([], np.float32))),
                np.array([facet.beam_dir], np.int32),
                np.array([j.get("interp", 0), mode], np.int32),
                j.get("kms_chan", np.array([], np.int32)),
                j.get("beam_chan", np.array([], np.int32)),
                np.array([], np.float32).reshape((0, 0)),
                j.get("apply_amp", 1), j.get("apply_phase", 1), j.get("scale", 0),
                j.get("calib_error", 0.0),
                sum_jones, sum_jones_chan,
                np.float32(j.get("reweight_snr", 0.0))]

    def _lsmear(self, facet):
        if self.decorr is None:
            return []
        uvw_dt, dt, dnu, ts, fs = self.decorr
        return [uvw_dt, dt, dnu, ts, fs, facet.l0, facet.m0]

    def facet_infos(self, i):
        f = self.facets[i]
        return np.float64(np.array([f.cu, f.cv, f.l0, f.m0, i]))

    def grid_args(self, i, grid, sumwt, dopsf=0, sum_jones=None, sum_jones_chan=None, vis=None):
        f = self.facets[i]
        return (grid, self.vis if vis is None else vis, self.uvw, self.flags, self.weights, sumwt,
                dopsf, f.wplanes, f.wplanes_conj, f.winfos, self.incr.astype(np.float64), self.freqs,
                [np.array([0, 5, 5, 0], np.int32), self.facet_infos(i)],
                self._ljones(f, grid.shape[0], sum_jones, sum_jones_chan, True),
                self.bda_grid,
                self.sparsification if self.sparsification is not None else np.array([], dtype=bool),
                [0, chan_equidistant(self.freqs), 1, 0],
                self._lsmear(f), np.int32(self.chan_map),
                np.array(self.corr_codes).astype(np.uint16),
                np.array(self.stokes_codes).astype(np.uint16))

    def degrid_args(self, i, grid, vis):
        f = self.facets[i]
        sumwt = np.zeros((grid.shape[0], grid.shape[1]), np.float64)
        rowinfos = np.array([0, self.uvw.shape[0]], np.int32)
        return (grid, vis, self.uvw, self.flags, sumwt, 0, f.wplanes_conj, f.wplanes, f.winfos,
                self.incr.astype(np.float64), self.freqs,
                [np.array([0, 5, 5, 0], np.int32), self.facet_infos(i), rowinfos],
                self._ljones(f, grid.shape[0], None, None, False),
                self.bda_degrid, np.array([], dtype=bool), [0, chan_equidistant(self.freqs), 1, 0],
                self._lsmear(f), np.int32(self.chan_map_degrid),
                np.array(self.corr_codes).astype(np.uint16),
                np.array(self.stokes_codes).astype(np.uint16))

    def new_grid(self, dtype=np.complex64, nband=None):
        return np.zeros((nband or self.nband, self.npol, self.npix, self.npix), dtype)


def make_cf_facet(l0, m0, npix, cell_arcsec, freqs, nw, sups, overs, wmax, sup0=None):
    wt = ClassWTermModified(Cell=cell_arcsec, Sup=int(sup0 or sups[0]), Nw=nw, wmax=wmax, Npix=npix,
                            Freqs=freqs, OverS=overs, lmShift=(l0, m0), sups=sups)
    return Facet(l0=l0, m0=m0, cu=float(wt.Cu), cv=float(wt.Cv), wplanes=wt.Wplanes,
                 wplanes_conj=wt.WplanesConj, winfos=wt.winfos(), sphe=wt.Sphe)


def random_problem(seed, nant=5, ntime=6, nchan=8, ncorr=4, npix=48, cell_arcsec=30.0, nfacet=1,
                   facet_spacing_px=40, nw=3, sups=(5, 7, 7), overs=5, nband=1, npol=1,
                   corr_codes=(9, 10, 11, 12), stokes_codes=(1,), jones=None, decorr=False,
                   bda="mixed", flag_frac=0.1, sparsify=False, freq0=1.2e9, freq1=1.6e9,
                   equidistant=True, wmax_scale=1.0, sort_by_baseline=False, uv_scale=1.0):
    """A small random problem whose uv coverage fits a `npix` grid (so every code path runs).

    jones: None, or dict(mode=1|2, interp=0|1, beam=bool, kms=bool, apply_amp, apply_phase, scale).
    bda: "pervis" (one block per row x chan), "mixed" (random multi-row / multi-chan blocks).
    uv_scale > 1 widens the array so that the longest baselines fall off the grid; wmax_scale < 1
    puts the largest |w| beyond the last w-plane (both drop blocks, as the reference does).
    """
    rng = np.random.default_rng(seed)
    cell = cell_arcsec * synth.ARCSEC
    # baselines short enough that |u|/lambda stays inside the grid
    umax_lambda = 0.42 * npix / 2 / (npix * cell)
    fmax = max(freq0, freq1)
    diam = umax_lambda * C_LIGHT / fmax * 2.2 * uv_scale
    if equidistant:
        freqs = np.linspace(freq0, freq1, nchan)
    else:
        freqs = np.sort(rng.uniform(freq0, freq1, nchan))
    obs = synth.make_observation("disk", nant, ntime, 60.0 * 20, freqs, 45.0, -30.7, seed,
                                 sort_by_baseline=sort_by_baseline, diameter_m=diam)
    R = obs.nrow
    vis = (rng.standard_normal((R, nchan, ncorr)) + 1j * rng.standard_normal((R, nchan, ncorr))).astype(np.complex64)
    flags = synth.random_flags(R, nchan, ncorr, flag_frac, seed + 3)
    weights = rng.uniform(0.5, 1.5, (R, nchan)).astype(np.float32)
    chan_map = np.minimum((np.arange(nchan) * nband) // nchan, nband - 1).astype(np.int32)
    wmax = float(np.max(np.abs(obs.uvw[:, 2]))) * wmax_scale + 1e-3
    facets = []
    nfx = int(np.ceil(np.sqrt(nfacet)))
    cen = synth.facet_centres(nfx, facet_spacing_px, cell)[:nfacet]
    if nfacet == 1:
        cen = np.zeros((1, 2))
    for i in range(nfacet):
        f = make_cf_facet(float(cen[i, 0]), float(cen[i, 1]), npix, cell_arcsec, freqs, nw, list(sups),
                          overs, wmax)
        f.kms_dir = i
        f.beam_dir = (i + 1) % max(1, nfacet)
        facets.append(f)
    if bda == "pervis":
        bmap = bdamod.bda_map_per_visibility(obs.a0, obs.a1, nchan)
    else:
        def fac(key):
            p, q = (int(x) for x in key.split(":"))
            h = (31 * p + q) % 7
            return [(1, 1), (2, 1), (1, 3), (3, 2), (2, 4), (4, 8), (6, 2)][h]
        bmap = bdamod.bda_map_factors(obs.a0, obs.a1, nchan, fac, chan_map)
    jd = None
    if jones is not None:
        nt = 3
        ntime_u = int(obs.tidx.max()) + 1
        tmap = np.minimum((obs.tidx * nt) // ntime_u, nt - 1).astype(np.int32)
        nchj = 2
        cmap = np.minimum((np.arange(nchan) * nchj) // nchan, nchj - 1).astype(np.int32)
        nd = max(2, nfacet + 1)
        jd = dict(mode=jones.get("mode", 1), interp=jones.get("interp", 0),
                  apply_amp=jones.get("apply_amp", 1), apply_phase=jones.get("apply_phase", 1),
                  scale=jones.get("scale", 0), calib_error=jones.get("calib_error", 0.0),
                  degrid_mode=jones.get("degrid_mode"))
        if jones.get("kms", True):
            jd["kms"] = synth.random_jones(nt, nd, nant, nchj, seed + 4, 0.2, 0.5)
            jd["kms_time"] = tmap
            jd["kms_chan"] = cmap
        if jones.get("beam", False):
            jd["beam"] = synth.random_jones(2, nd, nant, 1, seed + 5, 0.1, 0.3)
            jd["beam_time"] = np.minimum((obs.tidx * 2) // ntime_u, 1).astype(np.int32)
            jd["beam_chan"] = np.zeros(nchan, np.int32)
        if jd["interp"] == 1:
            c = rng.random(nd).astype(np.float32)
            c[rng.random(nd) < 0.3] = 0
            jd["coefs"] = c
    dec = None
    if decorr:
        uvw_dt = obs.uvw * 1e-3 + rng.standard_normal(obs.uvw.shape) * 1e-2
        dec = (np.ascontiguousarray(uvw_dt), 8.0, float(freqs[1] - freqs[0]) if nchan > 1 else 1e6, 1, 1)
    sp = None
    if sparsify:
        sp = rng.random(bdamod.nblocks_of(bmap)) < 0.7
    return Problem(uvw=obs.uvw, vis=vis, flags=flags, weights=weights, freqs=freqs, a0=obs.a0,
                   a1=obs.a1, chan_map=chan_map, chan_map_degrid=chan_map.copy(), bda_grid=bmap,
                   bda_degrid=bmap, facets=facets, npix=npix, nband=nband, npol=npol,
                   incr=np.array([-cell, cell]), corr_codes=list(corr_codes),
                   stokes_codes=list(stokes_codes), jones=jd, decorr=dec, sparsification=sp)
