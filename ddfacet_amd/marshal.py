"""Marshalling of the reference argument lists into the C-ABI structs.

``parse_grid_args`` / ``parse_degrid_args`` accept exactly the positional arguments of
``pyGridderWPol32|64`` and ``pyDeGridderWPol`` (DDFacet/Gridder/GridderSmearPols.cc:76-97,
212-233; call sites ClassDDEGridMachine.py:790-818, 1084-1108) and return a ``CallSpec`` of
contiguous numpy arrays and scalars.  ``fill_structs`` turns a spec into ``ddf_grid_desc`` /
``ddf_vis`` / ``ddf_facet`` through an address provider, so one spec serves the device path
(addresses of device copies) and the CPU oracle (host addresses).

Error behaviour: pybind11 ``array_t<T, c_style>`` silently copies a non-contiguous or
wrong-dtype argument, which would lose in-place outputs.  Here an output with the wrong dtype or
layout raises ``TypeError``; read-only inputs are converted, as pybind11 would.  Unsupported
correlation / Stokes combinations raise ``ValueError`` (reference: std::invalid_argument).
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import abi


def _as(a, dtype):
    return np.ascontiguousarray(np.asarray(a), dtype=dtype)


def _require_out(name, a, dtype):
    if not isinstance(a, np.ndarray):
        raise TypeError("%s must be a numpy array" % name)
    if a.dtype != np.dtype(dtype):
        raise TypeError("%s must be %s, got %s" % (name, np.dtype(dtype), a.dtype))
    if not a.flags.c_contiguous:
        raise TypeError("%s has to be C-contiguous" % name)
    return a


@dataclass
class CFPack:
    data: np.ndarray      # complex64, all planes
    offset: np.ndarray    # int64 [2*nw]
    support: np.ndarray   # int32 [nw]
    nw: int
    overs: int
    ref_wave: float
    wmax: float


def pack_cfs(lcfs, lcfs_conj, winfos):
    winfos = _as(winfos, np.float64)
    overs = int(np.floor(winfos[3]))
    nw_info = int(winfos[2])
    planes = list(lcfs) + list(lcfs_conj)
    nw = len(lcfs)
    if len(lcfs_conj) != nw:
        raise ValueError("Lcfs and LcfsConj must hold the same number of w-planes")
    if nw_info != nw:
        raise ValueError("WInfos[2]=%d does not match %d w-planes" % (nw_info, nw))
    sup = np.zeros(nw, np.int32)
    offs = np.zeros(2 * nw, np.int64)
    flat = []
    pos = 0
    for i, p in enumerate(planes):
        p = _as(p, np.complex64)
        if p.ndim != 2 or p.shape[0] != p.shape[1] or p.shape[0] % overs:
            raise ValueError("CF plane %d has shape %s, expected square (S*OverS)^2" % (i, p.shape))
        s = p.shape[0] // overs
        if s % 2 == 0:
            raise ValueError("CF support must be odd")
        if i < nw:
            sup[i] = s
        elif sup[i - nw] != s:
            raise ValueError("Wplanes[%d] and WplanesConj[%d] differ in support" % (i - nw, i - nw))
        offs[i] = pos
        flat.append(p.ravel())
        pos += p.size
    data = np.concatenate(flat) if flat else np.zeros(0, np.complex64)
    return CFPack(data, offs, sup, nw, overs, float(winfos[0]), float(winfos[1]))


@dataclass
class JonesSpec:
    mode: int
    interp_mode: int
    kms: Optional[np.ndarray]
    kms_time: Optional[np.ndarray]
    kms_chan: Optional[np.ndarray]
    kms_dir: int
    beam: Optional[np.ndarray]
    beam_time: Optional[np.ndarray]
    beam_chan: Optional[np.ndarray]
    beam_dir: int
    a0: np.ndarray
    a1: np.ndarray
    coefs: np.ndarray
    apply_amp: int
    apply_phase: int
    scale_jones: int
    calib_error: float
    sum_jones: np.ndarray
    sum_jones_chan: np.ndarray
    reweight_snr: float


def parse_jones(ljones, grid_side):
    if len(ljones) == 0:
        return None
    if len(ljones) != 20:
        raise ValueError("LJones must have 0 or 20 entries")
    kms = np.asarray(ljones[0])
    beam = np.asarray(ljones[2])
    use_kms = kms.ndim == 6 and kms.size != 0
    use_beam = beam.ndim == 6 and beam.size != 0
    if not (use_kms or use_beam):
        raise RuntimeError("Jones matricies specified but neither E or DD Jones are applied. This is a bug!")
    modes = _as(ljones[9], np.int32)
    sum_jones = ljones[17]
    sum_jones_chan = ljones[18]
    if grid_side:
        _require_out("SumJones (LJones[17])", sum_jones, np.float64)
        _require_out("SumJonesChan (LJones[18])", sum_jones_chan, np.float64)
    else:
        sum_jones = _as(sum_jones, np.float64)
        sum_jones_chan = _as(sum_jones_chan, np.float64)
    return JonesSpec(
        mode=int(modes[1]), interp_mode=int(modes[0]),
        kms=_as(kms, np.complex64) if use_kms else None,
        kms_time=_as(ljones[1], np.int32) if use_kms else None,
        kms_chan=_as(ljones[10], np.int32) if use_kms else None,
        kms_dir=int(_as(ljones[6], np.int32)[0]) if use_kms else 0,
        beam=_as(beam, np.complex64) if use_beam else None,
        beam_time=_as(ljones[3], np.int32) if use_beam else None,
        beam_chan=_as(ljones[11], np.int32) if use_beam else None,
        beam_dir=int(_as(ljones[8], np.int32)[0]) if use_beam else 0,
        a0=_as(ljones[4], np.int32), a1=_as(ljones[5], np.int32),
        coefs=_as(ljones[7], np.float32),
        apply_amp=int(bool(ljones[13])), apply_phase=int(bool(ljones[14])),
        scale_jones=int(bool(ljones[15])), calib_error=float(ljones[16]),
        sum_jones=sum_jones, sum_jones_chan=sum_jones_chan, reweight_snr=float(ljones[19]))


def _check_jones(j, nrow, nchan, nband):
    """Lengths the device code indexes with (the reference reads them unchecked)."""
    if j.a0.size != nrow or j.a1.size != nrow:
        raise RuntimeError("Antenna array is expected to have the same number of rows as the uvw array")
    for tab, tmap, cmap, what in ((j.kms, j.kms_time, j.kms_chan, "DD"), (j.beam, j.beam_time, j.beam_chan, "beam")):
        if tab is None:
            continue
        if tmap.size != nrow:
            raise ValueError("%s Jones time mapping must have one entry per row" % what)
        if cmap.size != nchan:
            raise ValueError("%s Jones channel mapping must have one entry per channel" % what)
    if j.mode not in (1, 2):
        raise ValueError("BDAJones mode must be 1 or 2")
    if j.sum_jones.size < 2 * nband or j.sum_jones_chan.size < 2 * nchan:
        raise ValueError("SumJones must be [2, nband] and SumJonesChan [2, nchan]")


@dataclass
class CallSpec:
    grid: np.ndarray
    grid_f64: int
    vis: np.ndarray
    uvw: np.ndarray
    flags: np.ndarray
    weights: Optional[np.ndarray]
    sumwt: Optional[np.ndarray]
    dopsf: int
    cf: CFPack
    incr: np.ndarray
    freqs: np.ndarray
    facet_infos: np.ndarray
    jones: Optional[JonesSpec]
    bda: np.ndarray
    nblocks: int
    sparsification: Optional[np.ndarray]
    chan_equidistant: int
    decorr: Optional[tuple]
    chan_map: np.ndarray
    corr_codes: np.ndarray
    stokes_codes: np.ndarray


def _parse_common(grid, vis, uvw, flags, dopsf, lcfs, lcfs_conj, winfos, increment, freqs, lmaps,
                  ljones, smear_mapping, sparsification, loptimisation, lsmearing, chan_mapping,
                  corr_fmt, out_stokes, grid_side):
    if vis.ndim != 3:
        raise ValueError("vis must be [nrow, nchan, ncorr]")
    uvw = _as(uvw, np.float64)
    flags = _as(flags, np.bool_)
    if flags.shape != vis.shape:
        raise ValueError("vis%s and flag%s should have the same shape" % (vis.shape, flags.shape))
    if uvw.ndim != 2 or uvw.shape[1] != 3 or uvw.shape[0] != vis.shape[0]:
        raise ValueError("uvw must be [nrow, 3]")
    freqs = _as(freqs, np.float64)
    if freqs.shape[0] != vis.shape[1]:
        raise ValueError("freqs must have one entry per channel")
    bda = _as(smear_mapping, np.int32)
    if bda.size < 2:
        raise ValueError("SmearMapping must hold at least its two-word header")
    nblocks = int(np.uint32(bda[0])) + (int(bda[1]) << 32)
    if bda.size < 2 + nblocks:
        raise ValueError("SmearMapping is shorter than its block-size table")
    sp = np.asarray(sparsification)
    if sp.size and grid_side:
        if sp.size != nblocks:
            raise ValueError("sparsification argument must be an array of length NTotBlocks")
        sp = _as(sp, np.bool_)
    else:
        sp = None
    chan_map = _as(chan_mapping, np.int32)
    if chan_map.shape[0] != vis.shape[1]:
        raise ValueError("ChanMapping must have one entry per channel")
    decorr = None
    if len(lsmearing):
        if len(lsmearing) != 7:
            raise ValueError("LSmearing must have 0 or 7 entries")
        decorr = (_as(lsmearing[0], np.float64), float(lsmearing[1]), float(lsmearing[2]),
                  int(bool(lsmearing[3])), int(bool(lsmearing[4])), float(lsmearing[5]),
                  float(lsmearing[6]))
    corr = _as(corr_fmt, np.uint16).ravel()
    stokes = _as(out_stokes, np.uint16).ravel()
    if corr.size > 4 or stokes.size > 4 or corr.size == 0 or stokes.size == 0:
        raise ValueError("Cannot convert input correlations to desired output Stokes parameters.")
    if corr.size != vis.shape[2]:
        raise ValueError("DataCorrFormat must list one code per correlation")
    jones = parse_jones(ljones, grid_side)
    if jones is not None:
        _check_jones(jones, vis.shape[0], vis.shape[1], grid.shape[0])
    if decorr is not None and decorr[0].shape != uvw.shape:
        raise ValueError("LSmearing[0] (uvw_dt) must be [nrow, 3]")
    return dict(
        uvw=uvw, flags=flags, dopsf=int(bool(dopsf)), cf=pack_cfs(lcfs, lcfs_conj, winfos),
        incr=_as(increment, np.float64), freqs=freqs, facet_infos=_as(lmaps[1], np.float64),
        jones=jones, bda=bda, nblocks=nblocks, sparsification=sp,
        chan_equidistant=int(bool(loptimisation[1])), decorr=decorr, chan_map=chan_map,
        corr_codes=corr, stokes_codes=stokes)


def parse_grid_args(args, grid_f64):
    (grid, vis, uvw, flags, weights, sumwt, dopsf, lcfs, lcfs_conj, winfos, increment, freqs, lmaps,
     ljones, smear_mapping, sparsification, loptimisation, lsmearing, chan_mapping, corr_fmt,
     out_stokes) = args
    _require_out("grid", grid, np.complex128 if grid_f64 else np.complex64)
    if grid.ndim != 4:
        raise ValueError("grid must be [nband, npol, ny, nx]")
    _require_out("sumwt", sumwt, np.float64)
    vis = _as(vis, np.complex64)
    common = _parse_common(grid, vis, uvw, flags, dopsf, lcfs, lcfs_conj, winfos, increment, freqs,
                           lmaps, ljones, smear_mapping, sparsification, loptimisation, lsmearing,
                           chan_mapping, corr_fmt, out_stokes, True)
    weights = _as(weights, np.float32)
    if weights.shape != vis.shape[:2]:
        raise ValueError("weights must be [nrow, nchan]")
    if sumwt.size < grid.shape[0] * grid.shape[1]:
        raise ValueError("sumwt must be [nband, npol]")
    return CallSpec(grid=grid, grid_f64=int(grid_f64), vis=vis, weights=weights, sumwt=sumwt, **common)


def parse_degrid_args(args):
    (grid, vis, uvw, flags, sumwt, dopsf, lcfs, lcfs_conj, winfos, increment, freqs, lmaps, ljones,
     smear_mapping, sparsification, loptimisation, lsmearing, chan_mapping, corr_fmt,
     out_stokes) = args
    grid = _as(grid, np.complex64)
    if grid.ndim != 4:
        raise ValueError("grid must be [nband, npol, ny, nx]")
    _require_out("vis", vis, np.complex64)
    common = _parse_common(grid, vis, uvw, flags, 0, lcfs, lcfs_conj, winfos, increment, freqs,
                           lmaps, ljones, smear_mapping, sparsification, loptimisation, lsmearing,
                           chan_mapping, corr_fmt, out_stokes, False)
    return CallSpec(grid=grid, grid_f64=0, vis=vis, weights=None, sumwt=None, **common)


def freq_means(freqs):
    """(mean frequency, mean wavelength) summed in channel order, as gridder.h:164-171 does."""
    wl = 0.0
    fm = 0.0
    for f in np.asarray(freqs, dtype=np.float64):
        wl += 299792458.0 / f
        fm += f
    return fm / len(freqs), wl / len(freqs)


def fill_structs(spec, addr, bda_start=None):
    """Build (ddf_grid_desc, ddf_vis, ddf_facet) for `spec`; addr(name, ndarray) -> int address
    (0 for None).  The caller keeps whatever addr allocates alive for the call."""
    g = abi.ddf_grid_desc()
    g.grid_f64 = spec.grid_f64
    g.nband, g.npol, g.ny, g.nx = [int(x) for x in spec.grid.shape]
    g.incr[0], g.incr[1] = float(spec.incr[0]), float(spec.incr[1])
    g.dopsf = spec.dopsf

    v = abi.ddf_vis()
    v.vis = addr("vis", spec.vis)
    v.uvw = addr("uvw", spec.uvw)
    v.flags = addr("flags", spec.flags)
    v.weights = addr("weights", spec.weights)
    v.freqs = addr("freqs", spec.freqs)
    v.chan_map = addr("chan_map", spec.chan_map)
    v.nrow, v.nchan, v.ncorr = [int(x) for x in spec.vis.shape]
    v.chan_equidistant = spec.chan_equidistant
    v.bda = addr("bda", spec.bda)
    v.bda_len = int(spec.bda.size)
    v.nblocks = int(spec.nblocks)
    v.bda_start = addr("bda_start", bda_start)
    v.sparsification = addr("sparsification", spec.sparsification)
    v.ncorr_codes = int(spec.corr_codes.size)
    for i, c in enumerate(spec.corr_codes):
        v.corr_codes[i] = int(c)
    v.nstokes = int(spec.stokes_codes.size)
    for i, c in enumerate(spec.stokes_codes):
        v.stokes_codes[i] = int(c)
    v.freq_mean, v.wavelength_mean = freq_means(spec.freqs)

    f = abi.ddf_facet()
    f.grid = addr("grid", spec.grid)
    f.sumwt = addr("sumwt", spec.sumwt)
    cf = spec.cf
    f.cf.data = addr("cf_data", cf.data)
    f.cf.offset = addr("cf_offset", cf.offset)
    f.cf.support = addr("cf_support", cf.support)
    f.cf.nw, f.cf.overs = cf.nw, cf.overs
    f.cf.ref_wave, f.cf.wmax = cf.ref_wave, cf.wmax
    f.cf.max_support = int(cf.support.max()) if cf.nw else 0
    fi = spec.facet_infos
    f.cu, f.cv, f.l0, f.m0 = float(fi[0]), float(fi[1]), float(fi[2]), float(fi[3])
    j = spec.jones
    if j is not None:
        f.jones.enabled = 1
        f.jones.mode = j.mode
        f.jones.interp_mode = j.interp_mode
        if j.kms is not None:
            f.jones.kms = addr("kms", j.kms)
            for i in range(4):
                f.jones.kms_dims[i] = int(j.kms.shape[i])
            f.jones.kms_time = addr("kms_time", j.kms_time)
            f.jones.kms_chan = addr("kms_chan", j.kms_chan)
            f.jones.kms_dir = j.kms_dir
        if j.beam is not None:
            f.jones.beam = addr("beam", j.beam)
            for i in range(4):
                f.jones.beam_dims[i] = int(j.beam.shape[i])
            f.jones.beam_time = addr("beam_time", j.beam_time)
            f.jones.beam_chan = addr("beam_chan", j.beam_chan)
            f.jones.beam_dir = j.beam_dir
        f.jones.a0 = addr("a0", j.a0)
        f.jones.a1 = addr("a1", j.a1)
        f.jones.coefs = addr("coefs", j.coefs if j.coefs.size else None)
        f.jones.ncoefs = int(j.coefs.size)
        f.jones.apply_amp = j.apply_amp
        f.jones.apply_phase = j.apply_phase
        f.jones.scale_jones = j.scale_jones
        f.jones.calib_error = j.calib_error
        f.jones.sum_jones = addr("sum_jones", j.sum_jones)
        f.jones.sum_jones_chan = addr("sum_jones_chan", j.sum_jones_chan)
        f.jones.reweight_snr = j.reweight_snr
    d = spec.decorr
    if d is not None:
        f.decorr.enabled = 1
        f.decorr.uvw_dt = addr("uvw_dt", d[0])
        f.decorr.dt, f.decorr.dnu = d[1], d[2]
        f.decorr.tsmear, f.decorr.fsmear = d[3], d[4]
        f.decorr.l0, f.decorr.m0 = d[5], d[6]
    return g, v, f
