"""Beam Jones producers: analytic primary-beam E-Jones tables and the Jones domain merge, with
the tables on the device in the layout of the ddf_jones inputs ([t][dir][ant][chan][2][2]
complex64), so the gridders read them without a host round trip.

* ``estimate_beam`` is ClassJones.EstimateBeam (DDFacet/Data/ClassJones.py:1014-1065) for the
  analytic beam models (``GMRT``: ClassGMRTBeam, DDFacet/Data/ClassGMRTBeam.py:50-173; ``ATCA``:
  ClassATCABeam with the "csiro" tables, DDFacet/Data/ClassATCABeam.py:45-245): one table entry per
  interval of the beam sample times (the models do not vary in time or between antennas), the
  channels of each Jones channel averaged, CenterNorm applied on request (ddf_beam_jones).  The
  per-channel model terms are formed here as calcCoefs does (band selection, the GMRT clipping
  point at the first positive stationary point of the band polynomial, float32 storage).
* ``beam_sample_times`` is getBeamSampleTimes (ClassGMRTBeam.py:57-64): every DtBeamMin minutes
  from the first time, then the last time + 1 s.
* ``merge_jones`` is ClassJonesDomains.MergeJones (DDFacet/Other/ClassJonesDomains.py:99-212,
  behind ClassJones.MergeJones): the merged frequency and time domains and the index maps are
  formed here from the small domain arrays, the cell products on the device (ddf_jones_merge).
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import check, lib
from .device import require_gpu, stream_handle

ARCMIN_PER_RAD = 180.0 / np.pi * 60

# Band polynomials, ascending powers of d nu / 1 GHz (d in arcmin): (f_low, f_high, coefficients)
GMRT_BANDS = (
    (250e6, 500e6, (1, 0, -2.939e-3, 0, 33.312e-7, 0, -16.659e-10, 0, 3.066e-13)),
    (550e6, 850e6, (1, 0, -3.190e-3, 0, 38.642e-7, 0, -20.471e-10, 0, 3.964e-13)),
    (1050e6, 1450e6, (1, 0, -2.608e-3, 0, 27.357e-7, 0, -13.091e-10, 0, 2.368e-13)),
)
# ATCA "csiro" band polynomials in (d nu0 * 1e-9)^k at the band frequencies nu0
ATCA_BANDS = (
    (1268e6, (1, 0, -1.154e-03, 0, +5.351e-07, 0, -1.265e-10, 0, +1.634e-14, 0, -1.099e-18, 0, +3.009e-23)),
    (1396e6, (1, 0, -1.092e-03, 0, +4.660e-07, 0, -9.875e-11, 0, +1.119e-14, 0, -6.504e-19, 0, +1.522e-23)),
    (1524e6, (1, 0, -1.048e-03, 0, +4.305e-07, 0, -8.824e-11, 0, +9.684e-15, 0, -5.446e-19, 0, +1.232e-23)),
    (1652e6, (1, 0, -1.033e-03, 0, +4.221e-07, 0, -8.692e-11, 0, +9.679e-15, 0, -5.571e-19, 0, +1.298e-23)),
    (1780e6, (1, 0, -1.003e-03, 0, +4.016e-07, 0, -8.170e-11, 0, +9.046e-15, 0, -5.199e-19, 0, +1.213e-23)),
    (1908e6, (1, 0, -9.794e-04, 0, +3.831e-07, 0, -7.562e-11, 0, +8.067e-15, 0, -4.440e-19, 0, +9.882e-24)),
    (2036e6, (1, 0, -9.692e-04, 0, +3.848e-07, 0, -7.809e-11, 0, +8.621e-15, 0, -4.927e-19, 0, +1.140e-23)),
    (2164e6, (1, 0, -9.919e-04, 0, +4.028e-07, 0, -8.251e-11, 0, +9.077e-15, 0, -5.122e-19, 0, +1.163e-23)),
    (2292e6, (1, 0, -1.007e-03, 0, +4.131e-07, 0, -8.480e-11, 0, +9.305e-15, 0, -5.220e-19, 0, +1.176e-23)),
    (2420e6, (1, 0, -1.016e-03, 0, +4.235e-07, 0, -8.817e-11, 0, +9.772e-15, 0, -5.521e-19, 0, +1.251e-23)),
    (2548e6, (1, 0, -1.033e-03, 0, +4.375e-07, 0, -9.223e-11, 0, +1.033e-14, 0, -5.885e-19, 0, +1.343e-23)),
    (2676e6, (1, 0, -1.035e-03, 0, +4.395e-07, 0, -9.269e-11, 0, +1.038e-14, 0, -5.919e-19, 0, +1.353e-23)),
    (2804e6, (1, 0, -1.064e-03, 0, +4.622e-07, 0, -9.913e-11, 0, +1.125e-14, 0, -6.502e-19, 0, +1.505e-23)),
    (2932e6, (1, 0, -1.081e-03, 0, +4.769e-07, 0, -1.045e-10, 0, +1.222e-14, 0, -7.325e-19, 0, +1.768e-23)),
)
# DDF_BEAM_GMRT / DDF_BEAM_ATCA of include/ddfgpu.h
KINDS = {"GMRT": 1, "ATCA": 2}


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def beam_sample_times(times, dt_beam_min):
    tmin, tmax = float(times[0]), float(times[-1]) + 1
    out = np.arange(tmin, tmax, dt_beam_min * 60).tolist()
    if tmax not in out:
        out.append(tmax)
    return np.array(out)


def freq_domains(chan_freqs, chan_width, nchan_jones=0):
    """[n, 2] equal frequency domains across the band (CalcFreqDomains); n = nchan when 0."""
    f = np.asarray(chan_freqs, np.float64).ravel()
    n = nchan_jones or f.size
    e = np.linspace(f.min() - chan_width / 2.0, f.max() + chan_width / 2.0, n + 1)
    return np.stack([e[:-1], e[1:]], axis=1)


def nearest_domain(chan_freqs, domains):
    """Jones channel of each channel: the domain with the nearest centre (first on a tie)."""
    c = domains.mean(axis=1)
    return np.argmin(np.abs(np.asarray(chan_freqs, np.float64)[:, None] - c[None, :]), axis=1)


def _gmrt_clip(coefs):
    """(x, P(x)) at the first positive real stationary point of P (roots in numpy's order)."""
    p = np.poly1d(np.asarray(coefs, np.float64)[::-1])
    r = np.roots(np.polyder(p))
    i = np.flatnonzero((np.abs(r.imag) < 1e-6) & (r.real > 0))[0]
    x = np.abs(r[i])
    return x, p(x)


def channel_terms(kind, chan_freqs):
    """Per-channel model terms (calcCoefs): ascending coefficients of the channel's band polynomial
    (ATCA: of the bands at or below and at or above the channel) as float64 arrays [nchan, ncoef],
    the two ATCA band frequencies [nchan, 2], and the null thresholds / values (float32)."""
    if kind not in KINDS:
        raise ValueError("beam model must be one of %s" % sorted(KINDS))
    f = np.asarray(chan_freqs, np.float64).ravel()
    n = f.size
    if kind == "GMRT":
        # the reference stores the selected coefficients in a float32 array
        c0 = np.zeros((n, 9), np.float32)
        xn = np.zeros(n, np.float32)
        yn = np.zeros(n, np.float32)
        clip = [_gmrt_clip(c) for _, _, c in GMRT_BANDS]
        for i, fi in enumerate(f):
            band = next((k for k, (lo, hi, _) in enumerate(GMRT_BANDS) if lo <= fi <= hi), None)
            if band is None:
                raise NotImplementedError("The GMRT beam is not modeled for channel %d @ %.3f MHz" % (i, fi / 1e6))
            c0[i] = GMRT_BANDS[band][2]
            xn[i] = clip[band][0] / (fi / 1e9)
            yn[i] = clip[band][1]
        return {"coef0": c0.astype(np.float64), "coef1": np.zeros((n, 9)), "band_f0": np.zeros((n, 2)),
                "x_null": xn, "y_null": yn}
    f0 = np.array([b[0] for b in ATCA_BANDS])
    nb = f0.size
    lo = np.array([max([0] + list(np.flatnonzero(f0 <= fi))) for fi in f], int)
    hi = np.array([min([nb - 1] + list(np.flatnonzero(f0 >= fi))) for fi in f], int)
    tab = np.array([b[1] for b in ATCA_BANDS], np.float64)
    return {"coef0": tab[lo], "coef1": tab[hi], "band_f0": np.stack([f0[lo], f0[hi]], axis=1),
            "x_null": np.full(n, 100.0, np.float32), "y_null": np.full(n, 0.0036, np.float32)}


class BeamModel:
    """Per-channel terms of an analytic beam model, on the device."""

    def __init__(self, kind, chan_freqs, device=None):
        terms = channel_terms(kind, chan_freqs)
        dev = device or require_gpu()
        self.kind = kind
        self.nchan, self.ncoef = terms["coef0"].shape
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in terms.items()}
        self.t["freqs"] = torch.from_numpy(np.asarray(chan_freqs, np.float64).ravel().copy()).to(dev)

    def desc(self, chan_to_jones, nchan_jones):
        d = abi.ddf_beam_desc()
        d.kind = KINDS[self.kind]
        d.nchan, d.ncoef, d.nchan_jones = self.nchan, self.ncoef, int(nchan_jones)
        for k in ("freqs", "coef0", "coef1", "band_f0", "x_null", "y_null"):
            setattr(d, k, self.t[k].data_ptr())
        d.chan_to_jones = chan_to_jones.data_ptr()
        return d


def estimate_beam(times_beam, ra, dec, ra_c, dec_c, chan_freqs, chan_width, na, kind="GMRT", nband=0,
                  center_norm=True, stream=None):
    """The DicoBeam of ClassJones.EstimateBeam: "Jones" complex64 [nt, ndir, na, nchan_jones, 2, 2] on
    the device, "t0" / "t1" / "tm", "FreqDomains" and "VisToJonesChanMapping" (numpy)."""
    dev = require_gpu()
    t = np.asarray(times_beam, np.float64)
    t0, t1 = t[:-1].copy(), t[1:].copy()
    dom = freq_domains(chan_freqs, chan_width, nband)
    mapping = nearest_domain(chan_freqs, dom)
    model = BeamModel(kind, chan_freqs, dev)
    c2j = torch.from_numpy(mapping.astype(np.int32)).to(dev)
    rad = torch.as_tensor(np.asarray(ra, np.float64), device=dev)
    decd = torch.as_tensor(np.asarray(dec, np.float64), device=dev)
    nd, nt, nchj = int(rad.numel()), int(t0.size), int(dom.shape[0])
    jones = torch.empty((nt, nd, int(na), nchj, 2, 2), dtype=torch.complex64, device=dev)
    d = model.desc(c2j, nchj)
    check(lib().ddf_beam_jones(C.byref(d), _p(rad), _p(decd), nd, float(ra_c), float(dec_c), nt, int(na),
                               int(bool(center_norm)), _p(jones), C.c_void_p(stream_handle(stream))))
    # the model tensors must live until the launch has run
    (stream or torch.cuda.current_stream(dev)).synchronize()
    return {"Jones": jones, "t0": t0, "t1": t1, "tm": (t0 + t1) / 2.0, "FreqDomains": dom,
            "VisToJonesChanMapping": mapping}


def merged_freq_domains(fd0, fd1):
    """The union of both tables' domain edges (edges within 1 Hz of an earlier one dropped), cut to
    the band both cover (domain centres in [max of minima, min of maxima))."""
    edges = np.unique(np.concatenate([np.ravel(fd0), np.ravel(fd1)]))
    keep = []
    for e in edges:
        if not keep or all(abs(e - k) >= 1.0 for k in keep):
            keep.append(e)
    keep = np.asarray(keep)
    dom = np.stack([keep[:-1], keep[1:]], axis=1)
    fm = dom.mean(axis=1)
    lo, hi = max(np.min(fd0), np.min(fd1)), min(np.max(fd0), np.max(fd1))
    return dom[(fm >= lo) & (fm < hi)]


def merged_time_domains(t0a, t1a, t0b, t1b):
    """Intervals from the earliest start, each ending at the next interval end of either table."""
    ends = np.asarray(t1a, np.float64), np.asarray(t1b, np.float64)
    cur = min(float(t0a[0]), float(t0b[0]))
    starts, stops = [], []
    while True:
        nxt = [e[e - cur > 0] for e in ends]
        cand = [(x[0] - cur) for x in nxt if x.size]
        if not cand:
            break
        dt = min(cand)
        starts.append(cur)
        stops.append(cur + dt)
        cur = cur + dt
    s, e = np.asarray(starts), np.asarray(stops)
    return s, e, (s + e) / 2.0


def merge_jones(d0, d1, stream=None):
    """ClassJonesDomains.MergeJones of two DicoJones (their "Jones" torch tensors on the device)."""
    dev = require_gpu()
    dom = merged_freq_domains(d0["FreqDomains"], d1["FreqDomains"])
    t0, t1, tm = merged_time_domains(d0["t0"], d0["t1"], d1["t0"], d1["t1"])
    it0 = np.argmin(np.abs(tm[:, None] - np.asarray(d0["tm"])[None, :]), axis=1)
    it1 = np.argmin(np.abs(tm[:, None] - np.asarray(d1["tm"])[None, :]), axis=1)
    fm = dom.mean(axis=1)

    def containing(fd):
        out = np.full(fm.size, -1, np.int32)
        for i, f in enumerate(fm):
            hit = np.flatnonzero((f >= fd[:, 0]) & (f < fd[:, 1]))
            if hit.size:
                out[i] = hit[0]
        return out

    ich0, ich1 = containing(np.asarray(d0["FreqDomains"])), containing(np.asarray(d1["FreqDomains"]))
    j0, j1 = d0["Jones"].contiguous(), d1["Jones"].contiguous()
    nt0, nd, na, nch0 = j0.shape[:4]
    nt1, _, _, nch1 = j1.shape[:4]
    out = torch.empty((tm.size, nd, na, fm.size, 2, 2), dtype=torch.complex64, device=dev)
    maps = [torch.from_numpy(np.ascontiguousarray(x, np.int32)).to(dev) for x in (it0, it1, ich0, ich1)]
    m = abi.ddf_jones_merge_desc()
    m.j0, m.j1, m.out = j0.data_ptr(), j1.data_ptr(), out.data_ptr()
    m.it0, m.it1, m.ich0, m.ich1 = (x.data_ptr() for x in maps)
    m.nt0, m.nch0, m.nt1, m.nch1 = int(nt0), int(nch0), int(nt1), int(nch1)
    m.nd, m.na, m.nt, m.nch = int(nd), int(na), int(tm.size), int(fm.size)
    check(lib().ddf_jones_merge(C.byref(m), C.c_void_p(stream_handle(stream))))
    (stream or torch.cuda.current_stream(dev)).synchronize()
    return {"Jones": out, "t0": t0, "t1": t1, "tm": tm, "FreqDomains": dom}
