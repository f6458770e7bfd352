"""Oracle (test infrastructure only): numpy restatement of the beam Jones producers.

Only tests/ and bench.py's CPU-baseline leg import this module; the product path runs
ddfacet_amd.beams on the device.

* ``gmrt_channel_model`` / ``raw_beam_gmrt`` follow ClassGMRTBeam.calcCoefs / GiveRawBeam
  (DDFacet/Data/ClassGMRTBeam.py:82-157): an even radial polynomial in d * nu / 1 GHz (d the
  angular distance from the phase centre in arcmin), per band of validity, clipped to its value at
  the first positive stationary point beyond it.
* ``atca_channel_model`` / ``raw_beam_atca`` follow ClassATCABeam.calcCoefs / GiveRawBeam
  (DDFacet/Data/ClassATCABeam.py:75-233, the "csiro" tables): the square root of a polynomial in
  d * nu0 * 1e-9 at the two bracketing band frequencies nu0, mixed with the weights |nu - nu0| of
  the reference, and the reference's value past 100 arcmin GHz.
* ``instrument_beam`` is GiveInstrumentBeam (ClassGMRTBeam.py:161-173): the mean of the channels of
  each Jones channel (``vis_to_jones_chan``, ClassGMRTBeam.py:69-80).
* ``estimate_beam`` is ClassJones.EstimateBeam (DDFacet/Data/ClassJones.py:1014-1065): one beam per
  interval of the beam sample times, at the interval centre, optionally normalised by the inverse of
  the beam at the phase centre (CenterNorm), stored as complex64 [t][dir][ant][chan][2][2].
* ``batch_inverse`` / ``batch_dot`` are ModLinAlg.BatchInverse / BatchDot
  (DDFacet/Array/ModLinAlg.py:50-134), the closed-form 2 x 2 inverse and product.
* ``merged_freq_domains`` / ``merge_jones`` are ClassJonesDomains.GetMergedFreqDomains /
  MergeJones (DDFacet/Other/ClassJonesDomains.py:99-212), the merge behind ClassJones.MergeJones
  (ClassJones.py:1067-1072): the union of the two frequency grids (edges closer than 1 Hz merged),
  clipped to the overlap; time intervals cut at every end of either table; each merged cell the
  product G0 G1 of the nearest-centre time and the containing channel of each table (identity where
  a table has no containing channel).
"""
import numpy as np

ARCMIN_PER_RAD = 180.0 / np.pi * 60


# ---- geometry -----------------------------------------------------------------------------------
def ang_dist(ra0, dec0, ra1, dec1):
    """Great-circle distance (rad), the cosine clipped to [-1, 1] (ClassGMRTBeam.AngDist)."""
    d = np.sin(dec0) * np.sin(dec1) + np.cos(dec0) * np.cos(dec1) * np.cos(ra0 - ra1)
    d = np.clip(np.asarray(d, np.float64), -1.0, 1.0)
    return np.arccos(d)


def freq_domains(chan_freqs, chan_width, nchan_jones):
    """[nchan_jones, 2] equal frequency domains over the band (ClassGMRTBeam.CalcFreqDomains)."""
    f = np.asarray(chan_freqs, np.float64).ravel()
    n = nchan_jones if nchan_jones else f.size
    edges = np.linspace(f.min() - chan_width / 2.0, f.max() + chan_width / 2.0, n + 1)
    return np.array([[edges[i], edges[i + 1]] for i in range(n)])


def vis_to_jones_chan(chan_freqs, domains):
    centre = (domains[:, 0] + domains[:, 1]) / 2.0
    d = np.abs(np.asarray(chan_freqs, np.float64).reshape(-1, 1) - centre.reshape(1, -1))
    return np.argmin(d, axis=1)


# ---- GMRT ---------------------------------------------------------------------------------------
GMRT_BANDS = {
    3: (250e6, 500e6, [1, 0, -2.939 / 1e3, 0, 33.312 / 1e7, 0, -16.659 / 1e10, 0, 3.066 / 1e13]),
    4: (550e6, 850e6, [1, 0, -3.190 / 1e3, 0, 38.642 / 1e7, 0, -20.471 / 1e10, 0, 3.964 / 1e13]),
    5: (1050e6, 1450e6, [1, 0, -2.608 / 1e3, 0, 27.357 / 1e7, 0, -13.091 / 1e10, 0, 2.368 / 1e13]),
}


def gmrt_band_null(coefs):
    """(x0, P(x0)): the first positive real root of P' in numpy's root order, P in ascending
    coefficients (ClassGMRTBeam.calcCoefs)."""
    c = np.array(coefs, np.float64)[::-1]
    p = np.poly1d(c)
    r = np.roots(np.polyder(p, m=1))
    ind = np.where((np.abs(r.imag) < 1e-6) & (r.real > 0.0))[0][0]
    xn = np.abs(r[ind])
    return xn, p(xn)


def gmrt_channel_model(chan_freqs):
    """Per channel: coefficients [nch, 9], x_null (arcmin) [nch], y_null [nch] (float32 as the
    reference stores them)."""
    f = np.asarray(chan_freqs, np.float64).ravel()
    nulls = {b: gmrt_band_null(c) for b, (_, _, c) in GMRT_BANDS.items()}
    C = np.zeros((f.size, 9), np.float32)
    xn = np.zeros(f.size, np.float32)
    yn = np.zeros(f.size, np.float32)
    for i, fi in enumerate(f):
        for b, (f0, f1, c) in GMRT_BANDS.items():
            if f0 <= fi <= f1:
                C[i] = c
                xn[i] = nulls[b][0] / (fi / 1e9)
                yn[i] = nulls[b][1]
                break
        else:
            raise NotImplementedError("The GMRT beam is not modeled at %.3f MHz" % (fi / 1e6))
    return C, xn, yn


def raw_beam_gmrt(d_arcmin, chan_freqs, model):
    """Real beam value [ndir, nch] (the diagonal of the reference's complex Jones)."""
    C, xn, yn = model
    f = np.asarray(chan_freqs, np.float64).ravel()
    d = np.asarray(d_arcmin, np.float64)
    out = np.zeros((d.size, f.size))
    for ich in range(f.size):
        dnu = d * f[ich] / 1e9
        b = np.polynomial.polynomial.polyval(dnu, C[ich])
        b[d > xn[ich]] = yn[ich]
        out[:, ich] = b
    return out


# ---- ATCA ("csiro") -----------------------------------------------------------------------------
ATCA_BANDS = [
    (1268e6, [1, 0, -1.154e-03, 0, +5.351e-07, 0, -1.265e-10, 0, +1.634e-14, 0, -1.099e-18, 0, +3.009e-23]),
    (1396e6, [1, 0, -1.092e-03, 0, +4.660e-07, 0, -9.875e-11, 0, +1.119e-14, 0, -6.504e-19, 0, +1.522e-23]),
    (1524e6, [1, 0, -1.048e-03, 0, +4.305e-07, 0, -8.824e-11, 0, +9.684e-15, 0, -5.446e-19, 0, +1.232e-23]),
    (1652e6, [1, 0, -1.033e-03, 0, +4.221e-07, 0, -8.692e-11, 0, +9.679e-15, 0, -5.571e-19, 0, +1.298e-23]),
    (1780e6, [1, 0, -1.003e-03, 0, +4.016e-07, 0, -8.170e-11, 0, +9.046e-15, 0, -5.199e-19, 0, +1.213e-23]),
    (1908e6, [1, 0, -9.794e-04, 0, +3.831e-07, 0, -7.562e-11, 0, +8.067e-15, 0, -4.440e-19, 0, +9.882e-24]),
    (2036e6, [1, 0, -9.692e-04, 0, +3.848e-07, 0, -7.809e-11, 0, +8.621e-15, 0, -4.927e-19, 0, +1.140e-23]),
    (2164e6, [1, 0, -9.919e-04, 0, +4.028e-07, 0, -8.251e-11, 0, +9.077e-15, 0, -5.122e-19, 0, +1.163e-23]),
    (2292e6, [1, 0, -1.007e-03, 0, +4.131e-07, 0, -8.480e-11, 0, +9.305e-15, 0, -5.220e-19, 0, +1.176e-23]),
    (2420e6, [1, 0, -1.016e-03, 0, +4.235e-07, 0, -8.817e-11, 0, +9.772e-15, 0, -5.521e-19, 0, +1.251e-23]),
    (2548e6, [1, 0, -1.033e-03, 0, +4.375e-07, 0, -9.223e-11, 0, +1.033e-14, 0, -5.885e-19, 0, +1.343e-23]),
    (2676e6, [1, 0, -1.035e-03, 0, +4.395e-07, 0, -9.269e-11, 0, +1.038e-14, 0, -5.919e-19, 0, +1.353e-23]),
    (2804e6, [1, 0, -1.064e-03, 0, +4.622e-07, 0, -9.913e-11, 0, +1.125e-14, 0, -6.502e-19, 0, +1.505e-23]),
    (2932e6, [1, 0, -1.081e-03, 0, +4.769e-07, 0, -1.045e-10, 0, +1.222e-14, 0, -7.325e-19, 0, +1.768e-23]),
]


def atca_channel_model(chan_freqs):
    """Per channel: the bracketing bands (band0 with nu0 <= nu, band1 with nu0 >= nu; band 0 and
    the last band when there is none)."""
    f = np.asarray(chan_freqs, np.float64).ravel()
    nb = len(ATCA_BANDS)
    b0 = np.zeros(f.size, int)
    b1 = np.zeros(f.size, int)
    for i, fi in enumerate(f):
        lo, hi = [0], [nb - 1]
        for k, (f0, _) in enumerate(ATCA_BANDS):
            if f0 >= fi:
                hi.append(k)
            if f0 <= fi:
                lo.append(k)
        b0[i], b1[i] = max(lo), min(hi)
    return b0, b1


def raw_beam_atca(d_arcmin, chan_freqs, model):
    b0s, b1s = model
    f = np.asarray(chan_freqs, np.float64).ravel()
    d = np.asarray(d_arcmin, np.float64)
    pv = np.polynomial.polynomial.polyval
    out = np.zeros((d.size, f.size))
    for ich in range(f.size):
        f00, c0 = ATCA_BANDS[b0s[ich]]
        f01, c1 = ATCA_BANDS[b1s[ich]]
        dnu = d * f[ich] * 1e-9
        dnu0 = d * f00 * 1e-9
        dnu1 = d * f01 * 1e-9
        df0 = np.abs(f[ich] - f00)
        df1 = np.abs(f[ich] - f01)
        if df0 == 0 and df1 == 0:
            b = np.sqrt(np.abs(pv(dnu, np.array(c0))))
            ynull = np.sqrt(np.abs(pv(100.0, np.array(c0))))
        else:
            beam0 = np.sqrt(np.abs(pv(dnu0, np.array(c0))))
            beam1 = np.sqrt(np.abs(pv(dnu1, np.array(c1))))
            b = (df0 * beam0 + df1 * beam1) / (df0 + df1)
            y0 = np.sqrt(np.abs(pv(100.0, np.array(c0))))
            y1 = np.sqrt(np.abs(pv(100.0, np.array(c1))))
            ynull = (df0 * y0 + df1 * y1) / (df0 + df1)
        ynull = np.float32(ynull)  # the reference keeps yNull in a float32 array
        m = dnu > np.float32(100.0)
        b[m] = ynull / (dnu[m] - ynull)
        out[:, ich] = b
    return out


# ---- Jones products -------------------------------------------------------------------------------
def batch_inverse(a):
    s = a.shape
    a = a.reshape(-1, 2, 2)
    a0, d0, b0, c0 = a[:, 0, 0], a[:, 1, 1], a[:, 0, 1], a[:, 1, 0]
    det = 1.0 / (a0 * d0 - b0 * c0)
    out = np.zeros_like(a)
    out[:, 0, 0] = d0 * det
    out[:, 0, 1] = -b0 * det
    out[:, 1, 0] = -c0 * det
    out[:, 1, 1] = a0 * det
    return out.reshape(s)


def batch_dot(a, b):
    s = a.shape
    a = a.reshape(-1, 2, 2)
    b = b.reshape(-1, 2, 2)
    out = np.zeros_like(a)
    out[:, 0, 0] = a[:, 0, 0] * b[:, 0, 0] + a[:, 0, 1] * b[:, 1, 0]
    out[:, 0, 1] = a[:, 0, 0] * b[:, 0, 1] + a[:, 0, 1] * b[:, 1, 1]
    out[:, 1, 0] = a[:, 1, 0] * b[:, 0, 0] + a[:, 1, 1] * b[:, 1, 0]
    out[:, 1, 1] = a[:, 1, 0] * b[:, 0, 1] + a[:, 1, 1] * b[:, 1, 1]
    return out.reshape(s)


def instrument_beam(raw, mapping, nchj, na):
    """complex128 [ndir, na, nchj, 2, 2] from the real raw beam [ndir, nch]."""
    nd = raw.shape[0]
    beam = np.zeros((nd, na, raw.shape[1], 2, 2), complex)
    beam[:, :, :, 0, 0] = raw[:, None, :]
    beam[:, :, :, 1, 1] = raw[:, None, :]
    mean = np.zeros((nd, na, nchj, 2, 2), complex)
    for j in range(nchj):
        ind = np.where(mapping == j)[0]
        mean[:, :, j] = np.mean(beam[:, :, ind], axis=2)
    return mean


def estimate_beam(times_beam, ra, dec, ra_c, dec_c, chan_freqs, chan_width, nband, na, kind="GMRT",
                  center_norm=True):
    t = np.asarray(times_beam, np.float64)
    t0, t1 = t[:-1].copy(), t[1:].copy()
    tm = (t0 + t1) / 2.0
    dom = freq_domains(chan_freqs, chan_width, nband)
    mapping = vis_to_jones_chan(chan_freqs, dom)
    model = gmrt_channel_model(chan_freqs) if kind == "GMRT" else atca_channel_model(chan_freqs)
    raw_f = raw_beam_gmrt if kind == "GMRT" else raw_beam_atca
    ra = np.asarray(ra, np.float64)
    dec = np.asarray(dec, np.float64)
    nd = ra.size
    jones = np.zeros((tm.size, nd, na, dom.shape[0], 2, 2), np.complex64)
    for it in range(tm.size):
        beam = instrument_beam(raw_f(ang_dist(ra, dec, ra_c, dec_c) * ARCMIN_PER_RAD, chan_freqs, model),
                               mapping, dom.shape[0], na)
        if center_norm:
            b0 = instrument_beam(raw_f(ang_dist(np.array([ra_c]), np.array([dec_c]), ra_c, dec_c) * ARCMIN_PER_RAD,
                                       chan_freqs, model), mapping, dom.shape[0], na)
            b0inv = batch_inverse(b0) * np.ones((nd, 1, 1, 1, 1), np.float32)
            beam = batch_dot(b0inv, beam)
        jones[it] = beam
    return {"Jones": jones, "t0": t0, "t1": t1, "tm": tm, "FreqDomains": dom, "VisToJonesChanMapping": mapping}


# ---- domain merge ---------------------------------------------------------------------------------
def merged_freq_domains(fd0, fd1):
    ff = np.array(sorted(set(np.asarray(fd0).ravel().tolist() + np.asarray(fd1).ravel().tolist())))
    skip = np.zeros(ff.size, bool)
    keep = []
    for i in range(ff.size):
        if skip[i]:
            continue
        skip[np.where(np.abs(ff - ff[i]) < 1.0)[0]] = True
        keep.append(ff[i])
    fmin = max(np.min(fd0), np.min(fd1))
    fmax = min(np.max(fd0), np.max(fd1))
    ff = np.array(keep)
    out = np.zeros((ff.size - 1, 2))
    out[:, 0] = ff[:-1]
    out[:, 1] = ff[1:]
    fm = out.mean(axis=1)
    return out[(fm >= fmin) & (fm < fmax)]


def merged_times(t0a, t1a, t0b, t1b):
    cur = min(t0a[0], t0b[0])
    T0, T1 = [], []
    while True:
        d0 = np.asarray(t1a) - cur
        d0 = d0[d0 > 0]
        d1 = np.asarray(t1b) - cur
        d1 = d1[d1 > 0]
        if d0.size == 0 and d1.size == 0:
            break
        dt = d1[0] if d0.size == 0 else d0[0] if d1.size == 0 else min(d0[0], d1[0])
        T0.append(cur)
        T1.append(cur + dt)
        cur = cur + dt
    T0, T1 = np.array(T0), np.array(T1)
    return T0, T1, (T0 + T1) / 2.0


def merge_jones(d0, d1):
    dom = merged_freq_domains(d0["FreqDomains"], d1["FreqDomains"])
    t0, t1, tm = merged_times(d0["t0"], d0["t1"], d1["t0"], d1["t1"])
    nt = tm.size
    _, nd, na, _, _, _ = d0["Jones"].shape
    fm = dom.mean(axis=1)
    out = np.zeros((nt, nd, na, fm.size, 2, 2), np.complex64)
    out[..., 0, 0] = 1.0
    out[..., 1, 1] = 1.0
    i0 = np.argmin(np.abs(tm.reshape(-1, 1) - np.asarray(d0["tm"]).reshape(1, -1)), axis=1)
    i1 = np.argmin(np.abs(tm.reshape(-1, 1) - np.asarray(d1["tm"]).reshape(1, -1)), axis=1)
    for ich in range(fm.size):
        c0 = np.where((fm[ich] >= d0["FreqDomains"][:, 0]) & (fm[ich] < d0["FreqDomains"][:, 1]))[0]
        if c0.size == 0:
            continue
        c1 = np.where((fm[ich] >= d1["FreqDomains"][:, 0]) & (fm[ich] < d1["FreqDomains"][:, 1]))[0]
        if c1.size == 0:
            continue
        for it in range(nt):
            out[it, :, :, ich] = batch_dot(d0["Jones"][i0[it], :, :, c0[0]], d1["Jones"][i1[it], :, :, c1[0]])
    return {"Jones": out, "t0": t0, "t1": t1, "tm": tm, "FreqDomains": dom}
