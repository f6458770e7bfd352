"""Synthetic observations in the layouts the gridder consumes (ClassVisServer chunk dict,
DDFacet/Data/ClassVisServer.py:512-608): uvw f64[R,3], vis c64[R,C,4], flags bool[R,C,4],
weights f32[R,C], A0/A1 int32[R], freqs f64[C], Jones c64[nt,nd,na,nch,2,2] with its time and
channel maps (DDFacet/Data/ClassJones.py:415-439, 1004-1012).

The reference reads uvw from the Measurement Set; here they come from earth rotation of
baseline vectors (standard (H, dec) rotation).  Seeds follow seed = 1000*cfg + k with
k = 0 array, 1 sky, 2 noise, 3 flags, 4 Jones.
"""
from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0
OMEGA_EARTH = 7.2921150e-5  # rad / s
ARCSEC = np.pi / 180.0 / 3600.0


def layout_disk(nant, diameter_m, rng):
    r = 0.5 * diameter_m * np.sqrt(rng.random(nant))
    t = 2 * np.pi * rng.random(nant)
    return np.stack([r * np.cos(t), r * np.sin(t), np.zeros(nant)], 1)


def layout_meerkat_like(nant, rng, core_frac=0.75, core_diameter=1000.0, arm_length=8000.0):
    """Dense core of `core_frac` of the antennas in a disk of `core_diameter`, the rest on three
    arms (120 degrees apart) log-spaced out to `arm_length`."""
    ncore = int(round(core_frac * nant))
    core = layout_disk(ncore, core_diameter, rng)
    narm = nant - ncore
    k = np.arange(narm)
    arm = k % 3
    radius = np.geomspace(0.6 * core_diameter, arm_length, narm)[k]
    ang = np.deg2rad(90.0 + 120.0 * arm) + 0.05 * rng.standard_normal(narm)
    arms = np.stack([radius * np.cos(ang), radius * np.sin(ang), np.zeros(narm)], 1)
    return np.concatenate([core, arms], 0)


def earth_rotation_uvw(enu, lat_rad, dec_rad, hour_angles):
    """uvw (metres) of every baseline (a0 < a1) at every hour angle, rows time-major
    (the Measurement Set order).  Returns uvw, a0, a1, time index."""
    e, n, u = enu[:, 0], enu[:, 1], enu[:, 2]
    x = -np.sin(lat_rad) * n + np.cos(lat_rad) * u
    y = e
    z = np.cos(lat_rad) * n + np.sin(lat_rad) * u
    xyz = np.stack([x, y, z], 1)
    na = enu.shape[0]
    p, q = np.triu_indices(na, 1)
    b = xyz[q] - xyz[p]  # [nbl, 3]
    h = np.asarray(hour_angles)[:, None]
    sd, cd = np.sin(dec_rad), np.cos(dec_rad)
    bu = np.sin(h) * b[None, :, 0] + np.cos(h) * b[None, :, 1]
    bv = -sd * np.cos(h) * b[None, :, 0] + sd * np.sin(h) * b[None, :, 1] + cd * b[None, :, 2]
    bw = cd * np.cos(h) * b[None, :, 0] - cd * np.sin(h) * b[None, :, 1] + sd * b[None, :, 2]
    uvw = np.stack([bu, bv, bw], -1).reshape(-1, 3)
    nt = h.shape[0]
    a0 = np.tile(p, nt).astype(np.int32)
    a1 = np.tile(q, nt).astype(np.int32)
    tidx = np.repeat(np.arange(nt), p.size).astype(np.int32)
    return np.ascontiguousarray(uvw), a0, a1, tidx


@dataclass
class Sky:
    flux: np.ndarray
    l: np.ndarray
    m: np.ndarray


def sky_loguniform(nsrc, half_width_rad, rng, smin=0.01, smax=1.0):
    flux = np.exp(rng.uniform(np.log(smin), np.log(smax), nsrc))
    l = rng.uniform(-half_width_rad, half_width_rad, nsrc)
    m = rng.uniform(-half_width_rad, half_width_rad, nsrc)
    return Sky(flux, l, m)


def sky_powerlaw(nsrc, half_width_rad, rng, smin=1e-3, smax=1.0, index=-1.5):
    """N(>S) proportional to S^index between smin and smax (inverse-CDF sampling)."""
    a = index
    uu = rng.random(nsrc)
    flux = (smin ** a - uu * (smin ** a - smax ** a)) ** (1.0 / a)
    l = rng.uniform(-half_width_rad, half_width_rad, nsrc)
    m = rng.uniform(-half_width_rad, half_width_rad, nsrc)
    return Sky(flux, l, m)


def predict_point_sources(uvw, freqs, sky, device="cpu", chunk_rows=65536):
    """V = sum_s S exp(+2 pi i nu/c (u l + v m + w (n - 1))) on XX and YY, XY = YX = 0, the
    predict convention of ClassDDEGridMachine.testGrid (ClassDDEGridMachine.py:229-271).
    FP64 evaluation (torch, on `device`), complex64 result [R, C, 4] as a torch tensor."""
    import torch
    R = uvw.shape[0]
    nchan = freqs.shape[0]
    out = torch.zeros((R, nchan, 4), dtype=torch.complex64, device=device)
    f = torch.as_tensor(freqs, dtype=torch.float64, device=device)
    l = torch.as_tensor(sky.l, dtype=torch.float64, device=device)
    m = torch.as_tensor(sky.m, dtype=torch.float64, device=device)
    n1 = torch.sqrt(1 - l * l - m * m) - 1
    s = torch.as_tensor(sky.flux, dtype=torch.float64, device=device)
    for r0 in range(0, R, chunk_rows):
        q = torch.as_tensor(uvw[r0:r0 + chunk_rows], dtype=torch.float64, device=device)
        geo = q[:, 0:1] * l[None] + q[:, 1:2] * m[None] + q[:, 2:3] * n1[None]  # [r, s]
        ph = (2 * np.pi / C_LIGHT) * geo[:, None, :] * f[None, :, None]  # [r, c, s]
        v = (torch.polar(torch.ones_like(ph), ph) * s[None, None, :]).sum(-1)
        out[r0:r0 + chunk_rows, :, 0] = v.to(torch.complex64)
        out[r0:r0 + chunk_rows, :, 3] = v.to(torch.complex64)
    return out


@dataclass
class Observation:
    uvw: np.ndarray
    a0: np.ndarray
    a1: np.ndarray
    tidx: np.ndarray
    times: np.ndarray
    freqs: np.ndarray
    dfreq: float
    nant: int
    extra: dict = field(default_factory=dict)

    @property
    def nrow(self):
        return self.uvw.shape[0]

    @property
    def nchan(self):
        return self.freqs.shape[0]


def make_observation(layout, nant, ntime, dt, freqs, dec_deg, lat_deg, seed, sort_by_baseline=False,
                     diameter_m=20000.0):
    rng = np.random.default_rng(seed)
    if layout == "disk":
        enu = layout_disk(nant, diameter_m, rng)
    elif layout == "meerkat":
        enu = layout_meerkat_like(nant, rng)
    else:
        raise ValueError(layout)
    times = (np.arange(ntime) - (ntime - 1) / 2.0) * dt
    ha = times * OMEGA_EARTH
    uvw, a0, a1, tidx = earth_rotation_uvw(enu, np.deg2rad(lat_deg), np.deg2rad(dec_deg), ha)
    if sort_by_baseline:
        order = np.lexsort((tidx, a1, a0))
        uvw, a0, a1, tidx = uvw[order].copy(), a0[order].copy(), a1[order].copy(), tidx[order].copy()
    freqs = np.asarray(freqs, dtype=np.float64)
    dfreq = float(freqs[1] - freqs[0]) if freqs.size > 1 else 1.0
    return Observation(uvw, a0, a1, tidx, times, freqs, dfreq, nant)


def random_flags(nrow, nchan, ncorr, frac, seed):
    """Flags drawn per (row, channel) and equalised over correlations (ClassMS.py:1165-1167)."""
    rng = np.random.default_rng(seed)
    f = rng.random((nrow, nchan)) < frac
    return np.ascontiguousarray(np.repeat(f[:, :, None], ncorr, axis=2))


def random_jones(nt, nd, na, nch, seed, amp_sigma=0.05, phase_sigma=0.3):
    """Identity plus complex perturbations: J = (I + a) * exp(i phi), a ~ N(0, amp_sigma),
    phi ~ N(0, phase_sigma), elementwise on the 2x2 matrices."""
    rng = np.random.default_rng(seed)
    eye = np.eye(2)[None, None, None, None]
    a = amp_sigma * rng.standard_normal((nt, nd, na, nch, 2, 2))
    phi = phase_sigma * rng.standard_normal((nt, nd, na, nch, 2, 2))
    return ((eye + a) * np.exp(1j * phi)).astype(np.complex64)


def facet_centres(nfx, facet_npix, cell_rad):
    """(l0, m0) of an nfx x nfx regular facet layout, facet index = iy * nfx + ix."""
    c = (np.arange(nfx) - (nfx - 1) / 2.0) * facet_npix * cell_rad
    ly, lx = np.meshgrid(c, c, indexing="ij")
    return np.stack([lx.ravel(), ly.ravel()], 1)
