"""Baseline-dependent-averaging (BDA) block maps in the reference int32 format.

Format (DDFacet/Data/ClassSmearMapping.py:83-135):
    [NTot_lo, NTot_hi, size_0 .. size_{NTot-1}, rec_0, rec_1, ...]
with rec_b = [chStart, chEnd, row, row, ...] and size_b = len(rec_b) = 2 + rows_b.  Per-baseline
block lists are concatenated in lexicographic order of the string key "a0:a1"
(ClassSmearMapping.py:113).

``bda_map_device`` builds the maps of the default block builder ``GiveBlocksRowsListBL_old``
(ClassSmearMapping.py:496-625, ``--Comp-BDAMode=1``) and of the phase-based builder
``GiveBlocksRowsListBL`` (ClassSmearMapping.py:348-491, ``--Comp-BDAMode=2``) on the GPU
(ddf_bda_build_*), in place of ``computeSmearMappingInBackground`` / ``collectSmearMapping``
(ClassSmearMapping.py:64-135, ClassVisServer.py:637-679).  The other builders are synthetic
layouts used by the benchmarks: one block per (row, channel) and fixed time x frequency
averaging factors per baseline.
"""
import numpy as np

C_BDA = 3e8  # the block builder uses 3e8, not 299792458 (ClassSmearMapping.py:508)


def pack(per_baseline):
    """per_baseline: dict key "a0:a1" -> (list of sizes, flat int list).  Returns the int32 map."""
    keys = sorted(per_baseline.keys())
    sizes = [np.asarray(per_baseline[k][0], np.int64) for k in keys]
    recs = [np.asarray(per_baseline[k][1], np.int64) for k in keys]
    nblk = int(sum(s.size for s in sizes))
    nrec = int(sum(r.size for r in recs))
    out = np.zeros(2 + nblk + nrec, np.int32)
    out[:2] = np.array([nblk & 0xFFFFFFFF, nblk >> 32], np.uint64).astype(np.uint32).view(np.int32)
    if nblk:
        out[2:2 + nblk] = np.concatenate(sizes)
        out[2 + nblk:] = np.concatenate(recs)
    return out


def nblocks_of(bda):
    return int(np.uint32(bda[0])) + (int(bda[1]) << 32)


def iter_blocks(bda):
    """Yield (chStart, chEnd, rows) of every block, in map order."""
    nb = nblocks_of(bda)
    sizes = bda[2:2 + nb]
    pos = 2 + nb
    for s in sizes:
        rec = bda[pos:pos + s]
        yield int(rec[0]), int(rec[1]), rec[2:]
        pos += s


def block_starts(bda):
    """Word offset of each block record (what ddf_bda_index computes on the device)."""
    nb = nblocks_of(bda)
    sizes = bda[2:2 + nb].astype(np.int64)
    starts = np.empty(nb, np.int64)
    if nb:
        starts[0] = 2 + nb
        np.cumsum(sizes[:-1], out=starts[1:])
        starts[1:] += 2 + nb
    return starts


def _baseline_rows(a0, a1):
    """Row indices of each ordered antenna pair present in the data (a0 != a1)."""
    out = {}
    order = np.lexsort((a1, a0))
    pairs = np.stack([a0[order], a1[order]], 1)
    if order.size == 0:
        return out
    brk = np.nonzero(np.any(pairs[1:] != pairs[:-1], axis=1))[0] + 1
    for seg in np.split(np.arange(order.size), brk):
        p, q = int(pairs[seg[0], 0]), int(pairs[seg[0], 1])
        if p == q:
            continue
        out["%d:%d" % (p, q)] = np.sort(order[seg])
    return out


def bda_map_per_visibility(a0, a1, nchan):
    """One block per (row, channel): no averaging."""
    per_bl = {}
    for key, ind in _baseline_rows(np.asarray(a0), np.asarray(a1)).items():
        n = ind.size
        rec = np.empty((n, nchan, 3), np.int64)
        rec[:, :, 0] = np.arange(nchan)[None, :]
        rec[:, :, 1] = np.arange(nchan)[None, :] + 1
        rec[:, :, 2] = ind[:, None]
        per_bl[key] = (np.full(n * nchan, 3, np.int64), rec.ravel())
    return pack(per_bl)


def bda_map_factors(a0, a1, nchan, factor_of_baseline, chan_map=None):
    """Fixed averaging per baseline: factor_of_baseline(key) -> (k_time, k_freq).  Channel blocks
    are also cut where chan_map changes, so a block never straddles two grid bands."""
    chan_map = np.zeros(nchan, np.int64) if chan_map is None else np.asarray(chan_map)
    per_bl = {}
    for key, ind in _baseline_rows(np.asarray(a0), np.asarray(a1)).items():
        kt, kf = factor_of_baseline(key)
        cuts = sorted(set(list(range(0, nchan, max(1, kf))) + [nchan] +
                          [int(c) for c in np.nonzero(np.diff(chan_map))[0] + 1]))
        sizes, flat = [], []
        for t0 in range(0, ind.size, max(1, kt)):
            rows = ind[t0:t0 + kt].tolist()
            for c0, c1 in zip(cuts[:-1], cuts[1:]):
                sizes.append(2 + len(rows))
                flat.extend([c0, c1] + rows)
        per_bl[key] = (sizes, flat)
    return pack(per_bl)


def _key_ranks(na):
    """rank[a0, a1]: position of "a0:a1" among the sorted string keys of every ordered antenna
    pair (the order collectSmearMapping concatenates the baselines in)."""
    keys = sorted("%d:%d" % (p, q) for p in range(na) for q in range(na) if p != q)
    rank = np.zeros((na, na), np.int32)
    for i, k in enumerate(keys):
        p, q = (int(x) for x in k.split(":"))
        rank[p, q] = i
    return rank


def bda_map_device(uvw, a0, a1, freqs, dfreq, decorr, cell_rad, facet_npix, chan_map, mode=1, na=None, stream=None):
    """The BDA map of bda_map_mode1 (mode 1) or bda_map_mode2 (mode 2) built on the device
    (ddf_bda_build_*), word for word the same int32 map.  uvw float64 [nrow, 3], a0 / a1 int32
    [nrow], chan_map int32 [nchan]: torch tensors on the GPU (or numpy arrays, uploaded);
    freqs / dfreq / decorr / cell_rad / facet_npix as the host builders take them.  Returns an
    int32 tensor on the GPU."""
    import ctypes as C

    import torch

    from . import abi
    from ._lib import check, lib
    from .device import require_gpu, stream_handle, to_device
    dev = require_gpu()
    t_uvw = to_device(np.ascontiguousarray(uvw, np.float64), dev) if not isinstance(uvw, torch.Tensor) else uvw
    t_a0 = to_device(np.ascontiguousarray(a0, np.int32), dev) if not isinstance(a0, torch.Tensor) else a0
    t_a1 = to_device(np.ascontiguousarray(a1, np.int32), dev) if not isinstance(a1, torch.Tensor) else a1
    t_cm = to_device(np.ascontiguousarray(chan_map, np.int32), dev) if not isinstance(chan_map, torch.Tensor) else chan_map
    for t, dt in ((t_uvw, torch.float64), (t_a0, torch.int32), (t_a1, torch.int32), (t_cm, torch.int32)):
        if t.dtype != dt or not t.is_contiguous() or t.device.type != "cuda":
            raise TypeError("uvw float64, a0 / a1 / chan_map int32, contiguous on the GPU")
    nrow = int(t_uvw.shape[0])
    freqs = np.asarray(freqs, np.float64)
    if na is None:
        na = int(max(int(t_a0.max()), int(t_a1.max()))) + 1 if nrow else 1
    decorr_in = 1.0 - decorr
    dphi = np.sqrt(6. * (1. - decorr_in))
    nu0 = np.max(freqs)
    s = abi.ddf_bda_spec()
    s.mode, s.nchan, s.na = int(mode), int(freqs.size), int(na)
    s.dfreq = float(dfreq)
    if mode == 1:
        fov_deg = cell_rad * facet_npix * (np.sqrt(2.) / 2.) * 180. / np.pi
        lrad = fov_deg * np.pi / 180
        s.duv_cut = float(C_BDA * dphi / (np.pi * lrad * nu0))
        s.k_dnu = float((C_BDA / np.pi) * dphi)
        s.l = float(lrad)
    elif mode == 2:
        lmax = cell_rad * facet_npix * (np.sqrt(2.) / 2.)
        s.lmax = float(lmax)
        s.nmax = float(abs(np.sqrt(1.0 - lmax ** 2) - 1.0))
        s.k_phase = float(2.0 * np.pi * nu0 / C_BDA)
        s.k_span = float((C_BDA / (2.0 * np.pi)) * dphi)
        s.dphi = float(dphi)
        s.per_row = 0 if (C_BDA * dphi / (2.0 * np.pi * nu0)) else 1
    else:
        raise ValueError("BDA mode must be 1 or 2")
    rank = _key_ranks(int(na))
    st = C.c_void_p(stream_handle(stream))
    L = lib()
    b = C.c_void_p()
    nw = C.c_int64()
    check(L.ddf_bda_build_create(C.byref(s), C.c_void_p(t_uvw.data_ptr()), C.c_void_p(t_a0.data_ptr()),
                                 C.c_void_p(t_a1.data_ptr()), nrow, C.c_void_p(t_cm.data_ptr()),
                                 rank.ctypes.data_as(C.c_void_p), C.byref(b), C.byref(nw), st))
    try:
        out = torch.empty(int(nw.value), dtype=torch.int32, device=dev)
        check(L.ddf_bda_build_write(b, C.c_void_p(out.data_ptr()), out.numel(), st))
        torch.cuda.current_stream(dev).synchronize() if stream is None else stream.synchronize()
    finally:
        L.ddf_bda_build_destroy(b)
    return out
