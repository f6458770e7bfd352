"""Delta-stamp decoding of the gridder's / degridder's integer placement maps.

A call with a one-block BDA map and CF planes whose every oversampling sub-kernel is zero except
at its centre, where it holds a code naming (plane, io, jo), touches exactly one grid cell: the
block's (locy, locx).  The gridder (PSF mode, unit weights) writes nvis * code there; the
degridder (a model grid holding 1 + i (x + N y) in cell (y, x)) subtracts code * (1 + i (x + N y))
from the block's visibilities.  Either way the cell, the band and the code reveal every integer
the placement formed (gridder.h:321-351, degridder.h:170-200): w-plane, conjugate choice, grid
cell, oversampling sub-kernel, band.  Facets keep their (Cu, Cv), which enter the placement,
but get l0 = m0 = 0, so the phasor is exactly 1; Jones and decorrelation are dropped (they do not
enter the placement).

Decoded rows are int32 [valid, iw, conj, locx, locy, io, jo, band], the layout of the oracle's
``cpuref.place``.
"""
import numpy as np

from ddfacet_amd.problems import Facet, Problem


def coded_planes(f):
    """Copies of facet f's CF lists with the centre of each sub-kernel set to its code
    1 + (plane * O + io) * O + jo (plane in put order: Wplanes k, WplanesConj nw + k)."""
    nw = len(f.wplanes)
    O = int(f.winfos[3])
    out = []
    for k, p in enumerate(list(f.wplanes) + list(f.wplanes_conj)):
        S = p.shape[0] // O
        a = np.zeros((O, O, S, S), np.complex64)
        io, jo = np.meshgrid(np.arange(O), np.arange(O), indexing="ij")
        a[:, :, S // 2, S // 2] = 1 + (k * O + io) * O + jo
        out.append(np.ascontiguousarray(a.reshape(S * O, S * O)))
    return out[:nw], out[nw:]


def decode_code(code, nw, O):
    c = int(code) - 1
    plane, rest = divmod(c, O * O)
    io, jo = divmod(rest, O)
    return plane % nw, int(plane >= nw), io, jo


def delta_problem(P):
    """P with coded CF lists, l0 = m0 = 0, Stokes I, no Jones, no decorrelation."""
    facets = []
    for f in P.facets:
        wp, wc = coded_planes(f)
        facets.append(Facet(l0=0.0, m0=0.0, cu=f.cu, cv=f.cv, wplanes=wp, wplanes_conj=wc, winfos=f.winfos))
    kw = {k: getattr(P, k) for k in P.__dataclass_fields__}
    kw.update(facets=facets, npol=1, stokes_codes=[1], jones=None, decorr=None,
              weights=np.ones_like(P.weights))
    return Problem(**kw)


def one_block_map(ch0, ch1, rows):
    return np.array([1, 0, 2 + len(rows), ch0, ch1] + [int(r) for r in rows], np.int32)


def blocks(bda):
    from ddfacet_amd import bda as bdamod
    return list(bdamod.iter_blocks(bda))


def grid_stamps(mod, D, fi):
    """Placement of every grid-map block of facet fi, decoded from the reference gridder."""
    f = D.facets[fi]
    nw, O = len(f.wplanes), int(f.winfos[3])
    full = D.bda_grid
    bl = blocks(full)
    out = np.zeros((len(bl), 8), np.int32)
    sp = D.sparsification
    for b, (c0, c1, rows) in enumerate(bl):
        D.bda_grid = one_block_map(c0, c1, rows)
        keep = D.sparsification
        D.sparsification = None if sp is None else np.array([sp[b]])
        g = D.new_grid(np.complex128)
        s = np.zeros((D.nband, D.npol))
        mod.pyGridderWPol64(*D.grid_args(fi, g, s, dopsf=1))
        D.sparsification = keep
        nz = np.argwhere(g != 0)
        if nz.size == 0:
            continue
        assert nz.shape[0] == 1, "a delta stamp touched %d cells" % nz.shape[0]
        band, _, y, x = (int(q) for q in nz[0])
        nvis = int(sum((~D.flags[int(r), c0:c1, 0].astype(bool)).sum() for r in rows if int(r) < D.uvw.shape[0]))
        code = g[band, 0, y, x].real / nvis
        assert abs(code - round(code)) < 1e-6 and g[band, 0, y, x].imag == 0
        iw, conj, io, jo = decode_code(round(code), nw, O)
        out[b] = (1, iw, conj, x, y, io, jo, band)
    D.bda_grid = full
    return out


def degrid_stamps(mod, D, fi):
    """Placement of every degrid-map block of facet fi, decoded from the reference degridder."""
    f = D.facets[fi]
    nw, O = len(f.wplanes), int(f.winfos[3])
    N = D.npix
    yy, xx = np.mgrid[0:N, 0:N]
    cell = (1 + 1j * (xx + N * yy)).astype(np.complex64)
    full = D.bda_degrid
    bl = blocks(full)
    out = np.zeros((len(bl), 8), np.int32)
    for b, (c0, c1, rows) in enumerate(bl):
        D.bda_degrid = one_block_map(c0, c1, rows)
        rr = [int(r) for r in rows if int(r) < D.uvw.shape[0]]
        if not rr:
            continue
        for band in range(D.nband):
            model = np.zeros((D.nband, 1, N, N), np.complex64)
            model[band, 0] = cell
            v = np.zeros(D.vis.shape, np.complex64)
            mod.pyDeGridderWPol(*D.degrid_args(fi, model, v))
            val = -v[rr[0], c0, 0]
            if val == 0:
                continue
            code = float(val.real)
            assert code == round(code)
            idx = float(val.imag) / code
            assert idx == round(idx)
            x, y = int(round(idx)) % N, int(round(idx)) // N
            iw, first_list, io, jo = decode_code(round(code), nw, O)
            # get() passes WplanesConj first (ClassDDEGridMachine.py:1091-1092): the put-order
            # conjugate list serves W > 0 here
            out[b] = (1, iw, 1 - first_list, x, y, io, jo, band)
            break
    D.bda_degrid = full
    return out
