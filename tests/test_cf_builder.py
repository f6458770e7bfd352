"""The host CF builder (ddfacet_amd/wterm.py, the restatement of ModCF.ClassWTermModified,
DDFacet/Imager/ModCF.py:374-597) pinned by what the mathematics fixes, independently of its own
code path:

* the w = 0 plane of the central facet is the brute-force FP64 DFT of the zero-padded taper,
  laid out [OverS][OverS][S][S];
* a w > 0 plane of an off-centre facet is the brute-force DFT of the w-screen built from the
  exact n(l, m) (not the fitted polynomial) with the facet's linear terms removed;
* the fitted n-term (Give_dn, ModCF.py:267-295) reproduces n(l0 + dl, m0 + dm) - n0 with
  (Cu, Cv) = the analytic gradient (-l0 / n, -m0 / n) and no linear part left in the fit;
* the central facet's planes are centro-symmetric, so WplanesConj = conj(Wplanes) there;
* physics: the residual phase of the oracle's degridded point source against the FP64 direct DFT
  carries none of the w-screen phase the facet's CFs apply, while CF tables with the screen
  dropped or its sign flipped leave all of it or twice it.
"""
import numpy as np
import pytest

from ddfacet_amd import wterm

ARC = np.pi / 180 / 3600


def _unreorg(a, overs):
    """Inverse of the [OverS][OverS][S][S] reorganisation: back to the oversampled plane."""
    n = a.shape[0]
    s = n // overs
    b = a.reshape(overs, overs, s, s)
    out = np.zeros((n, n), a.dtype)
    for i in range(overs):
        for j in range(overs):
            out[i::overs, j::overs] = b[i, j]
    return out


def _centred_dft(a):
    """K[p, q] = sum_xy a[x, y] exp(-2 pi i ((p - c)(x - c) + (q - c)(y - c)) / N) / N^2, c = N // 2."""
    n = a.shape[0]
    c = n // 2
    k = np.arange(n) - c
    E = np.exp(-2j * np.pi * np.outer(k, k) / n)
    return E @ a.astype(np.complex128) @ E.T / (n * n)


def _screen_grid(lrad, S):
    dx = 2 * lrad / S
    return np.mgrid[-lrad + dx / 2:lrad - dx / 2:S * 1j, -lrad + dx / 2:lrad - dx / 2:S * 1j]


def test_central_w0_plane_is_dft_of_zero_padded_taper():
    S, O = 7, 11
    wt = wterm.ClassWTermModified(Cell=2.0, Sup=S, Nw=4, wmax=2000.0, Npix=511, Freqs=np.array([1.4e9]),
                                  OverS=O, lmShift=(0.0, 0.0), sups=[S] * 4)
    taper = np.abs(wt.SpheM.make_sphe(S)[2])
    N = S * O
    pad = np.zeros((N, N))
    c = N // 2
    pad[c - S // 2:c + S // 2 + 1, c - S // 2:c + S // 2 + 1] = taper
    want = _centred_dft(pad)
    got = _unreorg(np.asarray(wt.Wplanes[0]), O)
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()
    # the peak sits at the centre of the oversampled plane: K(0) = sum of the taper / N^2
    assert abs(got[c, c] - taper.sum() / N ** 2) <= 1e-6 * abs(got[c, c])


@pytest.mark.parametrize("lm", [(1800 * ARC, -900 * ARC), (-2400 * ARC, 300 * ARC)])
def test_offcentre_plane_is_dft_of_exact_w_screen(lm):
    S, O, Nw, npix, cell, wmax = 9, 5, 3, 735, 1.5, 9000.0
    freqs = np.array([1.2e9, 1.67e9])
    l0, m0 = lm
    wt = wterm.ClassWTermModified(Cell=cell, Sup=S, Nw=Nw, wmax=wmax, Npix=npix, Freqs=freqs, OverS=O,
                                  lmShift=lm, sups=[S] * Nw)
    lrad = (npix / 2.0) * cell * ARC
    dl, dm = _screen_grid(lrad, S)
    n0 = np.sqrt(1 - l0 ** 2 - m0 ** 2)
    exact = np.sqrt(1 - (l0 + dl) ** 2 - (m0 + dm) ** 2) - n0 - wt.Cu * dl - wt.Cv * dm
    taper = np.abs(wt.SpheM.make_sphe(S)[2])
    N = S * O
    c = N // 2
    for i in (1, 2):
        w = np.linspace(0, wmax, Nw)[i]
        scr = np.exp(-2j * np.pi * (w / wt.RefWave) * exact) * taper
        pad = np.zeros((N, N), np.complex128)
        pad[c - S // 2:c + S // 2 + 1, c - S // 2:c + S // 2 + 1] = scr
        want = _centred_dft(pad)
        got = _unreorg(np.asarray(wt.Wplanes[i]), O)
        assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max(), i
        wantc = _centred_dft(np.conj(pad))
        gotc = _unreorg(np.asarray(wt.WplanesConj[i]), O)
        assert np.abs(gotc - wantc).max() <= 2e-5 * np.abs(wantc).max(), i


@pytest.mark.parametrize("lm", [(0.004, -0.002), (-0.011, 0.006), (0.0, 0.009)])
def test_n_term_fit_and_facet_gradient(lm):
    l0, m0 = lm
    rad = 0.003
    cv, cu, coef = wterm.give_dn(l0, m0, rad=rad, order=5)
    n0 = np.sqrt(1 - l0 ** 2 - m0 ** 2)
    np.testing.assert_allclose(cu, -l0 / n0, rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(cv, -m0 / n0, rtol=1e-7, atol=1e-12)
    c = np.asarray(coef).reshape(6, 6)
    assert c[0, 1] == 0 and c[1, 0] == 0
    rng = np.random.default_rng(1)
    dl, dm = rng.uniform(-rad, rad, (2, 200))
    n = np.sqrt(1 - (l0 + dl) ** 2 - (m0 + dm) ** 2) - n0
    fit = wterm.polyval2d(dl, dm, np.asarray(coef)) + cu * dl + cv * dm
    assert np.abs(fit - n).max() <= 1e-12


def test_central_facet_planes_are_centro_symmetric():
    O = 7
    wt = wterm.ClassWTermModified(Cell=1.5, Sup=7, Nw=4, wmax=9000.0, Npix=735, Freqs=np.array([1.67e9]),
                                  OverS=O, lmShift=(0.0, 0.0), sups=[7, 7, 9, 11])
    for p, q in zip(wt.Wplanes, wt.WplanesConj):
        k = _unreorg(np.asarray(p), O)
        kc = _unreorg(np.asarray(q), O)
        assert np.abs(k - k[::-1, ::-1]).max() <= 1e-6 * np.abs(k).max()
        assert np.abs(kc - np.conj(k)).max() <= 1e-6 * np.abs(k).max()


def _screen_phase_slope(screen, fi, dx, dy):
    """Regression slope of the residual phase arg(predicted / DFT) of a point source dx, dy px off
    facet fi's centre (3 x 3 facets of 81 px at 60" cells, padded to 147) on the w-screen phase
    2 pi (w / lambda) (n(l, m) - n0 - Cu dl - Cv dm) the facet's CFs are meant to apply.  The
    oversampling and taper errors of gridding are uncorrelated with that phase; a missing screen
    leaves slope -1, a screen of the wrong sign -2, the right one ~0.  CF tables: "exact" as built,
    "none" the w = 0 plane for every w, "flipped" Wplanes and WplanesConj exchanged."""
    from ddfacet_amd import configs
    from oracle import cpuref, dft
    from oracle import imaging as ref
    cell_as = 60.0
    P, geom = configs.small_cycle(seed=7, nchan=4, ntime=8, nant=9, cell_arcsec=cell_as)
    if screen != "exact":
        for f in P.facets:
            wt = wterm.ClassWTermModified(Cell=cell_as, Sup=7, Nw=4, wmax=float(f.winfos[1]), Npix=P.npix,
                                          Freqs=P.freqs, OverS=11, lmShift=(f.l0, f.m0), sups=[7, 7, 9, 9])
            if screen == "none":
                f.wplanes = [wt.Wplanes[0]] * 4
                f.wplanes_conj = [wt.WplanesConj[0]] * 4
            else:
                f.wplanes, f.wplanes_conj = wt.WplanesConj, wt.Wplanes
    X, Y = geom.pix[fi][0] + dx, geom.pix[fi][1] + dy
    model = np.zeros((1, 1, geom.m, geom.m), np.float32)
    model[0, 0, X, Y] = 1.0
    sw = geom.spatial_weights()
    fac = [dict(xc=geom.pix[i][0], yc=geom.pix[i][1], sw=sw[i], sphe=np.float32(f.sphe)) for i, f in enumerate(P.facets)]
    norm = ref.facet_norm(fac, geom.m)
    v = np.zeros_like(P.vis)
    for i, f in enumerate(fac):
        cpuref.degrid(*P.degrid_args(i, ref.model_grid(model, f, norm, 11), v))
    l, m = dft.image_pixel_lm([X], [Y], geom.m, P.extra["cell"])
    want = dft.predict(P.uvw, P.freqs, np.array([1.0]), l, m)
    f = P.facets[fi]
    n0 = np.sqrt(1 - f.l0 ** 2 - f.m0 ** 2)
    dn = np.sqrt(1 - l[0] ** 2 - m[0] ** 2) - n0 - f.cu * (l[0] - f.l0) - f.cv * (m[0] - f.m0)
    x = 2 * np.pi * (P.uvw[:, 2:3] * P.freqs[None, :] / 299792458.0) * dn
    ph = np.angle(-v[:, :, 0] / want)
    return float((ph * x).sum() / (x * x).sum()), float(np.abs(x).max())


@pytest.mark.parametrize("fi,dx,dy", [(0, 20, -15), (4, 25, 18)])
def test_w_screen_sign_and_presence_fixed_by_the_dft(fi, dx, dy):
    got = {k: _screen_phase_slope(k, fi, dx, dy) for k in ("exact", "none", "flipped")}
    print("residual phase / screen phase slope: %s (screen phase up to %.2f rad)"
          % ({k: round(v[0], 3) for k, v in got.items()}, got["exact"][1]))
    assert got["exact"][1] > 0.1
    assert abs(got["exact"][0]) < 0.1
    assert abs(got["none"][0] + 1) < 0.1
    assert abs(got["flipped"][0] + 2) < 0.1
