"""The benchmark configurations at their own sizes, against the oracle.

* The plain single facet (561 600 visibilities, one 512^2 grid at 2", 8 w-planes, S = 7,
  OverS = 11): the drop-in shim (plain kernels) and the batched plan against the reference
  gridder built from its own sources (oracle/_ref) where present and the oracle (cpu_ref)
  always: grid and degridded visibilities relative L2 <= 1e-5, touched cells identical, sumwt
  <= 1e-10.  The stitched dirty image against the FP64 direct DFT of the gridded visibilities is
  printed (inner half of the image; the unpadded facet aliases near its edges).
* The uneven batched workload (121 facets x 4 bands, baseline-dependent averaging 1-16): the
  autotuned plan's facet grids, weights and predicted visibilities against the oracle on seven
  facets: the corners, the centre, two edges and the facets with the fewest and the most gridded
  weight.
* The differentiable forward model (25 facets x 4 bands, 2025^2 model cube, 3.4 M
  visibilities): the dot-product adjoint test on random complex64 vectors, the chi^2 gradient
  against central finite differences on 64 random pixels of an FP64-precision plan (<= 1e-4),
  and the chi^2 gradient against the oracle gridder applied to the residuals (<= 1e-5).
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def _grid_oracles():
    from oracle import cpuref, refmod
    out = [("oracle", cpuref.grid, cpuref.degrid)]
    if refmod.available():
        m = refmod.module()
        out.append(("reference", m.pyGridderWPol32, m.pyDeGridderWPol))
    return out


def test_plain_single_facet_workload_against_reference():
    from cases import model_grid
    from ddfacet_amd import _pyGridderSmearPols as shim
    from ddfacet_amd import bda as bdamod
    from ddfacet_amd import configs
    from ddfacet_amd.facets import DeviceFacetSet
    from oracle import cpuref, dft
    from oracle import imaging as ref
    print("plain single-facet workload", flush=True)
    P = configs.cfg1(log=lambda *a: None)
    model = model_grid(P, 7)
    # predictions are compared from zero visibilities (degridding subtracts the model), so that
    # the FP32 rounding of the observed visibilities does not enter the comparison
    g_hip, s_hip = P.new_grid(), np.zeros((P.nband, P.npol))
    shim.pyGridderWPol32(*P.grid_args(0, g_hip, s_hip))
    v_hip = np.zeros_like(P.vis)
    shim.pyDeGridderWPol(*P.degrid_args(0, model, v_hip))

    fs = DeviceFacetSet(P)
    fs.models[0].copy_(torch.from_numpy(model))
    fs.plan_batched()
    fs.zero_grids()
    fs.grid_batched()
    vis_b = torch.zeros_like(fs.t["vis"])
    fs.degrid_batched(vis=vis_b)
    torch.cuda.synchronize()
    g_bat, s_bat = fs.grids[0].cpu().numpy(), fs.sumwt[0].cpu().numpy()
    v_bat = vis_b.cpu().numpy()
    fs.destroy_plan()

    for label, grid_fn, degrid_fn in _grid_oracles():
        g_ref, s_ref = P.new_grid(), np.zeros((P.nband, P.npol))
        grid_fn(*P.grid_args(0, g_ref, s_ref))
        v_ref = np.zeros_like(P.vis)
        degrid_fn(*P.degrid_args(0, model, v_ref))
        for path, g, s, v in (("plain", g_hip, s_hip, v_hip), ("batched", g_bat, s_bat, v_bat)):
            eg = rel_l2(g, g_ref)
            ev = rel_l2(v, v_ref)
            print("%s vs %s: grid rel L2 %.3e, degridded vis rel L2 %.3e" % (path, label, eg, ev), flush=True)
            assert eg <= 1e-5 and ev <= 1e-5, (path, label, eg, ev)
            assert np.array_equal(g != 0, g_ref != 0), (path, label)
            np.testing.assert_allclose(s, s_ref, rtol=1e-10)

    # the dirty image against the FP64 direct DFT of the visibilities the gridder kept
    pl = cpuref.place(*P.grid_args(0, P.new_grid(), np.zeros((P.nband, P.npol))))
    R, C = P.uvw.shape[0], P.freqs.size
    kept = np.zeros((R, C), bool)
    for b, (c0, c1, rows) in enumerate(bdamod.iter_blocks(P.bda_grid)):
        if pl[b, 0]:
            kept[rows, c0:c1] = True
    cell = P.extra["cell_arcsec"] * np.pi / 180 / 3600
    sphe = np.float32(P.facets[0].sphe).astype(np.float64)
    d = ref.grid_to_im(g_hip, int(P.facets[0].winfos[3]))[0, 0].real.astype(np.float64) / sphe / s_hip[0, 0]
    img = d[::-1, :].T
    # the sources in the inner half of the image (at their nearest pixels) and 64 random pixels there
    sky = P.extra["sky"]
    Xs = np.rint(sky.l / cell + P.npix // 2).astype(int)
    Ys = np.rint(sky.m / cell + P.npix // 2).astype(int)
    inner = (np.abs(Xs - P.npix // 2) < P.npix // 4) & (np.abs(Ys - P.npix // 2) < P.npix // 4)
    rng = np.random.default_rng(0)
    X = np.concatenate([Xs[inner], rng.integers(128, 384, 64)])
    Y = np.concatenate([Ys[inner], rng.integers(128, 384, 64)])
    l, m = dft.image_pixel_lm(X, Y, P.npix, cell)
    w = np.where(kept, P.weights, 0).astype(np.float32)
    want = dft.dirty(P.uvw, P.freqs, 0.5 * (P.vis[:, :, 0] + P.vis[:, :, 3]).astype(np.complex128), w,
                     P.flags[:, :, 0], l, m)
    err = np.abs(img[X, Y] - want).max() / np.abs(want).max()
    print("gridded fraction %.3f; dirty image vs FP64 DFT (%d inner sources + 64 inner pixels): max error / peak "
          "%.3e" % (kept.mean(), int(inner.sum()), err), flush=True)
    assert err < 0.1


def test_uneven_workload_autotuned_plan_against_oracle():
    from ddfacet_amd import autotune, configs, marshal
    from ddfacet_amd.facets import DeviceFacetSet
    from oracle import cpuref
    print("uneven 121-facet workload", flush=True)
    P = configs.cfg4(log=lambda *a: None)
    fs = DeviceFacetSet(P)
    gen = torch.Generator(device="cuda").manual_seed(41)
    fs.models.copy_(torch.randn(fs.models.shape, generator=gen, device="cuda", dtype=torch.complex64) * 1e-3)
    vis0 = fs.t["vis"].clone()
    space = {"tile": (16, 32), "chunk": (1024, 4096), "blocks_per_item": (32, 128)}
    best, _ = autotune.autotune(fs, space=space, reps=1)
    fs.t["vis"].copy_(vis0)
    print("autotuned plan:", best, flush=True)
    fs.zero_grids()
    fs.grid_batched()
    torch.cuda.synchronize()
    tot = fs.sumwt.sum(dim=(1, 2)).cpu().numpy()
    nfx = P.extra["nfx"]
    sample = sorted({0, nfx - 1, (fs.nf - 1) // 2, nfx // 2, fs.nf - 1, int(tot.argmin()), int(tot.argmax())})
    print("facets %s, gridded weight %s" % (sample, tot[sample]), flush=True)
    assert tot.max() > tot.min()

    specs, outs = [], []
    for fi in sample:
        g = P.new_grid()
        s = np.zeros((P.nband, P.npol))
        specs.append(marshal.parse_grid_args(P.grid_args(fi, g, s), False))
        outs.append((g, s))
    cpuref.grid_facets(specs, len(specs))
    for fi, (g, s) in zip(sample, outs):
        got = fs.grids[fi].cpu().numpy()
        assert rel_l2(got, g) <= 1e-5, (fi, rel_l2(got, g))
        assert np.array_equal(got != 0, g != 0), fi
        np.testing.assert_allclose(fs.sumwt[fi].cpu().numpy(), s, rtol=1e-10)

    mask = torch.zeros(fs.nf, dtype=torch.bool, device="cuda")
    mask[sample] = True
    fs.models[~mask] = 0
    vis = torch.zeros_like(fs.t["vis"])
    fs.degrid_batched(vis=vis)
    torch.cuda.synchronize()
    v = np.zeros_like(P.vis)
    dspecs = [marshal.parse_degrid_args(P.degrid_args(fi, fs.models[fi].cpu().numpy(), v)) for fi in sample]
    cpuref.degrid_facets(dspecs, len(dspecs))
    err = rel_l2(vis.cpu().numpy(), v)
    print("predicted visibilities of facets %s: rel L2 %.3e vs oracle" % (sample, err), flush=True)
    assert err <= 1e-5, err
    fs.destroy_plan()


@pytest.fixture(scope="module")
def forward_model():
    from ddfacet_amd import configs
    print("differentiable forward model, 25 facets x 4 bands", flush=True)
    P, geom = configs.cfg5(device="cuda", log=lambda *a: None)
    return P, geom


def _operator(P, geom, precision):
    from ddfacet_amd.autograd import DegridOperator
    from ddfacet_amd.cycle import MajorCycle
    return DegridOperator(MajorCycle(P, geom, precision=precision))


@pytest.mark.parametrize("precision,cs_bar,coherent_bar", [("mixed", 1e-8, 1e-6), ("fp64", 1e-9, 5e-9)])
def test_forward_model_dot_product_adjoint(forward_model, precision, cs_bar, coherent_bar):
    """<A x, y> = <x, A^H y> on random x and complex64 y, the mismatch relative to the
    Cauchy-Schwarz scale |A x| |y|: the inner product of random vectors is ~sqrt(n) below that
    scale (here 4e4 times), so a bar relative to it would measure that cancellation rather than the
    adjoint.  Then y = A x, where <A x, A x> = |A x|^2 has no cancellation, relative to it.
    Measured on MI355X: mixed 6e-11 to 1.5e-10 and 3e-8, FP64 4e-11 to 1.4e-10 and 7e-11 to 5e-10."""
    P, geom = forward_model
    op = _operator(P, geom, precision)
    gen = torch.Generator(device="cuda").manual_seed(5101)
    x = torch.randn((P.nband, P.npol, geom.m, geom.m), generator=gen, device="cuda")
    y = torch.randn(tuple(P.vis.shape), generator=gen, device="cuda", dtype=torch.complex64)
    ax = op.forward(x).to(torch.complex128)
    y64 = y.to(torch.complex128)
    lhs = torch.sum(ax.conj() * y64).real.item()
    rhs = torch.sum(x.double() * op.adjoint(y).double()).item()
    scale = (torch.linalg.vector_norm(ax) * torch.linalg.vector_norm(y64)).item()
    print("dot-product test (%s): <Ax,y> = %.10e, <x,A^H y> = %.10e, |Ax||y| = %.4e, rel %.2e" %
          (precision, lhs, rhs, scale, abs(lhs - rhs) / scale))
    assert abs(lhs - rhs) <= cs_bar * scale
    lhs2 = torch.sum(ax.conj() * ax).real.item()
    rhs2 = torch.sum(x.double() * op.adjoint(ax.to(torch.complex64)).double()).item()
    print("coherent (%s): <Ax,Ax> = %.12e, <x,A^H Ax> = %.12e, rel %.2e" % (precision, lhs2, rhs2, abs(lhs2 - rhs2) / lhs2))
    assert abs(lhs2 - rhs2) <= coherent_bar * lhs2
    op.cycle.fs.destroy_plan()


def test_forward_model_gradient_vs_fp64_finite_differences(forward_model):
    """Central differences of chi^2 on 64 random pixels (random bands) of an FP64-precision
    plan.  chi^2 is quadratic in the image, so the differences carry no truncation error for any
    step; the loss difference is summed per visibility in FP64 (sum w (|r+|^2 - |r-|^2)), which
    is (chi^2(x + h) - chi^2(x - h)) without the cancellation of two large sums.  Each pixel's
    error is relative to its derivative or to 1 % of the largest one, whichever is larger (the
    adjoint's FP32 grids round at ~1e-7 of the largest derivative)."""
    from ddfacet_amd.autograd import chi2
    P, geom = forward_model
    op = _operator(P, geom, "fp64")
    data = P.vis
    w = torch.from_numpy(P.weights).cuda().double()
    X, Y = P.extra["src_pix"]
    x0 = torch.zeros((P.nband, P.npol, geom.m, geom.m), dtype=torch.float32, device="cuda")
    for b in range(P.nband):
        x0[b, 0, X, Y] = torch.from_numpy(0.5 * P.extra["sky"].flux.astype(np.float32)).cuda()
    x = x0.clone().requires_grad_(True)
    chi2(x, op, data, w).backward()
    g = x.grad.double().cpu().numpy()
    rng = np.random.default_rng(5102)
    pix = np.stack([rng.integers(0, P.nband, 64), rng.integers(0, geom.m, 64), rng.integers(0, geom.m, 64)], 1)
    h = 1.0
    errs = []
    with torch.no_grad():
        for b, a, c in pix:
            xp = x0.clone()
            xp[b, 0, a, c] += h
            xm = x0.clone()
            xm[b, 0, a, c] -= h
            rp = op.forward(xp) - data
            rm = op.forward(xm) - data
            ap = rp.real.double() ** 2 + rp.imag.double() ** 2
            am = rm.real.double() ** 2 + rm.imag.double() ** 2
            fd = float((w[..., None] * (ap - am)).sum()) / (2 * h)
            errs.append(abs(fd - g[b, 0, a, c]) / max(abs(fd), 1e-2 * np.abs(g).max()))
    print("gradient vs central differences (FP64 plan, 64 pixels): max rel err %.3e" % max(errs))
    assert max(errs) <= 1e-4
    op.cycle.fs.destroy_plan()


def test_forward_model_chi2_gradient_vs_oracle_gridder_on_residuals(forward_model):
    """grad chi^2 = 2 Re(A^H W r) with r = A x - d.  The degridder is not the exact adjoint of
    the gridder: for W > 0 it gathers with WplanesConj where the adjoint would use conj(Wplanes
    mirrored), equal only for a centro-symmetric CF (the central facet).  Handing the gridder the
    lists conj(WplanesConj), conj(Wplanes) makes it the exact adjoint for every facet; with the
    Stokes 1/2 (I = (XX + YY) / 2 gridded, I written to XX and YY) and d|r|^2 = 2 Re(conj(r) dr),
    the gradient is 4 x the stitched dirty image of W r with the facet weights left out."""
    from ddfacet_amd import marshal
    from ddfacet_amd.autograd import chi2
    from ddfacet_amd.problems import Facet, Problem
    from oracle import cpuref
    from oracle import imaging as ref
    P, geom = forward_model
    op = _operator(P, geom, "mixed")
    data = P.vis
    w = torch.from_numpy(P.weights).cuda().double()
    X, Y = P.extra["src_pix"]
    x = torch.zeros((P.nband, P.npol, geom.m, geom.m), dtype=torch.float32, device="cuda")
    x[0, 0, X, Y] = 0.25
    x.requires_grad_(True)
    chi2(x, op, data, w).backward()
    g = x.grad.cpu().numpy()
    with torch.no_grad():
        resid = (op.forward(x.detach()) - data).cpu().numpy()
    op.cycle.fs.destroy_plan()

    facets = [Facet(l0=f.l0, m0=f.m0, cu=f.cu, cv=f.cv, wplanes=[np.conj(p) for p in f.wplanes_conj],
                    wplanes_conj=[np.conj(p) for p in f.wplanes], winfos=f.winfos, sphe=f.sphe) for f in P.facets]
    kw = {k: getattr(P, k) for k in P.__dataclass_fields__}
    kw.update(vis=resid.astype(np.complex64), facets=facets)
    H = Problem(**kw)
    specs, grids = [], []
    for fi in range(len(facets)):
        gr = H.new_grid()
        specs.append(marshal.parse_grid_args(H.grid_args(fi, gr, np.zeros((H.nband, H.npol))), False))
        grids.append(gr)
    cpuref.grid_facets(specs, 16)
    sw = geom.spatial_weights()
    overs = int(P.facets[0].winfos[3])
    fac = []
    for i, f in enumerate(P.facets):
        s = np.float32(f.sphe)
        fac.append(dict(xc=geom.pix[i][0], yc=geom.pix[i][1], sw=sw[i], sphe=s,
                        inv_sphe=np.float32(1.0 / s.astype(np.float64)), dirty=ref.grid_to_im(grids[i], overs),
                        sumwt=np.ones((H.nband, H.npol)), sum_jones=None))
    norm = ref.facet_norm(fac, geom.m)
    want = 4.0 * ref.facets_to_im(fac, H.nband, H.npol, geom.m, norm)
    err = rel_l2(g, want)
    print("chi2 gradient vs oracle gridder on residuals (25 facets x 4 bands): rel L2 %.3e" % err)
    assert err <= 1e-5
