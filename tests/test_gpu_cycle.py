"""One major cycle on the GPU (batched grid -> GridToIm -> stitch; model -> FFT -> degrid) against
the CPU pipeline built from the oracle (cpu_ref per facet + the numpy image side), against the
FP64 direct DFT, and the differentiable forward model (adjoint, finite differences, chi2
gradient against the reference gridder on the residuals)."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _cycle(jones=False, bda="pervis", precision="mixed", nband=1):
    from ddfacet_amd import configs
    from ddfacet_amd.cycle import MajorCycle
    P, geom = configs.small_cycle(jones=jones, bda=bda, nband=nband)
    return P, geom, MajorCycle(P, geom, precision=precision, tile=16)


def _oracle_facets(P, geom, dirty=None, sumwt=None, sj=None):
    sw = geom.spatial_weights()
    out = []
    for i, f in enumerate(P.facets):
        s = np.float32(f.sphe)
        out.append(dict(xc=geom.pix[i][0], yc=geom.pix[i][1], sw=sw[i], sphe=s,
                        inv_sphe=np.float32(1.0 / s.astype(np.float64)),
                        dirty=None if dirty is None else dirty[i], sumwt=None if sumwt is None else sumwt[i],
                        sum_jones=None if sj is None else sj[i]))
    return out


def _oracle_dirty(P, geom):
    from oracle import cpuref
    from oracle import imaging as ref
    dirty, sumwt, sj = [], [], []
    for i in range(len(P.facets)):
        g = P.new_grid()
        s = np.zeros((P.nband, P.npol))
        args = P.grid_args(i, g, s)
        cpuref.grid(*args)
        dirty.append(ref.grid_to_im(g, int(P.facets[i].winfos[3])))
        sumwt.append(s)
        sj.append(args[13][17] if args[13] else None)
    fac = _oracle_facets(P, geom, dirty, sumwt, sj if P.jones else None)
    norm = ref.facet_norm(fac, geom.m)
    return ref.facets_to_im(fac, P.nband, P.npol, geom.m, norm), norm


@pytest.mark.parametrize("jones", [False, True])
def test_residual_image_matches_oracle_pipeline(jones):
    P, geom, cyc = _cycle(jones=jones)
    img = cyc.residual_image().cpu().numpy()
    want, norm = _oracle_dirty(P, geom)
    np.testing.assert_allclose(cyc.norm.cpu().numpy(), norm, rtol=1e-6)
    assert rel_l2(img, want) <= 1e-5, rel_l2(img, want)


def test_dirty_image_vs_fp64_dft():
    """The stitched dirty image at the source pixels against the FP64 direct dirty image; the
    difference is the gridding (taper / w-kernel / oversampling) error, reported and bounded."""
    from oracle import dft
    P, geom, cyc = _cycle()
    img = cyc.residual_image().cpu().numpy()[0, 0]
    X, Y = P.extra["src_pix"]
    l, m = dft.image_pixel_lm(X, Y, geom.m, P.extra["cell"])
    vis_i = 0.5 * (P.vis[:, :, 0] + P.vis[:, :, 3]).astype(np.complex128)
    want = dft.dirty(P.uvw, P.freqs, vis_i, P.weights, P.flags[:, :, 0], l, m)
    got = img[X, Y]
    err = np.abs(got - want).max() / np.abs(want).max()
    print("dirty vs FP64 DFT: max rel err at sources %.3e" % err)
    assert err < 5e-2


@pytest.mark.parametrize("nband", [1, 2])
def test_predict_matches_oracle_pipeline(nband):
    from oracle import cpuref
    from oracle import imaging as ref
    P, geom, cyc = _cycle(jones=True, nband=nband)
    rng = np.random.default_rng(3)
    model = np.zeros((nband, 1, geom.m, geom.m), np.float32)
    X, Y = P.extra["src_pix"]
    for b in range(nband):
        model[b, 0, X, Y] = rng.uniform(0.5, 1.5, X.size).astype(np.float32)
    vis = torch.from_numpy(P.vis).cuda()
    cyc.predict(torch.from_numpy(model).cuda(), vis=vis)
    got = P.vis - vis.cpu().numpy()
    fac = _oracle_facets(P, geom)
    norm = ref.facet_norm(fac, geom.m)
    v = P.vis.copy()
    for i, f in enumerate(fac):
        mg = ref.model_grid(model, f, norm, int(P.facets[i].winfos[3]))
        cpuref.degrid(*P.degrid_args(i, mg, v))
    want = P.vis - v
    assert rel_l2(got, want) <= 1e-5, rel_l2(got, want)


def test_predict_vs_fp64_dft():
    """Degridded point-source model against the FP64 DFT predict of the same sources."""
    from oracle import dft
    P, geom, cyc = _cycle()
    X, Y = P.extra["src_pix"]
    model = np.zeros((1, 1, geom.m, geom.m), np.float32)
    model[0, 0, X, Y] = P.extra["sky"].flux.astype(np.float32)
    vis = torch.zeros(P.vis.shape, dtype=torch.complex64, device="cuda")
    cyc.predict(torch.from_numpy(model).cuda(), vis=vis)
    got = -vis.cpu().numpy()[:, :, 0]
    l, m = dft.image_pixel_lm(X, Y, geom.m, P.extra["cell"])
    want = dft.predict(P.uvw, P.freqs, model[0, 0, X, Y].astype(np.float64), l, m)
    err = rel_l2(got, want)
    print("predict vs FP64 DFT: rel L2 %.3e" % err)
    assert err < 5e-2


@pytest.mark.parametrize("nband", [1, 2])
def test_autograd_dot_product_adjoint(nband):
    from ddfacet_amd.autograd import DegridOperator
    P, geom, cyc = _cycle(jones=True, nband=nband)
    op = DegridOperator(cyc)
    rng = np.random.default_rng(11)
    x = torch.from_numpy(rng.standard_normal((nband, 1, geom.m, geom.m)).astype(np.float32)).cuda()
    y = torch.from_numpy((rng.standard_normal(P.vis.shape) + 1j * rng.standard_normal(P.vis.shape)).astype(np.complex64)).cuda()
    ax = op.forward(x).to(torch.complex128)
    y64 = y.to(torch.complex128)
    lhs = torch.sum(torch.conj(ax) * y64).real.item()
    rhs = torch.sum(x.double() * op.adjoint(y).double()).item()
    # relative to the Cauchy-Schwarz scale |A x| |y| (random vectors: their inner product is far
    # below it), then on y = A x, where <A x, A x> has no cancellation
    scale = (torch.linalg.vector_norm(ax) * torch.linalg.vector_norm(y64)).item()
    print("dot-product test: <Ax,y>=%.8e <x,A^H y>=%.8e |Ax||y|=%.4e rel %.2e" % (lhs, rhs, scale, abs(lhs - rhs) / scale))
    assert abs(lhs - rhs) <= 1e-8 * scale
    lhs2 = torch.sum(torch.conj(ax) * ax).real.item()
    rhs2 = torch.sum(x.double() * op.adjoint(ax.to(torch.complex64)).double()).item()
    print("coherent: <Ax,Ax>=%.12e <x,A^H Ax>=%.12e rel %.2e" % (lhs2, rhs2, abs(lhs2 - rhs2) / lhs2))
    assert abs(lhs2 - rhs2) <= 1e-6 * lhs2


def test_autograd_gradient_vs_finite_differences():
    """Central differences on 64 pixels of an FP64-precision plan, the loss difference summed per
    visibility in FP64 (chi^2 is quadratic in the image, so any step is free of truncation
    error).  Each pixel's error is relative to its derivative or to 1 % of the largest one,
    whichever is larger: the adjoint accumulates in FP32 grids, ~1e-7 of the largest derivative,
    which says nothing about a derivative a thousand times smaller."""
    from ddfacet_amd.autograd import DegridOperator, chi2
    P, geom, cyc = _cycle(precision="fp64")
    op = DegridOperator(cyc)
    rng = np.random.default_rng(12)
    data = torch.from_numpy(P.vis).cuda()
    w = torch.from_numpy(P.weights).cuda().double()
    x0 = np.zeros((1, 1, geom.m, geom.m), np.float32)
    X, Y = P.extra["src_pix"]
    x0[0, 0, X, Y] = 0.5
    x = torch.from_numpy(x0).cuda().requires_grad_(True)
    loss = chi2(x, op, data, w)
    loss.backward()
    g = x.grad.double().cpu().numpy()[0, 0]
    pix = list(zip(X, Y)) + [tuple(p) for p in rng.integers(geom.m // 4, 3 * geom.m // 4, (64 - X.size, 2))]
    h = 0.25
    errs = []
    for (a, b) in pix:
        xp = x0.copy(); xp[0, 0, a, b] += h
        xm = x0.copy(); xm[0, 0, a, b] -= h
        with torch.no_grad():
            rp = op.forward(torch.from_numpy(xp).cuda()) - data
            rm = op.forward(torch.from_numpy(xm).cuda()) - data
            d = (rp.real.double() ** 2 + rp.imag.double() ** 2) - (rm.real.double() ** 2 + rm.imag.double() ** 2)
            fd = float((w[..., None] * d).sum()) / (2 * h)
        errs.append(abs(fd - g[a, b]) / max(abs(fd), 1e-2 * np.abs(g).max()))
    k = int(np.argmax(errs))
    print("autograd vs central differences: max rel err %.3e (pixel %s, |g| / max|g| %.2e)"
          % (errs[k], pix[k], abs(g[pix[k]]) / np.abs(g).max()))
    assert max(errs) <= 1e-4


def test_chi2_gradient_vs_reference_gridder_on_residuals():
    """grad chi2 = 2 Re(A^H W r).  For a single central facet with unit weights, no flags and no
    Jones, A^H r equals the reference gridder's stitched dirty image of the residuals times
    2 sumwt (the degridder writes I to XX and YY, the gridder grids (XX+YY)/2); the CF of the
    central facet is centro-symmetric so Wplanes is the adjoint of WplanesConj."""
    from ddfacet_amd import configs
    from ddfacet_amd.autograd import DegridOperator, chi2
    from ddfacet_amd.cycle import MajorCycle
    from ddfacet_amd.problems import Problem
    P, geom = configs.small_cycle(nfx=1, m=147, n=147)
    cyc = MajorCycle(P, geom, tile=16)
    op = DegridOperator(cyc)
    data = torch.from_numpy(P.vis).cuda()
    w = torch.from_numpy(P.weights).cuda().double()
    x = torch.zeros((1, 1, geom.m, geom.m), dtype=torch.float32, device="cuda", requires_grad=True)
    loss = chi2(x, op, data, w)
    loss.backward()
    g = x.grad.cpu().numpy()[0, 0]
    resid = (op.forward(x.detach()) - data).cpu().numpy()
    P2 = Problem(**{k: getattr(P, k) for k in P.__dataclass_fields__ if k != "vis"}, vis=resid.astype(np.complex64))
    img, _ = _oracle_dirty(P2, geom)
    sumwt = float(P.weights.sum())
    want = 2.0 * 2.0 * sumwt * img[0, 0]
    sel = np.float32(P.facets[0].sphe) >= 1e-3
    err = rel_l2(g[sel], want[sel])
    print("chi2 gradient vs reference gridder on residuals: rel L2 %.3e" % err)
    assert err < 1e-4
