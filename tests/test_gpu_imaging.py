"""Image side of the major cycle on the GPU against its numpy restatement (oracle/imaging.py):
centred facet FFTs (rocFFT), FacetNorm, the facet stitch and the model extraction.
Tolerances: FFTs to FP32 FFT accuracy (rel L2 <= 1e-5); the elementwise stitch / norm kernels
form the same FP32 sums in the same order, so they are compared at rel L2 <= 1e-6."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import imaging as ref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [63, 64, 75])
@pytest.mark.parametrize("inverse", [0, 1])
def test_fft2c(n, inverse):
    from ddfacet_amd import imaging
    rng = np.random.default_rng(n + 10 * inverse)
    a = (rng.standard_normal((3, n, n)) + 1j * rng.standard_normal((3, n, n))).astype(np.complex64)
    t = torch.from_numpy(a).cuda()
    imaging.fft2c(t, inverse, 2.5)
    want = (ref.ifft2c(a.astype(np.complex128)) if inverse else ref.fft2c(a.astype(np.complex128))) * 2.5
    assert rel_l2(t.cpu().numpy(), want) <= 1e-5


def _setup(seed, m=153, nfx=3, n=75, nband=2, npol=2, jones=True):
    from ddfacet_amd.imaging import FacetGeometry
    from ddfacet_amd.wterm import SpheMachine
    geom = FacetGeometry(m, nfx, m // nfx, n, 1e-5, mixing_width=4.0)
    sw = geom.spatial_weights()
    _, _, sp = SpheMachine(7).make_sphe(n)
    sphe = np.repeat(np.float32(sp.real)[None], geom.nf, 0)
    inv = np.float32(1.0 / sphe.astype(np.float64))
    rng = np.random.default_rng(seed)
    dirty = (rng.standard_normal((geom.nf, nband, npol, n, n)) + 1j * rng.standard_normal((geom.nf, nband, npol, n, n))).astype(np.complex64)
    sumwt = rng.uniform(10, 20, (geom.nf, nband, npol))
    sj = rng.uniform(0.5, 2, (geom.nf, 2, nband)) if jones else None
    return geom, sw, sphe, inv, dirty, sumwt, sj


def _facets(geom, sw, sphe, inv, dirty=None, sumwt=None, sj=None):
    return [dict(xc=geom.pix[i][0], yc=geom.pix[i][1], sw=sw[i], sphe=sphe[i], inv_sphe=inv[i],
                 dirty=None if dirty is None else dirty[i], sumwt=None if sumwt is None else sumwt[i],
                 sum_jones=None if sj is None else sj[i]) for i in range(geom.nf)]


@pytest.mark.parametrize("jones", [True, False])
def test_facet_norm_and_stitch(jones):
    from ddfacet_amd import imaging
    geom, sw, sphe, inv, dirty, sumwt, sj = _setup(3, jones=jones)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()  # noqa: E731
    norm = torch.zeros((geom.m, geom.m), dtype=torch.float32, device="cuda")
    imaging.facet_norm(geom, d(sw), norm)
    norm_ref = ref.facet_norm(_facets(geom, sw, sphe, inv), geom.m)
    assert rel_l2(norm.cpu().numpy(), norm_ref) <= 1e-6
    img = torch.zeros((2, 2, geom.m, geom.m), dtype=torch.float32, device="cuda")
    imaging.stitch(geom, d(sphe), d(inv), d(sw), d(dirty), d(sumwt), None if sj is None else d(sj), norm, img)
    want = ref.facets_to_im(_facets(geom, sw, sphe, inv, dirty, sumwt, sj), 2, 2, geom.m, norm_ref)
    got = img.cpu().numpy()
    fin = np.isfinite(want)
    assert np.array_equal(fin, np.isfinite(got))
    assert rel_l2(got[fin], want[fin]) <= 1e-6


def test_model_to_grid():
    from ddfacet_amd import imaging
    geom, sw, sphe, inv, _, _, _ = _setup(4)
    rng = np.random.default_rng(9)
    image = np.zeros((2, 2, geom.m, geom.m), np.float32)
    idx = rng.integers(0, geom.m, (40, 2))
    image[:, :, idx[:, 0], idx[:, 1]] = rng.uniform(0.1, 1, (2, 2, 40)).astype(np.float32)
    image[1, 1] = 0  # an all-zero plane stays zero
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()  # noqa: E731
    norm = torch.zeros((geom.m, geom.m), dtype=torch.float32, device="cuda")
    imaging.facet_norm(geom, d(sw), norm)
    models = torch.zeros((geom.nf, 2, 2, geom.n, geom.n), dtype=torch.complex64, device="cuda")
    imaging.model_to_grid(geom, d(sphe), d(sw), d(image), norm, models, 11)
    norm_ref = ref.facet_norm(_facets(geom, sw, sphe, inv), geom.m)
    for i, f in enumerate(_facets(geom, sw, sphe, inv)):
        want = ref.model_grid(image, f, norm_ref, 11)
        got = models[i].cpu().numpy()
        assert rel_l2(got, want) <= 1e-5, (i, rel_l2(got, want))
        assert np.all(got[1, 1] == 0)
