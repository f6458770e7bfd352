"""Closed-form checks that pin the numpy restatement of the image side (oracle/imaging.py) and
the host-side geometry / CF helpers, independently of the code under test."""
import numpy as np
import pytest

from oracle import imaging as ref


def test_give_edges_hand_cases():
    # facet fully inside the image
    assert ref.give_edges(50, 60, 200, 10, 10, 21) == ([40, 61, 50, 71], [0, 21, 0, 21])
    # clipped on the low side: 5 - 10 = -5 -> main starts at 0, facet at 5
    assert ref.give_edges(5, 100, 200, 10, 10, 21) == ([0, 16, 90, 111], [5, 21, 0, 21])
    # clipped on the high side: 195 + 10 = 205 > 199 -> main ends at 200, facet at 21 - 6
    assert ref.give_edges(195, 100, 200, 10, 10, 21) == ([185, 200, 90, 111], [0, 15, 0, 21])


@pytest.mark.parametrize("n", [7, 8, 15])
def test_centred_fft_of_centred_delta_is_flat(n):
    a = np.zeros((n, n), complex)
    a[n // 2, n // 2] = 1.0
    # the centre pixel maps to the zero frequency: forward gives 1/n^2 everywhere
    np.testing.assert_allclose(ref.fft2c(a), np.full((n, n), 1.0 / n ** 2), atol=1e-15)
    np.testing.assert_allclose(ref.ifft2c(ref.fft2c(a)), a, atol=1e-14)


def test_centred_fft_shift_theorem():
    n = 9
    a = np.zeros((n, n), complex)
    a[n // 2 + 2, n // 2 - 1] = 1.0
    k = np.arange(n) - n // 2
    want = np.exp(-2j * np.pi * (2 * k[:, None] + (-1) * k[None, :]) / n) / n ** 2
    np.testing.assert_allclose(ref.fft2c(a), want, atol=1e-14)


def test_stitch_then_extract_single_facet_identity():
    """One facet covering the whole image with SW = Sphe = 1 and unit weights: the stitch is the
    flip of the facet image and the model extraction flips it back (x (OverS n)^2)."""
    n = m = 21
    rng = np.random.default_rng(0)
    dirty = rng.standard_normal((1, 1, n, n)).astype(np.complex64)
    ones = np.ones((n, n), np.float32)
    f = dict(xc=m // 2, yc=m // 2, sw=ones, sphe=ones, inv_sphe=ones, dirty=dirty, sumwt=np.ones((1, 1)),
             sum_jones=None)
    norm = ref.facet_norm([f], m)
    assert np.all(norm == 1)
    img = ref.facets_to_im([f], 1, 1, m, norm)
    np.testing.assert_allclose(img[0, 0], dirty[0, 0].real[::-1, :].T, rtol=0, atol=0)
    back = ref.give_model_tessel(img, f, norm, 3)
    np.testing.assert_allclose(back[0, 0], dirty[0, 0].real * np.float32((3 * n) ** 2), rtol=1e-7)


def test_stitch_normalisation_terms():
    """Weights, the Jones norm sqrt(SumJones0/SumJones1) and the Sphe cut enter as stated."""
    n = m = 11
    dirty = np.ones((1, 1, n, n), np.complex64)
    sphe = np.ones((n, n), np.float32)
    sphe[0, :] = 1e-4   # grid row 0 -> image column n-1 after the flip
    inv = np.full((n, n), 2.0, np.float32)
    f = dict(xc=5, yc=5, sw=np.ones((n, n), np.float32), sphe=sphe, inv_sphe=inv, dirty=dirty,
             sumwt=np.array([[4.0]]), sum_jones=np.array([[9.0], [1.0]]))
    norm = ref.facet_norm([f], m)
    img = ref.facets_to_im([f], 1, 1, m, norm)
    assert np.allclose(img[0, 0, :, : n - 1], 2.0 / (4.0 * 3.0))
    assert np.all(img[0, 0, :, n - 1] == 0)


def test_eval_sphe_known_values():
    from ddfacet_amd.wterm import eval_sphe
    # Schwab's approximation: psi(0) = P0/Q0 evaluated at nu^2 - 0.75^2, psi(1) = 0, 0 outside
    v0 = eval_sphe(np.array([0.0]))[0]
    d = -0.75 ** 2
    P = [8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1]
    Q = [1.0, 8.212018e-1, 2.078043e-1]
    want = sum(P[k] * d ** k for k in range(5)) / sum(Q[k] * d ** k for k in range(3))
    assert abs(v0 - want) < 1e-15
    assert eval_sphe(np.array([1.0]))[0] == 0.0
    assert eval_sphe(np.array([1.2]))[0] == 0.0
    x = np.linspace(0, 0.99, 50)
    assert np.all(np.diff(eval_sphe(x)) < 0)  # monotone decreasing taper


def test_reorg_is_a_permutation():
    from ddfacet_amd.wterm import reorg
    o, s = 5, 3
    a = np.arange((o * s) ** 2).reshape(o * s, o * s)
    b = reorg(a, o)
    assert sorted(b.ravel().tolist()) == list(range((o * s) ** 2))
    # plane (i, j) holds the sub-sampled kernel a[i::o, j::o]
    bb = b.reshape(o, o, s, s)
    np.testing.assert_array_equal(bb[2, 3], a[2::o, 3::o])


def test_conj_plane_identity():
    """fft2(conj(x)) = conj(mirror(fft2(x))): the WplanesConj of ModCF for a centro-symmetric taper."""
    from ddfacet_amd.wterm import _fft2
    rng = np.random.default_rng(1)
    x = rng.standard_normal((9, 9)) + 1j * rng.standard_normal((9, 9))
    lhs = _fft2(np.conj(x))
    rhs = np.conj(_fft2(x)[::-1, ::-1])
    np.testing.assert_allclose(lhs, rhs, atol=1e-14)


def test_estimate_npix_matches_workload_sizes():
    from ddfacet_amd.configs import estimate_npix
    assert estimate_npix(4096 / 9, 1.7) == (455, 735)
    assert estimate_npix(6144 / 11, 1.7)[1] == 975
    assert estimate_npix(2048 / 5, 1.7)[1] == 693


def _box_gauss_profile(n, half, sigma):
    """One axis of SW for a facet clear of the image edge, by direct summation: the polygon test
    on samples spaced s = n / (n - 1) pixels keeps the 2J + 1 central samples with J s < half;
    each output sample sums the Gaussian exp(-(d s)^2 / (2 sigma^2)) over them (d in samples).
    Near the box edge it follows the normal-CDF step of width sigma / s samples."""
    s = n / (n - 1.0)
    c = n // 2
    J = int(np.floor(half / s - 1e-12))
    x = np.arange(n) - c
    j = np.arange(-J, J + 1)
    f = np.exp(-((x[:, None] - j[None, :]) * s) ** 2 / (2 * sigma ** 2)).sum(1)
    return f / f[c]


@pytest.mark.parametrize("m,nfx,n,sigma", [(243, 3, 147, 4.0), (405, 5, 147, 2.5), (2025, 5, 693, 10.0)])
def test_spatial_weights_closed_form(m, nfx, n, sigma):
    """SW of the central facet (no image edge within reach) equals the separable box-Gaussian
    profile of the reference construction summed directly (no FFT, no padding); the step across
    the polygon edge has the normal-CDF width sigma / s; an image-corner facet is cut at the image
    edge."""
    from scipy.special import ndtr
    from ddfacet_amd.imaging import FacetGeometry
    g = FacetGeometry(m, nfx, m // nfx, n, 1e-5, mixing_width=sigma)
    sw = g.spatial_weights()
    fi = g.nf // 2
    f = _box_gauss_profile(n, (m // nfx) / 2.0, sigma)
    want = np.outer(f, f)
    want[want < 1e-3] = 0
    assert np.abs(sw[fi] - want).max() <= 2e-6
    c = n // 2
    assert sw[fi][c, c] == 1.0
    J0 = int(np.floor((m // nfx) / 2.0 * (n - 1) / n))
    x = np.arange(c + J0 - 3, c + J0 + 5) - c
    erf_step = ndtr((J0 + 0.5 - x) / (sigma * (n - 1) / n))
    np.testing.assert_allclose(sw[fi][c, c + x], erf_step, atol=3e-3)
    # edge samples of a regular facet sit symmetrically about its polygon edge
    J = int(np.floor((m // nfx) / 2.0 * (n - 1) / n))
    np.testing.assert_allclose(sw[fi][c, c + J] + sw[fi][c, c + J + 1], 1.0, atol=1e-3)
    # corner facet 0: samples outside the main image are masked before the convolution
    x0 = g.pix[0][0]
    outside = (np.arange(n) - c) * n / (n - 1.0) + x0 - m // 2 <= -m / 2.0
    assert outside.any()
    assert sw[0][c, outside].max() < 0.5 and sw[0][c, c] > 0.99
