"""The FP64 direct Fourier transforms of oracle/dft.py pinned by closed forms.

* A source at the phase centre gives V = S on every baseline and channel.
* Hand-computed phases: with nu = c (lambda = 1 m) the phase is 2 pi (u l + v m + w (n - 1)),
  chosen so that each term alone gives a quarter or half turn; the sign is the +2 pi i of
  ClassDDEGridMachine.testGrid (ClassDDEGridMachine.py:271).
* The dirty image of a noiseless single source equals its flux at its own position for any uv
  coverage, weights and flags; one visibility gives the dirty beam cos(2 pi u . dl) in closed form.
* Linearity in the sky, and the pixel -> (l, m) map of the main image.
"""
import numpy as np

from oracle import dft

C = 299792458.0


def test_phase_centre_source_is_flat():
    rng = np.random.default_rng(1)
    uvw = rng.normal(0, 3000, (50, 3))
    V = dft.predict(uvw, np.linspace(1e9, 2e9, 7), np.array([2.5]), np.array([0.0]), np.array([0.0]))
    np.testing.assert_allclose(V, 2.5, rtol=0, atol=1e-12)


def test_hand_computed_phases():
    nu = np.array([C])
    uvw = np.array([[100.0, 0.0, 0.0],     # u l = 100 * 0.0025 = 1/4 turn
                    [0.0, 50.0, 0.0],      # v m = 50 * 0.01 = 1/2 turn
                    [0.0, 0.0, 2.5],       # w (n - 1) = 2.5 * (-0.2) = -1/2 turn (l = 0.6)
                    [100.0, 50.0, 0.0]])
    flux = np.array([1.0])
    V = dft.predict(uvw[:1], nu, flux, np.array([0.0025]), np.array([0.0]))
    np.testing.assert_allclose(V[0, 0], 1j, atol=1e-12)
    V = dft.predict(uvw[1:2], nu, flux, np.array([0.0]), np.array([0.01]))
    np.testing.assert_allclose(V[0, 0], -1.0, atol=1e-12)
    V = dft.predict(uvw[2:3], nu, flux, np.array([0.6]), np.array([0.0]))
    np.testing.assert_allclose(V[0, 0], -1.0, atol=1e-12)
    V = dft.predict(uvw[3:4], nu, flux, np.array([0.0025]), np.array([0.01]))
    np.testing.assert_allclose(V[0, 0], -1j, atol=1e-12)
    # frequency scales the phase: nu = 2c doubles the quarter turn
    V = dft.predict(uvw[:1], 2 * nu, flux, np.array([0.0025]), np.array([0.0]))
    np.testing.assert_allclose(V[0, 0], -1.0, atol=1e-12)


def test_single_source_dirty_equals_flux_at_its_position():
    rng = np.random.default_rng(2)
    uvw = rng.normal(0, 2000, (80, 3))
    freqs = np.linspace(1.2e9, 1.6e9, 5)
    l, m = np.array([0.003]), np.array([-0.002])
    V = dft.predict(uvw, freqs, np.array([0.7]), l, m)
    w = rng.uniform(0.5, 2.0, (80, 5)).astype(np.float32)
    flags = rng.random((80, 5)) < 0.2
    np.testing.assert_allclose(dft.dirty(uvw, freqs, V, w, flags, l, m), 0.7, rtol=1e-12)


def test_one_visibility_dirty_beam():
    uvw = np.array([[700.0, -300.0, 40.0]])
    freqs = np.array([1.4e9])
    l0, m0 = np.array([0.001]), np.array([0.002])
    V = dft.predict(uvw, freqs, np.array([1.0]), l0, m0)
    lp, mp = np.array([0.0013, -0.002]), np.array([0.0025, 0.001])
    got = dft.dirty(uvw, freqs, V, np.ones((1, 1), np.float32), np.zeros((1, 1), bool), lp, mp)
    k = freqs[0] / C
    n = lambda a, b: np.sqrt(1 - a * a - b * b) - 1  # noqa: E731
    want = np.cos(2 * np.pi * k * (uvw[0, 0] * (l0 - lp) + uvw[0, 1] * (m0 - mp) + uvw[0, 2] * (n(l0, m0) - n(lp, mp))))
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_linearity_and_pixel_map():
    rng = np.random.default_rng(3)
    uvw = rng.normal(0, 1000, (20, 3))
    freqs = np.array([1.3e9, 1.5e9])
    l, m = np.array([0.001, -0.0007]), np.array([0.0004, 0.0012])
    both = dft.predict(uvw, freqs, np.array([1.5, 0.25]), l, m)
    one = dft.predict(uvw, freqs, np.array([1.5]), l[:1], m[:1])
    two = dft.predict(uvw, freqs, np.array([0.25]), l[1:], m[1:])
    np.testing.assert_allclose(both, one + two, rtol=1e-12)
    L, M = dft.image_pixel_lm(np.array([50, 57]), np.array([50, 40]), 101, 1e-4)
    np.testing.assert_allclose(L, [0.0, 7e-4])
    np.testing.assert_allclose(M, [0.0, -1e-3])
