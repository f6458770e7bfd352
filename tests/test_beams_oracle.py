"""The beam Jones producer oracle (oracle/beams.py) pinned by closed forms and hand-worked cases,
and the host side of ddfacet_amd.beams (model terms, domains) against it."""
import numpy as np
import pytest

from oracle import beams as ob


def test_ang_dist_known_separations():
    assert ob.ang_dist(0.3, -0.5, 0.3, -0.5) == 0.0
    assert ob.ang_dist(0.0, 0.0, np.pi / 2, 0.0) == pytest.approx(np.pi / 2, abs=1e-15)
    # from the pole every point is 90 deg - dec away
    assert ob.ang_dist(1.2, np.pi / 2, 0.1, 0.4) == pytest.approx(np.pi / 2 - 0.4, abs=1e-12)
    # a cosine rounded past 1 is clipped (no NaN)
    assert np.isfinite(ob.ang_dist(np.array([1e-9]), np.array([0.7]), 0.0, 0.7)).all()


def _poly(c, x):
    return sum(float(ck) * x ** k for k, ck in enumerate(c))


def test_gmrt_beam_closed_forms():
    f = np.array([1.1e9, 1.4e9, 320e6, 700e6])
    model = ob.gmrt_channel_model(f)
    C, xn, yn = model
    # unit response on axis
    assert np.array_equal(ob.raw_beam_gmrt(np.array([0.0]), f, model), np.ones((1, 4)))
    # the polynomial of the channel's band at d nu / 1 GHz, float32 coefficients
    d = np.array([5.0, 12.5])
    got = ob.raw_beam_gmrt(d, f, model)
    for j, fj in enumerate(f):
        for i, di in enumerate(d):
            if di <= xn[j]:
                assert got[i, j] == pytest.approx(_poly(C[j], di * fj / 1e9), rel=1e-12)
    # the clipping point is a stationary point of the band polynomial, the clipped value its value
    for b, (_, _, c) in ob.GMRT_BANDS.items():
        x0, y0 = ob.gmrt_band_null(c)
        dp = sum(k * ck * x0 ** (k - 1) for k, ck in enumerate(c) if k)
        assert x0 > 0 and abs(dp) < 1e-9 and y0 == pytest.approx(_poly(c, x0), rel=1e-12)
    far = ob.raw_beam_gmrt(np.array([1e4]), f, model)
    assert np.array_equal(far[0], yn.astype(np.float64))
    with pytest.raises(NotImplementedError):
        ob.gmrt_channel_model(np.array([950e6]))


def test_atca_beam_closed_forms():
    f = np.array([1396e6, 1460e6, 1000e6, 3.1e9])
    model = ob.atca_channel_model(f)
    b0, b1 = model
    assert list(b0) == [1, 1, 0, 13] and list(b1) == [1, 2, 0, 13]
    d = np.array([0.0, 20.0])
    got = ob.raw_beam_atca(d, f, model)
    assert np.allclose(got[0], 1.0, rtol=0, atol=1e-15)
    # on a band frequency: the square root of that band's polynomial at d nu * 1e-9
    c1 = ob.ATCA_BANDS[1][1]
    assert got[1, 0] == pytest.approx(np.sqrt(abs(_poly(c1, 20.0 * 1396e6 * 1e-9))), rel=1e-12)
    # between bands: the reference's mix, each band's term weighted by its own distance
    c2 = ob.ATCA_BANDS[2][1]
    w0, w1 = 1460e6 - 1396e6, 1524e6 - 1460e6
    v0 = np.sqrt(abs(_poly(c1, 20.0 * 1396e6 * 1e-9)))
    v1 = np.sqrt(abs(_poly(c2, 20.0 * 1524e6 * 1e-9)))
    assert got[1, 1] == pytest.approx((w0 * v0 + w1 * v1) / (w0 + w1), rel=1e-12)
    # past 100 arcmin GHz: y_null / (d nu - y_null), y_null the float32 value at 100
    far = ob.raw_beam_atca(np.array([200.0]), np.array([1396e6]), ob.atca_channel_model(np.array([1396e6])))
    yn = float(np.float32(np.sqrt(abs(_poly(c1, 100.0)))))
    dnu = 200.0 * 1396e6 * 1e-9
    assert far[0, 0] == pytest.approx(yn / (dnu - yn), rel=1e-12)


def test_batch_linear_algebra():
    rng = np.random.default_rng(1)
    a = (rng.normal(size=(5, 3, 2, 2)) + 1j * rng.normal(size=(5, 3, 2, 2))).astype(np.complex128)
    b = (rng.normal(size=(5, 3, 2, 2)) + 1j * rng.normal(size=(5, 3, 2, 2))).astype(np.complex128)
    assert np.allclose(ob.batch_dot(a, b), a @ b, rtol=1e-13, atol=1e-13)
    eye = np.broadcast_to(np.eye(2), a.shape)
    assert np.allclose(ob.batch_dot(ob.batch_inverse(a), a), eye, atol=1e-12)


def test_estimate_beam_center_norm_and_shapes():
    f = np.linspace(1.1e9, 1.4e9, 8)
    ra_c, dec_c = 1.0, -0.6
    ra = np.array([ra_c, ra_c + 0.01, ra_c - 0.02])
    dec = np.array([dec_c, dec_c + 0.005, dec_c - 0.01])
    t = np.array([0.0, 300.0, 600.0, 601.0])
    d = ob.estimate_beam(t, ra, dec, ra_c, dec_c, f, 40e6, 2, 4, "GMRT", center_norm=True)
    J = d["Jones"]
    assert J.shape == (3, 3, 4, 2, 2, 2) and J.dtype == np.complex64
    assert np.array_equal(d["t0"], t[:-1]) and np.array_equal(d["t1"], t[1:]) and np.array_equal(d["tm"], (t[:-1] + t[1:]) / 2)
    assert list(d["VisToJonesChanMapping"]) == [0] * 4 + [1] * 4
    # the phase centre normalised by itself is the identity
    assert np.array_equal(J[:, 0, :, :, 0, 0], np.ones((3, 4, 2), np.complex64))
    assert np.array_equal(J[..., 0, 1], np.zeros(J.shape[:4], np.complex64))
    # off axis: the mean of the channels' beams, below 1, the same for every time and antenna
    raw = ob.raw_beam_gmrt(ob.ang_dist(ra, dec, ra_c, dec_c) * ob.ARCMIN_PER_RAD, f, ob.gmrt_channel_model(f))
    assert np.allclose(J[0, 1, 0, 0, 0, 0].real, raw[1, :4].mean(), rtol=1e-6)
    assert (J[:, 1:, :, :, 0, 0].real < 1).all() and np.array_equal(J[0], J[2])


def test_merged_domains_hand_worked():
    fd0 = np.array([[0.0, 10.0], [10.0, 20.0]])
    fd1 = np.array([[5.0, 15.0], [15.0, 25.0]])
    assert np.array_equal(ob.merged_freq_domains(fd0, fd1), np.array([[5, 10], [10, 15], [15, 20.0]]))
    # edges within 1 Hz merge into the first
    fd2 = np.array([[0.5, 10.2], [10.2, 20.0]])
    assert np.array_equal(ob.merged_freq_domains(fd0, fd2), np.array([[0.0, 10.0], [10.0, 20.0]]))
    t0, t1, tm = ob.merged_times([0.0, 10.0], [10.0, 20.0], [0.0, 15.0], [15.0, 30.0])
    assert np.array_equal(t0, [0, 10, 15, 20]) and np.array_equal(t1, [10, 15, 20, 30])
    assert np.array_equal(tm, [5, 12.5, 17.5, 25])


def test_merge_jones_cells():
    rng = np.random.default_rng(2)
    J0 = (rng.normal(size=(2, 1, 2, 2, 2, 2)) + 1j * rng.normal(size=(2, 1, 2, 2, 2, 2))).astype(np.complex64)
    J1 = (rng.normal(size=(2, 1, 2, 2, 2, 2)) + 1j * rng.normal(size=(2, 1, 2, 2, 2, 2))).astype(np.complex64)
    d0 = {"Jones": J0, "t0": np.array([0.0, 10.0]), "t1": np.array([10.0, 20.0]), "tm": np.array([5.0, 15.0]),
          "FreqDomains": np.array([[0.0, 10.0], [10.0, 20.0]])}
    d1 = {"Jones": J1, "t0": np.array([0.0, 15.0]), "t1": np.array([15.0, 30.0]), "tm": np.array([7.5, 22.5]),
          "FreqDomains": np.array([[5.0, 15.0], [15.0, 25.0]])}
    m = ob.merge_jones(d0, d1)
    out = m["Jones"]
    assert out.shape == (4, 1, 2, 3, 2, 2)
    # merged time 12.5 (cell 1): nearest centres 15 (table 0, index 1) and 7.5 (table 1, index 0);
    # merged channel [10, 15) (cell 1): channel 1 of table 0, channel 0 of table 1
    assert np.allclose(out[1, :, :, 1], J0[1, :, :, 1] @ J1[0, :, :, 0], rtol=1e-5)
    # merged time 25 (cell 3) and channel [15, 20) (cell 2): indices (1, 1) and channels (1, 1)
    assert np.allclose(out[3, :, :, 2], J0[1, :, :, 1] @ J1[1, :, :, 1], rtol=1e-5)
    # with an identity second table the merge resamples the first
    d1i = dict(d1, Jones=np.broadcast_to(np.eye(2, dtype=np.complex64), J1.shape).copy())
    mi = ob.merge_jones(d0, d1i)["Jones"]
    assert np.array_equal(mi[0, :, :, 0], J0[0, :, :, 0])


def test_host_terms_and_domains_match_oracle():
    from ddfacet_amd import beams
    f = np.linspace(1.05e9, 1.45e9, 16)
    t = beams.channel_terms("GMRT", f)
    C, xn, yn = ob.gmrt_channel_model(f)
    assert np.array_equal(t["coef0"], C.astype(np.float64)) and np.array_equal(t["x_null"], xn)
    assert np.array_equal(t["y_null"], yn)
    f = np.linspace(1.1e9, 3.1e9, 37)
    t = beams.channel_terms("ATCA", f)
    b0, b1 = ob.atca_channel_model(f)
    assert np.array_equal(t["band_f0"][:, 0], [ob.ATCA_BANDS[i][0] for i in b0])
    assert np.array_equal(t["band_f0"][:, 1], [ob.ATCA_BANDS[i][0] for i in b1])
    assert np.array_equal(t["coef1"], np.array([ob.ATCA_BANDS[i][1] for i in b1], np.float64))
    dom = beams.freq_domains(f, 50e6, 5)
    assert np.array_equal(dom, ob.freq_domains(f, 50e6, 5))
    assert np.array_equal(beams.nearest_domain(f, dom), ob.vis_to_jones_chan(f, dom))
    rng = np.random.default_rng(3)
    for _ in range(20):
        e0 = np.sort(rng.uniform(0, 100, 4))
        e1 = np.sort(np.concatenate([rng.uniform(0, 100, 3), e0[:1] + rng.uniform(-0.9, 0.9, 1)]))
        fd0 = np.stack([e0[:-1], e0[1:]], 1)
        fd1 = np.stack([e1[:-1], e1[1:]], 1)
        assert np.array_equal(beams.merged_freq_domains(fd0, fd1), ob.merged_freq_domains(fd0, fd1))
        t0 = np.sort(rng.uniform(0, 100, 3))
        t1 = np.sort(rng.uniform(0, 100, 4))
        a = beams.merged_time_domains(t0[:-1], t0[1:], t1[:-1], t1[1:])
        b = ob.merged_times(t0[:-1], t0[1:], t1[:-1], t1[1:])
        assert all(np.array_equal(x, y) for x, y in zip(a, b))
    times = np.arange(0.0, 3600.0, 8.0)
    assert np.array_equal(beams.beam_sample_times(times, 5), np.concatenate([np.arange(0.0, 3593.0, 300.0), [3593.0]]))
