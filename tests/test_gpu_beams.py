"""Beam Jones producers on the device (ddf_beam_jones, ddf_jones_merge) against the oracle."""
import numpy as np
import pytest
import torch

from oracle import beams as ob

pytestmark = pytest.mark.gpu


def _sky(nd, seed, spread):
    rng = np.random.default_rng(seed)
    ra_c, dec_c = 2.1, -0.52
    ra = ra_c + rng.uniform(-spread, spread, nd)
    dec = dec_c + rng.uniform(-spread, spread, nd)
    ra[0], dec[0] = ra_c, dec_c
    return ra, dec, ra_c, dec_c


@pytest.mark.parametrize("kind,freqs,nband", [
    ("GMRT", np.linspace(1.06e9, 1.44e9, 48), 4),
    ("GMRT", np.linspace(560e6, 840e6, 16), 0),
    ("ATCA", np.linspace(1.1e9, 3.0e9, 64), 8),
    ("ATCA", np.array([1268e6, 1396e6, 1460e6, 2932e6, 3.05e9]), 0),
])
@pytest.mark.parametrize("center_norm", [True, False])
def test_beam_jones_matches_oracle(kind, freqs, nband, center_norm):
    from ddfacet_amd import beams
    print("case", kind, freqs.size, nband, center_norm, flush=True)
    # directions out to ~3 deg: past the GMRT clipping radius and the ATCA 100 arcmin GHz cut
    ra, dec, ra_c, dec_c = _sky(57, 11, 0.04)
    t = np.array([0.0, 300.0, 600.0, 900.0, 901.0])
    width = float(freqs[1] - freqs[0]) if freqs.size > 1 else 1e6
    d = beams.estimate_beam(t, ra, dec, ra_c, dec_c, freqs, width, 5, kind=kind, nband=nband,
                            center_norm=center_norm)
    r = ob.estimate_beam(t, ra, dec, ra_c, dec_c, freqs, width, nband, 5, kind, center_norm=center_norm)
    got = d["Jones"].cpu().numpy()
    assert got.shape == r["Jones"].shape
    np.testing.assert_allclose(got, r["Jones"], rtol=2e-6, atol=1e-7)
    assert np.array_equal(d["VisToJonesChanMapping"], r["VisToJonesChanMapping"])
    assert np.array_equal(d["FreqDomains"], r["FreqDomains"])
    assert np.array_equal(d["tm"], r["tm"])


def test_jones_merge_matches_oracle():
    from ddfacet_amd import beams
    rng = np.random.default_rng(5)
    nd, na = 7, 6
    f = np.linspace(1.06e9, 1.44e9, 32)
    # a DD-solution table (10-min intervals, 4 channels) and a beam table (5-min, 6 channels)
    t0a = np.arange(0.0, 3600.0, 600.0)
    fa = beams.freq_domains(f, f[1] - f[0], 4)
    Ja = (rng.normal(size=(t0a.size, nd, na, 4, 2, 2)) + 1j * rng.normal(size=(t0a.size, nd, na, 4, 2, 2)))
    ra, dec, ra_c, dec_c = _sky(nd, 3, 0.02)
    tb = beams.beam_sample_times(np.arange(0.0, 3600.0, 8.0), 5)
    beam = beams.estimate_beam(tb, ra, dec, ra_c, dec_c, f, f[1] - f[0], na, kind="GMRT", nband=6)
    da = {"Jones": torch.from_numpy(Ja.astype(np.complex64)).cuda(), "t0": t0a, "t1": t0a + 600.0,
          "tm": t0a + 300.0, "FreqDomains": fa}
    m = beams.merge_jones(da, beam)
    ref = ob.merge_jones(dict(da, Jones=Ja.astype(np.complex64)), dict(beam, Jones=beam["Jones"].cpu().numpy()))
    for k in ("t0", "t1", "tm", "FreqDomains"):
        assert np.array_equal(m[k], ref[k]), k
    np.testing.assert_allclose(m["Jones"].cpu().numpy(), ref["Jones"], rtol=2e-6, atol=2e-6)
    # channels of the merged grid outside one table keep the identity
    fb = np.array([[1.0e9, 1.2e9]])
    db = {"Jones": torch.ones((1, nd, na, 1, 2, 2), dtype=torch.complex64, device="cuda"), "t0": np.array([0.0]),
          "t1": np.array([3600.0]), "tm": np.array([1800.0]), "FreqDomains": fb}
    m2 = beams.merge_jones(da, db)
    r2 = ob.merge_jones(dict(da, Jones=Ja.astype(np.complex64)), dict(db, Jones=np.ones((1, nd, na, 1, 2, 2), np.complex64)))
    np.testing.assert_allclose(m2["Jones"].cpu().numpy(), r2["Jones"], rtol=2e-6, atol=2e-6)


def test_beam_jones_rejects_bad_input():
    import ctypes as C
    from ddfacet_amd import abi
    from ddfacet_amd._lib import lib
    d = abi.ddf_beam_desc()
    d.kind = 7
    assert lib().ddf_beam_jones(C.byref(d), None, None, 0, 0.0, 0.0, 0, 0, 0, None, None) < 0
    m = abi.ddf_jones_merge_desc()
    m.nt0 = 0
    assert lib().ddf_jones_merge(C.byref(m), None) < 0
