"""On-device CF tables (ddf_cf_build) against the host restatement of ModCF.ClassWTermModified
(ddfacet_amd/wterm.py): every w-plane and conjugate plane of several facets, centred and far off
centre, with the 81-facet workload's plane supports.  Both compute in FP64 and round to
complex64, so they agree to a few complex64 ulps of the plane's peak."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nw,sups,overs", [(32, None, 11), (4, (7, 7, 9, 11), 5), (1, (7,), 11)])
def test_device_cf_tables_match_host_builder(nw, sups, overs):
    from ddfacet_amd.cfdev import DeviceCFSet
    from ddfacet_amd.problems import make_cf_facet
    if sups is None:
        sups = [s if s % 2 else s + 1 for s in np.int64(np.linspace(7, 15, nw))]
    npix, cell = 735, 1.5
    freqs = np.linspace(0.9e9, 1.67e9, 16)
    wmax = 9000.0
    arc = np.pi / 180 / 3600
    lm = [(0.0, 0.0), (1800 * arc, -1200 * arc), (-2700 * arc, 2700 * arc)]
    dev = DeviceCFSet(lm, npix, cell, freqs, nw, sups, overs, wmax)
    torch.cuda.synchronize()
    for f, (l0, m0) in enumerate(lm):
        host = make_cf_facet(l0, m0, npix, cell, freqs, nw, list(sups), overs, wmax)
        planes = list(host.wplanes) + list(host.wplanes_conj)
        for k, want in enumerate(planes):
            got = dev.plane(f, k).cpu().numpy()
            want = np.asarray(want)
            assert got.shape == want.shape, (f, k)
            err = np.abs(got.astype(np.complex128) - want).max() / np.abs(want).max()
            assert err < 1e-6, (f, k, err)
