"""The 81-facet MeerKAT-like chunk at full size (77.4 M visibilities, 81 facets of 735^2, 32
w-planes with supports 7-15, per-visibility DD Jones, default BDA maps).

* The batched cycle against the oracle (cpu_ref, pinned to the reference gridder) on three
  facets, the corner 0, the edge 4 and the centre 40: facet grids relative L2 <= 1e-5 with the
  touched cells identical, sumwt / SumJones / SumJonesChan <= 1e-10, and the visibilities
  predicted from those facets' model grids <= 1e-5.
* The plan's integer placement maps of those facets (grid and degrid maps, 5.5 M blocks each)
  equal the oracle's bit for bit.
* The batched cycle against the plain per-facet kernels on facets 0, 40, 80, and over all 81
  facets for the predicted visibilities (<= 1e-5).
* Linearity, a size-independent property: gridding 2 x the visibilities gives exactly 2 x the
  facet grids (scaling by 2 is exact in FP32 and the batched passes sum in a fixed order).
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
SAMPLE = [0, 40, 80]
ORACLE_SAMPLE = [0, 4, 40]


@pytest.fixture(scope="module")
def chunk():
    from ddfacet_amd import configs
    from ddfacet_amd.facets import DeviceFacetSet
    print("81-facet chunk at full size", flush=True)
    P = configs.cfg2(with_vis=False, log=lambda *a: None)
    P.vis = configs.cfg2_visibilities(P.uvw, P.freqs, P.extra["sky"], device="cuda")
    fs = DeviceFacetSet(P)
    g = torch.Generator(device="cuda").manual_seed(5)
    fs.models.copy_(torch.randn(fs.models.shape, generator=g, device="cuda", dtype=torch.complex64) * 1e-3)
    fs.plan_batched()
    from ddfacet_amd.problems import Problem
    H = Problem(**{k: getattr(P, k) for k in P.__dataclass_fields__ if k != "vis"}, vis=fs.t["vis"].cpu().numpy())
    yield H, fs
    fs.destroy_plan()


def test_full_size_batched_cycle_against_oracle(chunk):
    from ddfacet_amd import marshal
    from oracle import cpuref
    H, fs = chunk
    vis_host = H.vis

    fs.zero_grids()
    fs.grid_batched()
    torch.cuda.synchronize()
    specs, outs = [], []
    for fi in ORACLE_SAMPLE:
        g = H.new_grid()
        s = np.zeros((H.nband, H.npol))
        args = H.grid_args(fi, g, s)
        specs.append(marshal.parse_grid_args(args, False))
        outs.append((g, s, args[13]))
    cpuref.grid_facets(specs, len(specs))
    for fi, (g, s, lj) in zip(ORACLE_SAMPLE, outs):
        got = fs.grids[fi].cpu().numpy()
        err = rel_l2(got, g)
        print("facet %d: grid rel L2 %.3e vs oracle" % (fi, err), flush=True)
        assert err <= 1e-5, (fi, err)
        assert np.array_equal(got != 0, g != 0), fi
        np.testing.assert_allclose(fs.sumwt[fi].cpu().numpy(), s, rtol=1e-10)
        np.testing.assert_allclose(fs.sum_jones[fi].cpu().numpy(), lj[17], rtol=1e-10)
        np.testing.assert_allclose(fs.sum_jones_chan[fi].cpu().numpy(), lj[18], rtol=1e-10, atol=1e-300)

    # prediction from the sampled facets' models only (the others zeroed), from zero visibilities
    keep = fs.models.clone()
    mask = torch.zeros(fs.nf, dtype=torch.bool, device=fs.models.device)
    mask[ORACLE_SAMPLE] = True
    fs.models[~mask] = 0
    vis = torch.zeros_like(fs.t["vis"])
    fs.degrid_batched(vis=vis)
    torch.cuda.synchronize()
    pred = (-vis).cpu().numpy()
    fs.models.copy_(keep)
    v = np.zeros_like(vis_host)
    dspecs = [marshal.parse_degrid_args(H.degrid_args(fi, fs.models[fi].cpu().numpy(), v)) for fi in ORACLE_SAMPLE]
    cpuref.degrid_facets(dspecs, len(dspecs))
    err = rel_l2(pred, -v)
    print("predicted visibilities of facets %s: rel L2 %.3e vs oracle" % (ORACLE_SAMPLE, err), flush=True)
    assert err <= 1e-5, err


def test_full_size_placement_maps_bit_exact(chunk):
    from oracle import cpuref
    P, fs = chunk
    got_g = fs.placements(degrid=False)
    got_d = fs.placements(degrid=True)
    g = P.new_grid()
    for fi in ORACLE_SAMPLE:
        want_g = cpuref.place(*P.grid_args(fi, g, np.zeros((P.nband, P.npol))))
        want_d = cpuref.place(*P.degrid_args(fi, g, P.vis), degrid=True)
        print("facet %d: %d / %d grid blocks placed, %d degrid" % (fi, want_g[:, 0].sum(), want_g.shape[0],
                                                                  want_d[:, 0].sum()), flush=True)
        assert want_g[:, 0].sum() > 0.5 * want_g.shape[0]
        np.testing.assert_array_equal(got_g[fi], want_g)
        np.testing.assert_array_equal(got_d[fi], want_d)


def test_full_size_batched_cycle_against_plain_kernels_and_linearity(chunk):
    P, fs = chunk
    vis0 = fs.t["vis"].clone()
    fs.zero_grids()
    fs.grid_batched()
    torch.cuda.synchronize()
    g_b = fs.grids[SAMPLE].clone()
    w_b = fs.sumwt[SAMPLE].clone()

    fs.t["vis"].mul_(2)
    fs.zero_grids()
    fs.grid_batched()
    torch.cuda.synchronize()
    assert torch.equal(fs.grids[SAMPLE], 2 * g_b)
    fs.t["vis"].copy_(vis0)

    fs.zero_grids()
    fs.grid_plain(facets=SAMPLE)
    torch.cuda.synchronize()
    for k, fi in enumerate(SAMPLE):
        assert rel_l2(g_b[k].cpu().numpy(), fs.grids[fi].cpu().numpy()) <= 1e-5, fi
        np.testing.assert_allclose(w_b[k].cpu().numpy(), fs.sumwt[fi].cpu().numpy(), rtol=1e-10)

    # predictions from zeros (degridding subtracts): batched against plain over all facets
    fs.t["vis"].zero_()
    fs.degrid_batched()
    torch.cuda.synchronize()
    pred_b = (-fs.t["vis"]).cpu().numpy()
    fs.t["vis"].zero_()
    fs.degrid_plain()
    torch.cuda.synchronize()
    pred_p = (-fs.t["vis"]).cpu().numpy()
    fs.t["vis"].copy_(vis0)
    assert rel_l2(pred_b, pred_p) <= 1e-5, rel_l2(pred_b, pred_p)
