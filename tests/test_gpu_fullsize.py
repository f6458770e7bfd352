"""The 81-facet MeerKAT-like chunk at full size (77.4 M visibilities, 81 facets of 735^2).

* The batched cycle against the plain per-facet kernels (each pinned to the reference golden
  vectors by test_gpu_plain.py): facet grids and weights of a sample of facets, and the
  visibilities predicted from all 81 facets, relative L2 <= 1e-5.
* Linearity, a size-independent property: gridding 2 x the visibilities gives exactly 2 x the
  facet grids (scaling by 2 is exact in FP32 and the batched passes sum in a fixed order).
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu
SAMPLE = [0, 40, 80]


def test_full_size_batched_cycle_against_plain_kernels_and_linearity():
    from ddfacet_amd import configs
    from ddfacet_amd.facets import DeviceFacetSet
    print("81-facet chunk at full size", flush=True)
    P = configs.cfg2(with_vis=False, log=lambda *a: None)
    P.vis = configs.cfg2_visibilities(P.uvw, P.freqs, P.extra["sky"], device="cuda")
    fs = DeviceFacetSet(P)
    g = torch.Generator(device="cuda").manual_seed(5)
    fs.models.copy_(torch.randn(fs.models.shape, generator=g, device="cuda", dtype=torch.complex64) * 1e-3)
    vis0 = fs.t["vis"].clone()
    fs.plan_batched()

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
    assert rel_l2(pred_b, pred_p) <= 1e-5, rel_l2(pred_b, pred_p)
    fs.destroy_plan()
