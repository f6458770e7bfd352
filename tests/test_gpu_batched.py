"""Batched all-facet path (ddf_plan_*) against the golden vectors of the reference gridder.

One plan grids every facet of a case and degrids the sum of every facet's model; the result must
equal one reference call per facet: per-facet grids / sumwt / Jones sums (gridding) and
vis - sum_f model_f (degridding), relative L2 <= 1e-5, touched grid cells identical.
"""
import os

import numpy as np
import pytest
import torch

from cases import CASES, model_grid, problem
from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _golden(name):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return np.load(os.path.join(here, name + ".npz"), allow_pickle=False)


def _facet_set(P):
    from ddfacet_amd.facets import DeviceFacetSet
    fs = DeviceFacetSet(P)
    for fi in range(fs.nf):
        fs.models[fi].copy_(torch.from_numpy(model_grid(P, 100 + fi)))
    return fs


@pytest.mark.parametrize("name", [c[0] for c in CASES])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batched_grid_and_degrid_match_reference(name, precision):
    print("case", name, precision, flush=True)
    P = problem(name)
    G = _golden(name)
    fs = _facet_set(P)
    info = fs.plan_batched(precision=precision, tile=16)
    assert info["grid_records"] > 0
    for psf in (0, 1):
        fs.zero_grids()
        fs.grid_batched(dopsf=psf)
        torch.cuda.synchronize()
        for fi in range(fs.nf):
            key = "f%d_psf%d" % (fi, psf)
            g = fs.grids[fi].cpu().numpy()
            ref = G["grid_" + key]
            assert rel_l2(g, ref) <= TOL, (key, rel_l2(g, ref))
            assert np.array_equal(g != 0, ref != 0), key
            np.testing.assert_allclose(fs.sumwt[fi].cpu().numpy(), G["sumwt_" + key], rtol=1e-10)
            if P.jones is not None:
                np.testing.assert_allclose(fs.sum_jones[fi].cpu().numpy(), G["sumjones_" + key], rtol=1e-10)
                np.testing.assert_allclose(fs.sum_jones_chan[fi].cpu().numpy(), G["sumjoneschan_" + key],
                                           rtol=1e-10, atol=1e-300)
    vis = fs.t["vis"].clone()
    fs.degrid_batched(vis=vis)
    torch.cuda.synchronize()
    model_gpu = P.vis - vis.cpu().numpy()
    model_ref = sum(P.vis - G["vis_f%d" % fi] for fi in range(fs.nf))
    assert rel_l2(model_gpu, model_ref) <= TOL, rel_l2(model_gpu, model_ref)
    fs.destroy_plan()


def test_batched_cycle_is_repeatable():
    """Two cycles with the same plan give the same grids (plan reuse across major cycles)."""
    P = problem("multifacet")
    fs = _facet_set(P)
    fs.plan_batched()
    fs.zero_grids()
    fs.grid_batched()
    g1 = fs.grids.clone()
    fs.zero_grids()
    fs.grid_batched()
    assert rel_l2(fs.grids.cpu().numpy(), g1.cpu().numpy()) <= 1e-6
