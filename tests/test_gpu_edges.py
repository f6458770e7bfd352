"""Edge cases of the batched all-facet path against the oracle (reference restatement), facet by
facet: every visibility flagged, an empty BDA map, and uvw so long that every stamp falls outside
its facet grid (all records dropped).  Gridding must give the oracle's grids and weights (zero
where nothing is gridded); degridding (which ignores flags) the oracle's visibilities."""
import numpy as np
import pytest
import torch

from cases import model_grid, problem
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _run(P):
    from ddfacet_amd.facets import DeviceFacetSet
    from oracle import cpuref
    fs = DeviceFacetSet(P)
    models = [model_grid(P, 100 + fi) for fi in range(fs.nf)]
    for fi in range(fs.nf):
        fs.models[fi].copy_(torch.from_numpy(models[fi]))
    info = fs.plan_batched(tile=16)
    fs.zero_grids()
    fs.grid_batched()
    vis = fs.t["vis"].clone()
    fs.degrid_batched(vis=vis)
    torch.cuda.synchronize()
    v_ref = P.vis.copy()
    for fi in range(fs.nf):
        g_ref, s_ref = P.new_grid(), np.zeros((P.nband, P.npol))
        cpuref.grid(*P.grid_args(fi, g_ref, s_ref))
        g = fs.grids[fi].cpu().numpy()
        if np.any(g_ref):
            assert rel_l2(g, g_ref) <= 1e-5, fi
        else:
            assert not np.any(g), fi
        np.testing.assert_allclose(fs.sumwt[fi].cpu().numpy(), s_ref, rtol=1e-10, atol=0)
        cpuref.degrid(*P.degrid_args(fi, models[fi], v_ref))
    v = vis.cpu().numpy()
    pred, pred_ref = P.vis - v, P.vis - v_ref
    if np.any(pred_ref):
        assert rel_l2(pred, pred_ref) <= 1e-5
    else:
        assert np.array_equal(v, P.vis)
    fs.destroy_plan()
    return info


def test_all_visibilities_flagged():
    P = problem("multifacet")
    P.flags[:] = True
    info = _run(P)
    # a block without an unflagged visibility has no placement in gridding (its means are
    # undefined); degridding ignores flags
    assert info["grid_records"] == 0 and info["degrid_records"] > 0


def test_empty_bda_map():
    P = problem("multifacet")
    P.bda_grid = np.zeros(2, np.int32)
    P.bda_degrid = np.zeros(2, np.int32)
    info = _run(P)
    assert info["grid_records"] == 0 and info["degrid_records"] == 0


def test_all_stamps_outside_the_facets():
    P = problem("multifacet")
    P.uvw[:, :2] *= 1e3
    info = _run(P)
    assert info["grid_records"] == 0 and info["degrid_records"] == 0
