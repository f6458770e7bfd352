"""Edge cases of the batched all-facet path against the oracle (reference restatement), facet by
facet: every visibility flagged, an empty BDA map, uvw so long that every stamp falls outside its
facet grid (all records dropped), and BDA blocks cut by Jones-channel edges.  Gridding must give the oracle's grids and weights (zero
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


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_blocks_cut_by_jones_channel_edges(precision):
    """Per-visibility Jones (mode 1) with DD Jones channels of 2 channels and beam Jones channels
    cut at channels 5 and 7, over BDA blocks of 1-8 channels: the gridding walk takes blocks cut
    into 2-4 segments segment by segment (channels 8-15: 4 segments) and falls back to row-major
    past 4 segments (channels 0-7: 6 segments).  Grids, weights and both Jones sums against the
    oracle, facet by facet."""
    from ddfacet_amd import synth
    from ddfacet_amd.facets import DeviceFacetSet
    from ddfacet_amd.problems import random_problem
    from oracle import cpuref
    P = random_problem(41, nchan=16, nfacet=4, ntime=8, jones=dict(mode=1, beam=True))
    nd = P.jones["kms"].shape[1]
    nant = P.jones["kms"].shape[2]
    P.jones["kms"] = synth.random_jones(3, nd, nant, 8, 4141, 0.2, 0.5)
    P.jones["kms_chan"] = (np.arange(16) // 2).astype(np.int32)
    P.jones["beam"] = synth.random_jones(2, nd, nant, 3, 4142, 0.1, 0.3)
    P.jones["beam_chan"] = np.array([0] * 5 + [1] * 2 + [2] * 9, np.int32)
    print("case jones-channel edges", precision, flush=True)
    fs = DeviceFacetSet(P)
    fs.plan_batched(precision=precision, tile=16)
    fs.zero_grids()
    fs.grid_batched()
    torch.cuda.synchronize()
    for fi in range(fs.nf):
        g_ref, s_ref = P.new_grid(), np.zeros((P.nband, P.npol))
        sj, sjc = np.zeros((2, P.nband)), np.zeros((2, P.freqs.size))
        cpuref.grid(*P.grid_args(fi, g_ref, s_ref, sum_jones=sj, sum_jones_chan=sjc))
        g = fs.grids[fi].cpu().numpy()
        assert np.any(g_ref), fi
        assert rel_l2(g, g_ref) <= 1e-5, (fi, rel_l2(g, g_ref))
        assert np.array_equal(g != 0, g_ref != 0), fi
        np.testing.assert_allclose(fs.sumwt[fi].cpu().numpy(), s_ref, rtol=1e-10, atol=0)
        np.testing.assert_allclose(fs.sum_jones[fi].cpu().numpy(), sj, rtol=1e-10)
        np.testing.assert_allclose(fs.sum_jones_chan[fi].cpu().numpy(), sjc, rtol=1e-10, atol=1e-300)
    fs.destroy_plan()
