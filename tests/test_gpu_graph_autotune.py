"""The batched major cycle captured in a HIP graph, and the plan autotuner.

Graph: every replay (with eager device work and synchronisations between replays) must give the
facet grids, weights and degridded visibilities of an eager cycle.  The LDS accumulation order
of the tile passes can differ between runs, so the bar is 1e-6 relative L2, not bit equality.
Autotuner: every candidate plan gives the same cycle (same bar), and the chosen plan is one of
them.
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _facet_set(seed=21):
    from ddfacet_amd.facets import DeviceFacetSet
    from ddfacet_amd.problems import random_problem
    P = random_problem(seed, nfacet=4, npix=64, nchan=8, jones=dict(mode=1, interp=0, kms=True),
                       bda="mixed")
    fs = DeviceFacetSet(P)
    g = torch.Generator(device="cuda").manual_seed(seed)
    fs.models.copy_(torch.randn(fs.models.shape, generator=g, device="cuda", dtype=torch.complex64))
    return fs


def _eager(fs, vis0):
    fs.t["vis"].copy_(vis0)
    fs.zero_grids()
    fs.grid_batched()
    fs.degrid_batched()
    torch.cuda.synchronize()
    return fs.grids.clone(), fs.sumwt.clone(), fs.t["vis"].clone()


def test_graph_replays_match_eager_cycle():
    fs = _facet_set()
    fs.plan_batched(tile=16)
    vis0 = fs.t["vis"].clone()
    g_ref, w_ref, v_ref = _eager(fs, vis0)
    fs.t["vis"].copy_(vis0)
    torch.cuda.synchronize()
    graph = fs.capture_cycle()
    for k in range(4):
        fs.t["vis"].copy_(vis0)
        scratch = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(scratch)
        assert rel_l2(fs.grids.cpu().numpy(), g_ref.cpu().numpy()) < 1e-6, k
        assert rel_l2(fs.sumwt.cpu().numpy(), w_ref.cpu().numpy()) < 1e-12, k
        assert rel_l2(fs.t["vis"].cpu().numpy(), v_ref.cpu().numpy()) < 1e-6, k
    fs.destroy_plan()


def test_autotuner_candidates_agree_and_best_is_kept():
    from ddfacet_amd import autotune
    fs = _facet_set(seed=22)
    vis0 = fs.t["vis"].clone()
    space = {"tile": (8, 16, 32), "chunk": (256, 2048), "blocks_per_item": (16, 128)}
    outs = []
    for c in autotune.candidates(space, halo=3):
        fs.plan_batched(**c)
        outs.append(_eager(fs, vis0))
    for g, w, v in outs[1:]:
        assert rel_l2(g.cpu().numpy(), outs[0][0].cpu().numpy()) < 1e-6
        assert rel_l2(w.cpu().numpy(), outs[0][1].cpu().numpy()) < 1e-12
        assert rel_l2(v.cpu().numpy(), outs[0][2].cpu().numpy()) < 1e-6
    fs.t["vis"].copy_(vis0)
    best, results = autotune.autotune(fs, space=space, reps=1)
    assert best in [c for c, _ in results]
    assert len(results) == len(autotune.candidates(space, halo=3))
    assert fs.plan_info()["tile"] == best["tile"]
    fs.destroy_plan()
