"""Precision study on the precision workload (cfg3 in ddfacet_amd/configs.py): the batched major
cycle in each arithmetic variant against the FP64 direct Fourier transform of the same point
sources, and against the FP64 variant.

Variants (ddf_plan_opts.precision):
  mixed  FP64 per-facet uvw rotation and per-row phase base, FP32 sincos at anchors with the
         per-channel phasor advanced in FP32 and re-anchored every 16 channels, FP64 block sums,
         complex64 grids accumulated in FP32
  fp64   FP64 phase, sincos and block sums (complex64 grids and CFs, as the reference's
         pyGridderWPol32 / pyDeGridderWPol)
  fp32   phase formed and reduced in FP32

The error against the DFT is dominated by the gridding approximation (prolate taper, w-kernel
sampling with S = 7, oversampling), the same for every variant; the error against the fp64
variant isolates the arithmetic.  The eight off-centre facets carry phase rotations of up to
~2000 rad on the longest baselines, where FP32 phase arithmetic shows.  The table is printed (run with -s) and bounded.
"""
import json
import os

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu

VARIANTS = ("mixed", "fp64", "fp32")
_cache = {}


def _workload(chan_block=1):
    if chan_block not in _cache:
        from ddfacet_amd import configs
        _cache[chan_block] = configs.cfg3(chan_block=chan_block, log=lambda *a: None)
    return _cache[chan_block]


def _run(precision, chan_block=1):
    from ddfacet_amd.cycle import MajorCycle
    P, geom = _workload(chan_block)
    cyc = MajorCycle(P, geom, precision=precision)
    X, Y = P.extra["src_pix"]
    model = np.zeros((1, 1, geom.m, geom.m), np.float32)
    model[0, 0, X, Y] = P.extra["sky"].flux.astype(np.float32)
    vis = torch.zeros(P.vis.shape, dtype=torch.complex64, device="cuda")
    cyc.predict(torch.from_numpy(model).cuda(), vis=vis)
    pred = -vis.cpu().numpy()[:, :, 0].astype(np.complex128)
    img = cyc.residual_image().cpu().numpy()[0, 0].astype(np.float64)
    cyc.fs.destroy_plan()
    return pred, img


def test_precision_study_vs_fp64_dft():
    from oracle import dft
    P, geom = _workload()
    X, Y = P.extra["src_pix"]
    l, m = dft.image_pixel_lm(X, Y, geom.m, P.extra["cell"])
    res = {v: _run(v) for v in VARIANTS}
    # visibilities whose BDA block lands on the grid (the others are dropped by every variant)
    on_grid = res["fp64"][0] != 0
    # predicted visibilities compared on every third row (the DFT dominates the test's run time)
    sub = np.arange(0, P.uvw.shape[0], 3)
    want_pred = dft.predict(P.uvw[sub], P.freqs, P.extra["sky"].flux.astype(np.float32).astype(np.float64), l, m)
    vis_i = 0.5 * (P.vis[:, :, 0] + P.vis[:, :, 3]).astype(np.complex128)
    want_dirty = dft.dirty(P.uvw, P.freqs, vis_i, P.weights * on_grid, P.flags[:, :, 0], l, m)
    table = {"workload": "cfg3: 27 ant 20 km disk, 100 x 10 s, 256 ch 0.9-1.67 GHz, 3 x 3 facets 297 -> 525 px "
                         "at 2\", 20 sources 4-12 px from the facet boundaries",
             "on_grid_fraction": float(on_grid.mean())}
    og = on_grid[sub]
    for v in VARIANTS:
        pred, img = res[v]
        ps = pred[sub]
        got_d = img[X, Y]
        table[v] = {
            "predict_rel_l2_vs_dft": rel_l2(ps[og], want_pred[og]),
            "predict_max_vs_dft": float(np.abs(ps - want_pred)[og].max() / np.abs(want_pred[og]).max()),
            "dirty_rel_l2_vs_dft": rel_l2(got_d, want_dirty),
            "dirty_max_vs_dft": float(np.abs(got_d - want_dirty).max() / np.abs(want_dirty).max()),
            "predict_rel_l2_vs_fp64": rel_l2(pred, res["fp64"][0]),
            "dirty_rel_l2_vs_fp64": rel_l2(img, res["fp64"][1]),
        }
    print("precision study: " + json.dumps(table))
    out = os.environ.get("DDF_RESULTS_DIR", "")
    if out and os.path.isdir(out):
        with open(os.path.join(out, "precision_cfg3.json"), "w") as fh:
            json.dump(table, fh, indent=1)
    # the gridding approximation (S = 7 w-kernels up to the longest baselines, sources at the facet
    # boundaries) bounds the agreement with the DFT; the bar only catches gross errors
    for v in VARIANTS:
        assert table[v]["predict_rel_l2_vs_dft"] < 0.1, (v, table[v])
        assert table[v]["dirty_max_vs_dft"] < 2e-2, (v, table[v])
    assert table["mixed"]["predict_rel_l2_vs_fp64"] < 1e-5
    assert table["mixed"]["dirty_rel_l2_vs_fp64"] < 1e-5
    assert table["fp32"]["predict_rel_l2_vs_fp64"] < 1e-3
    assert table["fp32"]["dirty_rel_l2_vs_fp64"] < 1e-3


def test_precision_variants_on_channel_blocks_vs_fp64():
    """Blocks of one row x 64 channels: every lane walks 64 channels, so the mixed variant
    advances its phasor in FP32 over up to 15 channels between anchors.  Against the FP64 variant
    (same blocks, so the averaging is identical) the arithmetic difference is isolated."""
    res = {v: _run(v, chan_block=64) for v in VARIANTS}
    table = {}
    for v in VARIANTS:
        pred, img = res[v]
        table[v] = {"predict_rel_l2_vs_fp64": rel_l2(pred, res["fp64"][0]),
                    "dirty_rel_l2_vs_fp64": rel_l2(img, res["fp64"][1])}
    print("precision on 64-channel blocks: " + json.dumps(table))
    out = os.environ.get("DDF_RESULTS_DIR", "")
    if out and os.path.isdir(out):
        with open(os.path.join(out, "precision_cfg3_blocks64.json"), "w") as fh:
            json.dump(table, fh, indent=1)
    assert table["mixed"]["predict_rel_l2_vs_fp64"] < 1e-5
    assert table["mixed"]["dirty_rel_l2_vs_fp64"] < 1e-5
    assert table["fp32"]["predict_rel_l2_vs_fp64"] < 1e-3
    assert table["fp32"]["dirty_rel_l2_vs_fp64"] < 1e-3
