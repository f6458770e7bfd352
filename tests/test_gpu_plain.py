"""HIP plain kernels (one facet per launch) through the drop-in shim, against the golden vectors
of the reference gridder and against the CPU oracle.

Tolerances: relative L2 <= 1e-5 on grids and on degridded model visibilities (the parity bar
for floating point); sumwt / SumJones are FP64 sums compared at 1e-10.  The set of grid cells a case
touches must match the reference exactly: it is fixed by the integer placement maps.
"""
import os

import numpy as np
import pytest

from cases import CASES, model_grid, problem
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _golden(name):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return np.load(os.path.join(here, name + ".npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def shim():
    from ddfacet_amd import _pyGridderSmearPols as m
    return m


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_grid_matches_reference(shim, name):
    print("case", name, flush=True)
    P = problem(name)
    G = _golden(name)
    for fi in range(len(P.facets)):
        for psf in (0, 1):
            g = P.new_grid()
            s = np.zeros((P.nband, P.npol))
            args = P.grid_args(fi, g, s, dopsf=psf)
            shim.pyGridderWPol32(*args)
            key = "f%d_psf%d" % (fi, psf)
            ref = G["grid_" + key]
            assert rel_l2(g, ref) <= TOL, (key, rel_l2(g, ref))
            assert np.array_equal(g != 0, ref != 0), key
            np.testing.assert_allclose(s, G["sumwt_" + key], rtol=1e-10)
            if args[13]:
                np.testing.assert_allclose(args[13][17], G["sumjones_" + key], rtol=1e-10, atol=1e-300)
                np.testing.assert_allclose(args[13][18], G["sumjoneschan_" + key], rtol=1e-10, atol=1e-300)


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_degrid_matches_reference(shim, name):
    print("case", name, flush=True)
    P = problem(name)
    G = _golden(name)
    for fi in range(len(P.facets)):
        v = P.vis.copy()
        out = shim.pyDeGridderWPol(*P.degrid_args(fi, model_grid(P, 100 + fi), v))
        assert out is v
        model_hip = P.vis - v
        model_ref = P.vis - G["vis_f%d" % fi]
        assert np.abs(model_ref).max() > 0
        assert rel_l2(model_hip, model_ref) <= TOL, rel_l2(model_hip, model_ref)


def test_grid64_matches_oracle(shim):
    from oracle import cpuref
    P = problem("jones1")
    g1 = P.new_grid(np.complex128)
    s1 = np.zeros((P.nband, P.npol))
    shim.pyGridderWPol64(*P.grid_args(0, g1, s1))
    g2 = P.new_grid(np.complex128)
    s2 = np.zeros((P.nband, P.npol))
    cpuref.grid(*P.grid_args(0, g2, s2), f64=True)
    assert rel_l2(g1, g2) <= 1e-6
    np.testing.assert_allclose(s1, s2, rtol=1e-10)


def test_grid_accumulates_in_place(shim):
    P = problem("base")
    g = P.new_grid()
    s = np.zeros((P.nband, P.npol))
    shim.pyGridderWPol32(*P.grid_args(0, g, s))
    g1, s1 = g.copy(), s.copy()
    shim.pyGridderWPol32(*P.grid_args(0, g, s))
    assert rel_l2(g, 2 * g1) <= 1e-6
    np.testing.assert_allclose(s, 2 * s1, rtol=1e-12)


def test_accumulate_weights(shim):
    rng = np.random.default_rng(5)
    grid = np.zeros(1000)
    w = rng.random(5000).astype(np.float32)
    w[rng.random(5000) < 0.2] = 0
    idx = rng.integers(0, 1000, 5000)
    shim.pyAccumulateWeightsOntoGrid(grid, w, idx)
    ref = np.zeros(1000)
    np.add.at(ref, idx[w != 0], w[w != 0].astype(np.float64))
    np.testing.assert_allclose(grid, ref, rtol=1e-12)
