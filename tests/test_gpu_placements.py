"""The batched plan's integer placement maps, bit for bit.

For every parity case and facet, the plan's per-(block, facet) placement records
(ddf_plan_placements: w-plane, conjugate choice, grid cell, oversampling sub-kernel, band, and
which blocks are dropped) of the grid and degrid maps equal the maps the reference gridder /
degridder reveal through delta stamps (tests/golden/delta_stamps.npz) and the oracle's
(cpuref.place), in the mixed and the FP64 precision plans.
"""
import os

import numpy as np
import pytest

from cases import CASES, problem

pytestmark = pytest.mark.gpu


def _stamps():
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return np.load(os.path.join(here, "delta_stamps.npz"), allow_pickle=False)


@pytest.mark.parametrize("name", [c[0] for c in CASES])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_plan_placements_equal_reference_delta_stamps(name, precision):
    from ddfacet_amd.facets import DeviceFacetSet
    from oracle import cpuref
    print("case", name, precision, flush=True)
    P = problem(name)
    D = _stamps()
    fs = DeviceFacetSet(P)
    fs.plan_batched(precision=precision, tile=16)
    got_g = fs.placements(degrid=False)
    got_d = fs.placements(degrid=True)
    fs.destroy_plan()
    np.testing.assert_array_equal(got_g, D[name + "_grid"])
    np.testing.assert_array_equal(got_d, D[name + "_degrid"])
    g = P.new_grid()
    for fi in range(len(P.facets)):
        np.testing.assert_array_equal(got_g[fi], cpuref.place(*P.grid_args(fi, g, np.zeros((P.nband, P.npol)))))
        np.testing.assert_array_equal(got_d[fi], cpuref.place(*P.degrid_args(fi, g, P.vis.copy()), degrid=True))
