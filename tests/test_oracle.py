"""The CPU oracle (oracle/cpu_ref.cc) pinned against the reference golden vectors."""
import numpy as np
import pytest

from cases import CASES, model_grid, problem
from conftest import rel_l2
from oracle import cpuref


def _golden(name):
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return np.load(os.path.join(here, name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_oracle_matches_reference_golden(name):
    from golden.make_golden import digest
    P = problem(name)
    G = _golden(name)
    assert str(G["digest"]) == digest(P), "problem inputs drifted from the golden vectors"
    for fi in range(len(P.facets)):
        for psf in (0, 1):
            g = P.new_grid()
            s = np.zeros((P.nband, P.npol))
            args = P.grid_args(fi, g, s, dopsf=psf)
            cpuref.grid(*args)
            key = "f%d_psf%d" % (fi, psf)
            assert rel_l2(g, G["grid_" + key]) <= 1e-12
            np.testing.assert_allclose(s, G["sumwt_" + key], rtol=1e-12)
            if args[13]:
                np.testing.assert_allclose(args[13][17], G["sumjones_" + key], rtol=1e-12, atol=1e-300)
                np.testing.assert_allclose(args[13][18], G["sumjoneschan_" + key], rtol=1e-12, atol=1e-300)
        v = P.vis.copy()
        cpuref.degrid(*P.degrid_args(fi, model_grid(P, 100 + fi), v))
        assert rel_l2(P.vis - v, P.vis - G["vis_f%d" % fi]) <= 1e-12


def test_golden_cases_exercise_gridding():
    """Every case grids a non-trivial share of its visibilities (guards against vacuous parity)."""
    for name, _ in CASES:
        G = _golden(name)
        assert np.count_nonzero(G["grid_f0_psf0"]) > 50, name
        assert np.abs(G["sumwt_f0_psf0"]).sum() > 0, name


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_oracle_placement_maps_equal_reference_delta_stamps(name):
    """The oracle's integer placement maps (w-plane, conjugate choice, grid cell, oversampling
    sub-kernel, band, dropped blocks) of every block and facet, grid and degrid maps, equal the
    ones the reference gridder / degridder reveal through delta stamps, bit for bit."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    D = np.load(os.path.join(here, "delta_stamps.npz"), allow_pickle=False)
    P = problem(name)
    for fi in range(len(P.facets)):
        g = P.new_grid()
        got_g = cpuref.place(*P.grid_args(fi, g, np.zeros((P.nband, P.npol))))
        got_d = cpuref.place(*P.degrid_args(fi, g, P.vis.copy()), degrid=True)
        np.testing.assert_array_equal(got_g, D[name + "_grid"][fi])
        np.testing.assert_array_equal(got_d, D[name + "_degrid"][fi])
