"""Parity cases shared by the oracle tests, the golden generator and the GPU tests.

Each case is (name, kwargs of ddfacet_amd.problems.random_problem).  Together they cover the
reference code paths: every correlation / Stokes dispatch family, Jones modes 1 and 2 with DD
and beam tables, interpolation mode 1, decorrelation, several bands and polarisations,
sparsification, non-equidistant channels, multi-row / multi-channel BDA blocks and one block
per visibility, several facets (non-zero l0, m0, Cu, Cv), 1/2/4 correlations, supports
of 17-23 (stamps of more than 256 cells), and blocks dropped off the grid or beyond the last
w-plane.
"""
CASES = [
    ("base", dict(seed=11)),
    ("pervis", dict(seed=12, bda="pervis")),
    ("jones1", dict(seed=13, jones=dict(mode=1))),
    ("jones2_beam", dict(seed=14, jones=dict(mode=2, beam=True))),
    ("jones1_beam_only", dict(seed=15, jones=dict(mode=1, kms=False, beam=True))),
    ("jones1_interp1", dict(seed=16, nfacet=4, jones=dict(mode=1, interp=1))),
    ("jones1_noamp_nophase", dict(seed=17, jones=dict(mode=1, apply_amp=0, apply_phase=0))),
    ("decorr_bands_pols", dict(seed=18, decorr=True, nband=2, npol=2, stokes_codes=(1, 2))),
    ("iquv", dict(seed=19, npol=4, stokes_codes=(1, 2, 3, 4))),
    ("circ_iquv", dict(seed=20, npol=4, corr_codes=(5, 6, 7, 8), stokes_codes=(1, 2, 3, 4))),
    ("xx_yy", dict(seed=21, ncorr=2, corr_codes=(9, 12), npol=2, stokes_codes=(1, 2))),
    ("rr_ll", dict(seed=22, ncorr=2, corr_codes=(5, 8), npol=2, stokes_codes=(1, 4))),
    ("stokes_i_in", dict(seed=23, ncorr=1, corr_codes=(1,), stokes_codes=(1,))),
    ("sparsify", dict(seed=24, sparsify=True)),
    ("nonequi", dict(seed=25, equidistant=False)),
    ("multifacet", dict(seed=26, nfacet=4, nw=4, sups=(5, 5, 7, 9), overs=7)),
    ("overs11", dict(seed=27, nw=2, sups=(7, 9), overs=11, npix=64)),
    ("jones2_multiband", dict(seed=28, nband=2, jones=dict(mode=2), nchan=12)),
    ("iu_quirk", dict(seed=29, npol=2, stokes_codes=(1, 3))),
    ("qu", dict(seed=30, npol=2, stokes_codes=(2, 3))),
    ("wide_support", dict(seed=31, nant=6, nfacet=2, nw=4, sups=(17, 19, 21, 23), overs=5, npix=96,
                          jones=dict(mode=1))),
    ("drops", dict(seed=32, nant=7, nfacet=2, uv_scale=1.6, wmax_scale=0.6, jones=dict(mode=1))),
]


def problem(name):
    from ddfacet_amd import problems
    kw = dict(dict(CASES)[name])
    seed = kw.pop("seed")
    return problems.random_problem(seed, **kw)


def model_grid(P, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    g = P.new_grid()
    return (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)).astype(np.complex64)
