"""Writes the golden vectors of the gridder / degridder with the reference module.

The reference native gridder is compiled from its own sources (make -f oracle/Makefile.ref ->
oracle/_ref/, git-ignored) and driven here with the argument lists of tests/cases.py, built as
ClassDDEGridMachine.put / .get build them.  For every case and facet the outputs of
pyGridderWPol32 (image and PSF mode: grid, sumwt, SumJones, SumJonesChan) and of
pyDeGridderWPol (vis after the subtraction of a seeded random model grid) are stored with a
digest of the inputs, in tests/golden/<case>.npz.  The inputs themselves are regenerated from
the seeds by the tests.

    python tests/golden/make_golden.py [case ...]      (default: every case)
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from cases import CASES, model_grid, problem  # noqa: E402


def digest(P):
    h = hashlib.sha256()
    for a in (P.uvw, P.vis, P.flags, P.weights, P.freqs, P.bda_grid, P.chan_map):
        h.update(np.ascontiguousarray(a).tobytes())
    for f in P.facets:
        for p in f.wplanes + f.wplanes_conj:
            h.update(np.ascontiguousarray(p).tobytes())
    if P.jones is not None:
        for k in sorted(P.jones):
            v = P.jones[k]
            if isinstance(v, np.ndarray):
                h.update(v.tobytes())
    return h.hexdigest()


def run_case(mod, name):
    P = problem(name)
    out = {"digest": np.array(digest(P))}
    for fi in range(len(P.facets)):
        for psf in (0, 1):
            g = P.new_grid()
            s = np.zeros((P.nband, P.npol))
            args = P.grid_args(fi, g, s, dopsf=psf)
            mod.pyGridderWPol32(*args)
            out["grid_f%d_psf%d" % (fi, psf)] = g
            out["sumwt_f%d_psf%d" % (fi, psf)] = s
            if args[13]:
                out["sumjones_f%d_psf%d" % (fi, psf)] = args[13][17]
                out["sumjoneschan_f%d_psf%d" % (fi, psf)] = args[13][18]
        model = model_grid(P, 100 + fi)
        v = P.vis.copy()
        mod.pyDeGridderWPol(*P.degrid_args(fi, model, v))
        out["vis_f%d" % fi] = v
    return out


def main():
    from oracle import refmod
    mod = refmod.module()
    here = os.path.dirname(os.path.abspath(__file__))
    names = sys.argv[1:] or [c[0] for c in CASES]
    for name in names:
        out = run_case(mod, name)
        np.savez_compressed(os.path.join(here, name + ".npz"), **out)
        print("wrote", name)


if __name__ == "__main__":
    main()
