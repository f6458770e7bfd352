"""Writes the integer placement maps of the reference gridder and degridder, revealed by delta
stamps (tests/delta.py), for every parity case of tests/cases.py and every facet, into
tests/golden/delta_stamps.npz: "<case>_grid" and "<case>_degrid", int32 [nfacets, nblocks, 8]
rows [valid, iw, conj, locx, locy, io, jo, band].

The reference module is the one compiled from its own sources (make -f oracle/Makefile.ref).

    python tests/golden/make_delta_stamps.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from cases import CASES, problem  # noqa: E402
from delta import degrid_stamps, delta_problem, grid_stamps  # noqa: E402


def main():
    from oracle import refmod
    mod = refmod.module()
    out = {}
    for name, _ in CASES:
        P = problem(name)
        D = delta_problem(P)
        g = [grid_stamps(mod, D, fi) for fi in range(len(D.facets))]
        d = [degrid_stamps(mod, D, fi) for fi in range(len(D.facets))]
        out[name + "_grid"] = np.stack(g)
        out[name + "_degrid"] = np.stack(d)
        print(name, "grid valid %d / %d, degrid valid %d / %d"
              % (out[name + "_grid"][..., 0].sum(), out[name + "_grid"][..., 0].size,
                 out[name + "_degrid"][..., 0].sum(), out[name + "_degrid"][..., 0].size))
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "delta_stamps.npz"), **out)


if __name__ == "__main__":
    main()
