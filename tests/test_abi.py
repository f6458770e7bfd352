"""The C ABI: the library loads, exports every entry point include/ddfgpu.h declares, the ctypes
mirror (ddfacet_amd/abi.py) has the C layout, and argument errors come back as error codes
before any device work (so these run without a GPU)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ddfgpu.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ddf_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_entry_point():
    from ddfacet_amd import _lib
    L = _lib.lib()
    names = _declared()
    assert "ddf_grid" in names and "ddf_plan_create" in names and "ddf_stitch" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.ddf_version() == 1


def test_oracle_library_loads():
    from oracle import cpuref
    L = cpuref.lib()
    for n in ("ref_grid", "ref_degrid", "ref_grid_facets", "ref_degrid_facets"):
        assert hasattr(L, n)


STRUCTS = ["ddf_cf", "ddf_jones", "ddf_decorr", "ddf_vis", "ddf_facet", "ddf_grid_desc", "ddf_plan_opts",
           "ddf_image_facet", "ddf_cf_spec", "ddf_flag_select", "ddf_bda_spec", "ddf_beam_desc",
           "ddf_jones_merge_desc"]


def test_ctypes_layout_matches_c_header():
    from ddfacet_amd import abi
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){"]
    for s in STRUCTS:
        cls = getattr(abi, s)
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (s, s))
        for fname, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (s, fname, s, fname))
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-O0", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.split() for l in out if l.strip())
    for s in STRUCTS:
        cls = getattr(abi, s)
        assert int(got[s]) == C.sizeof(cls), s
        for fname, _ in cls._fields_:
            assert int(got["%s.%s" % (s, fname)]) == getattr(cls, fname).offset, (s, fname)


def _structs_for(corr, stokes):
    from ddfacet_amd import abi
    g = abi.ddf_grid_desc()
    g.nband, g.npol, g.nx, g.ny = 1, len(stokes), 16, 16
    v = abi.ddf_vis()
    v.nrow, v.nchan, v.ncorr = 1, 1, len(corr)
    dummy = np.zeros(16, np.int64)
    for name in ("vis", "uvw", "flags", "weights", "freqs", "chan_map", "bda"):
        setattr(v, name, dummy.ctypes.data)
    v.bda_len, v.nblocks = 2, 0
    v.ncorr_codes = len(corr)
    for i, c in enumerate(corr):
        v.corr_codes[i] = c
    v.nstokes = len(stokes)
    for i, c in enumerate(stokes):
        v.stokes_codes[i] = c
    f = abi.ddf_facet()
    f.grid = f.sumwt = dummy.ctypes.data
    f.cf.data = f.cf.offset = f.cf.support = dummy.ctypes.data
    f.cf.nw, f.cf.overs, f.cf.max_support = 1, 1, 1
    return g, v, f, dummy


# (corr codes, stokes codes, accepted by the gridder, accepted by the degridder):
# GridderSmearPols.cc:128-206 and :261-342
DISPATCH = [
    ((9, 10, 11, 12), (1,), True, True),
    ((9, 10, 11, 12), (1, 3), True, True),     # {I,U}: the gridder runs the IQ policy
    ((9, 10, 11, 12), (2, 3), True, True),
    ((9, 10, 11, 12), (3, 4), False, False),
    ((5, 6, 7, 8), (2,), False, False),
    ((5, 6, 7, 8), (1, 4), True, True),
    ((9, 12), (1, 2), True, True),
    ((9, 12), (1, 4), False, False),
    ((5, 8), (1, 4), True, True),
    ((1,), (1,), True, True),
    ((1,), (1, 2), False, False),
    ((9, 10, 11, 12), (2,), True, True),
    ((5, 6, 7, 8), (1, 2, 3), True, True),
    ((13, 14, 15, 16), (1,), False, False),    # unknown correlation codes
]


@pytest.mark.parametrize("corr,stokes,ok_grid,ok_degrid", DISPATCH)
def test_dispatch_table_through_the_c_abi(corr, stokes, ok_grid, ok_degrid):
    """Supported combinations pass validation (and return OK with no blocks to process);
    unsupported ones return DDF_EINVAL with the reference message, before any device call."""
    from ddfacet_amd import _lib, abi
    L = _lib.lib()
    g, v, f, _keep = _structs_for(corr, stokes)
    rc = L.ddf_grid(C.byref(g), C.byref(v), C.byref(f), None)
    assert (rc == abi.DDF_OK) == ok_grid, (rc, L.ddf_last_error())
    if not ok_grid:
        assert rc == abi.DDF_EINVAL
        assert b"Only accepts" in L.ddf_last_error() or b"Cannot convert" in L.ddf_last_error()
    rc = L.ddf_degrid(C.byref(g), C.byref(v), C.byref(f), None)
    assert (rc == abi.DDF_OK) == ok_degrid, (rc, L.ddf_last_error())


def test_jones_without_tables_is_a_runtime_error():
    from ddfacet_amd import _lib, abi
    L = _lib.lib()
    g, v, f, _keep = _structs_for((9, 10, 11, 12), (1,))
    f.jones.enabled = 1
    f.jones.mode = 1
    assert L.ddf_grid(C.byref(g), C.byref(v), C.byref(f), None) == abi.DDF_ERUNTIME
    assert b"neither E or DD Jones" in L.ddf_last_error()


def test_shim_rejects_wrong_output_types():
    """Outputs with the wrong dtype are rejected instead of silently copied (GridderSmearPols.cc
    takes array_t<T, c_style>, which would copy and drop the in-place result)."""
    from cases import problem
    from ddfacet_amd import marshal
    P = problem("base")
    g = P.new_grid().astype(np.complex128)
    with pytest.raises(TypeError):
        marshal.parse_grid_args(P.grid_args(0, g, np.zeros((1, 1))), False)
    g = P.new_grid()
    with pytest.raises(TypeError):
        marshal.parse_grid_args(P.grid_args(0, g, np.zeros((1, 1), np.float32)), False)
    v = P.vis.astype(np.complex128)
    with pytest.raises(TypeError):
        marshal.parse_degrid_args(P.degrid_args(0, P.new_grid(), v))
    args = list(P.grid_args(0, P.new_grid(), np.zeros((1, 1))))
    args[15] = np.ones(3, bool)  # sparsification of the wrong length
    with pytest.raises(ValueError):
        marshal.parse_grid_args(tuple(args), False)
