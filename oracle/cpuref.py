"""TEST INFRASTRUCTURE: ctypes driver of the C++ oracle (oracle/cpu_ref.cc).

``grid(*args, f64=False)`` and ``degrid(*args)`` take the reference positional argument lists of
pyGridderWPol32|64 / pyDeGridderWPol (DDFacet/Gridder/GridderSmearPols.cc:76-97, 212-233) and
update the numpy outputs in place, exactly like the reference module.  ``place(*args, degrid)``
returns the per-block integer placement maps of the same call.
"""
import ctypes as C
import os

import numpy as np

from ddfacet_amd import abi, marshal

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libddfcpuref.so")
        if not os.path.exists(path):
            raise RuntimeError("CPU oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        for name in ("ref_grid", "ref_degrid"):
            fn = getattr(L, name)
            fn.argtypes = [C.POINTER(abi.ddf_grid_desc), C.POINTER(abi.ddf_vis), C.POINTER(abi.ddf_facet)]
            fn.restype = C.c_int
        for name in ("ref_grid_facets", "ref_degrid_facets"):
            fn = getattr(L, name)
            fn.argtypes = [C.POINTER(abi.ddf_grid_desc), C.POINTER(abi.ddf_vis), C.POINTER(abi.ddf_facet),
                           C.c_int, C.c_int]
            fn.restype = C.c_int
        L.ref_place.argtypes = [C.POINTER(abi.ddf_grid_desc), C.POINTER(abi.ddf_vis), C.POINTER(abi.ddf_facet),
                                C.c_int, C.c_void_p]
        L.ref_place.restype = C.c_int
        L.ref_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def host_addr(_name, a):
    return 0 if a is None else a.ctypes.data


def _check(rc):
    if rc == abi.DDF_EINVAL:
        raise ValueError(lib().ref_last_error().decode())
    if rc != abi.DDF_OK:
        raise RuntimeError(lib().ref_last_error().decode())


def grid(*args, f64=False):
    spec = marshal.parse_grid_args(args, f64)
    g, v, f = marshal.fill_structs(spec, host_addr)
    _check(lib().ref_grid(C.byref(g), C.byref(v), C.byref(f)))


def degrid(*args):
    spec = marshal.parse_degrid_args(args)
    g, v, f = marshal.fill_structs(spec, host_addr)
    _check(lib().ref_degrid(C.byref(g), C.byref(v), C.byref(f)))
    return spec.vis


def place(*args, degrid=False):
    """Integer placement maps of every BDA block for the facet of a pyGridderWPol (degrid=False)
    or pyDeGridderWPol (degrid=True) argument list: int32 [nblocks, 8] = valid, iw, conj
    (W <= 0), locx, locy, io, jo (oversampling sub-kernel), band (gridder.h:321-351,
    degridder.h:170-200)."""
    spec = marshal.parse_degrid_args(args) if degrid else marshal.parse_grid_args(args, False)
    g, v, f = marshal.fill_structs(spec, host_addr)
    out = np.zeros((max(int(v.nblocks), 1), 8), np.int32)
    _check(lib().ref_place(C.byref(g), C.byref(v), C.byref(f), int(bool(degrid)), out.ctypes.data))
    return out[:int(v.nblocks)]


def grid_facets(specs, nthreads):
    """Grid several facets (CallSpecs sharing one visibility set) on `nthreads` threads."""
    structs = [marshal.fill_structs(s, host_addr) for s in specs]
    fa = (abi.ddf_facet * len(specs))(*[st[2] for st in structs])
    g, v, _ = structs[0]
    _check(lib().ref_grid_facets(C.byref(g), C.byref(v), fa, len(specs), nthreads))


def degrid_facets(specs, nthreads):
    structs = [marshal.fill_structs(s, host_addr) for s in specs]
    fa = (abi.ddf_facet * len(specs))(*[st[2] for st in structs])
    g, v, _ = structs[0]
    _check(lib().ref_degrid_facets(C.byref(g), C.byref(v), fa, len(specs), nthreads))
