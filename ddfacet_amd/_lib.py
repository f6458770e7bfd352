"""Loader of libddfgpu.so (the HIP library built from ddfacet_amd/csrc).

Fails loudly: there is no CPU fallback on the product path.  ``lib()`` raises if the library is
missing or was built for another ABI version.
"""
import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libddfgpu.so")
_LIB = None


class DDFError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise DDFError("libddfgpu.so is not built (run __graft_entry__.build() or make -C ddfacet_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    L.ddf_version.restype = C.c_int
    L.ddf_last_error.restype = C.c_char_p
    L.ddf_init.argtypes = [C.c_int]
    L.ddf_init.restype = C.c_int
    L.ddf_bda_index.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]
    L.ddf_bda_index.restype = C.c_int
    for name in ("ddf_grid", "ddf_degrid"):
        fn = getattr(L, name)
        fn.argtypes = [C.POINTER(abi.ddf_grid_desc), C.POINTER(abi.ddf_vis), C.POINTER(abi.ddf_facet), C.c_void_p]
        fn.restype = C.c_int
    L.ddf_accumulate_weights.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    L.ddf_accumulate_weights.restype = C.c_int
    _bind_optional(L)
    L.ddf_weights_briggs_work_size.argtypes = [C.c_int]
    L.ddf_weights_briggs_work_size.restype = C.c_int64
    if L.ddf_version() != abi.DDF_ABI_VERSION:
        raise DDFError("libddfgpu.so ABI %d, expected %d" % (L.ddf_version(), abi.DDF_ABI_VERSION))
    _LIB = L
    return L


def _bind_optional(L):
    """Entry points added beside the plain ones (batched major cycle etc.)."""
    for name, args in getattr(abi, "EXTRA_SIGNATURES", {}).items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int


def check(rc):
    if rc == abi.DDF_OK:
        return
    msg = lib().ddf_last_error().decode(errors="replace")
    if rc == abi.DDF_EINVAL:
        raise ValueError(msg)
    raise DDFError(msg)
