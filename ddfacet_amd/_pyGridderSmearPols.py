"""Drop-in replacement of DDFacet.cbuild.Gridder._pyGridderSmearPols3x on the GPU.

Exports the five names of the reference module (DDFacet/Gridder/GridderSmearPols.cc:349-370)
with the same positional argument lists, so the call sites in ClassDDEGridMachine.put / .get
(DDFacet/Imager/ClassDDEGridMachine.py:790-818, 1084-1108), ClassFacetMachine.py:52-55 and
ClassWeighting.py:374 work unchanged:

    pyGridderWPol32 / pyGridderWPol64   grid (accumulated), sumwt, SumJones, SumJonesChan in place
    pyDeGridderWPol                     vis -= model in place, returns vis
    pyAccumulateWeightsOntoGrid[NoSem]  grid[index] += weights (weights != 0)
    pySetSemaphores / pyDeleteSemaphore no-ops: the device reduction needs no row locks

numpy arguments are staged to the GPU for the call and the outputs copied back; torch GPU
tensors are used in place.  Each call is synchronous, like the reference.
"""
import ctypes as C

import numpy as np
import torch

from . import marshal
from ._lib import check, lib
from .device import Staging, stream_handle


def _outputs_back(st, spec, names):
    for name, arr in names:
        if arr is not None and isinstance(arr, np.ndarray):
            st.download(name, arr)


def _grid(args, f64):
    spec = marshal.parse_grid_args(args, f64)
    st = Staging()
    g, v, f = marshal.fill_structs(spec, st.addr)
    check(lib().ddf_grid(C.byref(g), C.byref(v), C.byref(f), C.c_void_p(stream_handle())))
    torch.cuda.synchronize()
    outs = [("grid", spec.grid), ("sumwt", spec.sumwt)]
    if spec.jones is not None:
        outs += [("sum_jones", spec.jones.sum_jones), ("sum_jones_chan", spec.jones.sum_jones_chan)]
    _outputs_back(st, spec, outs)


def pyGridderWPol32(*args):
    _grid(args, False)


def pyGridderWPol64(*args):
    _grid(args, True)


def pyDeGridderWPol(*args):
    spec = marshal.parse_degrid_args(args)
    st = Staging()
    g, v, f = marshal.fill_structs(spec, st.addr)
    check(lib().ddf_degrid(C.byref(g), C.byref(v), C.byref(f), C.c_void_p(stream_handle())))
    torch.cuda.synchronize()
    st.download("vis", spec.vis)
    return spec.vis


def _accumulate(grid, weights, index):
    if not isinstance(grid, np.ndarray) or grid.dtype != np.float64 or not grid.flags.c_contiguous:
        raise TypeError("grid must be a C-contiguous float64 array")
    w = np.ascontiguousarray(weights, dtype=np.float32).ravel()
    idx = np.ascontiguousarray(index, dtype=np.int64).ravel()
    if idx.size < w.size:
        raise ValueError("index must have one entry per weight")
    sel = w != 0
    if sel.any() and (idx[sel].min() < 0 or idx[sel].max() >= grid.size):
        raise ValueError("weight index outside the grid")
    st = Staging()
    ga = st.addr("grid", grid)
    check(lib().ddf_accumulate_weights(C.c_void_p(ga), grid.size, C.c_void_p(st.addr("w", w)),
                                       C.c_void_p(st.addr("idx", idx)), w.size, C.c_void_p(stream_handle())))
    torch.cuda.synchronize()
    st.download("grid", grid)


def pyAccumulateWeightsOntoGrid(grid, weights, index):
    _accumulate(grid, weights, index)


def pyAccumulateWeightsOntoGridNoSem(grid, weights, index):
    _accumulate(grid, weights, index)


def pySetSemaphores(names):
    """Accepted for interface compatibility (Semaphores.cc:41-49); nothing to lock on the GPU."""
    return None


def pyDeleteSemaphore():
    return None
