"""Chunk staging on the device: what ClassMS does to a chunk between the read and the gridders,
and the upload of the next chunk overlapped with the major cycle on the current one.

* ``sort_baseline_time`` / ``apply_sort``: the baseline-time order of ClassMS.ReadData
  (DDFacet/Data/ClassMS.py:575-586), ``sorted(zip(A0, A1, TIME, range(nrow)))``, as an index and
  as a row gather of every per-row array (ddf_sort_baseline_time, ddf_gather_rows).
* ``update_flags``: ClassMS.UpdateFlags (ClassMS.py:1155-1233) on the device: the flag
  preparation of ``flagprep.prepare_flags_device`` (autocorrelations, NaNs, 1e9 fill,
  equalisation over correlations), then the uv-range and time-range cuts and the flagging of
  antennas with more than 90 % of their data flagged or named by the caller
  (ddf_flag_selection).
* ``ChunkPipeline``: the reference reads chunk n + 1 while the imager works on chunk n
  (ClassDeconvMachine.py:1320-1337, ClassVisServer.startChunkLoadInBackground); here the chunk's
  host arrays (pinned) go to a second set of device buffers on a copy stream while the compute
  stream runs the cycle on the first, with events ordering the two.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import check, lib
from .device import require_gpu, stream_handle
from .flagprep import prepare_flags_device

# MJD seconds -> unix seconds (40587 days), for TimeRange selections on MS TIME values
MJD_UNIX_OFFSET_S = 3506716800.0


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def sort_baseline_time(a0, a1, times, stream=None):
    """int32 index of the rows in baseline-time order (device tensors a0 / a1 int32, times f64)."""
    n = int(a0.numel())
    for t, dt in ((a0, torch.int32), (a1, torch.int32), (times, torch.float64)):
        if t.dtype != dt or t.device.type != "cuda" or not t.is_contiguous() or t.numel() != n:
            raise TypeError("a0 / a1 int32 and times float64, [nrow], contiguous on the GPU")
    index = torch.empty(n, dtype=torch.int32, device=a0.device)
    check(lib().ddf_sort_baseline_time(_p(a0), _p(a1), _p(times), n, _p(index), C.c_void_p(stream_handle(stream))))
    return index


def gather_rows(src, index, stream=None):
    """src[index] along the first axis (a new tensor), by a device gather of whole rows."""
    if src.device.type != "cuda" or not src.is_contiguous():
        raise TypeError("src must be contiguous on the GPU")
    if index.dtype != torch.int32 or index.numel() != src.shape[0]:
        raise TypeError("index must be int32 [nrow]")
    dst = torch.empty_like(src)
    row_bytes = src[0].numel() * src.element_size() if src.shape[0] else 1
    check(lib().ddf_gather_rows(_p(src), _p(dst), int(src.shape[0]), int(row_bytes), _p(index),
                                C.c_void_p(stream_handle(stream))))
    return dst


def apply_sort(index, arrays, stream=None):
    """Every tensor of the dict `arrays` (first axis = rows) in the order of `index`."""
    return {k: gather_rows(v, index, stream) for k, v in arrays.items()}


def update_flags(flags, uvw, data, a0, a1, times=None, uv_range_km=None, time_range_unix=None, named_antennas=None,
                 threshold=0.9, na=None, stream=None):
    """ClassMS.UpdateFlags on device tensors, in place.  flags uint8 / bool [nrow, nchan, ncorr],
    data complex64 of the same shape or None, uvw f64 [nrow, 3], a0 / a1 int32, times f64 (MS TIME,
    MJD seconds) when a time range is given.  Returns (per-antenna flagged fraction f64 [na],
    flagged-antenna mask bool [na]) on the device."""
    flags = prepare_flags_device(flags, a0, a1, data, stream)
    nrow, nchan, ncorr = flags.shape
    if na is None:
        na = int(max(int(a0.max()), int(a1.max()))) + 1 if nrow else 1
    s = abi.ddf_flag_select()
    s.na = int(na)
    if uv_range_km is not None:
        d0, d1 = uv_range_km
        s.use_uv_range = 1
        s.uv_min_m2 = float(d0 ** 2 * 1e6)
        s.uv_max_m2 = float(d1 ** 2 * 1e6)
    if time_range_unix is not None:
        if times is None:
            raise ValueError("a time range needs the row times")
        s.use_time_range = 1
        s.time_min, s.time_max = float(time_range_unix[0]), float(time_range_unix[1])
        s.time_offset = MJD_UNIX_OFFSET_S
    s.ant_threshold = float(threshold)
    named = None
    if named_antennas is not None:
        named = np.zeros(int(na), np.uint8)
        named[np.asarray(list(named_antennas), np.int64)] = 1
        s.named_antennas = named.ctypes.data_as(C.POINTER(C.c_uint8))
    frac = torch.empty(int(na), dtype=torch.float64, device=flags.device)
    flagged = torch.empty(int(na), dtype=torch.uint8, device=flags.device)
    check(lib().ddf_flag_selection(_p(flags), _p(uvw), _p(times), _p(a0), _p(a1), nrow, nchan, ncorr, C.byref(s),
                                   _p(frac), _p(flagged), C.c_void_p(stream_handle(stream))))
    if named is not None:
        torch.cuda.current_stream(flags.device).synchronize() if stream is None else stream.synchronize()
    return frac, flagged.bool()


class ChunkPipeline:
    """Double-buffered chunk uploads overlapped with the major cycle.

    ``submit(i, host)`` copies chunk i's host arrays (a dict of pinned CPU tensors, or numpy
    arrays, which are pinned first) into device slot i % 2 on the copy stream and records an
    event; ``wait(i)`` makes the compute stream wait for it and returns that slot's device tensors.
    Before slot s is written again, the copy stream waits for the compute work recorded with
    ``release(i)`` on chunk i - 2's slot.  Streams have the default priority."""

    def __init__(self, device=None):
        self.dev = device or require_gpu()
        self.copy = torch.cuda.Stream(device=self.dev)
        self.slots = [None, None]
        self.ready = [None, None]
        self.free = [None, None]

    @staticmethod
    def pinned(host):
        out = {}
        for k, v in host.items():
            t = v if isinstance(v, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(v))
            if t.dtype == torch.bool:
                t = t.view(torch.uint8)
            out[k] = t if t.is_pinned() else t.pin_memory()
        return out

    def submit(self, i, host):
        s = i % 2
        if self.slots[s] is None or any(self.slots[s][k].shape != host[k].shape for k in host):
            self.slots[s] = {k: torch.empty(tuple(v.shape), dtype=v.dtype, device=self.dev) for k, v in host.items()}
        with torch.cuda.stream(self.copy):
            if self.free[s] is not None:
                self.copy.wait_event(self.free[s])
            for k, v in host.items():
                self.slots[s][k].copy_(v, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.copy)
        self.ready[s] = ev
        return ev

    def wait(self, i, stream=None):
        s = i % 2
        st = stream or torch.cuda.current_stream(self.dev)
        st.wait_event(self.ready[s])
        for t in self.slots[s].values():
            t.record_stream(st)
        return self.slots[s]

    def release(self, i, stream=None):
        s = i % 2
        st = stream or torch.cuda.current_stream(self.dev)
        ev = torch.cuda.Event()
        ev.record(st)
        self.free[s] = ev
