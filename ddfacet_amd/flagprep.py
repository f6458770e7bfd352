"""Flag preparation that the gridders rely on (DDFacet/Data/ClassMS.py:1155-1167).

The gridder tests the flag of correlation 0 only (gridder.h:229); that is valid because the chunk
reader flags autocorrelations, flags NaN visibilities (and overwrites every flagged value with
1e9), and then flags all correlations of a (row, channel) when any of them is flagged.

``prepare_flags_device`` runs these steps on a chunk resident in HBM (``ddf_prepare_flags``): a
thread per (row, channel) flags autocorrelations and NaNs, fills flagged values with 1e9 and
equalises the flags over the correlations, in place.
"""
import ctypes as C


def prepare_flags_device(flags, a0, a1, data=None, stream=None):
    """The flag preparation on device tensors, in place: flags uint8 or bool [nrow, nchan, ncorr] (CUDA),
    data complex64 of the same shape or None, a0 / a1 int32 [nrow] (CUDA)."""
    import torch
    from ._lib import check, lib
    from .device import stream_handle
    if flags.device.type != "cuda" or flags.dim() != 3 or not flags.is_contiguous():
        raise TypeError("flags must be a contiguous [nrow, nchan, ncorr] tensor on the GPU")
    if flags.dtype == torch.bool:
        flags = flags.view(torch.uint8)
    if flags.dtype != torch.uint8:
        raise TypeError("flags must be uint8 or bool")
    nrow, nchan, ncorr = flags.shape
    if data is not None and (data.dtype != torch.complex64 or tuple(data.shape) != (nrow, nchan, ncorr)
                             or not data.is_contiguous() or data.device != flags.device):
        raise TypeError("data must be a contiguous complex64 tensor of the flags' shape on the same GPU")
    for t in (a0, a1):
        if t.dtype != torch.int32 or t.numel() != nrow or t.device != flags.device:
            raise TypeError("a0 / a1 must be int32 [nrow] on the GPU")
    check(lib().ddf_prepare_flags(C.c_void_p(flags.data_ptr()),
                                  C.c_void_p(data.data_ptr()) if data is not None else None,
                                  C.c_void_p(a0.data_ptr()), C.c_void_p(a1.data_ptr()), nrow, nchan, ncorr,
                                  C.c_void_p(stream_handle(stream))))
    return flags
