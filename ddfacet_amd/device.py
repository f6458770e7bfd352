"""Device buffers for the C ABI: PyTorch-ROCm tensors as plain device memory.

``Staging`` uploads numpy arrays (or passes torch device tensors through), hands their device
addresses to ``marshal.fill_structs`` and keeps them alive for the duration of a call.
"""
import numpy as np
import torch

_BOOL_AS = np.uint8


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("ddfacet_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def to_device(a, device):
    if isinstance(a, torch.Tensor):
        if a.device.type != "cuda":
            raise TypeError("tensor must live on the GPU")
        if not a.is_contiguous():
            raise TypeError("tensor must be contiguous")
        return a
    a = np.ascontiguousarray(a)
    if a.dtype == np.bool_:
        a = a.view(_BOOL_AS)
    return torch.from_numpy(a).to(device, non_blocking=False)


class Staging:
    def __init__(self, device=None):
        self.device = device or require_gpu()
        self.buffers = {}

    def addr(self, name, a):
        if a is None:
            return 0
        t = to_device(a, self.device)
        self.buffers[name] = t
        return t.data_ptr()

    def download(self, name, out):
        """Copy device buffer `name` back into numpy array `out` (in place)."""
        t = self.buffers[name]
        host = t.cpu().numpy()
        if out.dtype == np.bool_:
            host = host.view(np.bool_)
        out[...] = host.reshape(out.shape)
