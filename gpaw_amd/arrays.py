"""Device-array plumbing: torch tensors as the HBM container.

The kernels take raw device pointers; torch (ROCm) only allocates memory and
names the stream.  Host (numpy) arrays given to the drop-in classes are
staged through device memory (the compute still runs on the GPU).
"""
from __future__ import annotations

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

_NP2T = {}
if torch is not None:
    _NP2T = {np.dtype(np.float64): torch.float64,
             np.dtype(np.complex128): torch.complex128,
             np.dtype(np.float32): torch.float32}


def torch_dtype(dtype):
    return _NP2T[np.dtype(dtype)]


def np_dtype(t):
    for k, v in _NP2T.items():
        if v == t:
            return k
    raise TypeError(f'unsupported dtype {t}')


def is_device(a) -> bool:
    return torch is not None and isinstance(a, torch.Tensor)


def require_gpu():
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError('gpaw_amd needs a visible MI355X (gfx950) GPU; '
                           'there is no CPU compute path')


def device():
    require_gpu()
    return torch.device('cuda', torch.cuda.current_device())


def stream_handle():
    """hipStream_t of torch's current stream (ordering with torch ops)."""
    return torch.cuda.current_stream().cuda_stream


def ptr(t) -> int:
    return t.data_ptr()


def dtype_of(a):
    return np_dtype(a.dtype) if is_device(a) else np.dtype(a.dtype)


def check_contiguous(a, name='array'):
    ok = a.is_contiguous() if is_device(a) else a.flags.c_contiguous
    if not ok:
        raise ValueError(f'{name} must be C-contiguous')


def empty(shape, dtype=float, xp=None):
    return torch.empty(tuple(shape), dtype=torch_dtype(dtype),
                       device=device())


def zeros(shape, dtype=float, xp=None):
    return torch.zeros(tuple(shape), dtype=torch_dtype(dtype),
                       device=device())


def to_device(a):
    """numpy -> device tensor (copy); device tensors pass through."""
    if is_device(a):
        return a
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a).to(device())


def to_host(t):
    if is_device(t):
        return t.detach().cpu().numpy()
    return np.asarray(t)


class Staging:
    """Run device code on host arrays: upload inputs, download outputs."""

    def __init__(self, *arrays):
        self.host = [a for a in arrays]
        self.any_host = any(a is not None and not is_device(a)
                            for a in arrays)
        self.dev = [None if a is None else to_device(a) for a in arrays]

    def download(self, *idx):
        for i in idx:
            h = self.host[i]
            if h is not None and not is_device(h):
                h[...] = self.dev[i].cpu().numpy()
