"""Device helpers for the GPU parity tests."""
import contextlib
import sys

import numpy as np

from gpaw_amd import _lib


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to('cuda')


def host(t):
    return t.detach().cpu().numpy()


def say(*args):
    print(*args, flush=True, file=sys.stderr)


@contextlib.contextmanager
def general_path(on=True):
    """Route every stencil through the general kernel while active."""
    lib = _lib.load()
    lib.gfd_set_tuning(b'force_general', int(on))
    try:
        yield
    finally:
        lib.gfd_set_tuning(b'force_general', 0)
