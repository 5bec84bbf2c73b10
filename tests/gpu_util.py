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


PATHS = {'auto': 0, 'general': 1, 'xmarch': 2, 'row': 3}


@contextlib.contextmanager
def kernel_path(name='general'):
    """Route every stencil through one kernel family while active (a
    family that cannot take a grid falls back to the general kernel)."""
    lib = _lib.load()
    lib.gfd_set_tuning(b'force_path', PATHS[name])
    try:
        yield
    finally:
        lib.gfd_set_tuning(b'force_path', 0)


def general_path(on=True):
    return kernel_path('general' if on else 'auto')


@contextlib.contextmanager
def xfer_path(name='march'):
    """Route transfers through the x-marching kernel ('march', default;
    the timed kernel choice is off meanwhile, so it is not bypassed) or the
    one-point-per-thread / 1-D pass kernels ('legacy') while active."""
    lib = _lib.load()
    lib.gfd_set_tuning(b'xfer_march', 1 if name == 'march' else 0)
    if name == 'march':
        lib.gfd_set_tuning(b'autotune', 0)
    try:
        yield
    finally:
        lib.gfd_set_tuning(b'xfer_march', 1)
        lib.gfd_set_tuning(b'autotune', 1)


BATCH_KERNELS = {'auto': 0, 'regs': 1, 'march': 2}


@contextlib.contextmanager
def batch_kernel(name='march'):
    """Limit the batched stencil tuner to one kernel ('regs' register
    tiles, 'march' x-march tiles) for batches tuned while active."""
    lib = _lib.load()
    _lib.check(lib.gfd_set_tuning(b'batch_kernel', BATCH_KERNELS[name]),
               'batch_kernel')
    try:
        yield
    finally:
        lib.gfd_set_tuning(b'batch_kernel', 0)


@contextlib.contextmanager
def batch_xfer(code):
    """Batched transfers of batches tuned while active: 1 one point per
    thread, 2 x-march tiles, or a forced x-march tile code."""
    lib = _lib.load()
    _lib.check(lib.gfd_set_tuning(b'batch_xfer', code), 'batch_xfer')
    try:
        yield
    finally:
        lib.gfd_set_tuning(b'batch_xfer', 0)
