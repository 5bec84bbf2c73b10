"""Grid transformers on the GPU (drop-in for gpaw.transformers).

``Transformer(gdin, gdout, nn, dtype, xp)`` restricts (gdin twice as fine as
gdout) or interpolates (gdout twice as fine) with a K = 2 nn tap stencil.
The padding arithmetic is the reference's (gpaw/transformers.py:27-64): the
same numbers are handed to the C ABI (gfd_transformer_new), which turns them
into the interior-index offsets of its wrap-based kernels.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .arrays import (Staging, check_contiguous, dtype_of, is_device, ptr,
                     stream_handle)
from .arrays import empty as dev_empty


def transformer_paddings(gdin, gdout, nn):
    """(interpolate, pad_cd, neighborpad_cd, skip_cd) as in
    gpaw/transformers.py:27-56 (raises ValueError like the reference)."""
    pad_cd = np.empty((3, 2), int)
    neighborpad_cd = np.empty((3, 2), int)
    skip_cd = np.empty((3, 2), int)
    if (gdin.N_c == 2 * gdout.N_c).all():
        pad_cd[:, 0] = 2 * nn - 1 - 2 * gdout.beg_c + gdin.beg_c
        pad_cd[:, 1] = 2 * nn - 2 + 2 * gdout.end_c - gdin.end_c
        neighborpad_cd[:, 0] = 2 * nn - 2 + 2 * gdout.beg_c - gdin.beg_c
        neighborpad_cd[:, 1] = 2 * nn - 1 - 2 * gdout.end_c + gdin.end_c
        interpolate = False
    else:
        assert (gdout.N_c == 2 * gdin.N_c).all()
        pad_cd[:, 0] = nn - 1 - gdout.beg_c // 2 + gdin.beg_c
        pad_cd[:, 1] = nn + gdout.end_c // 2 - gdin.end_c
        neighborpad_cd[:, 0] = nn + gdout.beg_c // 2 - gdin.beg_c
        neighborpad_cd[:, 1] = nn - 1 - gdout.end_c // 2 + gdin.end_c
        skip_cd[:, 0] = gdout.beg_c % 2
        skip_cd[:, 1] = gdout.end_c % 2
        interpolate = True
        inpoints = (gdin.n_c[0] + 2 * nn - 1) * (gdin.n_c[1] + 2 * nn - 1)
        outpoints = gdout.n_c[0] * gdout.n_c[1]
        if inpoints > outpoints:
            points = ' x '.join([str(N) for N in gdin.N_c])
            raise ValueError('Cannot construct interpolator.  Grid %s '
                             'may be too small' % points)
    if not interpolate:
        skip_cd[:] = 0
    assert (pad_cd.ravel() >= 0).all()
    return interpolate, pad_cd, neighborpad_cd, skip_cd


class _Transformer:
    def __init__(self, gdin, gdout, nn=1, dtype=float, xp=np):
        self.gdin = gdin
        self.gdout = gdout
        self.nn = nn
        assert 1 <= nn <= 4
        self.dtype = dtype
        self.xp = xp
        (self.interpolate, self.pad_cd, self.neighborpad_cd,
         self.skip_cd) = transformer_paddings(gdin, gdout, nn)
        self.ngpin = tuple(int(n) for n in gdin.n_c)
        self.ngpout = tuple(int(n) for n in gdout.n_c)
        assert dtype in [float, complex]
        self._h = None
        self._lib = None
        self.transformer = self  # reference attribute: the C object

    def _handle(self):
        if self._h is None:
            lib = _lib.load()
            a1, p_in = _lib.longs(self.gdin.n_c)
            a2, p_out = _lib.longs(self.gdout.n_c)
            a3, p_pad = _lib.longs(self.pad_cd)
            a4, p_npad = _lib.longs(self.neighborpad_cd)
            a5, p_skip = _lib.longs(self.skip_cd)
            a6, p_nb = _lib.longs(self.gdin.neighbor_cd)
            h = _lib.ctypes.c_void_p()
            _lib.check(lib.gfd_transformer_new(
                p_in, p_out, 2 * self.nn, p_pad, p_npad, p_skip, p_nb,
                int(self.dtype == float), int(self.interpolate),
                _lib.ctypes.byref(h)), 'Transformer')
            self._lib = lib
            self._h = h
        return self._h

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            self._lib.gfd_transformer_free(h)
            self._h = None

    def _check(self, input, output, phases):
        if tuple(input.shape[-3:]) != self.ngpin:
            raise ValueError(f'input grid {tuple(input.shape[-3:])} != '
                             f'{self.ngpin}')
        if tuple(output.shape[-3:]) != self.ngpout:
            raise ValueError(f'output grid {tuple(output.shape[-3:])} != '
                             f'{self.ngpout}')
        if tuple(input.shape[:-3]) != tuple(output.shape[:-3]):
            raise ValueError('batch shapes differ')
        want = np.dtype(self.dtype)
        if dtype_of(input) != want or dtype_of(output) != want:
            raise TypeError(f'arrays must be {want}')
        check_contiguous(input, 'input')
        check_contiguous(output, 'output')

    def apply(self, input, output=None, phases=None):
        if output is None:
            if is_device(input):
                output = dev_empty(tuple(input.shape[:-3]) + self.ngpout,
                                   self.dtype)
            else:
                output = self.gdout.empty(tuple(input.shape[:-3]),
                                          dtype=self.dtype)
        self._check(input, output, phases)
        st = Staging(input, output)
        keep, ph = (None, None)
        if self.dtype != float:
            keep, ph = _lib.phases_arg(np.ones((3, 2), complex)
                                       if phases is None else phases)
        nin = int(np.prod(input.shape[:-3], dtype=np.int64))
        h = self._handle()
        _lib.check(self._lib.gfd_transformer_apply(
            h, ptr(st.dev[0]), ptr(st.dev[1]), nin, ph, stream_handle()),
            'Transformer.apply')
        st.download(1)
        return output

    def get_async_sizes(self):
        return (0, 1)


def Transformer(gdin, gdout, nn=1, dtype=float, xp=np):
    if nn != 9:
        return _Transformer(gdin, gdout, nn, dtype, xp)

    class T:
        nn = 1

        def apply(self, input, output, phases=None):
            output[:] = input

    return T()
