"""Weighted finite-difference operators on the GPU.

``WeightedFDOperator(operators)`` mirrors gpaw/wfd_operators.py:7-119: the
compound operator A = sum_i weights_i * operators_i with per-point weights
(the dielectric Poisson operator of the solvation code uses a Laplacian and
the three gradients weighted by eps and its gradient).  The weights are set
with ``set_weights`` (device tensors, or numpy arrays uploaded once) and are
borrowed, not copied, as in the reference.

Kernels: ``gfd_woperator_*`` (csrc/wstencil.hip), restating
c/bmgs/wfd.cpp (apply) and c/bmgs/wrelax.cpp (weighted Jacobi).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .arrays import (Staging, check_contiguous, dtype_of, is_device, ptr,
                     stream_handle, torch)


class WeightedFDOperator:
    def __init__(self, operators):
        """Compound operator from a list of FDOperators on one grid
        (gpaw/wfd_operators.py:8-62): every stencil gets its centre as the
        first point (a zero coefficient is prepended where it has none)."""
        self.gd = operators[0].gd
        self.shape = tuple(int(n) for n in self.gd.n_c)
        self.coef_ps = []
        self.offset_pcs = []
        for op in operators:
            assert op.gd is self.gd or tuple(op.gd.n_c) == self.shape
            assert operators[0].dtype == op.dtype
            off = np.asarray(op.offset_pc, int).reshape((-1, 3))
            coef = np.asarray(op.coef_p, float)
            zero = np.flatnonzero((off == 0).all(axis=1))
            if len(zero):
                assert zero[0] == 0, 'the centre must be the first point'
                self.offset_pcs.append(np.ascontiguousarray(off))
                self.coef_ps.append(np.ascontiguousarray(coef))
            else:
                self.offset_pcs.append(
                    np.ascontiguousarray(np.vstack(([[0, 0, 0]], off))))
                self.coef_ps.append(
                    np.ascontiguousarray(np.hstack(([0.0], coef))))
        self.nweights = len(operators)
        self.cfd = all(op.cfd for op in operators)
        self.mp = max(int(op.mp) for op in operators)
        self.dtype = operators[0].dtype
        self.npoints = max(op.npoints for op in operators)
        self.weights = None
        self.operator = None
        self.description = 'Weighted Finite Difference Operator\n      '
        self.description += '\n      '.join(op.description
                                             for op in operators)
        self._h = None
        self._lib = None
        self._keep = None

    def set_weights(self, weights):
        """Set the operator weights (gpaw/wfd_operators.py:63-79): a list of
        nweights float64 arrays of the grid's shape."""
        assert len(weights) == self.nweights
        dev = []
        for w in weights:
            if tuple(w.shape) != self.shape or dtype_of(w) != np.float64:
                raise ValueError('weights: float64 arrays of the grid shape')
            if is_device(w):
                check_contiguous(w, 'weight')
                dev.append(w)
            else:
                dev.append(torch.from_numpy(np.ascontiguousarray(w)).to(
                    'cuda'))
        self._free()
        lib = _lib.load()
        M = np.array(self.shape) + 2 * self.mp
        stride = np.array([M[1] * M[2], M[2], 1])
        nc = np.array([len(c) for c in self.coef_ps], dtype=np.int32)
        coef = np.ascontiguousarray(np.concatenate(self.coef_ps))
        off = np.ascontiguousarray(
            np.concatenate([o @ stride for o in self.offset_pcs]),
            dtype=np.int64)
        c_a, c_p = _lib.doubles(coef)
        o_a, o_p = _lib.longs(off)
        n_a, n_p = _lib.longs(self.gd.n_c)
        nb_a, nb_p = _lib.longs(self.gd.neighbor_cd)
        wp = (_lib.ctypes.c_void_p * self.nweights)(*[ptr(w) for w in dev])
        wpp = _lib.ctypes.cast(wp, _lib.ctypes.c_void_p)
        h = _lib.ctypes.c_void_p()
        _lib.check(lib.gfd_woperator_new(
            self.nweights, nc.ctypes.data_as(_lib.ctypes.c_void_p), c_p, o_p,
            n_p, int(self.mp), nb_p, int(self.dtype == float), wpp,
            _lib.ctypes.byref(h)), 'WOperator')
        self._lib = lib
        self._h = h
        self._keep = dev
        self.weights = weights
        self.operator = self

    def _free(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            self._lib.gfd_woperator_free(h)
        self._h = None

    def __del__(self):
        self._free()

    def _handle(self):
        if self._h is None:
            raise RuntimeError('set_weights() first')
        return self._h

    def apply(self, in_xg, out_xg, phase_cd=None):
        """out = sum_w weights_w * operator_w(in) over the batch."""
        if tuple(in_xg.shape) != tuple(out_xg.shape) or \
                tuple(in_xg.shape[-3:]) != self.shape:
            raise ValueError('input / output shape mismatch')
        want = np.dtype(self.dtype)
        if dtype_of(in_xg) != want or dtype_of(out_xg) != want:
            raise TypeError(f'arrays must be {want}')
        check_contiguous(in_xg, 'input')
        check_contiguous(out_xg, 'output')
        keep, ph = (None, None)
        if self.dtype == complex:
            keep, ph = _lib.phases_arg(np.ones((3, 2), complex)
                                       if phase_cd is None else phase_cd)
        h = self._handle()
        st = Staging(in_xg, out_xg)
        nin = int(np.prod(in_xg.shape[:-3], dtype=np.int64))
        _lib.check(self._lib.gfd_woperator_apply(
            h, ptr(st.dev[0]), ptr(st.dev[1]), nin, ph, stream_handle()),
            'WOperator.apply')
        st.download(1)

    def relax(self, relax_method, f_g, s_g, n, w=None):
        """n sweeps for A f = s: relax_method 2 weighted Jacobi, 1
        lexicographic Gauss-Seidel (w unused), as c/bmgs/wrelax.cpp."""
        if self.dtype != float:
            raise TypeError('relax: real operators only')
        if tuple(f_g.shape) != self.shape or tuple(s_g.shape) != self.shape:
            raise ValueError('relax: grid shape mismatch')
        if dtype_of(f_g) != np.float64 or dtype_of(s_g) != np.float64:
            raise TypeError('relax: float64 arrays only')
        check_contiguous(f_g, 'f_g')
        check_contiguous(s_g, 's_g')
        if w is None:
            w = 1.0
        h = self._handle()
        st = Staging(f_g, s_g)
        _lib.check(self._lib.gfd_woperator_relax(
            h, int(relax_method), ptr(st.dev[0]), ptr(st.dev[1]), int(n),
            float(w), stream_handle()), 'WOperator.relax')
        st.download(0)

    def get_diagonal_element(self):
        """max over the grid of |sum_w weight_w c_{w,0}|
        (c/woperators.cpp:341-378)."""
        out = _lib.ctypes.c_double(0.0)
        _lib.check(self._lib.gfd_woperator_get_diagonal_element(
            self._handle(), _lib.ctypes.byref(out), stream_handle()),
            'WOperator.get_diagonal_element')
        return out.value

    def get_async_sizes(self):
        return (0, 1, 1)


__all__ = ['WeightedFDOperator']
