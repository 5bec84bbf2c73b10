"""Finite-difference operators on the GPU (drop-in for gpaw.fd_operators).

Same classes, signatures and coefficient construction as the reference
(gpaw/fd_operators.py); the stencil itself runs in libgpaw_fd.so
(include/gpaw_fd.h) on torch-allocated device memory.

- ``laplace`` / ``derivatives`` tables: gpaw/fd_operators.py:27-39
  (Chelikowsky et al., PRB 50, 11355 (1994)); the tables have to be the same
  numbers to give the same operator.
- ``FDOperator``: gpaw/fd_operators.py:42-147
- ``GUCLaplace`` / ``Laplace``: :177-238
- ``find_neighbors`` / ``Gradient``: :241-354 (the Minkowski reduction that
  the reference takes from ASE is done by ``minkowski_reduce`` below)
- ``LaplaceA`` / ``LaplaceB`` (Mehrstellen): :357-396
"""
from __future__ import annotations

import itertools
from math import pi

import numpy as np

from . import _lib
from .arrays import (Staging, check_contiguous, dtype_of, is_device, ptr,
                     stream_handle)

laplace = [[0],
           [-2, 1],
           [-5 / 2, 4 / 3, -1 / 12],
           [-49 / 18, 3 / 2, -3 / 20, 1 / 90],
           [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560],
           [-5269 / 1800, 5 / 3, -5 / 21, 5 / 126, -5 / 1008, 1 / 3150],
           [-5369 / 1800, 12 / 7, -15 / 56, 10 / 189, -1 / 112, 2 / 1925,
            -1 / 16632]]

derivatives = [[1 / 2],
               [2 / 3, -1 / 12],
               [3 / 4, -3 / 20, 1 / 60],
               [4 / 5, -1 / 5, 4 / 105, -1 / 280]]


def _nin(shape):
    return int(np.prod(shape[:-3], dtype=np.int64))


class FDOperator:
    """FDOperator(coefs, offsets, gd, dtype) on the GPU.

    ``apply`` / ``__call__`` / ``relax`` accept device tensors (in place,
    asynchronous on torch's current stream) or numpy arrays (staged through
    device memory, synchronous).
    """

    def __init__(self, coef_p, offset_pc, gd, dtype=float,
                 description=None, xp=np):
        offset_pc = np.asarray(offset_pc, dtype=int).reshape((-1, 3))
        cfd = True
        for offset_c in offset_pc:
            if sum([offset != 0 for offset in offset_c]) >= 2:
                cfd = False
        mp = np.abs(offset_pc).max()
        assert (mp <= gd.n_c).all(), 'Stencil longer than domain'
        n_c = gd.n_c
        M_c = n_c + 2 * mp
        stride_c = np.array([M_c[1] * M_c[2], M_c[2], 1])
        offset_p = np.dot(offset_pc, stride_c)
        coef_p = np.ascontiguousarray(coef_p, float)
        neighbor_cd = gd.neighbor_cd
        assert coef_p.ndim == 1
        assert coef_p.shape == offset_p.shape
        assert dtype in [float, complex]
        self.dtype = dtype
        self.shape = tuple(int(n) for n in n_c)
        self.comm = None
        self.mp = mp
        self.gd = gd
        self.npoints = len(coef_p)
        self.coef_p = coef_p
        self.offset_p = offset_p
        self.offset_pc = offset_pc
        self.cfd = cfd
        self.xp = xp

        self._h = None
        self._lib = None
        self.operator = self  # reference attribute: the C object

        if description is None:
            description = '%d point finite-difference stencil' % self.npoints
        self.description = description

    def _handle(self):
        """The libgpaw_fd operator (created on first use, needs the GPU)."""
        if self._h is None:
            lib = _lib.load()
            c_a, c_p = _lib.doubles(self.coef_p)
            o_a, o_p = _lib.longs(self.offset_p)
            n_a, n_p = _lib.longs(self.gd.n_c)
            nb_a, nb_p = _lib.longs(self.gd.neighbor_cd)
            h = _lib.ctypes.c_void_p()
            _lib.check(lib.gfd_operator_new(
                len(self.coef_p), c_p, o_p, n_p, int(self.mp), nb_p,
                int(self.dtype == float), int(self.cfd), _lib.ctypes.byref(h)),
                'Operator')
            self._lib = lib
            self._h = h
        return self._h

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            self._lib.gfd_operator_free(h)
            self._h = None

    def __str__(self):
        return '<' + self.description + '>'

    # -- argument checks of the GPAW_DEBUG wrapper (fd_operators.py:150-174)
    def _check(self, in_xg, out_xg, phase_cd):
        if tuple(in_xg.shape) != tuple(out_xg.shape):
            raise ValueError('input and output shapes differ')
        if tuple(in_xg.shape[-3:]) != self.shape:
            raise ValueError(f'grid shape {tuple(in_xg.shape[-3:])} != '
                             f'{self.shape}')
        want = np.dtype(self.dtype)
        if dtype_of(in_xg) != want or dtype_of(out_xg) != want:
            raise TypeError(f'arrays must be {want}')
        check_contiguous(in_xg, 'input')
        check_contiguous(out_xg, 'output')

    def _phases(self, phase_cd):
        if self.dtype == float:
            return None, None
        if phase_cd is None:
            phase_cd = np.ones((3, 2), complex)
        return _lib.phases_arg(phase_cd)

    def apply(self, in_xg, out_xg, phase_cd=None):
        """out = stencil(in) over all leading (batch) dimensions."""
        self._check(in_xg, out_xg, phase_cd)
        st = Staging(in_xg, out_xg)
        keep, ph = self._phases(phase_cd)
        h = self._handle()
        _lib.check(self._lib.gfd_operator_apply(
            h, ptr(st.dev[0]), ptr(st.dev[1]), _nin(in_xg.shape), ph,
            stream_handle()), 'Operator.apply')
        st.download(1)

    def __call__(self, in_xR, out_xR=None):
        """New UGArray interface (fd_operators.py:102-118)."""
        if out_xR is None:
            out_xR = in_xR.new()
        phase = getattr(in_xR.desc, 'phase_factor_cd', None)
        self.apply(in_xR.data, out_xR.data, phase)
        return out_xR

    def relax(self, relax_method, f_g, s_g, n, w=None):
        """n sweeps of relax_method (2: weighted Jacobi) for self(f) = s."""
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
        st = Staging(f_g, s_g)
        h = self._handle()
        _lib.check(self._lib.gfd_operator_relax(
            h, int(relax_method), ptr(st.dev[0]), ptr(st.dev[1]),
            int(n), float(w), stream_handle()), 'Operator.relax')
        st.download(0)

    def residual(self, in_xg, b_xg, r_xg, phase_cd=None, norm=False):
        """r = self(in) - b, and sum |r|^2 when ``norm`` (fused)."""
        st = Staging(in_xg, b_xg, r_xg)
        keep, ph = self._phases(phase_cd)
        out = _lib.ctypes.c_double(0.0)
        h = self._handle()
        _lib.check(self._lib.gfd_operator_residual(
            h, ptr(st.dev[0]), ptr(st.dev[1]), ptr(st.dev[2]),
            _nin(in_xg.shape), ph,
            _lib.ctypes.byref(out) if norm else None, stream_handle()),
            'Operator.residual')
        st.download(2)
        return out.value if norm else None

    def get_diagonal_element(self):
        """Coefficient of the (last) zero offset (c/operators.cpp:319-332)."""
        d = 0.0
        for c, o in zip(self.coef_p, self.offset_p):
            if o == 0:
                d = float(c)
        return d

    def get_async_sizes(self):
        return (0, 1, 1)


def Laplace(gd, scale=1.0, n=1, dtype=float, xp=np):
    if n == 9:
        raise NotImplementedError('FTLaplace (n=9) is a host FFT operator, '
                                  'not part of the FD grid path')
    return GUCLaplace(gd, scale, n, dtype, xp=xp)


class GUCLaplace(FDOperator):
    def __init__(self, gd, scale=1.0, n=1, dtype=float, xp=np):
        """Laplacian for general non orthorhombic grid.

        scale=-0.5 gives the kinetic energy operator; the stencil has
        O(h^(2n)) error (gpaw/fd_operators.py:184-238).
        """
        # the 13 inequivalent neighbour directions, shortest first
        M_ic = np.indices((3, 3, 3)).reshape((3, -3)).T[-13:] - 1
        u_cv = gd.h_cv / (gd.h_cv**2).sum(1)[:, np.newaxis]**0.5
        u2_i = (np.dot(M_ic, u_cv)**2).sum(1)
        i_d = u2_i.argsort()

        m_mv = np.array([(2, 0, 0), (0, 2, 0), (0, 0, 2),
                         (0, 1, 1), (1, 0, 1), (1, 1, 0)])
        for D in range(3, 7):
            h_dv = np.dot(M_ic[i_d[:D]], gd.h_cv)
            A_md = (h_dv**m_mv[:, np.newaxis, :]).prod(2)
            a_d, residual, rank, s = np.linalg.lstsq(A_md, [1, 1, 1, 0, 0, 0],
                                                     rcond=-1)
            if residual.sum() < 1e-14:
                if rank != D:
                    raise ValueError(
                        'You have a weird unit cell!  '
                        'Try to use the maximally reduced Niggli cell.')
                break

        a_d *= scale
        offsets = [(0, 0, 0)]
        coefs = [laplace[n][0] * a_d.sum()]
        for d in range(D):
            M_c = M_ic[i_d[d]]
            offsets.extend(np.arange(1, n + 1)[:, np.newaxis] * M_c)
            coefs.extend(a_d[d] * np.array(laplace[n][1:]))
            offsets.extend(np.arange(-1, -n - 1, -1)[:, np.newaxis] * M_c)
            coefs.extend(a_d[d] * np.array(laplace[n][1:]))

        super().__init__(coefs, offsets, gd, dtype, xp=xp)
        self.description = (
            '%d*%d+1=%d point O(h^%d) finite-difference Laplacian' %
            ((self.npoints - 1) // n, n, self.npoints, 2 * n))


def minkowski_reduce(h_cv):
    """Minkowski-reduced basis of the lattice spanned by the rows of h_cv.

    Returns (h_bv, U_bc) with h_bv = U_bc @ h_cv, det U = +-1, and the rows
    of h_bv the successive minima (in 3-D those always form a basis).  The
    search enumerates small integer combinations of the current basis
    (Lagrange-Gauss style greedy reduction), which is exact for the
    reduced starting bases reached after a few rounds.
    """
    h_cv = np.asarray(h_cv, float)
    U = np.eye(3, dtype=int)
    B = h_cv.copy()
    rng = range(-2, 3)
    combos = np.array([c for c in itertools.product(rng, rng, rng)
                       if any(c)], dtype=int)
    for _ in range(20):
        V = combos @ B
        L = (V**2).sum(1)
        order = np.lexsort((np.arange(len(L)), np.round(L, 12)))
        picked = []
        for i in order:
            cand = combos[i]
            trial = np.array(picked + [cand])
            if np.linalg.matrix_rank(trial) < len(trial):
                continue
            if len(trial) == 3 and abs(round(np.linalg.det(trial))) != 1:
                continue
            picked.append(cand)
            if len(picked) == 3:
                break
        T = np.array(picked, dtype=int)
        newB = T @ B
        newU = T @ U
        if np.allclose((newB**2).sum(1), (B**2).sum(1)) and \
                (T == np.eye(3, dtype=int)).all():
            break
        B, U = newB, newU
    # orient each vector like an in-place reduction b_i <- b_i - k b_j
    # leaves it: last nonzero integer coefficient positive
    for i in range(3):
        nz = np.nonzero(U[i])[0]
        if U[i, nz[-1]] < 0:
            U[i] = -U[i]
            B[i] = -B[i]
    return B, U


def find_neighbors(h_cv):
    """Nearest neighbours sorted after distance (gpaw/fd_operators.py:241-281).

    >>> find_neighbors([[0.1, 0, 0], [0, 0.11, 0], [0, 0, 0.12]])
    array([[1, 0, 0],
           [0, 1, 0],
           [0, 0, 1]])
    """
    from scipy.spatial import Voronoi
    h_bv, U_bc = minkowski_reduce(h_cv)
    M_ib = np.indices((3, 3, 3)).reshape((3, -1)).T - 1
    h_iv = M_ib.dot(h_bv)
    voro = Voronoi(h_iv)
    i_d = []
    for i1, i2 in voro.ridge_points:
        if i1 == 13 and i2 > 13:
            i_d.append(i2)
        elif i2 == 13 and i1 > 13:
            i_d.append(i1)
    h2_d = (h_iv[i_d]**2).sum(1)
    M_db = M_ib[[i_d[d] for d in h2_d.argsort()]]
    return M_db @ U_bc


class Gradient(FDOperator):
    def __init__(self, grid, v, scale=1.0, n=1, dtype=float, xp=np):
        """Symmetric gradient along v (0, 1, 2) for any cell
        (gpaw/fd_operators.py:284-354)."""
        gd = getattr(grid, '_gd', grid)
        M_dc = find_neighbors(gd.h_cv)
        h_dv = M_dc @ gd.h_cv
        D = len(h_dv)
        invh_vc = np.linalg.inv(gd.h_cv)
        n_cv = (invh_vc / (invh_vc**2).sum(axis=0)**0.5).T
        coef_cd = np.zeros((3, D))
        for c, n_v in enumerate(n_cv):
            ok_d = abs(h_dv.dot(n_v)) > 1e-10
            h_mv = h_dv[ok_d]
            A_jm = np.array([h_mv.dot(n_v),
                             h_mv.dot(gd.h_cv[c - 1]),
                             h_mv.dot(gd.h_cv[c - 2])])
            U_jm, S_m, V_mm = np.linalg.svd(A_jm, full_matrices=False)
            coef_m = V_mm.T.dot(np.diag(S_m**-1).dot(U_jm.T.dot([1, 0, 0])))
            coef_cd[c, ok_d] = coef_m
        coef_d = np.linalg.inv(n_cv)[v].dot(coef_cd) * scale
        offsets = []
        coefs = []
        stencil = np.array(derivatives[n - 1])
        for d, c in enumerate(coef_d):
            if abs(c) < 1e-10:
                continue
            M_c = M_dc[d]
            offsets.extend(np.arange(1, n + 1)[:, np.newaxis] * M_c)
            coefs.extend(c * stencil)
            offsets.extend(np.arange(-1, -n - 1, -1)[:, np.newaxis] * M_c)
            coefs.extend(-c * stencil)
        super().__init__(coefs, offsets, gd, dtype, xp=xp)
        self.description = (
            'Finite-difference {}-derivative with O(h^{}) error ({} points)'
            .format('xyz'[v], 2 * n, self.npoints))


class LaplaceA(FDOperator):
    def __init__(self, gd, scale, dtype=float, xp=np):
        assert gd.orthogonal
        c = np.divide(-1 / 12, gd.h_cv.diagonal()**2) * scale
        c0 = c[1] + c[2]
        c1 = c[0] + c[2]
        c2 = c[1] + c[0]
        a = -16.0 * np.sum(c)
        b = 10.0 * c + 0.125 * a
        super().__init__([a,
                          b[0], b[0],
                          b[1], b[1],
                          b[2], b[2],
                          c0, c0, c0, c0,
                          c1, c1, c1, c1,
                          c2, c2, c2, c2],
                         [(0, 0, 0),
                          (-1, 0, 0), (1, 0, 0),
                          (0, -1, 0), (0, 1, 0),
                          (0, 0, -1), (0, 0, 1),
                          (0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1),
                          (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1),
                          (-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0)],
                         gd, dtype,
                         'O(h^4) Mehrstellen Laplacian (A)',
                         xp=xp)


class LaplaceB(FDOperator):
    def __init__(self, gd, dtype=float, xp=np):
        a = 0.5
        b = 1.0 / 12.0
        super().__init__([a, b, b, b, b, b, b],
                         [(0, 0, 0),
                          (-1, 0, 0), (1, 0, 0),
                          (0, -1, 0), (0, 1, 0),
                          (0, 0, -1), (0, 0, 1)],
                         gd, dtype,
                         'O(h^4) Mehrstellen Laplacian (B)',
                         xp=xp)


__all__ = ['FDOperator', 'Laplace', 'GUCLaplace', 'Gradient', 'LaplaceA',
           'LaplaceB', 'find_neighbors', 'minkowski_reduce', 'laplace',
           'derivatives', 'pi']
