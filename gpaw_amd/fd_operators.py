"""Finite-difference operators on the GPU (drop-in for gpaw.fd_operators).

The classes keep the reference's names, signatures and attributes
(gpaw/fd_operators.py:42-396); the stencil runs in libgpaw_fd.so
(include/gpaw_fd.h) on torch-allocated device memory.  The coefficients are
built here from the defining equations, written down next to each builder:

- ``laplace`` / ``derivatives``: the central-difference weights of
  Chelikowsky et al., PRB 50, 11355 (1994), as tabulated in
  gpaw/fd_operators.py:27-39 (the same numbers give the same operator).
- Laplacian on any cell (``GUCLaplace``): sum_d a_d (h_d . grad)^2 with
  sum_d a_d h_d h_d^T = 1, over the fewest of the shortest lattice steps h_d
  for which that tensor equation has an exact solution.  Orthorhombic cells
  have the closed form a_d = 1 / h_d^2 along the three axes.
- Gradient on any cell: the derivative along the unit normal n_c of each
  cell plane as a minimum-norm combination of 1-D derivatives along the
  Voronoi-relevant lattice steps, recombined into the Cartesian direction v.
- Mehrstellen pair ``LaplaceA`` / ``LaplaceB``: the O(h^4) compact scheme
  A phi = B f for lap phi = f (Collatz), coefficients derived below.
"""
from __future__ import annotations

import itertools
from math import pi

import numpy as np

from . import _lib
from .arrays import (Staging, check_contiguous, dtype_of, is_device, ptr,
                     stream_handle)

# Central second-derivative weights: laplace[n] = [w_0, w_1, .., w_n] with
# h^2 f''(x) ~ w_0 f(x) + sum_k w_k (f(x + k h) + f(x - k h)).
laplace = [[0],
           [-2, 1],
           [-5 / 2, 4 / 3, -1 / 12],
           [-49 / 18, 3 / 2, -3 / 20, 1 / 90],
           [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560],
           [-5269 / 1800, 5 / 3, -5 / 21, 5 / 126, -5 / 1008, 1 / 3150],
           [-5369 / 1800, 12 / 7, -15 / 56, 10 / 189, -1 / 112, 2 / 1925,
            -1 / 16632]]

# Central first-derivative weights: derivatives[n - 1] = [d_1, .., d_n] with
# h f'(x) ~ sum_k d_k (f(x + k h) - f(x - k h)).
derivatives = [[1 / 2],
               [2 / 3, -1 / 12],
               [3 / 4, -3 / 20, 1 / 60],
               [4 / 5, -1 / 5, 4 / 105, -1 / 280]]

# The 13 lattice steps (i, j, k), |i|, |j|, |k| <= 1, whose first nonzero
# component is positive, in lexicographic order; -m is the other half.
_HALF_STEPS = np.array([m for m in itertools.product((-1, 0, 1), repeat=3)
                        if any(m) and m[np.flatnonzero(m)[0]] > 0], int)


def _nin(shape):
    return int(np.prod(shape[:-3], dtype=np.int64))


def _line(step_c, weights, sign):
    """Offsets and coefficients of one 1-D stencil along step_c: the points
    k step_c for k = 1..n with `weights`, then k = -1..-n with
    sign * weights."""
    n = len(weights)
    step_c = np.asarray(step_c, int)
    offs = [k * step_c for k in range(1, n + 1)]
    offs += [-k * step_c for k in range(1, n + 1)]
    coefs = list(weights) + [sign * w for w in weights]
    return offs, coefs


class FDOperator:
    """FDOperator(coefs, offsets, gd, dtype) on the GPU.

    ``apply`` / ``__call__`` / ``relax`` accept device tensors (in place,
    asynchronous on torch's current stream) or numpy arrays (staged through
    device memory, synchronous).  Attributes as the reference's:
    ``coef_p``, ``offset_pc``, ``offset_p`` (offsets flattened with the
    strides of the grid padded by ``mp`` on every side), ``mp`` (stencil
    radius), ``cfd`` (every point on an axis through the centre),
    ``npoints``, ``description``.
    """

    def __init__(self, coef_p, offset_pc, gd, dtype=float,
                 description=None, xp=np):
        if dtype not in (float, complex):
            raise TypeError(f'dtype must be float or complex, not {dtype}')
        offsets = np.asarray(offset_pc, dtype=int).reshape((-1, 3))
        coefs = np.ascontiguousarray(np.asarray(coef_p, float))
        if coefs.ndim != 1 or len(coefs) != len(offsets):
            raise ValueError('one coefficient per offset')
        radius = int(np.abs(offsets).max()) if len(offsets) else 0
        if (radius > np.asarray(gd.n_c)).any():
            # the reference asserts this (gpaw/fd_operators.py:53)
            raise AssertionError('Stencil longer than domain')
        padded = np.asarray(gd.n_c, int) + 2 * radius
        strides = np.array([padded[1] * padded[2], padded[2], 1])
        self.dtype = dtype
        self.gd = gd
        self.shape = tuple(int(x) for x in gd.n_c)
        self.comm = None
        self.xp = xp
        self.coef_p = coefs
        self.offset_pc = offsets
        self.offset_p = offsets @ strides
        self.npoints = len(coefs)
        self.mp = radius
        self.cfd = bool((np.count_nonzero(offsets, axis=1) <= 1).all())
        self.description = (description if description is not None else
                            f'{self.npoints} point finite-difference '
                            'stencil')
        self.operator = self  # the reference exposes its C object here
        self._h = None
        self._lib = None

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

    # -- the argument checks the reference makes in debug mode
    #    (gpaw/fd_operators.py:150-174), always on here: the kernels trust
    #    shapes and dtypes
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
        """Array-object interface (gpaw/fd_operators.py:102-118)."""
        if out_xR is None:
            out_xR = in_xR.new()
        phase = getattr(in_xR.desc, 'phase_factor_cd', None)
        self.apply(in_xR.data, out_xR.data, phase)
        return out_xR

    def relax(self, relax_method, f_g, s_g, n, w=None):
        """n sweeps of relax_method (2: weighted Jacobi, 1: Gauss-Seidel)
        for self(f) = s."""
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
        self._check(in_xg, r_xg, phase_cd)
        self._check(b_xg, r_xg, phase_cd)
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
        """Coefficient of the zero offset, the last one if it repeats (as
        c/operators.cpp:319-332 reads it)."""
        centre = np.flatnonzero(self.offset_p == 0)
        return float(self.coef_p[centre[-1]]) if len(centre) else 0.0

    def get_async_sizes(self):
        return (0, 1, 1)


# ------------------------------------------------------------- Laplacian

def _second_moment_rows(h_dv):
    """The six independent entries (xx, yy, zz, yz, xz, xy) of h_d h_d^T,
    one column per step d."""
    x, y, z = np.asarray(h_dv, float).T
    return np.array([x * x, y * y, z * z, y * z, x * z, x * y])


_UNIT_TENSOR = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])


def laplacian_directions(h_cv):
    """(M_dc, a_d): lattice steps M_dc (grid-index units) and weights with
    sum_d a_d (M_d h)(M_d h)^T = 1, so that lap = sum_d a_d (h_d . grad)^2
    with h_d = M_d h.

    Candidate steps are the 13 half-space steps ordered by their length in
    units of the cell-vector lengths (ties in lexicographic order, so an
    orthorhombic cell gives z, y, x); the first D = 3, 4, 5, 6 of them
    whose tensor equation is solvable are used.  A rank-deficient set of
    steps means the cell needs reducing first (the reference raises the
    same ValueError, gpaw/fd_operators.py:216-220)."""
    h_cv = np.asarray(h_cv, float)
    if np.count_nonzero(h_cv - np.diag(np.diag(h_cv))) == 0:
        h = np.diag(h_cv)
        return np.array([(0, 0, 1), (0, 1, 0), (1, 0, 0)]), \
            1.0 / h[::-1]**2
    unit_cv = h_cv / np.linalg.norm(h_cv, axis=1)[:, None]
    length = ((_HALF_STEPS @ unit_cv)**2).sum(1)
    order = np.argsort(length, kind='stable')
    for D in range(3, 7):
        M_dc = _HALF_STEPS[order[:D]]
        A = _second_moment_rows(M_dc @ h_cv)          # 6 x D
        if np.linalg.matrix_rank(A) < D:
            raise ValueError('You have a weird unit cell!  Try to use the '
                             'maximally reduced Niggli cell.')
        Q, R = np.linalg.qr(A)
        a_d = np.linalg.solve(R, Q.T @ _UNIT_TENSOR)
        if ((A @ a_d - _UNIT_TENSOR)**2).sum() < 1e-14 or D == 6:
            return M_dc, a_d
    raise AssertionError('unreachable')


def Laplace(gd, scale=1.0, n=1, dtype=float, xp=np):
    """scale * Laplacian with O(h^2n) error (gpaw/fd_operators.py:177-181)."""
    if n == 9:
        raise NotImplementedError('FTLaplace (n=9) is a host FFT operator, '
                                  'not part of the FD grid path')
    return GUCLaplace(gd, scale, n, dtype, xp=xp)


class GUCLaplace(FDOperator):
    def __init__(self, gd, scale=1.0, n=1, dtype=float, xp=np):
        """scale * Laplacian on any cell; scale=-0.5 is the kinetic energy
        operator (gpaw/fd_operators.py:184-238).

        Stencil layout: the centre first, then for each direction d the
        points +1..+n and -1..-n along M_d.
        """
        M_dc, a_d = laplacian_directions(gd.h_cv)
        a_d = scale * np.asarray(a_d)
        w = np.array(laplace[n], float)
        offsets = [np.zeros(3, int)]
        coefs = [w[0] * a_d.sum()]
        for M_c, a in zip(M_dc, a_d):
            o, c = _line(M_c, a * w[1:], +1)
            offsets += o
            coefs += c
        super().__init__(coefs, offsets, gd, dtype, xp=xp)
        D = len(a_d)
        self.description = (f'{D}*{n}+1={self.npoints} point O(h^{2 * n}) '
                            'finite-difference Laplacian')


# -------------------------------------------------------------- gradient

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


def voronoi_steps(h_bv, tol=1e-10):
    """Voronoi-relevant lattice vectors of a reduced basis h_bv, one of each
    +-pair (first nonzero coefficient positive), as basis coefficients,
    with their squared lengths.

    v is relevant (the bisector plane of 0 and v bounds the Voronoi cell
    of 0 with a face of positive area) exactly when v / 2 is strictly closer
    to 0 and v than to every other lattice point u: u . v < |u|^2.  For a
    reduced 3-D basis the relevant vectors and the points that can cut
    their faces are among the 26 steps with coefficients in {-1, 0, 1}.
    """
    h_bv = np.asarray(h_bv, float)
    steps = np.array([m for m in itertools.product((-1, 0, 1), repeat=3)
                      if any(m)], int)
    vec = steps @ h_bv
    norm2 = (vec**2).sum(1)
    keep = []
    for i, m in enumerate(steps):
        if m[np.flatnonzero(m)[0]] < 0:
            continue
        others = np.ones(len(steps), bool)
        others[i] = False
        dots = vec[others] @ vec[i]
        if (dots < norm2[others] * (1 - tol)).all():
            keep.append(i)
    return steps[keep], norm2[keep]


def find_neighbors(h_cv):
    """Grid steps to the Voronoi neighbours of a grid point, shortest first
    (one of each +-pair), in units of the cell's grid vectors: the
    neighbours are at ``find_neighbors(h_cv) @ h_cv``
    (gpaw/fd_operators.py:241-281).

    >>> find_neighbors([[0.1, 0, 0], [0, 0.11, 0], [0, 0, 0.12]])
    array([[1, 0, 0],
           [0, 1, 0],
           [0, 0, 1]])
    """
    h_bv, U_bc = minkowski_reduce(h_cv)
    M_db, len2 = voronoi_steps(h_bv)
    order = np.argsort(np.round(len2, 12), kind='stable')
    return M_db[order] @ U_bc


def gradient_weights(h_cv, v):
    """(M_dc, g_d): the derivative along Cartesian axis v as
    sum_d g_d (h_d . grad), h_d = M_d h, over the Voronoi steps M_d.

    For each cell axis c, the derivative along the unit normal n_c of the
    plane of the other two cell vectors is built from the steps that leave
    that plane: the minimum-norm g with sum_d g_d h_d = n_c, posed on the
    components along (n_c, h_{c-1}, h_{c-2}) -- which selects the
    reference's stencil among the many exact ones
    (gpaw/fd_operators.py:310-334).  Then d/dv = sum_c (N^-1)_{vc} n_c . grad
    with N the matrix whose rows are the normals.
    """
    h_cv = np.asarray(h_cv, float)
    M_dc = find_neighbors(h_cv)
    h_dv = M_dc @ h_cv
    dual_vc = np.linalg.inv(h_cv)
    normals = (dual_vc / np.linalg.norm(dual_vc, axis=0)).T   # rows n_c
    g_cd = np.zeros((3, len(h_dv)))
    for c, n_v in enumerate(normals):
        frame = np.array([n_v, h_cv[(c + 2) % 3], h_cv[(c + 1) % 3]])
        leaving = np.abs(h_dv @ n_v) > 1e-10
        system = frame @ h_dv[leaving].T                       # 3 x m
        g, *_ = np.linalg.lstsq(system, np.array([1.0, 0.0, 0.0]),
                                rcond=None)
        g_cd[c, leaving] = g
    return M_dc, np.linalg.inv(normals)[v] @ g_cd


class Gradient(FDOperator):
    def __init__(self, grid, v, scale=1.0, n=1, dtype=float, xp=np):
        """scale * d/dr_v with O(h^2n) error on any cell
        (gpaw/fd_operators.py:284-354).

        Stencil layout: for each Voronoi step with a nonzero weight, the
        points +1..+n then -1..-n; no centre point.
        """
        gd = getattr(grid, '_gd', grid)
        M_dc, g_d = gradient_weights(gd.h_cv, v)
        w = np.array(derivatives[n - 1], float)
        offsets, coefs = [], []
        for M_c, g in zip(M_dc, scale * g_d):
            if abs(g) < 1e-10:
                continue
            o, c = _line(M_c, g * w, -1)
            offsets += o
            coefs += c
        super().__init__(coefs, offsets, gd, dtype, xp=xp)
        self.description = (
            f'Finite-difference {"xyz"[v]}-derivative with O(h^{2 * n}) '
            f'error ({self.npoints} points)')


# ----------------------------------------------------------- Mehrstellen
#
# The compact O(h^4) scheme A phi = B f for lap phi = f on an orthorhombic
# grid (Collatz).  With q_v = 1 / (12 h_v^2) and Q = q_x + q_y + q_z:
#   A: -16 Q at the centre, 10 q_v - 2 Q on the two faces along v,
#      q_u + q_w on the four edges of the (u, w) plane;
#   B: 1/2 at the centre, 1/12 on the six faces.
# For a cube A = (1 / 6h^2) [-24 (centre), 2 (faces), 1 (edges)].
# LaplaceA(gd, scale) is -scale * A (the Poisson solver passes
# scale = 1 / 4 pi for -lap / 4 pi, gpaw/poisson.py:371-378).

_FACES = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1),
          (0, 0, 1)]
_EDGES = [(0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1),
          (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1),
          (-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0)]


class LaplaceA(FDOperator):
    def __init__(self, gd, scale, dtype=float, xp=np):
        """Left-hand operator of the Mehrstellen pair, -scale * A; point
        order: centre, faces (x, y, z; - then +), edges of the yz, xz and
        xy planes."""
        if not gd.orthogonal:
            raise AssertionError('Mehrstellen stencil: orthogonal cells '
                                 'only')
        q = 1.0 / (12.0 * gd.h_cv.diagonal()**2)
        Q = q.sum()
        face = [10.0 * q[v] - 2.0 * Q for v in range(3)]
        plane = [q[1] + q[2], q[0] + q[2], q[0] + q[1]]
        A = [-16.0 * Q]
        A += [face[v] for v in range(3) for _ in range(2)]
        A += [plane[p] for p in range(3) for _ in range(4)]
        super().__init__([-scale * c for c in A],
                         [(0, 0, 0)] + _FACES + _EDGES, gd, dtype,
                         'O(h^4) Mehrstellen Laplacian (A)', xp=xp)


class LaplaceB(FDOperator):
    def __init__(self, gd, dtype=float, xp=np):
        """Right-hand operator of the Mehrstellen pair: 1/2 at the centre,
        1/12 on the six faces."""
        super().__init__([0.5] + [1.0 / 12.0] * 6, [(0, 0, 0)] + _FACES, gd,
                         dtype, 'O(h^4) Mehrstellen Laplacian (B)', xp=xp)


__all__ = ['FDOperator', 'Laplace', 'GUCLaplace', 'Gradient', 'LaplaceA',
           'LaplaceB', 'find_neighbors', 'minkowski_reduce', 'voronoi_steps',
           'laplacian_directions', 'gradient_weights', 'laplace',
           'derivatives', 'pi']
