"""Pins of the oracle (and of the drop-in coefficient / padding glue) that do
not depend on the reference build: the reference's own exact tests and
closed forms of the mathematics.

- polynomial exactness of GUCLaplace / Gradient   gpaw/test/fd_ops/test_laplace.py:15-103
- exact gradient vectors and edges                gpaw/test/fd_ops/test_gradient.py:27-110
- Bloch plane wave: lap e^{ikr} = -k^2 e^{ikr}    gpaw/test/core/test_grid.py:49-63
- integral conservation of the transfers          gpaw/test/fd_ops/test_transformations.py,
                                                  gpaw/test/fd_ops/test_non_periodic.py
- restriction is 1/8 the adjoint of interpolation (dense matrices)
- one Jacobi sweep equals its dense-matrix form
- the converged periodic V-cycle equals the spectral solution
  ifft(fft(rho) / lambda(k)), lambda(k) = sum_p c_p exp(i k.o_p)
"""
from math import cos, pi, sin

import numpy as np
import pytest

import oracle
from gpaw_amd.fd_operators import FDOperator, GUCLaplace, Gradient
from gpaw_amd.grid_descriptor import GridDescriptor
from gpaw_amd.poisson import FDPoissonSolver
from gpaw_amd.transformers import Transformer
from helpers import spec, tspec, vcycle

CELLS = [
    ('distorted hexagonal', 4,
     [(1, 0, 0), (1.02 * cos(pi / 3 - 0.02), 1.02 * sin(pi / 3 - 0.02), 0),
      (0, 0, 1.0)]),
    ('hexagonal', 4, [(1, 0, 0), (0.5, 3**0.5 / 2, 0), (0, 0, 1.1)]),
    ('fcc', 6, [(0, 1, 1), (1, 0, 1), (1, 1, 0)]),
    ('fcc-alternative', 6, [(1, 0, 0), (0.5, 3**0.5 / 2, 0),
                            (0.5, 3**0.5 / 6, (2 / 3)**0.5)]),
    ('bcc', 4, [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]),
    ('sc', 3, [1.1, 1.02, 1.03]),
    ('distorted sc', 6, [(1, 0, 0), (0.01, 1, 0), (0, 0.02, 1)]),
    ('rocksalt', 6, [(2 * np.sqrt(1 / 3), np.sqrt(1 / 8), -np.sqrt(1 / 24)),
                     (2 * np.sqrt(1 / 3), -np.sqrt(1 / 8), -np.sqrt(1 / 24)),
                     (2 * np.sqrt(1 / 3), 0, np.sqrt(1 / 6))]),
    ('nasty', 6, [(1, 0, 0), (0.0001, 1.03, 0), (0.0001, 0.0001, 1.0)]),
    ('Mike', 6, 5 * np.array([(5.565 / 28, 0, 0),
                              (0.0001 / 28, 5.565 / 28, 0),
                              (0.0001 / 24, 0.0001 / 24, 4.684 / 24)])),
    ('MnO', 6, [(1, 0.5, 0.5), (0.5, 1, 0.5), (0.5, 0.5, 1)]),
]


def centre_value(op, a):
    b = oracle.fd_apply(op.coef_p, op.offset_pc, op.gd.n_c,
                        op.gd.neighbor_cd, a)
    n = (op.gd.n_c[0] - 1) // 2 + 0
    return b


@pytest.mark.parametrize('name,D,cell', CELLS, ids=[c[0] for c in CELLS])
def test_polynomial_exactness(name, D, cell):
    for n in range(1, 5):
        N = 2 * n + 2
        gd = GridDescriptor((N, N, N), cell)
        r_gv = gd.get_grid_point_coordinates().transpose((1, 2, 3, 0))
        r_gv = r_gv - gd.cell_cv.sum(0) / 2
        lap = GUCLaplace(gd, n=n)
        grad_v = [Gradient(gd, v, n=n) for v in range(3)]
        assert lap.npoints == D * 2 * n + 1
        for m in range(0, 2 * n + 1):
            for ix in range(m + 1):
                for iy in range(m - ix + 1):
                    iz = m - ix - iy
                    a_g = (r_gv**(ix, iy, iz)).prod(3)
                    r = 2.0 if (m == 2 and max(ix, iy, iz) == 2) else 0.0
                    b = oracle.fd_apply(lap.coef_p, lap.offset_pc, gd.n_c,
                                        gd.neighbor_cd, a_g)
                    assert abs(b[n + 1, n + 1, n + 1] - r) < 2e-12
                    for v in range(3):
                        g = grad_v[v]
                        b = oracle.fd_apply(g.coef_p, g.offset_pc, gd.n_c,
                                            gd.neighbor_cd, a_g)
                        r = 1 if (m == 1 and [ix, iy, iz][v] == 1) else 0
                        assert abs(b[n + 1, n + 1, n + 1] - r) < 4e-12


def _apply(op, a, phase=None):
    return oracle.fd_apply(op.coef_p, op.offset_pc, op.gd.n_c,
                           op.gd.neighbor_cd, a, phase)


def test_gradient_values():
    gd = GridDescriptor((8, 1, 1), (8.0, 1.0, 1.0))
    a = gd.zeros()
    a[:, 0, 0] = np.arange(gd.beg_c[0], gd.end_c[0])
    d = _apply(Gradient(gd, v=0), a)
    assert not (d[:, 0, 0] - [-3, 1, 1, 1, 1, 1, 1, -3]).any()
    back = FDOperator([-1, 1], [[-1, 0, 0], [0, 0, 0]], gd)
    d = _apply(back, a)
    assert not (d[:, 0, 0] - [-7, 1, 1, 1, 1, 1, 1, 1]).any()

    gd = GridDescriptor((1, 8, 1), (1.0, 8.0, 1.0), (1, 0, 1))
    a = gd.zeros()
    a[0, :, 0] = np.arange(gd.beg_c[1], gd.end_c[1]) - 1
    d = _apply(Gradient(gd, v=1), a)
    assert d[0, 0, 0] == 0.5 and np.sum(d[0, :, 0]) == 3.0

    gd = GridDescriptor((1, 7, 1), ((1.0, 0.0, 0.0), (5.0, 5.0, 0.0),
                                    (0.0, 0.0, 0.7)))
    a = gd.zeros()
    a[0, :, 0] = np.arange(gd.beg_c[1], gd.end_c[1]) - 1
    d = _apply(Gradient(gd, v=1), a)
    assert abs(d[0, 0, 0] - -3.5) < 1e-12 and abs(np.sum(d[0, :, 0])) < 1e-12


@pytest.mark.parametrize('cell,size', [([8.7, 5.7, 7.7, 28, 120, 95],
                                        [60, 40, 54]),
                                       ([2, 2, 2, 120.1, 120.1, 90],
                                        [20, 20, 20])])
def test_gradient_strange_cell(cell, size):
    """x-derivative of x in skew cells (test_gradient.py:10-24)."""
    from numpy import cos, radians, sin, sqrt
    a, b, c, al, be, ga = cell
    al, be, ga = radians(al), radians(be), radians(ga)
    cy = (cos(al) - cos(be) * cos(ga)) / sin(ga)
    cell_cv = np.array([[a, 0, 0], [b * cos(ga), b * sin(ga), 0],
                        [c * cos(be), c * cy,
                         c * sqrt(1 - cos(be)**2 - cy**2)]])
    gd = GridDescriptor(size, cell_cv)
    grad = Gradient(gd, v=0)
    x = gd.get_grid_point_coordinates()[0]
    d = _apply(grad, np.ascontiguousarray(x))
    n1, n2, n3 = np.array(size) // 2
    assert d[n1, n2, n3] == pytest.approx(1.0)


def test_bloch_plane_wave():
    a = 2.5
    gd = GridDescriptor((24, 8, 8), (a, a, a))
    kpt_c = np.array([1 / 3, 0, 0])
    r = gd.get_grid_point_coordinates()
    k_v = 2 * pi * kpt_c / a
    f = np.exp(1j * np.einsum('v,vxyz->xyz', k_v, r))
    phase_cd = np.exp(2j * pi * np.array([1, -1])[None, :] * kpt_c[:, None])
    lap = GUCLaplace(gd, n=2, dtype=complex)
    g = _apply(lap, f, phase_cd)
    k = 2 * pi / a / 3
    assert abs(g.conj() * f + k**2).max() < 1e-6


def test_transfer_integrals():
    n = 20
    gd = GridDescriptor((n, n, n))
    rng = np.random.RandomState(8)
    a = gd.empty()
    a[:] = rng.random(a.shape)
    ac = a + 1j * rng.random(a.shape)
    phase = np.ones((3, 2), complex)
    for gd2 in (gd.refine(), gd.coarsen()):
        for k in [2, 4, 6, 8]:
            t = Transformer(gd, gd2, k // 2)
            s = tspec(t)
            b = oracle.transform(s.interpolate, s.k, s.size_in, s.size_out,
                                 s.pad_cd, s.npad_cd, s.skip_cd,
                                 s.neighbor_cd, a)
            assert abs(gd.integrate(a) - gd2.integrate(b)) < 1e-14
            t = Transformer(gd, gd2, k // 2, complex)
            b = oracle.transform(s.interpolate, s.k, s.size_in, s.size_out,
                                 s.pad_cd, s.npad_cd, s.skip_cd,
                                 s.neighbor_cd, ac, phase)
            assert abs(gd.integrate(ac) - gd2.integrate(b)) < 1e-14


def test_transfer_integrals_zero_boundary():
    n = 20
    gd1 = GridDescriptor((n, n, n), (8.0, 8.0, 8.0), pbc_c=0)
    rng = np.random.default_rng(8)
    a1 = gd1.zeros()
    a1[:] = rng.random(a1.shape)
    gd2 = gd1.refine()
    s = tspec(Transformer(gd1, gd2))
    a2 = oracle.transform(s.interpolate, s.k, s.size_in, s.size_out, s.pad_cd,
                          s.npad_cd, s.skip_cd, s.neighbor_cd, a1)
    assert abs(gd1.integrate(a1) - gd2.integrate(a2)) < 1e-10
    for sl in (0, -1):
        a2[sl] = 0.0
        a2[:, sl] = 0.0
        a2[:, :, sl] = 0.0
    s = tspec(Transformer(gd2, gd1))
    a1 = oracle.transform(s.interpolate, s.k, s.size_in, s.size_out, s.pad_cd,
                          s.npad_cd, s.skip_cd, s.neighbor_cd, a2)
    assert abs(gd1.integrate(a1) - gd2.integrate(a2)) < 1e-10


def _dense(fn, nin):
    cols = []
    for i in range(nin):
        e = np.zeros(nin)
        e[i] = 1.0
        cols.append(fn(e).ravel())
    return np.array(cols).T


@pytest.mark.parametrize('nn', [1, 2, 3, 4])
def test_restriction_is_adjoint_of_interpolation(nn):
    N = (12, 10, 14) if nn <= 2 else (20, 16, 18)
    fine = GridDescriptor(N, (3.0, 2.5, 3.5))
    coarse = fine.coarsen()
    R = tspec(Transformer(fine, coarse, nn))
    I = tspec(Transformer(coarse, fine, nn))
    tr = lambda s, shape: (lambda e: oracle.transform(  # noqa: E731
        s.interpolate, s.k, s.size_in, s.size_out, s.pad_cd, s.npad_cd,
        s.skip_cd, s.neighbor_cd, e.reshape(shape)))
    Rm = _dense(tr(R, tuple(fine.n_c)), int(np.prod(fine.n_c)))
    Im = _dense(tr(I, tuple(coarse.n_c)), int(np.prod(coarse.n_c)))
    assert np.abs(Rm - Im.T / 8).max() < 1e-15
    # interpolation reproduces the coarse values at even fine points
    ev = Im.reshape(tuple(fine.n_c) + (-1,))[::2, ::2, ::2]
    assert (ev.reshape(Im.shape[1], -1) == np.eye(Im.shape[1])).all()


def test_jacobi_sweep_dense():
    gd = GridDescriptor((6, 5, 7), (2.0, 1.8, 2.3), (True, False, True))
    from gpaw_amd.fd_operators import Laplace
    op = Laplace(gd, -0.25 / pi, 2)
    n = int(np.prod(gd.n_c))
    A = _dense(lambda e: _apply(op, e.reshape(tuple(gd.n_c))), n)
    c0 = op.coef_p[0]
    rng = np.random.default_rng(3)
    f = rng.standard_normal(tuple(gd.n_c))
    s = rng.standard_normal(tuple(gd.n_c))
    w = 2 / 3
    expect = (1 - w) * f.ravel() + w * (s.ravel() -
                                        (A - c0 * np.eye(n)) @ f.ravel()) / c0
    f2 = f.copy()
    oracle.fd_relax(2, op.coef_p, op.offset_pc, gd.n_c, gd.neighbor_cd, f2,
                    s, 1, w)
    assert np.abs(f2.ravel() - expect).max() < 1e-13 * np.abs(expect).max()


def spectral_solution(op, rho):
    N = np.array(rho.shape)
    k = [2 * pi * np.fft.fftfreq(n) for n in N]
    K = np.meshgrid(*k, indexing='ij')
    lam = np.zeros(rho.shape, complex)
    for c, o in zip(op.coef_p, op.offset_pc):
        lam += c * np.exp(1j * (K[0] * o[0] + K[1] * o[1] + K[2] * o[2]))
    rq = np.fft.fftn(rho)
    lam[0, 0, 0] = 1.0
    rq[0, 0, 0] = 0.0
    return np.fft.ifftn(rq / lam).real


@pytest.mark.parametrize('nn', [1, 3])
def test_vcycle_converges_to_spectral_solution(nn):
    gd = GridDescriptor((16, 16, 16), (5.0, 5.0, 5.0))
    rng = np.random.default_rng(5)
    rho = rng.standard_normal(tuple(gd.n_c))
    rho -= rho.mean()
    s = FDPoissonSolver(nn=nn, eps=1e-24)
    s.set_grid_descriptor(gd)
    V = vcycle(s)
    phi = np.zeros(tuple(gd.n_c))
    niter = V.solve_neutral(phi, rho, 1e-24, 200, True)
    assert niter < 100
    ref = spectral_solution(s.operators[0], rho)
    ref -= ref.mean()
    assert np.abs(phi - ref).max() < 1e-9 * np.abs(ref).max()


def test_mehrstellen_pair_exact_to_degree_5():
    """LaplaceA p = LaplaceB lap(p) (scale -1) at interior points for every
    monomial of degree <= 5 on an anisotropic grid: the defining property
    of the O(h^4) compact scheme.  Degree 6 must fail (the check is
    sensitive)."""
    from gpaw_amd.fd_operators import LaplaceA, LaplaceB
    gd = GridDescriptor((14, 16, 12), (14 * 0.31, 16 * 0.27, 12 * 0.35),
                        False)
    A = LaplaceA(gd, -1.0)
    B = LaplaceB(gd)
    x, y, z = gd.get_grid_point_coordinates()
    inner = (slice(2, -2),) * 3

    def d2(e, v):
        return e * (e - 1) * v**(e - 2) if e >= 2 else 0 * v

    def mismatch(a, b, c):
        p = x**a * y**b * z**c
        lap = (d2(a, x) * y**b * z**c + x**a * d2(b, y) * z**c +
               x**a * y**b * d2(c, z))
        Ap = oracle.fd_apply(A.coef_p, A.offset_pc, gd.n_c, gd.neighbor_cd,
                             p)[inner]
        Bl = oracle.fd_apply(B.coef_p, B.offset_pc, gd.n_c, gd.neighbor_cd,
                             lap)[inner]
        # against the size of the terms the stencils sum
        scale = (np.abs(A.coef_p).sum() * np.abs(p).max() +
                 np.abs(B.coef_p).sum() * np.abs(lap).max())
        return np.abs(Ap - Bl).max() / scale

    for a in range(6):
        for b in range(6 - a):
            for c in range(6 - a - b):
                assert mismatch(a, b, c) < 1e-12, (a, b, c)
    assert mismatch(6, 0, 0) > 1e-9
    assert mismatch(2, 4, 0) > 1e-9


@pytest.mark.parametrize('cell,pairs', [
    ([1.1, 1.02, 1.03], 3),                        # simple cubic: 6 faces
    ([(0, 1, 1), (1, 0, 1), (1, 1, 0)], 6),        # fcc: 12 faces
    ([(-1, 1, 1), (1, -1, 1), (1, 1, -1)], 7),     # bcc: 14 faces
    ([(1, 0, 0), (0.5, 3**0.5 / 2, 0), (0, 0, 1.1)], 4)])  # hexagonal: 8
def test_voronoi_face_counts(cell, pairs):
    """Number of faces of the Voronoi cell (the Wigner-Seitz cell) of the
    textbook lattices, and every neighbour step at most as long as the
    next."""
    from gpaw_amd.fd_operators import find_neighbors
    h = np.asarray(cell, float)
    h = np.diag(h) if h.ndim == 1 else h
    M = find_neighbors(h)
    assert len(M) == pairs
    L = ((M @ h)**2).sum(1)
    assert (np.diff(L) >= -1e-12).all()


def test_laplacian_tensor_identity():
    """sum_d a_d h_d h_d^T = 1 for the directions and weights of every
    test cell."""
    from gpaw_amd.fd_operators import laplacian_directions
    for name, D, cell in CELLS:
        gd = GridDescriptor((8, 8, 8), cell)
        M, a = laplacian_directions(gd.h_cv)
        h = M @ gd.h_cv
        T = np.einsum('d,dv,dw->vw', a, h, h)
        assert len(a) == D, name
        assert np.abs(T - np.eye(3)).max() < 1e-12, name
