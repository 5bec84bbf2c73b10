"""Host-side logic of the drop-in classes (no GPU): coefficient tables and
their order, transformer paddings, multigrid hierarchy, grid layout, error
behaviour."""
import warnings
from math import pi

import numpy as np
import pytest

from gpaw_amd.fd_operators import (FDOperator, Gradient, Laplace,
                                   find_neighbors, minkowski_reduce)
from gpaw_amd.grid_descriptor import BadGridError, GridDescriptor
from gpaw_amd.poisson import FDPoissonSolver, PoissonConvergenceError
from gpaw_amd.transformers import Transformer, transformer_paddings


def test_laplace_nn3_closed_form():
    gd = GridDescriptor((12, 10, 14), (3.0, 3.3, 3.6))
    lap = Laplace(gd, -0.5, 3)
    h = gd.h_cv.diagonal()
    a = -0.5 / h**2
    c = [3 / 2, -3 / 20, 1 / 90]
    assert lap.npoints == 19 and lap.cfd
    assert lap.coef_p[0] == pytest.approx(-49 / 18 * a.sum(), rel=1e-14)
    # directions z, y, x; +1..+3 then -1..-3 each
    k = 1
    for axis in (2, 1, 0):
        for sign in (1, -1):
            for j in range(3):
                off = np.zeros(3, int)
                off[axis] = sign * (j + 1)
                assert (lap.offset_pc[k] == off).all()
                assert lap.coef_p[k] == pytest.approx(a[axis] * c[j],
                                                      rel=1e-13)
                k += 1
    assert lap.get_diagonal_element() == lap.coef_p[0]


def test_gradient_nn3_closed_form():
    gd = GridDescriptor((12, 10, 14), (3.0, 3.3, 3.6))
    g = Gradient(gd, 0, 1.0, 3)
    hx = gd.h_cv[0, 0]
    assert g.npoints == 6
    expect = np.array([3 / 4, -3 / 20, 1 / 60]) / hx
    assert np.allclose(g.coef_p[:3], expect, rtol=1e-13)
    assert np.allclose(g.coef_p[3:], -expect, rtol=1e-13)
    assert (g.offset_pc[:, 1:] == 0).all()
    assert g.get_diagonal_element() == 0.0


def test_find_neighbors_doc():
    assert (find_neighbors([[0.1, 0, 0], [0, 0.11, 0], [0, 0, 0.12]]) ==
            np.eye(3, dtype=int)).all()
    M = find_neighbors([[0.1, 0, 0], [0.2, 0.11, 0], [0, 0, 0.12]])
    assert M.tolist() == [[1, 0, 0], [-2, 1, 0], [0, 0, 1]]


def test_minkowski_reduce():
    h = np.array([[1.0, 0, 0], [7.3, 1.1, 0], [3.2, -4.1, 0.9]])
    B, U = minkowski_reduce(h)
    assert abs(round(np.linalg.det(U))) == 1
    assert np.allclose(U @ h, B)
    L = (B**2).sum(1)
    assert L.max() < 2.5 and (np.diff(L) >= -1e-12).all()


def test_grid_layout():
    gd = GridDescriptor((8, 6, 10), (4.0, 3.0, 5.0), (True, False, True))
    assert gd.n_c.tolist() == [8, 5, 10]
    assert gd.beg_c.tolist() == [0, 1, 0]
    assert gd.neighbor_cd.tolist() == [[0, 0], [-1, -1], [0, 0]]
    assert gd.dv == pytest.approx(60.0 / 480)
    assert gd.coarsen().N_c.tolist() == [4, 3, 5]
    with pytest.raises(ValueError):
        gd.coarsen().coarsen()
    with pytest.raises(BadGridError):
        GridDescriptor((1, 4, 4), pbc_c=False)


def test_transformer_paddings():
    fine = GridDescriptor((16, 16, 16))
    I, pad, npad, skip = transformer_paddings(fine, fine.coarsen(), 3)
    assert not I and pad[0].tolist() == [5, 4] and npad[0].tolist() == [4, 5]
    zfine = GridDescriptor((16, 16, 16), pbc_c=False)
    I, pad, npad, skip = transformer_paddings(zfine.coarsen(), zfine, 1)
    assert I and pad[0].tolist() == [1, 1] and skip[0].tolist() == [1, 0]
    I, pad, npad, skip = transformer_paddings(zfine, zfine.coarsen(), 1)
    assert pad[0].tolist() == [0, 0]
    with pytest.raises(ValueError, match='too small'):
        Transformer(GridDescriptor((8, 6, 10)),
                    GridDescriptor((16, 12, 20)), 4)


def test_poisson_hierarchy():
    gd = GridDescriptor((48, 48, 48), (18.0,) * 3)
    s = FDPoissonSolver(nn=3, eps=2e-10)
    s.set_grid_descriptor(gd)
    s._init()
    assert s.levels == 4
    assert [o.gd.N_c[0] for o in s.operators] == [48, 24, 12, 6, 3]
    assert s.presmooths == [2, 4, 4, 4, 8]
    assert s.postsmooths == [1, 4, 4, 4, 8]
    assert s.weights == [2 / 3, 1, 1, 1, 1]
    assert [o.npoints for o in s.operators] == [19, 7, 7, 7, 7]
    assert s.step == 0.66666666 / s.operators[0].get_diagonal_element()
    d = s.get_description()
    assert 'Jacobi multigrid solver, 5 levels' in d
    assert '3 x 3 x 3' in d
    assert s.todict() == {'name': 'fd', 'nn': 3, 'relax': 'J',
                          'eps': 2e-10}


def test_poisson_warns_on_coarse_grid():
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter('always')
        s = FDPoissonSolver()
        s.set_grid_descriptor(GridDescriptor((74, 74, 74)))
    assert any('cannot coarsen' in str(x.message) for x in w)


def test_errors():
    gd = GridDescriptor((4, 4, 4))
    with pytest.raises(AssertionError, match='Stencil longer'):
        FDOperator([1.0, 1.0], [[0, 0, 0], [5, 0, 0]], gd)
    with pytest.raises(NotImplementedError):
        FDPoissonSolver(relax='SOR')
    assert issubclass(PoissonConvergenceError, Exception)
    with pytest.raises(NotImplementedError):
        GridDescriptor((8, 8, 8), comm=type('C', (), {'size': 2})())
    with pytest.raises(NotImplementedError):
        Laplace(gd, 1.0, 9)


def test_madelung_constants():
    """Madelung potentials of point-charge lattices in a neutralising
    background (textbook Wigner-lattice constants, in units of the cubic
    lattice constant a), independent of the Ewald splitting."""
    from gpaw_amd.poisson import madelung
    a = 3.0
    assert madelung(np.eye(3) * a) * a == pytest.approx(2.8372974794806,
                                                        rel=1e-12)
    fcc = np.array([(0, 1, 1), (1, 0, 1), (1, 1, 0)]) * a / 2
    bcc = np.array([(-1, 1, 1), (1, -1, 1), (1, 1, -1)]) * a / 2
    assert madelung(fcc) * a == pytest.approx(4.5848622, rel=1e-7)
    assert madelung(bcc) * a == pytest.approx(3.6392334, rel=1e-7)
    hexa = np.array([(1, 0, 0), (0.5, 3**0.5 / 2, 0), (0, 0, 1.6)]) * 2.0
    vals = [madelung(hexa, G=G) for G in (0.6, 1.0, 1.7)]
    assert max(vals) - min(vals) < 1e-13 * abs(vals[0])


def test_gaussian_charge():
    """The compensating Gaussian holds unit charge and its potential solves
    lap phi = -4 pi rho (checked with the 19-point stencil inside the
    box, to its O(h^6) truncation error)."""
    import oracle
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.poisson import gaussian_charge
    gd = GridDescriptor((48, 48, 48), (6.0, 6.0, 6.0), False)
    rho, phi = gaussian_charge(gd)
    assert gd.integrate(rho) == pytest.approx(1.0, rel=1e-8)
    lap = Laplace(gd, 1.0, 3)
    lp = oracle.fd_apply(lap.coef_p, lap.offset_pc, gd.n_c, gd.neighbor_cd,
                         phi)
    inner = (slice(8, -8),) * 3
    err = np.abs(lp[inner] + 4 * np.pi * rho[inner]).max()
    assert err < 1e-3 * 4 * np.pi * rho.max()
    # off-centre charge: the width follows the nearest face
    rho2, _ = gaussian_charge(gd, centre=(2.0, 3.0, 3.5))
    assert gd.integrate(rho2) == pytest.approx(1.0, rel=1e-6)
