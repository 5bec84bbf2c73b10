"""Weighted FD operators (gpaw/wfd_operators.py; c/bmgs/wfd.cpp,
wrelax.cpp) against the oracle, which equals the reference WOperator bit
for bit (test_oracle_ref.py).  The operator of the dielectric Poisson
solver: eps * Laplace + sum_v d_v eps * Gradient_v, nn=3."""
import numpy as np
import pytest

import oracle
from gpaw_amd import _lib
from gpaw_amd.fd_operators import Gradient, Laplace
from gpaw_amd.grid_descriptor import GridDescriptor
from gpaw_amd.wfd_operators import WeightedFDOperator
from gpu_util import dev, host, say
from helpers import bloch_phases, rand, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-13


def dielectric_op(gd, dtype=float, nn=3):
    ops = [Laplace(gd, 1.0, nn, dtype)]
    ops += [Gradient(gd, v, 1.0, nn, dtype) for v in range(3)]
    return WeightedFDOperator(ops)


import contextlib


@contextlib.contextmanager
def wx_path(on):
    """Weighted operators through the weighted x-march kernel (on) or the
    one-point-per-thread kernels (off) while active."""
    lib = _lib.load()
    lib.gfd_set_tuning(b'wx_march', int(on))
    try:
        yield
    finally:
        lib.gfd_set_tuning(b'wx_march', 1)


def weights_of(rng, n):
    eps = 1.0 + rng.uniform(0, 78, n)
    return [eps] + [rng.standard_normal(n) for _ in range(3)]


@pytest.mark.parametrize('wx', [1, 0])
@pytest.mark.parametrize('pbc', [True, False, (True, False, True)])
@pytest.mark.parametrize('dtype', [float, complex])
def test_apply(pbc, dtype, wx):
    rng = np.random.default_rng(11)
    gd = GridDescriptor((16, 12, 20), (4.0, 3.0, 5.0), pbc)
    n = tuple(int(x) for x in gd.n_c)
    A = dielectric_op(gd, dtype)
    w = weights_of(rng, n)
    A.set_weights([dev(x) for x in w])
    a = rand(rng, (3,) + n, dtype)
    ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
    ref = oracle.wfd_apply(A.coef_ps, A.offset_pcs, np.array(w), n,
                           gd.neighbor_cd, a, ph)
    say('wfd apply', pbc, dtype.__name__, 'wx', wx)
    with wx_path(wx):
        d = dev(a)
        out = d.new_empty(d.shape)
        A.apply(d, out, ph)
        assert relerr(host(out), ref) < TOL
        # numpy arrays staged through the device give the same result
        outh = np.empty_like(a)
        A.apply(a, outh, ph)
        assert relerr(outh, ref) < TOL


@pytest.mark.parametrize('wx', [1, 0])
@pytest.mark.parametrize('pbc', [True, False])
def test_relax_and_diagonal(pbc, wx):
    rng = np.random.default_rng(12)
    gd = GridDescriptor((16, 16, 12), (4.0, 4.0, 3.0), pbc)
    n = tuple(int(x) for x in gd.n_c)
    A = dielectric_op(gd)
    w = weights_of(rng, n)
    A.set_weights([dev(x) for x in w])
    f = rng.standard_normal(n)
    src = rng.standard_normal(n)
    for nrelax in (1, 4):
        fr = f.copy()
        oracle.wfd_relax(2, A.coef_ps, A.offset_pcs, np.array(w), n,
                         gd.neighbor_cd, fr, src, nrelax, 0.6)
        say('wfd relax', pbc, nrelax, 'wx', wx)
        fd = dev(f)
        with wx_path(wx):
            A.relax(2, fd, dev(src), nrelax, 0.6)
        assert relerr(host(fd), fr) < TOL
    diag = abs(sum(wt * c[0] for wt, c in zip(w, A.coef_ps))).max()
    assert A.get_diagonal_element() == pytest.approx(diag, rel=1e-14)
    # lexicographic Gauss-Seidel, bit for bit (c/bmgs/wrelax.cpp:26-58)
    for nrelax in (1, 2):
        fr = f.copy()
        oracle.wfd_relax(1, A.coef_ps, A.offset_pcs, np.array(w), n,
                         gd.neighbor_cd, fr, src, nrelax, 0.6)
        say('wfd gauss-seidel', pbc, nrelax)
        fd = dev(f)
        A.relax(1, fd, dev(src), nrelax, 0.6)
        assert (host(fd) == fr).all(), relerr(host(fd), fr)


@pytest.mark.parametrize('nn,shape,pbc', [
    (3, (40, 48, 96), True), (1, (24, 36, 64), (False, True, True)),
    (2, (33, 20, 50), (True, False, False))])
def test_weighted_march_larger_grids(nn, shape, pbc):
    """The weighted x-march kernel on grids with several y tiles and x
    chunks (and ragged ones), apply over a batch and Jacobi sweeps."""
    rng = np.random.default_rng(13 + nn)
    gd = GridDescriptor(shape, tuple(0.3 * x for x in shape), pbc)
    n = tuple(int(x) for x in gd.n_c)
    A = dielectric_op(gd, float, nn)
    w = weights_of(rng, n)
    A.set_weights([dev(x) for x in w])
    a = rand(rng, (2,) + n)
    ref = oracle.wfd_apply(A.coef_ps, A.offset_pcs, np.array(w), n,
                           gd.neighbor_cd, a)
    say('wfd march', nn, shape, pbc)
    d = dev(a)
    out = d.new_empty(d.shape)
    A.apply(d, out)
    assert relerr(host(out), ref) < TOL
    f = rng.standard_normal(n)
    src = rng.standard_normal(n)
    fr = f.copy()
    oracle.wfd_relax(2, A.coef_ps, A.offset_pcs, np.array(w), n,
                     gd.neighbor_cd, fr, src, 3, 0.6)
    fd = dev(f)
    A.relax(2, fd, dev(src), 3, 0.6)
    assert relerr(host(fd), fr) < TOL
