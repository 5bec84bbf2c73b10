"""FFT / DST Poisson solver (FastPoissonSolver) on hipFFT.

Pinned by the equation itself.  The solver diagonalises the stencil in the
Fourier basis of periodic axes and the DST-I basis of zero-boundary axes,
i.e. it solves the periodic stencil on the oddly extended grid
(gpaw/poisson.py:721-771, 929-968).  For nn = 1 (and on fully periodic
grids) that is exactly the FD zero-boundary operator, so the result must
satisfy L phi = rho with the oracle's operator and equal the converged
multigrid solution; for nn > 1 on zero-boundary axes the equation holds for
the oracle's periodic stencil applied to the odd extension (the reference
test checks the DST path only for nn = 1, test_fastpoisson.py:77-93)."""
from math import pi

import numpy as np
import pytest

import oracle
from gpaw_amd.fd_operators import Laplace
from gpaw_amd.fft_poisson import FastPoissonSolver
from gpaw_amd.grid_descriptor import GridDescriptor
from gpaw_amd.poisson import FDPoissonSolver, PoissonSolver
from gpu_util import dev, host, say
from helpers import rand, relerr

pytestmark = pytest.mark.gpu

CASES = [((32, 24, 40), True, 3), ((24, 32, 16), False, 3),
         ((32, 16, 24), (True, False, True), 2), ((16, 24, 32), False, 1),
         ((20, 28, 36), (False, True, True), 3)]


@pytest.mark.parametrize('N,pbc,nn', CASES)
def test_equation_and_multigrid(N, pbc, nn):
    rng = np.random.default_rng(sum(N) + nn)
    gd = GridDescriptor(N, np.array(N) * 0.3, pbc)
    rho = rand(rng, tuple(gd.n_c))
    if gd.pbc_c.all():
        rho -= rho.mean()
    s = FastPoissonSolver(nn=nn)
    s.set_grid_descriptor(gd)
    phi = dev(np.zeros(tuple(gd.n_c)))
    say('fft poisson', N, pbc, nn)
    assert s.solve_neutral(phi, dev(rho)) == 1
    p = host(phi)
    lap = Laplace(gd, -0.25 / pi, nn)
    # the periodic stencil on the odd extension of every zero axis
    ext, sl = p, []
    for c in range(3):
        if gd.pbc_c[c]:
            sl.append(slice(None))
            continue
        z = np.zeros_like(np.take(ext, [0], axis=c))
        ext = np.concatenate([z, ext, z, -np.flip(ext, axis=c)], axis=c)
        sl.append(slice(1, 1 + gd.n_c[c]))
    m = np.array(ext.shape)
    lp = oracle.fd_apply(lap.coef_p, lap.offset_pc, m,
                         np.array([[0, 0]] * 3), ext)[tuple(sl)]
    assert relerr(lp, rho) < 1e-11
    if nn > 1 and not gd.pbc_c.all():
        return
    lp = oracle.fd_apply(lap.coef_p, lap.offset_pc, gd.n_c, gd.neighbor_cd,
                         p)
    assert relerr(lp, rho) < 1e-11
    # the converged multigrid solution
    mg = FDPoissonSolver(nn=nn, eps=1e-26 * float((rho**2).sum()) * gd.dv,
                         maxiter=500)
    mg.set_grid_descriptor(gd)
    q = dev(np.zeros(tuple(gd.n_c)))
    mg.solve_neutral(q, dev(rho))
    q = host(q)
    if gd.pbc_c.all():
        p, q = p - p.mean(), q - q.mean()
    assert relerr(p, q) < 1e-9


def test_periodic_against_numpy_fft():
    rng = np.random.default_rng(77)
    N = (24, 20, 28)
    gd = GridDescriptor(N, (6.0, 5.0, 7.0))
    rho = rand(rng, N)
    rho -= rho.mean()
    lap = Laplace(gd, -0.25 / pi, 3)
    k = np.meshgrid(*[np.fft.fftfreq(n) for n in N], indexing='ij')
    lam = np.zeros(N)
    for c, o in zip(lap.coef_p, lap.offset_pc):
        lam += c * (np.cos(2 * pi * sum(kk * oo for kk, oo in zip(k, o)))
                    - 1.0)
    inv = np.where(np.abs(lam) > 1e-10, 1.0 / np.where(lam == 0, 1, lam), 0)
    ref = np.fft.ifftn(np.fft.fftn(rho) * inv).real
    s = PoissonSolver('fast', nn=3)
    s.set_grid_descriptor(gd)
    phi = dev(np.zeros(N))
    s.solve(phi, dev(rho))
    assert relerr(host(phi), ref) < 1e-12
    assert 'FFT axes' in s.get_description()
    assert s.todict()['name'] == 'fast'
