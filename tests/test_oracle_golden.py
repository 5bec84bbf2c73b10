"""The oracle against the reference's golden vectors (bit for bit).

The fixtures were written by tests/golden/gen_golden.py from the reference
C sources (c/bmgs, c/bc.cpp, c/operators.cpp, c/transformers.cpp,
c/fd_preconditioner.cpp) compiled in the build container.
"""
import numpy as np
import pytest

import oracle
from gpaw_amd.fd_operators import Laplace
from gpaw_amd.grid_descriptor import GridDescriptor
from gpaw_amd.poisson import FDPoissonSolver
from gpaw_amd.preconditioner import Preconditioner
from gpaw_amd import workloads
from helpers import case, case_names, spec, tspec, vcycle


def _phase(c):
    return None if c['phase'].size == 0 else c['phase']


@pytest.mark.parametrize('name', ['lap3_pbc_real', 'lap3_zero_real',
                                  'lap2_mixed_cplx', 'lap1_pbc_cplx',
                                  'lap4_pbc_real', 'lap6_pbc_real',
                                  'grad3x_pbc_real', 'grad3y_zero_real',
                                  'grad2z_pbc_cplx', 'guc_hex_lap2_real',
                                  'guc_fcc_lap1_cplx', 'mehrstellenA_real'])
def test_apply(golden, name):
    c = case(golden('fd'), 'apply', name)
    out = oracle.fd_apply(c['coef'], c['offset'], c['n_c'], c['neighbor'],
                          c['in'], _phase(c))
    assert (out == c['out']).all()


def test_all_apply_cases_listed(golden):
    assert len(case_names(golden('fd'), 'apply')) == 12


@pytest.mark.parametrize('name', ['lap3_pbc_w23', 'lap1_zero_w1',
                                  'lap2_mixed_w07'])
def test_relax(golden, name):
    c = case(golden('fd'), 'relax', name)
    f = c['f'].copy()
    oracle.fd_relax(2, c['coef'], c['offset'], c['n_c'], c['neighbor'], f,
                    c['src'], int(c['sweeps']), float(c['w']))
    assert (f == c['out']).all()


def test_transfers(golden):
    npz = golden('fd')
    names = case_names(npz, 'transfer')
    assert len(names) == 16
    for name in names:
        c = case(npz, 'transfer', name)
        out = oracle.transform(int(c['interpolate']), int(c['k']),
                               c['size_in'], c['size_out'], c['pad'],
                               c['npad'], c['skip'], c['neighbor'], c['in'],
                               _phase(c))
        assert (out == c['out']).all(), name


@pytest.mark.parametrize('name', ['real_pbc', 'cplx_bloch', 'real_zero'])
def test_precond(golden, name):
    c = case(golden('precond'), 'precond', name)
    dtype = complex if c['residuals'].dtype == np.complex128 else float
    gd = GridDescriptor(tuple(c['N']), tuple(c['cell']), tuple(c['pbc']))
    P = Preconditioner(gd, Laplace(gd, -0.5, 3, dtype), dtype, block=2)
    assert P.step == float(c['step'])
    d0 = oracle.fd_precond(spec(P.kin0), spec(P.kin1), spec(P.kin2),
                           tspec(P.restrictor_object0),
                           tspec(P.restrictor_object1),
                           tspec(P.interpolator_object1),
                           tspec(P.interpolator_object2), c['residuals'],
                           P.step, _phase(c))
    assert (d0 == c['out']).all()


def test_poisson_zero16(golden):
    c = case(golden('poisson'), 'poisson', 'zero16')
    gd = GridDescriptor(tuple(c['N']), (float(c['cell']),) * 3, False)
    s = FDPoissonSolver(nn=3, eps=1e-30)
    s.set_grid_descriptor(gd)
    V = vcycle(s)
    phi = np.zeros(tuple(gd.n_c))
    V.phis[0] = phi
    V.rhos[0][:] = c['rho']
    errs = [V.iterate2() for _ in range(6)]
    assert (phi == c['phi']).all()
    assert errs == pytest.approx(list(c['errors']), rel=1e-12)


def test_poisson_pbc48(golden):
    c = case(golden('poisson'), 'poisson', 'pbc48')
    gd, rho = workloads.poisson_48()
    s = FDPoissonSolver(nn=3, eps=2e-10)
    s.set_grid_descriptor(gd)
    V = vcycle(s)
    background = gd.integrate(rho) / gd.dv / np.prod(gd.N_c)
    phi = np.zeros(tuple(gd.n_c))
    niter = V.solve_neutral(phi, rho - background, 2e-10, 1000, True)
    assert niter == int(c['niter']) == 13
    # numpy's dot may sum in another order with other BLAS thread counts
    assert V.last_error == pytest.approx(float(c['error']), rel=1e-12)
    assert (phi[::4, ::4, ::4] == c['phi_sub']).all()
    assert float((phi**2).sum()) == pytest.approx(float(c['phi_sum2']),
                                                  rel=1e-14)
