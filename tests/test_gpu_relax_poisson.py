"""Jacobi relaxation and the native multigrid Poisson driver against the
reference fixtures and the oracle V-cycle."""
import numpy as np
import pytest

import oracle
from gpaw_amd import workloads
from gpaw_amd.fd_operators import Laplace
from gpaw_amd.grid_descriptor import GridDescriptor
from gpaw_amd.poisson import FDPoissonSolver, PoissonConvergenceError
from gpu_util import dev, host, kernel_path, say
from helpers import case, case_names, rand, relerr, vcycle

pytestmark = pytest.mark.gpu
TOL = 1e-13


@pytest.mark.parametrize('path', ['row', 'xmarch', 'general'])
def test_relax_golden(golden, path):
    npz = golden('fd')
    with kernel_path(path):
        for name in case_names(npz, 'relax'):
            c = case(npz, 'relax', name)
            pbc = tuple(bool(x >= 0) for x in c['neighbor'][:, 0])
            N = tuple(int(n) + (0 if p else 1) for n, p in zip(c['n_c'], pbc))
            gd = GridDescriptor(N, (1.0, 1.0, 1.0), pbc)
            from gpaw_amd.fd_operators import FDOperator
            op = FDOperator(c['coef'], c['offset'], gd)
            f = dev(c['f'])
            say('golden relax', name, path)
            op.relax(2, f, dev(c['src']), int(c['sweeps']), float(c['w']))
            assert relerr(host(f), c['out']) < TOL, name


def test_relax_sweep_counts():
    rng = np.random.default_rng(3)
    gd = GridDescriptor((20, 24, 28), (4.0, 4.4, 5.0), (True, False, True))
    op = Laplace(gd, -0.25 / np.pi, 3)
    f0 = rand(rng, tuple(gd.n_c))
    s = rand(rng, tuple(gd.n_c))
    for n in (0, 1, 2, 5):
        ref = f0.copy()
        oracle.fd_relax(2, op.coef_p, op.offset_pc, gd.n_c, gd.neighbor_cd,
                        ref, s, n, 2 / 3)
        for path in ('row', 'xmarch', 'general'):
            f = dev(f0)
            say('relax sweeps', n, path)
            with kernel_path(path):
                op.relax(2, f, dev(s), n, 2 / 3)
            assert relerr(host(f), ref) < TOL


GS_CASES = [
    # (kind, N, pbc): small grids run in one workgroup, the 72 x 64 one
    # with one launch per wavefront step
    ('nn1', (12, 10, 14), True),
    ('nn3', (14, 12, 10), (True, False, True)),
    ('nn3', (72, 64, 16), True),
    ('mehrstellen', (12, 14, 10), False),
    ('mehrstellen', (72, 66, 8), (False, True, True)),
    ('hexagonal nn2', (12, 12, 10), True),
]


def _gs_operator(kind, gd):
    from gpaw_amd.fd_operators import GUCLaplace, LaplaceA
    if kind == 'mehrstellen':
        return LaplaceA(gd, 0.25 / np.pi)
    if kind.startswith('hexagonal'):
        return GUCLaplace(gd, -0.25 / np.pi, 2)
    return Laplace(gd, -0.25 / np.pi, int(kind[2]))


@pytest.mark.parametrize('kind,N,pbc', GS_CASES)
def test_gauss_seidel_bit_exact(kind, N, pbc):
    """Lexicographic Gauss-Seidel (relax method 1) in wavefront order is the
    reference's sweep bit for bit: same neighbour values, same sums."""
    rng = np.random.default_rng(40 + GS_CASES.index((kind, N, pbc)))
    cell = (np.array([(1, 0, 0), (0.5, 3**0.5 / 2, 0), (0, 0, 1.1)]) *
            np.array(N)[:, None] * 0.3 if kind.startswith('hexagonal')
            else np.array(N) * 0.3)
    gd = GridDescriptor(N, cell, pbc)
    op = _gs_operator(kind, gd)
    f0 = rand(rng, tuple(gd.n_c))
    src = rand(rng, tuple(gd.n_c))
    for n in (1, 3):
        ref = f0.copy()
        oracle.fd_relax(1, op.coef_p, op.offset_pc, gd.n_c, gd.neighbor_cd,
                        ref, src, n, 0.7)
        f = dev(f0)
        say('gauss-seidel', kind, N, pbc, n)
        op.relax(1, f, dev(src), n, 0.7)
        got = host(f)
        assert (got == ref).all(), relerr(got, ref)


@pytest.mark.parametrize('nn,pbc', [(3, False), (1, True), ('M', False),
                                    ('M', True)])
def test_vcycles_gauss_seidel_and_mehrstellen(nn, pbc):
    """FDPoissonSolver(relax='GS') and nn='M' (B rho, then V-cycles on
    LaplaceA) against the oracle V-cycle with the same hierarchy, four
    cycles from phi = 0."""
    relax = 'J' if nn == 'M' and pbc else 'GS'
    rng = np.random.default_rng(17)
    gd = GridDescriptor((16, 16, 24), (4.0, 4.0, 6.0), pbc)
    rho = rand(rng, tuple(gd.n_c))
    if gd.pbc_c.all():
        rho -= rho.mean()
    s = FDPoissonSolver(nn=nn, relax=relax, eps=1e-30)
    s.set_grid_descriptor(gd)
    V = vcycle(s)
    rhs = rho
    if nn == 'M':
        B = s.B
        rhs = oracle.fd_apply(B.coef_p, B.offset_pc, gd.n_c, gd.neighbor_cd,
                              rho)
    V.phis[0] = np.zeros(tuple(gd.n_c))
    V.rhos[0][:] = rhs
    phi = dev(np.zeros(tuple(gd.n_c)))
    drhs = dev(rhs)
    for i in range(4):
        e_ref = V.iterate2()
        say('V-cycle', nn, relax, pbc, i)
        e = s.iterate2(s.step, phi=phi, rho=drhs)
        assert relerr(host(phi), V.phis[0]) < TOL
        assert e == pytest.approx(e_ref, rel=1e-10)
    # the full solve converges and B is applied inside solve_neutral
    phi2 = dev(np.zeros(tuple(gd.n_c)))
    s.eps = 1e-20
    n = s.solve_neutral(phi2, dev(rho))
    assert 1 < n < 200
    r = dev(np.zeros(tuple(gd.n_c)))
    assert s.operators[0].residual(phi2, drhs, r, norm=True) * gd.dv <= 1e-20


def test_vcycle_golden_zero16(golden):
    c = case(golden('poisson'), 'poisson', 'zero16')
    gd = GridDescriptor(tuple(c['N']), (float(c['cell']),) * 3, False)
    s = FDPoissonSolver(nn=3, eps=1e-30)
    s.set_grid_descriptor(gd)
    phi = dev(np.zeros(tuple(gd.n_c)))
    rho = dev(c['rho'])
    errs = []
    for i in range(6):
        say('zero16 V-cycle', i)
        errs.append(s.iterate2(s.step if hasattr(s, 'step') else 1.0,
                               phi=phi, rho=rho))
    assert relerr(host(phi), c['phi']) < TOL
    assert errs == pytest.approx(list(c['errors']), rel=1e-11)


def test_vcycle_matches_oracle_mixed_bc():
    rng = np.random.default_rng(21)
    gd = GridDescriptor((32, 16, 24), (6.0, 3.0, 4.5), (True, False, True))
    rho = rand(rng, tuple(gd.n_c))
    s = FDPoissonSolver(nn=2, eps=1e-30)
    s.set_grid_descriptor(gd)
    V = vcycle(s)
    phi_ref = np.zeros(tuple(gd.n_c))
    V.phis[0] = phi_ref
    V.rhos[0][:] = rho
    phi = dev(np.zeros(tuple(gd.n_c)))
    drho = dev(rho)
    for i in range(4):
        say('mixed-bc V-cycle', i)
        e_ref = V.iterate2()
        e = s.iterate2(1.0, phi=phi, rho=drho)
        assert e == pytest.approx(e_ref, rel=1e-10)
        assert relerr(host(phi), phi_ref) < 1e-12


def test_poisson_48_solve(golden):
    c = case(golden('poisson'), 'poisson', 'pbc48')
    gd, rho = workloads.poisson_48()
    s = FDPoissonSolver(nn=3, eps=2e-10)
    s.set_grid_descriptor(gd)
    phi = dev(np.zeros(tuple(gd.n_c)))
    say('poisson_48 solve')
    niter = s.solve(phi, dev(rho))
    assert niter == int(c['niter'])
    assert s.last_error == pytest.approx(float(c['error']), rel=1e-9)
    assert relerr(host(phi)[::4, ::4, ::4], c['phi_sub']) < TOL
    # warm start from the converged potential
    assert s.solve(phi, dev(rho)) == 1
    # host arrays work too (staged through the device)
    phi_h = np.zeros(tuple(gd.n_c))
    assert s.solve(phi_h, rho) == int(c['niter'])
    assert relerr(phi_h[::4, ::4, ::4], c['phi_sub']) < TOL


def test_poisson_convergence_error():
    gd, rho = workloads.poisson_48()
    s = FDPoissonSolver(nn=3, eps=1e-30, maxiter=3)
    s.set_grid_descriptor(gd)
    with pytest.raises(PoissonConvergenceError):
        s.solve(dev(np.zeros(tuple(gd.n_c))), dev(rho))


def _blob(gd, centre, b, Q):
    r_vG = gd.get_grid_point_coordinates() - np.asarray(centre)[:, None,
                                                                 None, None]
    r = np.sqrt(np.maximum((r_vG**2).sum(0), 1e-12))
    from scipy.special import erf
    rho = Q * (b / np.pi)**1.5 * np.exp(-b * r * r)
    return rho, Q * erf(np.sqrt(b) * r) / r


@pytest.mark.parametrize('charge_center', [False, True])
def test_charged_zero_boundary(charge_center):
    """A charged Gaussian blob in a box with zero boundaries: the
    compensating Gaussian carries the charge, the V-cycles solve for the
    neutral rest, and the result is the free-space potential
    Q erf(sqrt(b) r) / r of the blob (to the stencil's accuracy)."""
    gd = GridDescriptor((48, 48, 48), (12.0, 12.0, 12.0), False)
    centre = (6.6, 5.5, 6.2) if charge_center else (6.0, 6.0, 6.0)
    rho, exact = _blob(gd, centre, 1.5, 2.0)
    s = FDPoissonSolver(nn=3, eps=1e-18, use_charge_center=charge_center)
    s.set_grid_descriptor(gd)
    phi = dev(np.zeros(tuple(gd.n_c)))
    say('charged zero-boundary solve, charge centre', charge_center)
    n = s.solve(phi, dev(rho))
    assert 1 < n < 200
    err = relerr(host(phi), exact)
    say(f'  error against the free-space potential {err:.3g}')
    assert err < 2e-3
    # zero_initial_phi restarts from 0 and gives the same potential
    n2 = s.solve(phi, dev(rho), zero_initial_phi=True)
    assert relerr(host(phi), exact) < 2e-3 and n2 >= 1


def test_charged_periodic_madelung():
    """Charged periodic cell: the uniform background takes the charge out;
    with use_charged_periodic_corrections the potential is shifted by
    charge x Madelung constant of the cell."""
    from gpaw_amd.poisson import madelung
    gd = GridDescriptor((32, 32, 32), (8.0, 8.0, 9.0), True)
    rho, _ = _blob(gd, (4.0, 4.0, 4.5), 2.0, -1.5)
    q = gd.integrate(rho)
    out = []
    for corr in (False, True):
        s = FDPoissonSolver(nn=3, eps=1e-20,
                            use_charged_periodic_corrections=corr)
        s.set_grid_descriptor(gd)
        phi = dev(np.zeros(tuple(gd.n_c)))
        say('charged periodic, Madelung correction', corr)
        s.solve(phi, dev(rho))
        out.append(host(phi))
    shift = out[1] - out[0]
    assert np.abs(shift - q * madelung(gd.cell_cv)).max() < 1e-10
