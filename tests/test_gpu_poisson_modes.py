"""Graph-replayed vs eager V-cycles, and the mixed-precision solver against
the FP64 solution and the oracle (spectral closed form)."""
import numpy as np
import pytest

from gpaw_amd import _lib, workloads
from gpaw_amd.grid_descriptor import GridDescriptor
from gpaw_amd.poisson import FDPoissonSolver
from gpu_util import dev, host, say
from helpers import rand, relerr

pytestmark = pytest.mark.gpu


def _solver(gd, **kw):
    s = FDPoissonSolver(**kw)
    s.set_grid_descriptor(gd)
    s._handle()
    return s


def test_graph_replays_match_eager():
    gd, rho = workloads.poisson_48()
    s = _solver(gd, nn=3, eps=2e-10)
    lib = _lib.load()
    out = {}
    for mode in (0, 1, 1):  # eager, graph, graph again (replays)
        lib.gfd_poisson_set_graph(s._native, mode)
        phi = dev(np.zeros(tuple(gd.n_c)))
        say('poisson graph mode', mode)
        n = s.solve(phi, dev(rho))
        out.setdefault(mode, []).append((n, host(phi)))
    (n0, p0), = out[0]
    for n1, p1 in out[1]:
        assert n1 == n0
        assert (p1 == p0).all()


def test_mixed_precision_reaches_tolerance():
    rng = np.random.default_rng(8)
    gd = GridDescriptor((64, 48, 32), (12.0, 9.0, 6.0))
    rho = rand(rng, tuple(gd.n_c))
    rho -= rho.mean()
    s = _solver(gd, nn=3, eps=1e-30)
    drho = dev(rho)
    # a 3-level hierarchy ends on 16 x 12 x 8 points, where 8 Jacobi
    # sweeps leave the smoothest error: it converges, in many more steps
    for levels, nmax in ((None, 40), (4, 40), (3, 100)):
        phi = dev(np.zeros(tuple(gd.n_c)))
        say('mixed levels', levels)
        n, rel = s.solve_mixed(phi, drho, 1e-10, levels=levels)
        assert rel <= 1e-10 and n <= nmax
        # the residual of the result, recomputed in FP64
        r = dev(np.zeros(tuple(gd.n_c)))
        n2 = s.operators[0].residual(phi, drho, r, norm=True)
        assert (n2 / float((rho**2).sum()))**0.5 <= 1.001e-10
    # same answer as the FP64 V-cycle driven to the same residual
    phi64 = dev(np.zeros(tuple(gd.n_c)))
    s.eps = 1e-20 * float((rho**2).sum()) * gd.dv
    s.solve_neutral(phi64, drho)
    a, b = host(phi), host(phi64)
    a -= a.mean()
    b -= b.mean()
    assert relerr(a, b) < 1e-8


@pytest.mark.parametrize('pbc', [True, (True, False, True)])
def test_refinement_with_fp64_inner_cycles(pbc):
    """The refinement solver with the same inner cycles in FP64
    (fp64_inner): it reaches the tolerance, graph replay and eager launches
    agree bit for bit, and its solution is the FP32-inner one and the
    plain FP64 V-cycle one to the accuracy both were driven to."""
    rng = np.random.default_rng(81)
    gd = GridDescriptor((64, 48, 32), (12.0, 9.0, 6.0), pbc)
    rho = rand(rng, tuple(gd.n_c))
    if gd.pbc_c.all():
        rho -= rho.mean()
    s = _solver(gd, nn=3, eps=1e-30)
    drho = dev(rho)
    kw = dict(levels=4, pre=2, post=2, coarse_sweeps=8, w0=0.8, inner=2)
    out = {}
    for fp64 in (False, True):
        for graph in (True, False):
            _lib.load().gfd_poisson_set_graph(s._native, int(graph))
            phi = dev(np.zeros(tuple(gd.n_c)))
            say('refinement fp64_inner', fp64, 'graph', graph, pbc)
            n, rel = s.solve_mixed(phi, drho, 1e-10, fp64_inner=fp64, **kw)
            assert rel <= 1e-10
            r = dev(np.zeros(tuple(gd.n_c)))
            n2 = s.operators[0].residual(phi, drho, r, norm=True)
            assert (n2 / float((rho**2).sum()))**0.5 <= 1.001e-10
            out[fp64, graph] = (n, host(phi))
    _lib.load().gfd_poisson_set_graph(s._native, 1)
    for fp64 in (False, True):
        assert out[fp64, True][0] == out[fp64, False][0]
        assert (out[fp64, True][1] == out[fp64, False][1]).all()
    a, b = out[False, True][1], out[True, True][1]
    if gd.pbc_c.all():
        a, b = a - a.mean(), b - b.mean()
    assert relerr(a, b) < 1e-8
    # FP64 inner cycles need no more refinement steps than FP32 ones
    assert out[True, True][0] <= out[False, True][0]


@pytest.mark.parametrize('shape,pbc', [((48, 48, 48), True),
                                       ((40, 24, 32), False),
                                       ((64, 48, 32), (True, False, True))])
def test_coarse_lds_matches_global(shape, pbc):
    """The LDS-resident coarse levels against the per-level launches: the
    same V-cycle up to rounding, FP64 and mixed precision."""
    rng = np.random.default_rng(31)
    gd = GridDescriptor(shape, np.array(shape) * 0.3, pbc)
    rho = rand(rng, tuple(gd.n_c))
    if gd.pbc_c.all():
        rho -= rho.mean()
    lib = _lib.load()
    res = {}
    try:
        for on in (0, 1):
            _lib.check(lib.gfd_set_tuning(b'coarse_lds', on), 'tuning')
            s = _solver(gd, nn=3, eps=1e-24 * float((rho**2).sum()) * gd.dv,
                        maxiter=300)
            phi = dev(np.zeros(tuple(gd.n_c)))
            say('coarse lds', on, shape, pbc)
            n = s.solve_neutral(phi, dev(rho))
            phi32 = dev(np.zeros(tuple(gd.n_c)))
            m, rel = s.solve_mixed(phi32, dev(rho), 1e-10)
            res[on] = (n, host(phi), m, rel, host(phi32))
    finally:
        lib.gfd_set_tuning(b'coarse_lds', 1)
    (n0, p0, m0, r0, q0), (n1, p1, m1, r1, q1) = res[0], res[1]
    assert abs(n1 - n0) <= 1
    assert relerr(p1, p0) < 1e-10
    assert r0 <= 1e-10 and r1 <= 1e-10 and abs(m1 - m0) <= 2
    a, b = q1 - q1.mean(), q0 - q0.mean()
    assert relerr(a, b) < 1e-8
