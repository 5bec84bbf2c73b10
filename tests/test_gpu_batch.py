"""Batched uneven grids: one launch per stage over every grid, against the
oracle grid by grid; the batched mixed-precision solve against the
single-grid FP64 solve."""
import numpy as np
import pytest

import oracle
from gpaw_amd import workloads
from gpaw_amd.batch_poisson import BatchedPoissonSolver
from gpaw_amd.grid_descriptor import GridDescriptor
from gpu_util import batch_kernel, batch_xfer, dev, host, say
from helpers import relerr, spec, tspec

pytestmark = pytest.mark.gpu
TOL = 1e-13

SHAPES = [((24, 16, 40), True), ((40, 32, 16), False), ((16, 48, 24), True),
          ((32, 24, 32), False), ((48, 16, 16), (True, False, True))]


def batch(levels=3):
    gds = [GridDescriptor(N, np.array(N) * 0.35, pbc) for N, pbc in SHAPES]
    b = BatchedPoissonSolver(nn=3, levels=levels)
    b.set_grid_descriptors(gds)
    return b


@pytest.fixture(params=['regs', 'march'])
def kernel(request):
    """Batches tuned under either batched stencil kernel."""
    with batch_kernel(request.param):
        yield request.param


def test_apply_per_level(kernel):
    import torch
    b = batch()
    tiles = b.tune()
    say('batched apply tiles', kernel, tiles)
    assert all(t['kernel'] == kernel for t in tiles.values())
    _check_apply(b)


@pytest.mark.parametrize('nt,py,zmax,xc,pf', [
    (256, 1, 64, 8, 1), (256, 2, 32, 5, 2), (512, 4, 128, 16, 1),
    (512, 4, 128, 7, 2), (128, 2, 16, 3, 2), (64, 1, 16, 40, 1)])
def test_apply_forced_tiles(nt, py, zmax, xc, pf):
    """Every x-march variant (rows per thread, loads in flight, z parts
    narrower than a row, ragged x chunks) on the same five grids."""
    from gpaw_amd import _lib
    lib = _lib.load()
    code = nt | py << 12 | zmax << 16 | xc << 28 | pf << 40
    _lib.check(lib.gfd_set_tuning(b'batch_tile', code), 'batch_tile')
    try:
        b = batch()
        tiles = b.tune()
    finally:
        lib.gfd_set_tuning(b'batch_tile', 0)
    say('forced tiles', tiles)
    assert all(t['kernel'] == 'march' and t['py'] == py and t['pf'] == pf
               for t in tiles.values())
    _check_apply(b)


def _check_apply(b):
    import torch
    rng = np.random.default_rng(21)
    for level in range(b.L):
        ops = [s.operators[level] for s in b.solvers]
        arrays = [rng.uniform(-1, 1, tuple(op.gd.n_c)) for op in ops]
        flat = b.pack([dev(a) for a in arrays])
        out = torch.empty_like(flat)
        say('batched apply level', level)
        b.apply(level, flat, out)
        for op, a, got in zip(ops, arrays, b.unpack(out, level)):
            ref = oracle.fd_apply(op.coef_p, op.offset_pc, op.gd.n_c,
                                  op.gd.neighbor_cd, a)
            assert relerr(host(got), ref) < TOL
        # FP32 against the same oracle at float32 accuracy
        out32 = torch.empty_like(flat, dtype=torch.float32)
        b.apply(level, flat.float(), out32)
        for op, a, got in zip(ops, arrays, b.unpack(out32, level)):
            ref = oracle.fd_apply(op.coef_p, op.offset_pc, op.gd.n_c,
                                  op.gd.neighbor_cd,
                                  a.astype(np.float32).astype(float))
            assert relerr(host(got).astype(float), ref) < 2e-6


def _xcode(nt, py, zmax, xc):
    return nt | py << 12 | zmax << 16 | xc << 28


XFER_MODES = [1, 2, _xcode(256, 1, 64, 4), _xcode(512, 2, 16, 3),
              _xcode(128, 4, 128, 16), _xcode(64, 1, 8, 5)]


@pytest.mark.parametrize('mode', XFER_MODES)
def test_transfers(mode):
    """One point per thread (1), the timed x-march tile (2) and forced
    x-march tiles (z parts narrower than a row, ragged x chunks)."""
    import torch
    with batch_xfer(mode):
        b = batch()
        b.tune()
    rng = np.random.default_rng(22)
    for level in range(b.L - 1):
        rs = [s.restrictors[level] for s in b.solvers]
        its = [s.interpolators[level] for s in b.solvers]
        fine = [rng.uniform(-1, 1, tuple(t.gdin.n_c)) for t in rs]
        coarse = [rng.uniform(-1, 1, tuple(t.gdin.n_c)) for t in its]
        say('batched transfers level', level)
        out = torch.empty(b.level_points(level + 1), dtype=torch.float64,
                          device='cuda')
        b.transfer(level, False, b.pack([dev(a) for a in fine]), out)
        for t, a, got in zip(rs, fine, b.unpack(out, level + 1)):
            s = tspec(t)
            ref = oracle.transform(False, s.k, s.size_in, s.size_out,
                                   s.pad_cd, s.npad_cd, s.skip_cd,
                                   s.neighbor_cd, a)
            assert relerr(host(got), ref) < TOL
        base = [rng.uniform(-1, 1, tuple(t.gdout.n_c)) for t in its]
        flat_base = b.pack([dev(a) for a in base])
        out = torch.empty_like(flat_base)
        b.transfer(level, True, b.pack([dev(a) for a in coarse]), out,
                   flat_base)
        for t, a, bb, got in zip(its, coarse, base, b.unpack(out, level)):
            s = tspec(t)
            ref = bb - oracle.transform(True, s.k, s.size_in, s.size_out,
                                        s.pad_cd, s.npad_cd, s.skip_cd,
                                        s.neighbor_cd, a)
            assert relerr(host(got), ref) < TOL


def test_mixed_solve_matches_fp64(kernel):
    from gpaw_amd.poisson import FDPoissonSolver
    problems = workloads.uneven_48(seed=9, nproblems=6)
    gds = [p[0] for p in problems]
    b = BatchedPoissonSolver(nn=3, levels=4)
    b.set_grid_descriptors(gds)
    say('batched mixed solve tiles', b.tune())
    rhos = [dev(p[1]) for p in problems]
    phis = [dev(np.zeros(tuple(gd.n_c))) for gd in gds]
    n, rel = b.solve_mixed(phis, rhos, 1e-10)
    say('refinements', n, 'relres', rel)
    # grids whose 4th level is still ~20 points per side converge slowly
    # under Jacobi on the coarsest level; the bound is max_outer
    assert (rel <= 1e-10).all() and n < 100
    # graph replay and eager steps give the same iterates
    phis2 = [dev(np.zeros(tuple(gd.n_c))) for gd in gds]
    n2, rel2 = b.solve_mixed(phis2, rhos, 1e-10, graph=False)
    assert n2 == n
    for p, q in zip(phis, phis2):
        assert (host(p) == host(q)).all()
    for gd, (_, rho), phi in zip(gds, problems, phis):
        s = FDPoissonSolver(nn=3, eps=1e-24 * float((rho**2).sum()) * gd.dv,
                            maxiter=400)
        s.set_grid_descriptor(gd)
        ref = dev(np.zeros(tuple(gd.n_c)))
        s.solve_neutral(ref, dev(rho))
        a, r = host(phi), host(ref)
        if gd.pbc_c.all():
            a = a - a.mean()
            r = r - r.mean()
        assert relerr(a, r) < 1e-8


def _precond_ref(kin, rst, itp, res, step):
    return oracle.fd_precond(spec(kin[0]), spec(kin[1]), spec(kin[2]),
                             tspec(rst[0]), tspec(rst[1]), tspec(itp[0]),
                             tspec(itp[1]), res, step)


@pytest.mark.parametrize('xfer', [0, 2])
def test_batched_preconditioner(kernel, xfer):
    """The FP32 chain on five uneven grids in one set of launches, grid by
    grid against the FP64 oracle (float32 rounding over ~10 passes: bar
    2e-6, as for the single-grid FP32 chain) and against the single-grid
    FP32 preconditioner."""
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.preconditioner import (BatchedPreconditioner,
                                         Preconditioner)
    rng = np.random.default_rng(23)
    gds = [GridDescriptor(N, np.array(N) * (0.3 + 0.02 * i), pbc)
           for i, (N, pbc) in enumerate(SHAPES)]
    with batch_xfer(xfer):
        P = BatchedPreconditioner(gds)
        # the first call tunes the batch (transfers included)
        P([dev(np.zeros(tuple(gd.n_c))) for gd in gds])
    assert len(set(P.steps)) > 1  # per-grid steps
    res = [rng.uniform(-1, 1, tuple(gd.n_c)) for gd in gds]
    say('batched preconditioner', len(gds), 'grids')
    out = P([dev(r) for r in res])
    for g, (gd, r, o) in enumerate(zip(gds, res, out)):
        kin, rst, itp = P.levels[g]
        ref = _precond_ref(kin, rst, itp, r, P.steps[g])
        err = relerr(host(o), ref)
        assert 1e-12 < err < 2e-6, (g, err)
        one = Preconditioner(gd, Laplace(gd, -0.5, 3), float, fp32=True)
        assert one.step == P.steps[g]
        single = host(one(dev(r), None))
        assert relerr(host(o), single) < 2e-6, g
    # a second call on other residuals reuses the tables
    res2 = [2 * r for r in res]
    out2 = P([dev(r) for r in res2])
    for g, o in enumerate(out2):
        assert relerr(host(o), 2 * host(out[g])) < 1e-6
    # the packed layout gives the same bits as the per-grid call
    import torch
    flat = P.pack([dev(r) for r in res])
    fout = torch.empty_like(flat)
    P.apply_packed(flat, fout)
    assert (host(fout) == host(P.pack(out))).all()
    with pytest.raises(TypeError):
        P.apply_packed(flat[1:], fout)
