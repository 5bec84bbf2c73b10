"""Kinetic preconditioner (fd_precond chain) parity."""
import numpy as np
import pytest

import oracle
from gpaw_amd.fd_operators import Laplace
from gpaw_amd.grid_descriptor import GridDescriptor
from gpaw_amd.preconditioner import Preconditioner
from gpu_util import dev, host, say, xfer_path
from helpers import bloch_phases, case, rand, relerr, spec, tspec

pytestmark = pytest.mark.gpu
TOL = 1e-13


class K:
    def __init__(self, phase_cd):
        self.phase_cd = phase_cd


@pytest.mark.parametrize('name', ['real_pbc', 'cplx_bloch', 'real_zero'])
def test_golden(golden, name):
    c = case(golden('precond'), 'precond', name)
    dtype = complex if c['residuals'].dtype == np.complex128 else float
    gd = GridDescriptor(tuple(c['N']), tuple(c['cell']), tuple(c['pbc']))
    P = Preconditioner(gd, Laplace(gd, -0.5, 3, dtype), dtype, block=2)
    ph = None if c['phase'].size == 0 else c['phase']
    say('golden precond', name)
    out = host(P(dev(c['residuals']), K(ph)))
    assert relerr(out, c['out']) < TOL


def test_blocks_and_host_arrays():
    rng = np.random.default_rng(5)
    gd = GridDescriptor((24, 16, 32), (4.0, 3.0, 6.0))
    dtype = complex
    P = Preconditioner(gd, Laplace(gd, -0.5, 3, dtype), dtype, block=3)
    res = rand(rng, (7,) + tuple(gd.n_c), dtype)
    ph = bloch_phases(rng, gd.pbc_c)
    ref = oracle.fd_precond(spec(P.kin0), spec(P.kin1), spec(P.kin2),
                            tspec(P.restrictor_object0),
                            tspec(P.restrictor_object1),
                            tspec(P.interpolator_object1),
                            tspec(P.interpolator_object2), res, P.step, ph)
    say('precond blocks')
    out = dev(np.zeros_like(res))
    P(dev(res), K(ph), out=out)
    assert relerr(host(out), ref) < TOL
    out_h = P(res, K(ph), out=np.zeros_like(res))
    assert relerr(out_h, ref) < TOL
    one = host(P(dev(res[2]), K(ph)))
    assert relerr(one, ref[2]) < TOL


@pytest.mark.parametrize('path', ['march', 'legacy'])
@pytest.mark.parametrize('dtype,pbc', [(complex, True), (float, True),
                                       (float, False)])
def test_fp32_chain(dtype, pbc, path):
    """FP32 chain against the FP64 oracle: float32 rounding over ~10
    passes, so a relative max-norm error of a few 1e-7 (bar 2e-6)."""
    rng = np.random.default_rng(6)
    gd = GridDescriptor((32, 24, 48), (5.0, 4.0, 7.0), pbc)
    P = Preconditioner(gd, Laplace(gd, -0.5, 3, dtype), dtype, block=4,
                       fp32=True)
    res = rand(rng, (6,) + tuple(gd.n_c), dtype)
    ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
    ref = oracle.fd_precond(spec(P.kin0), spec(P.kin1), spec(P.kin2),
                            tspec(P.restrictor_object0),
                            tspec(P.restrictor_object1),
                            tspec(P.interpolator_object1),
                            tspec(P.interpolator_object2), res, P.step, ph)
    say('precond fp32', dtype.__name__, pbc, path)
    with xfer_path(path):
        out = host(P(dev(res), K(ph)))
    assert out.dtype == res.dtype
    err = relerr(out, ref)
    assert err < 2e-6
    assert err > 1e-12        # it did run in FP32


FUSED_CASES = [((32, 24, 48), True, 0), ((16, 20, 24), False, 0),
               ((24, 16, 32), (True, False, True), 0),
               ((20, 28, 16), True, 1 | 4 << 8),
               ((28, 16, 24), (False, True, True), 2 | 4 << 8 | 5 << 16),
               ((8, 12, 8), True, 0), ((12, 8, 40), (True, True, False), 0)]


@pytest.mark.parametrize('fp32', [True, False])
@pytest.mark.parametrize('dtype', [complex, float])
@pytest.mark.parametrize('shape,pbc,tile', FUSED_CASES)
def test_fused_fine_level(shape, pbc, tile, dtype, fp32):
    """The chain's fine level fused (interpolation + update in one pass,
    d0' never stored) against the same chain as two passes and against the
    FP64 oracle, FP32 and FP64 chains (the FP64 one at its 1e-13 bar);
    complex data with Bloch phases and real data, periodic, zero and mixed
    boundaries, forced tiles (rows per thread, thread rows, ragged x
    chunks)."""
    from gpaw_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(16)
    gd = GridDescriptor(shape, np.array(shape) * 0.2, pbc)
    P = Preconditioner(gd, Laplace(gd, -0.5, 3, dtype), dtype, block=3,
                       fp32=fp32)
    res = rand(rng, (5,) + tuple(gd.n_c), dtype)
    ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
    ref = oracle.fd_precond(spec(P.kin0), spec(P.kin1), spec(P.kin2),
                            tspec(P.restrictor_object0),
                            tspec(P.restrictor_object1),
                            tspec(P.interpolator_object1),
                            tspec(P.interpolator_object2), res, P.step, ph)
    say('fused fine level', 'fp32' if fp32 else 'fp64', dtype.__name__,
        shape, pbc, hex(tile))
    try:
        _lib.check(lib.gfd_set_tuning(b'precond_fine_tile', tile), 'tile')
        fused = host(P(dev(res), K(ph)))
        _lib.check(lib.gfd_set_tuning(b'precond_fine', 0), 'off')
        two = host(P(dev(res), K(ph)))
    finally:
        lib.gfd_set_tuning(b'precond_fine', 1)
        lib.gfd_set_tuning(b'precond_fine_tile', 0)
    err2 = relerr(fused, two)
    say('  fused vs two passes', err2, ' vs oracle', relerr(fused, ref))
    bar = 2e-6 if fp32 else TOL
    assert err2 < (1e-6 if fp32 else 1e-15)
    assert relerr(fused, ref) < bar
    assert relerr(two, ref) < bar
