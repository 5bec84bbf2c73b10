"""HIP stencil (FDOperator.apply) parity: against the reference fixtures and
the oracle.  Tolerance: 1e-13 relative max-norm (FP64)."""
import numpy as np
import pytest

import oracle
from gpaw_amd.fd_operators import FDOperator, GUCLaplace, Gradient, Laplace
from gpaw_amd.grid_descriptor import GridDescriptor
from gpu_util import dev, host, kernel_path, say
from helpers import bloch_phases, case, case_names, rand, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-13


def gpu_apply(op, a, phase=None):
    d = dev(a)
    out = d.new_empty(d.shape)
    op.apply(d, out, phase)
    return host(out)


def ref_apply(op, a, phase=None):
    return oracle.fd_apply(op.coef_p, op.offset_pc, op.gd.n_c,
                           op.gd.neighbor_cd, a, phase)


@pytest.mark.parametrize('path', ['row', 'xmarch', 'general'])
def test_golden(golden, path):
    npz = golden('fd')
    with kernel_path(path):
        for name in case_names(npz, 'apply'):
            c = case(npz, 'apply', name)
            say('golden apply', name, path)
            real = c['in'].dtype == np.float64
            gd_n = tuple(c['n_c'])
            # build through the public constructor on a matching layout
            pbc = tuple(bool(x >= 0) for x in c['neighbor'][:, 0])
            N = tuple(n + (0 if p else 1) for n, p in zip(gd_n, pbc))
            gd = GridDescriptor(N, (1.0, 1.0, 1.0), pbc)
            op = FDOperator(c['coef'], c['offset'], gd,
                            float if real else complex)
            ph = None if c['phase'].size == 0 else c['phase']
            out = gpu_apply(op, c['in'], ph)
            assert relerr(out, c['out']) < TOL, name


SHAPES = [(12, 10, 14), (13, 17, 23), (5, 7, 9), (40, 48, 56), (3, 4, 5),
          (24, 96, 20)]
BCS = [(True, True, True), (False, False, False), (True, False, True),
       (False, True, False)]


@pytest.mark.parametrize('nn', [1, 2, 3, 4])
@pytest.mark.parametrize('dtype', [float, complex])
def test_laplace_random(nn, dtype):
    rng = np.random.default_rng(nn * 10 + (dtype == complex))
    for N in SHAPES:
        for pbc in BCS:
            try:
                gd = GridDescriptor(N, np.array(N) * 0.21, pbc)
            except ValueError:
                continue
            if (gd.n_c < nn).any():
                continue
            op = Laplace(gd, -0.5, nn, dtype)
            a = rand(rng, (2,) + tuple(gd.n_c), dtype)
            ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
            say('laplace', nn, dtype.__name__, N, pbc)
            ref = ref_apply(op, a, ph)
            for path in ('row', 'xmarch', 'general'):
                with kernel_path(path):
                    assert relerr(gpu_apply(op, a, ph), ref) < TOL, path


@pytest.mark.parametrize('v', [0, 1, 2])
def test_gradient(v):
    rng = np.random.default_rng(v)
    for N, pbc in [((12, 10, 14), True), ((13, 9, 11), False)]:
        gd = GridDescriptor(N, (3.0, 2.7, 3.3), pbc)
        for dtype in (float, complex):
            op = Gradient(gd, v, 1.0, 3, dtype)
            a = rand(rng, (3,) + tuple(gd.n_c), dtype)
            ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
            say('gradient', v, N, pbc, dtype.__name__)
            ref = ref_apply(op, a, ph)
            for path in ('row', 'xmarch', 'general'):
                with kernel_path(path):
                    assert relerr(gpu_apply(op, a, ph), ref) < TOL, path


def test_guc_and_general_offsets():
    rng = np.random.default_rng(9)
    hexa = [(3.0, 0, 0), (1.5, 3**0.5 * 1.5, 0), (0, 0, 3.3)]
    for N, cell, nn in [((12, 12, 10), hexa, 2),
                        ((8, 8, 8), [(0, 2, 2), (2, 0, 2), (2, 2, 0)], 3)]:
        gd = GridDescriptor(N, cell)
        for dtype in (float, complex):
            op = GUCLaplace(gd, -0.5, nn, dtype)
            assert not op.cfd
            a = rand(rng, (2,) + tuple(gd.n_c), dtype)
            ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
            say('guc', N, nn, dtype.__name__)
            assert relerr(gpu_apply(op, a, ph), ref_apply(op, a, ph)) < TOL


def test_batch_dims_inplace_and_host_arrays():
    rng = np.random.default_rng(4)
    gd = GridDescriptor((16, 12, 20), (3.0, 2.5, 4.0))
    op = Laplace(gd, 1.0, 3)
    a = rand(rng, (2, 3) + tuple(gd.n_c))
    ref = ref_apply(op, a.reshape((6,) + tuple(gd.n_c))).reshape(a.shape)
    say('batch dims')
    assert relerr(gpu_apply(op, a), ref) < TOL
    d = dev(a)
    op.apply(d, d)  # in place, like the reference's padded copy allows
    assert relerr(host(d), ref) < TOL
    out = np.zeros_like(a)
    op.apply(a, out)  # numpy arrays are staged through the device
    assert relerr(out, ref) < TOL


def test_argument_errors():
    gd = GridDescriptor((8, 8, 8))
    op = Laplace(gd, 1.0, 2)
    import torch
    a = torch.zeros((8, 8, 8), dtype=torch.float64, device='cuda')
    with pytest.raises(ValueError):
        op.apply(a, torch.zeros((8, 8, 9), dtype=torch.float64,
                                device='cuda'))
    with pytest.raises(TypeError):
        op.apply(a.to(torch.complex128), a.to(torch.complex128))
    with pytest.raises(ValueError):
        op.apply(a.transpose(0, 2), a)


def test_residual_norm():
    rng = np.random.default_rng(12)
    gd = GridDescriptor((24, 20, 16), (4.0, 3.5, 3.0))
    op = Laplace(gd, -0.25 / np.pi, 3)
    phi = rand(rng, tuple(gd.n_c))
    rho = rand(rng, tuple(gd.n_c))
    r_ref = ref_apply(op, phi) - rho
    d = dev(phi)
    r = d.new_empty(d.shape)
    say('residual norm')
    for path in ('row', 'xmarch', 'general'):
        with kernel_path(path):
            n2 = op.residual(d, dev(rho), r, norm=True)
            assert relerr(host(r), r_ref) < TOL
            assert n2 == pytest.approx(float((r_ref**2).sum()), rel=1e-12)


@pytest.mark.parametrize('nt', [1, 3])
@pytest.mark.parametrize('dtype', [float, complex])
def test_row_kernel_nontemporal(nt, dtype):
    """The non-temporal load/store variants of the full-row kernel (stores
    are on by default above 512 MB) at test sizes, forced on."""
    from gpaw_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(60 + nt)
    gd = GridDescriptor((24, 32, 16), (3.0, 4.0, 2.0), (True, False, True))
    op = Laplace(gd, -0.5, 3, dtype)
    a = rand(rng, (5,) + tuple(gd.n_c), dtype)
    ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
    ref = ref_apply(op, a, ph)
    try:
        _lib.check(lib.gfd_set_tuning(b'xr_nt', nt), 'tuning')
        with kernel_path('row'):
            say('row kernel nt', nt, dtype.__name__)
            out = gpu_apply(op, a, ph)
    finally:
        lib.gfd_set_tuning(b'xr_nt', -1)
    assert relerr(out, ref) < TOL


def test_row_kernel_default_tiles_large_batch():
    """A batch big enough for the large-batch tile choice (several x chunks
    per grid, two rows per thread) against the oracle on two bands."""
    rng = np.random.default_rng(61)
    gd = GridDescriptor((48, 48, 48), (9.0, 9.0, 9.0))
    op = Laplace(gd, 1.0, 3)
    nb = 300   # 33M points
    import torch
    d = torch.rand((nb, 48, 48, 48), dtype=torch.float64, device='cuda')
    out = torch.empty_like(d)
    say('large batch row tiles')
    op.apply(d, out)
    for band in (0, nb - 1):
        ref = ref_apply(op, host(d[band]))
        assert relerr(host(out[band]), ref) < TOL


@pytest.mark.parametrize('pbc', [True, False, (True, False, True)])
def test_row_kernel_wide_complex_tiles(pbc):
    """Full-row tiles with two complex points per thread (forced through
    the tile override), plain apply with Bloch phases and the H apply
    epilogue, against the oracle."""
    import torch
    from gpaw_amd import _lib
    from gpaw_amd.bands import apply_hamiltonian
    lib = _lib.load()
    rng = np.random.default_rng(90)
    per = (pbc,) * 3 if isinstance(pbc, bool) else pbc
    # interior sizes (12, 12, 16) whatever the boundaries
    N = tuple(m + (0 if p else 1) for m, p in zip((12, 12, 16), per))
    gd = GridDescriptor(N, (3.0, 3.0, 4.0), pbc)
    n = tuple(int(x) for x in gd.n_c)
    assert n == (12, 12, 16)
    op = Laplace(gd, -0.5, 3, complex)
    a = rand(rng, (4,) + n, complex)
    ph = bloch_phases(rng, gd.pbc_c)
    ref = ref_apply(op, a, ph)
    vt = rng.standard_normal(n)
    for nseg, py in [(2, 1), (3, 1), (4, 1), (2, 2)]:
        if n[1] % (nseg * py) or (n[2] // 2) * nseg > 512:
            continue
        code = n[2] | (nseg << 12) | (py << 24) | (5 << 32) | (2 << 56)
        try:
            _lib.check(lib.gfd_set_tuning(b'xr_tile', code), 'tuning')
            with kernel_path('row'):
                say('wide c128 tile', n, pbc, nseg, py)
                out = gpu_apply(op, a, ph)
                d = dev(a)
                h = d.new_empty(d.shape)
                apply_hamiltonian(op, torch.tensor(vt, device=d.device), d,
                                  h, ph)
        finally:
            lib.gfd_set_tuning(b'xr_tile', 0)
        assert relerr(out, ref) < TOL, (nseg, py)
        assert relerr(host(h), ref + a * vt) < TOL, (nseg, py)
