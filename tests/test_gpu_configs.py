"""Each benchmark workload at its own size, checked on the GPU:

- 256 x 96^3 FP64 bands: Laplace nn=3 and Gradient-x nn=3 through the
  tuner, bands 0 / 127 / 255 against the oracle; every full-row tile the
  tuner can pick at 96^3 forced in turn; restrict / interpolate nn=3 on
  4 x 96^3 <-> 48^3 Gaussian densities;
- the 192^3 Poisson problem with FP32 V-cycles inside FP64 refinement, at
  4 levels and at the full hierarchy: residual recomputed in FP64, the
  spectral closed form ifft(fft(rho) / lambda(k)) and the FP64 solution;
- all 48 uneven problems solved as one batch against per-grid FP64 solves;
- one k-point of 1024 complex128 bands on 128^3 with Bloch phases: fused
  residuals on sampled bands against the oracle, the overlap on a 32 x 32
  block against numpy, orthonormality after the Cholesky rotation, the
  FP32 preconditioner against the oracle's FP64 chain.

Tolerances: 1e-13 relative max-norm for FP64 work; the precision forms as
stated per test."""
from math import pi

import numpy as np
import pytest

import oracle
from gpaw_amd import _lib, workloads
from gpaw_amd.fd_operators import Gradient, Laplace
from gpu_util import dev, host, kernel_path, say
from helpers import relerr, spec, tspec

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _ref(op, a, phase_cd=None):
    return oracle.fd_apply(op.coef_p, op.offset_pc, op.gd.n_c,
                           op.gd.neighbor_cd, a, phase_cd)


@pytest.fixture(scope='module')
def bands_96():
    import torch
    gd = workloads.bands_96_gd()
    g = torch.Generator(device='cuda')
    g.manual_seed(1234)
    a = torch.rand((256, 96, 96, 96), dtype=torch.float64, device='cuda',
                   generator=g) * 2 - 1
    yield gd, a
    del a
    torch.cuda.empty_cache()


@pytest.mark.parametrize('kind', ['laplace', 'gradient_x'])
def test_bands_96_full_size(bands_96, kind):
    import torch
    gd, a = bands_96
    op = Laplace(gd, 1.0, 3) if kind == 'laplace' else Gradient(gd, 0, 1.0, 3)
    out = torch.empty_like(a)
    say(f'{kind} 256 x 96^3 (tuner on)')
    op.apply(a, out)      # first call tunes
    out.fill_(np.nan)
    op.apply(a, out)
    torch.cuda.synchronize()
    for band in (0, 127, 255):
        ref = _ref(op, host(a[band]))
        assert relerr(host(out[band]), ref) < TOL, band
    # every band: the periodic operator of a field sums to zero
    s = out.sum(dim=(1, 2, 3)).abs().max().item()
    assert s < 1e-12 * out.abs().sum(dim=(1, 2, 3)).min().item()


def _candidates(R, n, dma):
    lib = _lib.load()
    arr, p = _lib.longs(n)
    cnt = lib.gfd_row_tile_candidates(R, 0, p, int(dma), None, 0)
    assert cnt > 0
    codes = np.zeros(cnt, np.int64)
    lib.gfd_row_tile_candidates(R, 0, p, int(dma),
                                codes.ctypes.data_as(_lib._lp), cnt)
    return codes


@pytest.mark.parametrize('kind', ['laplace', 'gradient_x'])
def test_every_row_tile_at_96(bands_96, kind):
    """Each tile the tuner may pick at 96^3 (both row kernels), forced on
    16 bands, against the oracle on the device."""
    import torch
    gd, a_all = bands_96
    a = a_all[:16]
    op = Laplace(gd, 1.0, 3) if kind == 'laplace' else Gradient(gd, 0, 1.0, 3)
    ref = dev(_ref(op, host(a)))
    scale = ref.abs().max().item()
    codes = _candidates(3, gd.n_c, True)
    ndma = int(sum((int(c) >> 48) & 0xf > 0 for c in codes))
    say(f'{kind}: {len(codes)} tiles ({ndma} LDS-DMA)')
    assert ndma > 0
    lib = _lib.load()
    out = torch.empty_like(a)
    bad = []
    try:
        with kernel_path('row'):
            for code in codes:
                _lib.check(lib.gfd_set_tuning(b'xr_tile', int(code)),
                           'tuning')
                out.fill_(np.nan)
                op.apply(a, out)
                err = (out - ref).abs().max().item() / scale
                if not err < TOL:
                    bad.append((hex(int(code)), err))
    finally:
        lib.gfd_set_tuning(b'xr_tile', 0)
    assert not bad, bad


def test_density_96_transfers():
    from gpaw_amd.transformers import Transformer
    gd, dens = workloads.density_96()
    coarse = gd.coarsen()
    R = Transformer(gd, coarse, 3)
    I = Transformer(coarse, gd, 3)
    say('restrict / interpolate nn=3, 4 x 96^3 <-> 48^3')
    c = R.apply(dev(dens))
    s = tspec(R)
    refc = oracle.transform(False, s.k, s.size_in, s.size_out, s.pad_cd,
                            s.npad_cd, s.skip_cd, s.neighbor_cd, dens)
    assert relerr(host(c), refc) < TOL
    f = I.apply(dev(refc))
    s = tspec(I)
    reff = oracle.transform(True, s.k, s.size_in, s.size_out, s.pad_cd,
                            s.npad_cd, s.skip_cd, s.neighbor_cd, refc)
    assert relerr(host(f), reff) < TOL
    # restriction conserves the integral of each density
    for i in range(4):
        assert abs(refc[i].sum() * 8 - dens[i].sum()) < 1e-10 * \
            np.abs(dens[i]).sum()


def _spectral(op, rho):
    """ifft(fft(rho) / lambda(k)), lambda(k) = sum_p c_p exp(i k.o_p), the
    k = 0 mode removed (periodic grid)."""
    N = rho.shape
    k = [2 * pi * np.fft.fftfreq(n) for n in N]
    lam = np.zeros(N, complex)
    for c, o in zip(op.coef_p, op.offset_pc):
        f = [np.exp(1j * k[v] * o[v]) for v in range(3)]
        lam += c * f[0][:, None, None] * f[1][None, :, None] * \
            f[2][None, None, :]
    rq = np.fft.fftn(rho)
    lam[0, 0, 0] = 1.0
    rq[0, 0, 0] = 0.0
    return np.fft.ifftn(rq / lam).real


@pytest.fixture(scope='module')
def poisson_192():
    from gpaw_amd.poisson import FDPoissonSolver
    gd, rho = workloads.poisson_192()
    rn2 = float((rho**2).sum())
    s = FDPoissonSolver(nn=3, eps=1e-24 * rn2 * gd.dv, maxiter=1000)
    s.set_grid_descriptor(gd)
    say('192^3: FP64 reference-schedule solve')
    drho = dev(rho)
    phi64 = dev(np.zeros(tuple(gd.n_c)))
    n64 = s.solve_neutral(phi64, drho)
    ref = _spectral(s.operators[0], rho)
    ref -= ref.mean()
    return gd, rho, drho, s, host(phi64), n64, ref


@pytest.mark.parametrize('levels,coarse', [(4, 64), (None, 8)])
def test_poisson_192_mixed(poisson_192, levels, coarse):
    import torch
    gd, rho, drho, s, phi64, n64, ref = poisson_192
    rn2 = float((rho**2).sum())
    # the FP64 V-cycles reach the spectral solution
    a = phi64 - phi64.mean()
    assert relerr(a, ref) < 1e-9
    phi = torch.zeros_like(drho)
    say(f'192^3 mixed, levels {levels}, coarse sweeps {coarse}')
    n, rel = s.solve_mixed(phi, drho, 1e-10, levels=levels, pre=2, post=2,
                           coarse_sweeps=coarse, w0=0.8, inner=2,
                           max_outer=200)
    say(f'  {n} refinement steps, relres {rel:.3g} (FP64: {n64} cycles)')
    assert rel <= 1e-10
    r = torch.empty_like(drho)
    n2 = s.operators[0].residual(phi, drho, r, norm=True)
    assert (n2 / rn2)**0.5 <= 1.001e-10
    b = host(phi)
    b -= b.mean()
    assert relerr(b, ref) < 1e-9
    assert relerr(b, a) < 1e-9


def test_uneven_48_all_problems():
    import torch
    from gpaw_amd.batch_poisson import BatchedPoissonSolver
    from gpaw_amd.poisson import FDPoissonSolver
    problems = workloads.uneven_48()
    assert len(problems) == 48
    gds = [g for g, _ in problems]
    rhos = [dev(r) for _, r in problems]
    b = BatchedPoissonSolver(nn=3, levels=4, w0=0.8)
    b.set_grid_descriptors(gds)
    say('48 uneven problems: tune')
    b.tune()
    phis = [torch.zeros_like(r) for r in rhos]
    n, rel = b.solve_mixed(phis, rhos, 1e-10, coarse_sweeps=16)
    say(f'  batched: {n} refinement steps, max relres {rel.max():.3g}')
    assert (rel <= 1e-10).all()
    worst = 0.0
    for i, (gd, rho) in enumerate(problems):
        s = FDPoissonSolver(nn=3, eps=1e-24 * float((rho**2).sum()) * gd.dv,
                            maxiter=600)
        s.set_grid_descriptor(gd)
        ref = torch.zeros_like(rhos[i])
        s.solve_neutral(ref, rhos[i])
        x, y = host(phis[i]), host(ref)
        if gd.pbc_c.all():
            x, y = x - x.mean(), y - y.mean()
        e = relerr(x, y)
        worst = max(worst, e)
        assert e < 1e-8, (i, tuple(gd.N_c), bool(gd.pbc_c.all()), e)
    say(f'  worst error against the FP64 solves: {worst:.3g}')


def test_uneven_48_preconditioner():
    """The preconditioner half of the 48-problem workload: the batched FP32
    chain on every grid against the FP64 oracle grid by grid."""
    from gpaw_amd.preconditioner import BatchedPreconditioner
    problems = workloads.uneven_48()
    gds = [g for g, _ in problems]
    P = BatchedPreconditioner(gds)
    rng = np.random.default_rng(4807)
    res = [rng.uniform(-1, 1, tuple(g.n_c)) for g in gds]
    say('48 uneven problems: batched preconditioner')
    out = P([dev(r) for r in res])
    worst = 0.0
    for g, (r, o) in enumerate(zip(res, out)):
        kin, rst, itp = P.levels[g]
        ref = oracle.fd_precond(spec(kin[0]), spec(kin[1]), spec(kin[2]),
                                tspec(rst[0]), tspec(rst[1]), tspec(itp[0]),
                                tspec(itp[1]), r, P.steps[g])
        e = relerr(host(o), ref)
        worst = max(worst, e)
        assert e < 2e-6, (g, tuple(gds[g].N_c), e)
    say(f'  worst error against the FP64 oracle: {worst:.3g}')


@pytest.fixture(scope='module')
def eigen_128():
    import torch
    from gpaw_amd.bands import apply_hamiltonian, band_dots
    gd = workloads.eigen_128_gd()
    ph = workloads.eigen_128_kpoints()[0]
    vt_h = workloads.eigen_128_potential(gd)
    vt = dev(vt_h)
    kin = Laplace(gd, -0.5, 3, complex)
    g = torch.Generator(device='cuda')
    g.manual_seed(1006)
    say('1024 complex bands of 128^3')
    psi = torch.randn((1024, 128, 128, 128), dtype=torch.complex128,
                      device='cuda', generator=g)
    R = torch.empty_like(psi)
    apply_hamiltonian(kin, vt, psi, R, ph)
    eps = band_dots(psi, R, gd.dv) / band_dots(psi, psi, gd.dv)
    yield gd, ph, vt_h, vt, kin, psi, R, eps
    del psi, R
    torch.cuda.empty_cache()


def test_eigen_128_residuals(eigen_128):
    from gpaw_amd.bands import apply_hamiltonian
    gd, ph, vt_h, vt, kin, psi, R, eps = eigen_128
    say('fused residuals (H - eps_n) psi_n, 1024 bands')
    apply_hamiltonian(kin, vt, psi, R, ph, eps=eps)
    e = host(eps)
    for band in (0, 511, 1023):
        x = host(psi[band])
        ref = _ref(kin, x, ph) + (vt_h - e[band]) * x
        assert relerr(host(R[band]), ref) < TOL, band


def test_eigen_128_overlap_block(eigen_128):
    import torch
    from gpaw_amd.bands import overlap
    gd, ph, vt_h, vt, kin, psi, R, eps = eigen_128
    say('overlap S = dv psi psi^H, 1024 bands')
    S = overlap(psi, gd.dv, lower_only=False)
    idx = np.sort(np.random.default_rng(3).choice(1024, 32, replace=False))
    P = host(psi[torch.from_numpy(idx).cuda()]).reshape(32, -1)
    ref = gd.dv * (P @ P.conj().T)
    got = host(S)[np.ix_(idx, idx)]
    assert relerr(got, ref) < TOL


def test_eigen_128_orthonormalize(eigen_128):
    import torch
    from gpaw_amd.bands import orthonormalize, overlap
    gd, ph, vt_h, vt, kin, psi, R, eps = eigen_128
    say('Cholesky orthonormalisation, 1024 bands')
    out, _ = orthonormalize(psi, gd.dv, out=R)
    S = overlap(out, gd.dv, lower_only=False)
    err = (S - torch.eye(1024, dtype=S.dtype, device=S.device)).abs().max()
    assert err.item() <= 1e-12


def test_eigen_128_subspace_diagonalize(eigen_128):
    """Subspace diagonalisation on the full band set: H = psi^H H psi,
    eigenvalues against numpy's on the same matrix, three rotated bands
    against the rotation recomputed with torch from the returned vectors."""
    import torch
    from gpaw_amd.bands import (apply_hamiltonian, subspace_diagonalize,
                                subspace_matrix)
    gd, ph, vt_h, vt, kin, psi, R, eps = eigen_128
    say('subspace diagonalisation, 1024 bands')
    apply_hamiltonian(kin, vt, psi, R, ph)
    H = subspace_matrix(psi, R, gd.dv)
    w = np.linalg.eigvalsh(host(H), UPLO='L')
    e, out = subspace_diagonalize(H, psi, out=R)
    assert np.abs(host(e) - w).max() < 1e-12 * np.abs(w).max()
    V = H  # rows: eigenvectors
    A = psi.reshape(1024, -1)
    for m in (0, 511, 1023):
        ref = (V[m:m + 1] @ A).reshape(psi.shape[1:])
        assert relerr(host(out[m]), host(ref)) < 1e-12, m


def test_eigen_128_fp32_preconditioner(eigen_128):
    from gpaw_amd.preconditioner import Preconditioner
    gd, ph, vt_h, vt, kin, psi, R, eps = eigen_128

    class K:
        phase_cd = ph

    say('FP32 preconditioner, 2 bands against the FP64 oracle chain')
    P = Preconditioner(gd, kin, complex, block=128, fp32=True)
    x = psi[:2].contiguous()
    got = host(P(x, K))
    ref = oracle.fd_precond(spec(P.kin0), spec(P.kin1), spec(P.kin2),
                            tspec(P.restrictor_object0),
                            tspec(P.restrictor_object1),
                            tspec(P.interpolator_object1),
                            tspec(P.interpolator_object2), host(x), P.step,
                            ph)
    err = relerr(got, ref)
    say(f'  FP32 chain error {err:.3g}')
    assert 1e-12 < err < 2e-6
