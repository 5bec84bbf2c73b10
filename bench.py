"""Benchmark of the GPAW FD grid hot path on one MI355X.

Headline (``value``): batched 6th-order (nn=3, 19-point) FD Laplacian over
256 real float64 bands on a 96^3 periodic grid, in GPts/s (grid points of
the whole batch per second), with the kernel's fraction of the 8 TB/s HBM
roofline.  One step = one Laplace apply over all 256 bands.

Also reported in ``stages`` (same run): Gradient-x nn=3 on the same bands,
restrict / interpolate nn=3 on 4 densities 96^3 <-> 48^3, the FP64 multigrid
Poisson solve of the 48^3 four-Gaussian problem (eps 2e-10), the kinetic
preconditioner on 32 complex bands, the weighted (dielectric) operator on
16 densities of 96^3, the 192^3 Poisson time to a 1e-10
relative residual (all-FP64 V-cycles against FP32 V-cycles inside FP64
refinement), the 48 uneven Poisson problems solved as one batch, and the FD
eigensolver step on 1024 complex bands of 128^3 at 4 k-points (fused
residuals, the RMM-DIIS block step with the FP32 preconditioner, FP64 MFMA
overlap, Cholesky orthonormalisation).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--no-stages] [--no-eigen]
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ('batched FD stencil GPts/s + %HBM peak; '
          'Poisson time-to-1e-10 mixed vs FP64')
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--no-stages', action='store_true')
    p.add_argument('--bands', type=int, default=256)
    p.add_argument('--no-eigen', action='store_true')
    p.add_argument('--pmc', action='store_true',
                   help='counter runs: 3 headline applies and 3 copies of '
                        'the same 3.62 GB (calibration), no timing line')
    p.add_argument('--eigen-bands', type=int, default=1024)
    p.add_argument('--stage', choices=['poisson192', 'uneven48', 'eigen',
                                       'precond128', 'density96'],
                   help='run only this stage and print its result (for '
                        'profiling)')
    return p.parse_args()


def timed(fn, steps, warmup):
    """Average ms per call of fn() between events on torch's current stream
    (the stream every libgpaw_fd call is launched on)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return e0.elapsed_time(e1) / steps, wall


def traffic_from_profiles(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, calibrated on a copy kernel), if a summary is committed."""
    path = os.path.join(ROOT, 'profiles', f'pmc_{name}.md')
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith('traffic_bytes_per_launch:'):
            return float(line.split(':', 1)[1])
    return None


def stage_poisson_192(dev, tol=1e-10, inner=2, coarse=8, w0=0.8, pre=3,
                      post=4):
    """Time to |L phi - rho| / |rho| <= tol from phi = 0 on the 192^3
    problem: all-FP64 V-cycles (the reference schedule: V(2,1) with weight
    2/3 on the fine level, V(4,4) below, every level) against FP32
    V(pre,post)-cycles (Jacobi weight w0 on the fine level, 1 below; the four
    levels 192^3 .. 24^3 as grid-wide launches, 12^3 .. 3^3 in one
    LDS-resident workgroup with `coarse` sweeps on 3^3; `inner` cycles per
    step) inside FP64 residual refinement.  The same refinement with the
    inner cycles in FP64 (fp64_same_schedule) isolates what FP32 gains;
    `levels4` repeats both on the four-level cycles alone (64 sweeps on
    24^3 in place of the coarser levels)."""
    import torch
    from gpaw_amd import workloads
    from gpaw_amd.poisson import FDPoissonSolver
    gd, rho = workloads.poisson_192()
    drho = torch.from_numpy(rho).to(dev)
    rn2 = float((rho**2).sum())
    s = FDPoissonSolver(nn=3, eps=tol**2 * rn2 * gd.dv, maxiter=1000)
    s.set_grid_descriptor(gd)

    def wall(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3, r

    phi64 = torch.zeros_like(drho)
    s.solve_neutral(phi64, drho)              # capture + warm
    t64 = []
    for _ in range(3):
        phi64.zero_()
        ms, n64 = wall(lambda: s.solve_neutral(phi64, drho))
        t64.append(ms)
    rel64 = (s.last_error / (rn2 * gd.dv))**0.5
    def refine(levels, fp64_inner, cs=coarse, pre=pre, post=post):
        """best-of-3 time of the refinement solve on one schedule"""
        ph = torch.zeros_like(drho)
        kw = dict(levels=levels, pre=pre, post=post, inner=inner,
                  coarse_sweeps=cs, w0=w0, fp64_inner=fp64_inner)
        s.solve_mixed(ph, drho, tol, **kw)
        t = []
        for _ in range(3):
            ph.zero_()
            ms, (n, rel) = wall(lambda: s.solve_mixed(ph, drho, tol, **kw))
            t.append(ms)
        return min(t), n, rel, ph

    tmx, nmx, relmx, phi = refine(None, False)
    t64s, n64s, rel64s, phi64s = refine(None, True)
    # the configuration's 4-level cycles (coarsest 24^3: 64 sweeps there;
    # the coarse sweeps bound their convergence, so V(2,2) is cheaper)
    t4, n4, rel4, phi4 = refine(4, False, 64, 2, 2)
    t4d, n4d, rel4d, _ = refine(4, True, 64, 2, 2)
    a = phi - phi.mean()
    b = phi64 - phi64.mean()
    err = float((a - b).abs().max() / b.abs().max())
    c4 = phi4 - phi4.mean()
    err4 = float((c4 - b).abs().max() / b.abs().max())
    # the FFT solver on the same problem (exact for the FD operator)
    from math import pi
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.fft_poisson import FastPoissonSolver
    fs = FastPoissonSolver(nn=3)
    fs.set_grid_descriptor(gd)
    pf = torch.zeros_like(drho)
    fs.solve_neutral(pf, drho)
    tf = []
    for _ in range(5):
        ms, _ = wall(lambda: fs.solve_neutral(pf, drho))
        tf.append(ms)
    r = torch.empty_like(drho)
    relf = (Laplace(gd, -0.25 / pi, 3).residual(pf, drho, r, norm=True) /
            rn2)**0.5
    c = pf - pf.mean()
    errf = float((c - b).abs().max() / b.abs().max())
    return {'fp64_ms': round(min(t64), 3), 'fp64_vcycles': n64,
            'fp64_relres': rel64, 'mixed_ms': round(tmx, 3),
            'mixed_refinements': nmx, 'mixed_relres': relmx,
            'speedup_mixed_vs_fp64': round(min(t64) / tmx, 3),
            'mixed_err_vs_fp64': err,
            'fp64_same_schedule_ms': round(t64s, 3),
            'fp64_same_schedule_refinements': n64s,
            'fp64_same_schedule_relres': rel64s,
            'speedup_precision_alone': round(t64s / tmx, 3),
            'levels4': {'mixed_ms': round(t4, 3), 'refinements': n4,
                        'relres': rel4, 'err_vs_fp64': err4,
                        'fp64_same_schedule_ms': round(t4d, 3),
                        'fp64_same_schedule_refinements': n4d,
                        'fp64_same_schedule_relres': rel4d,
                        'speedup_precision_alone': round(t4d / t4, 3),
                        'speedup_vs_fp64_reference_schedule':
                            round(min(t64) / t4, 3),
                        'cycle': 'V(2,2) over 4 levels (192^3 .. 24^3), '
                                 '64 sweeps on 24^3'},
            'fft_ms': round(min(tf), 3), 'fft_relres': relf,
            'fft_err_vs_fp64': errf,
            'levels': s.levels + 1,
            'mixed_cycle': 'V(%d,%d) FP32, w0=%g, %d per step, 4 grid-wide '
                           'levels + LDS coarse solve' % (pre, post, w0,
                                                          inner)}


def stage_uneven_48(dev, tol=1e-10):
    """Config 4: 48 independent Poisson problems of uneven shape (sides
    from {40, .., 160}, alternately periodic / zero boundary), solved
    together (FP32 V(3,3) in batched launches + per-grid LDS coarse levels
    with Chebyshev iteration on the coarsest, FP64 refinement, one graph
    per step) against the 48 FP64 reference-schedule solves one after the
    other."""
    import torch
    from gpaw_amd import workloads
    from gpaw_amd.batch_poisson import BatchedPoissonSolver
    from gpaw_amd.poisson import FDPoissonSolver
    problems = workloads.uneven_48()
    gds = [g for g, _ in problems]
    rhos = [torch.from_numpy(r).to(dev) for _, r in problems]
    npts = sum(int(np.prod(g.n_c)) for g in gds)
    b = BatchedPoissonSolver(nn=3, levels=4, w0=0.8)
    b.set_grid_descriptors(gds)
    t0 = time.perf_counter()
    b.tune()
    tune_s = time.perf_counter() - t0

    def wall(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3, r

    phis = [torch.zeros_like(r) for r in rhos]
    b.solve_mixed(phis, rhos, tol, coarse_sweeps=16)
    best = None
    for _ in range(2):
        phis = [torch.zeros_like(r) for r in rhos]
        ms, (n, rel) = wall(lambda: b.solve_mixed(phis, rhos, tol,
                                                  coarse_sweeps=16))
        best = ms if best is None else min(best, ms)
    solvers = []
    for g, r in zip(gds, rhos):
        s = FDPoissonSolver(nn=3, eps=tol**2 * float((r**2).sum()) * g.dv,
                            maxiter=1000)
        s.set_grid_descriptor(g)
        solvers.append(s)
    refs = [torch.zeros_like(r) for r in rhos]

    def run64():
        its = []
        for s, r, p in zip(solvers, rhos, refs):
            p.zero_()
            its.append(s.solve_neutral(p, r))
        return its

    run64()
    ms64, its = wall(run64)

    # the batched schedule on one grid at a time: FP32 inner cycles
    # (batching alone) and FP64 inner cycles (precision alone)
    # (every grid's coarsest level, up to 20^3 points, in the LDS-resident
    # coarse solve as in the batch: Chebyshev steps there)
    from gpaw_amd import _lib
    lib = _lib.load()

    def run_seq(fp64_inner):
        ps = [torch.zeros_like(r) for r in rhos]
        ns = []
        lib.gfd_set_tuning(b'coarse_max_points', 8192)
        try:
            for s, r, p in zip(solvers, rhos, ps):
                n, _ = s.solve_mixed(p, r, tol, pre=3, post=3,
                                     coarse_sweeps=16, w0=0.8, inner=1,
                                     fp64_inner=fp64_inner)
                ns.append(n)
        finally:
            lib.gfd_set_tuning(b'coarse_max_points', 0)
        return ns

    run_seq(False)
    ms_seq32, ns32 = wall(lambda: run_seq(False))
    run_seq(True)
    ms_seq64, ns64 = wall(lambda: run_seq(True))
    err = 0.0
    for g, p, q in zip(gds, phis, refs):
        a, c = p, q
        if g.pbc_c.all():
            a, c = a - a.mean(), c - c.mean()
        err = max(err, float((a - c).abs().max() / c.abs().max()))
    return {'batched_ms': round(best, 2), 'refinements': n,
            'max_relres': float(rel.max()), 'tune_s': round(tune_s, 2),
            'fp64_sequential_ms': round(ms64, 2),
            'fp64_vcycles_total': int(sum(its)),
            'speedup': round(ms64 / best, 2), 'max_err_vs_fp64': err,
            'same_schedule_sequential_fp32_ms': round(ms_seq32, 2),
            'same_schedule_sequential_fp64_ms': round(ms_seq64, 2),
            'same_schedule_refinements_fp32_fp64': [int(sum(ns32)),
                                                    int(sum(ns64))],
            'speedup_batching_alone': round(ms_seq32 / best, 2),
            'speedup_precision_alone_sequential': round(ms_seq64 / ms_seq32,
                                                        2),
            'preconditioner': stage_uneven_48_precond(dev, gds),
            'problems': len(gds), 'points': npts,
            'coarse_solve': 'per-grid LDS levels, 16 Chebyshev steps on '
                            'each grid\'s coarsest level',
            'tiles': {str(k): v for k, v in b.tiles().items()}}


def stage_uneven_48_precond(dev, gds, reps=10):
    """The preconditioner half of the 48-problem workload: the FP32 chain
    on every grid in one set of launches (BatchedPreconditioner) against
    the single-grid FP32 preconditioner called grid by grid.  batched_ms:
    residuals and results in the packed batch layout; batched_list_ms:
    per-grid arrays gathered and scattered around it.  Traffic model
    per fine point: FP64 residual read twice (restriction, last pass), FP64
    result written once, FP32 q0 written and read once (40 B), plus the
    coarse levels (~1/8 and 1/64 of that)."""
    import torch
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.preconditioner import BatchedPreconditioner, Preconditioner
    rng = np.random.default_rng(4807)
    res = [torch.from_numpy(rng.uniform(-1, 1, tuple(g.n_c))).to(dev)
           for g in gds]
    npts = sum(int(np.prod(g.n_c)) for g in gds)
    P = BatchedPreconditioner(gds)
    out = P(res)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = P(res, out)
    torch.cuda.synchronize()
    ms_list = (time.perf_counter() - t0) * 1e3 / reps
    # the chain alone, on residuals kept in the packed batch layout
    flat = P.pack(res)
    fout = torch.empty_like(flat)
    P.apply_packed(flat, fout)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        P.apply_packed(flat, fout)
    torch.cuda.synchronize()
    ms_b = (time.perf_counter() - t0) * 1e3 / reps
    singles = [Preconditioner(g, Laplace(g, -0.5, 3), float, fp32=True)
               for g in gds]
    outs = [p(r, None) for p, r in zip(singles, res)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        outs = [p(r, None) for p, r in zip(singles, res)]
    torch.cuda.synchronize()
    ms_s = (time.perf_counter() - t0) * 1e3 / reps
    diff = max(float((a - b).abs().max() / b.abs().max())
               for a, b in zip(out, outs))
    gb = npts * 40 * (1 + 1 / 8 + 1 / 64) / 1e9
    return {'batched_ms': round(ms_b, 3),
            'batched_list_ms': round(ms_list, 3),
            'sequential_fp32_ms': round(ms_s, 3),
            'speedup_batching': round(ms_s / ms_b, 2),
            'GPts/s': round(npts / ms_b / 1e6, 1),
            'model_GB': round(gb, 3),
            'hbm_frac_model': round(gb / ms_b / 8.0, 4),
            'max_diff_vs_single_grid': diff}


def subspace_step(gd, kin, vt, phase, psi, hpsi, out):
    """The subspace diagonalisation of one k-point on the full band set
    (gpaw/new/pwfd/wave_functions.py:289-368, grid part): H = dv psi^H H psi
    (FP64 MFMA, every tile), rocSOLVER zheevd, psi' = V psi (dense FP64
    MFMA rotation).  MFMA fractions count 8 n^2 X flop per product.  Checks:
    eigenvalues against numpy on the same matrix."""
    import torch
    from gpaw_amd.bands import (apply_hamiltonian, subspace_diagonalize,
                                subspace_matrix)
    nb = psi.shape[0]
    X = int(np.prod(psi.shape[1:]))
    apply_hamiltonian(kin, vt, psi, hpsi, phase)
    Hm = torch.empty((nb, nb), dtype=torch.complex128, device=psi.device)
    ms_h, _ = timed(lambda: subspace_matrix(psi, hpsi, gd.dv, Hm), 3, 1)
    H0 = Hm.clone()
    eps = torch.empty(nb, dtype=torch.float64, device=psi.device)
    small = torch.zeros((nb, 1, 1, 128), dtype=torch.complex128,
                        device=psi.device)
    small_out = torch.empty_like(small)

    def eig_only():
        Hm.copy_(H0)
        subspace_diagonalize(Hm, small, small_out, eps)

    def full():
        Hm.copy_(H0)
        subspace_diagonalize(Hm, psi, out, eps)

    ms_eig, _ = timed(eig_only, 2, 1)
    ms_full, _ = timed(full, 2, 1)
    ms_rot = max(ms_full - ms_eig, 1e-6)
    w = np.linalg.eigvalsh(H0.cpu().numpy(), UPLO='L')
    e = eps.cpu().numpy()
    flop = 8.0 * nb * nb * X
    return {'subspace_matrix_ms': round(ms_h, 3),
            'subspace_matrix_mfma_frac': round(flop / ms_h / 1e9 / 78.6, 4),
            'zheevd_ms': round(ms_eig, 3),
            'rotation_ms': round(ms_rot, 3),
            'rotation_mfma_frac': round(flop / ms_rot / 1e9 / 78.6, 4),
            'total_ms': round(ms_h + ms_full, 3),
            'eigenvalue_err_vs_numpy': float(np.abs(e - w).max() /
                                             np.abs(w).max())}


def stage_precond_128(dev, nbands=128, reps=5):
    """The FP32 kinetic preconditioner on one 128-band block of the
    1024-band workload (128^3 c128, first k-point): reps applications,
    then the same with the FP64 chain.
    Traffic model per fine point: complex128 residual read twice and result
    written once (48 B; the fine level is one fused pass, its interpolated
    d0' never stored), the coarse levels ~1/7 of the two-pass fine work
    (64 B / 7)."""
    import torch
    from gpaw_amd import workloads
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.preconditioner import Preconditioner
    gd = workloads.eigen_128_gd()
    ph = workloads.eigen_128_kpoints()[0]
    kin = Laplace(gd, -0.5, 3, complex)
    P = Preconditioner(gd, kin, complex, block=nbands, fp32=True)
    g = torch.Generator(device=dev)
    g.manual_seed(1128)
    R = torch.randn((nbands, 128, 128, 128), dtype=torch.complex128,
                    device=dev, generator=g)
    out = torch.empty_like(R)

    class K:
        phase_cd = ph
    ms, _ = timed(lambda: P(R, K, out=out), reps, 1)
    pts = nbands * 128**3
    gb = pts * (48 + 64 / 7) / 1e9
    # the FP64 chain (the reference's precision) on the same block
    del P
    P64 = Preconditioner(gd, kin, complex, block=nbands)
    ms64, _ = timed(lambda: P64(R, K, out=out), reps, 1)
    return {'ms': round(ms, 3), 'bands': nbands,
            'ms_per_1024_bands': round(ms * 1024 / nbands, 2),
            'GPts/s': round(pts / ms / 1e6, 1), 'model_GB': round(gb, 2),
            'hbm_frac_model': round(gb / ms / 8.0, 4),
            'fp64_chain_ms': round(ms64, 3),
            'fp64_chain_ms_per_1024_bands': round(ms64 * 1024 / nbands, 2)}


def stage_density_96(dev, reps=20):
    """nn=3 restriction 96^3 -> 48^3 and interpolation back on the four
    densities of the batched-stencil workload."""
    import torch
    from gpaw_amd import workloads
    from gpaw_amd.transformers import Transformer
    gdd, dens = workloads.density_96()
    coarse = gdd.coarsen()
    R = Transformer(gdd, coarse, 3)
    I = Transformer(coarse, gdd, 3)
    d = torch.from_numpy(dens).to(dev)
    c = torch.empty((4, 48, 48, 48), dtype=torch.float64, device=dev)
    f = torch.empty_like(d)
    ms_r, _ = timed(lambda: R.apply(d, c), reps, 3)
    ms_i, _ = timed(lambda: I.apply(c, f), reps, 3)
    fine_pts = 4 * 96**3
    return {'restrict_ms': round(ms_r, 4),
            'restrict_GPts/s_fine': round(fine_pts / ms_r / 1e6, 1),
            'interpolate_ms': round(ms_i, 4),
            'interpolate_GPts/s_fine': round(fine_pts / ms_i / 1e6, 1)}


def stage_eigen_128(dev, nbands, steps=2, warmup=1, block=128):
    """One RMM-DIIS iteration per k-point (4 Bloch k-points) on nbands
    complex bands of 128^3: the residuals R = (-1/2 lap(nn=3) + V - eps_n)
    psi_n in one fused stencil pass (eps_n the band energies), the block
    step (two FP32 multigrid preconditioner applications, one more fused
    residual, the step lengths and the updates), S = dv psi psi^H (FP64
    MFMA, lower tiles) and psi <- chol(S)^-1 psi."""
    import torch
    from gpaw_amd import workloads
    from gpaw_amd.bands import apply_hamiltonian, band_dots, orthonormalize
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.preconditioner import Preconditioner
    from gpaw_amd.rmmdiis import BlockStep
    gd = workloads.eigen_128_gd()
    phases = workloads.eigen_128_kpoints()
    vt = torch.from_numpy(workloads.eigen_128_potential(gd)).to(dev)
    kin = Laplace(gd, -0.5, 3, complex)
    P = Preconditioner(gd, kin, complex, block=block, fp32=True)
    g = torch.Generator(device=dev)
    g.manual_seed(1006)
    shape = (nbands, 128, 128, 128)
    psi = torch.randn(shape, dtype=torch.complex128, device=dev,
                      generator=g)
    hpsi = torch.empty_like(psi)
    S = torch.zeros((nbands, nbands), dtype=torch.complex128, device=dev)
    # band energies eps_n = <psi_n|H|psi_n> / <psi_n|psi_n> (first k-point)
    apply_hamiltonian(kin, vt, psi, hpsi, phases[0])
    eps = band_dots(psi, hpsi, gd.dv) / band_dots(psi, psi, gd.dv)
    ms_dots, _ = timed(lambda: band_dots(psi, hpsi, gd.dv), 3, 1)
    rmm = BlockStep(kin, vt, P)

    class Kpt:
        phase_cd = None

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    acc = np.zeros(4)

    bufs = [psi, hpsi]

    def step(record):
        for k in range(len(phases)):
            psi, R = bufs
            Kpt.phase_cd = phases[k]
            ev[0].record()
            apply_hamiltonian(kin, vt, psi, R, phases[k], eps=eps)
            ev[1].record()
            rmm(psi, R, eps, phases[k])
            ev[2].record()
            # orthonormalised bands land in R (free after the block step);
            # S is computed inside orthonormalize
            orthonormalize(psi, gd.dv, out=R, S=S)
            bufs.reverse()
            ev[3].record()
            torch.cuda.synchronize()
            if record:
                acc[0] += ev[0].elapsed_time(ev[1])
                acc[1] += ev[1].elapsed_time(ev[2])
                acc[2] += ev[2].elapsed_time(ev[3])

    for _ in range(warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    psi, R = bufs
    # the overlap and the preconditioner alone
    from gpaw_amd.bands import overlap
    ms_s, _ = timed(lambda: overlap(psi, gd.dv, True, S), 3, 1)
    Kpt.phase_cd = phases[-1]
    ms_pc, _ = timed(lambda: P(R, Kpt, out=rmm._PR), 3, 1)
    # checks at full size: the rotated bands are orthonormal (their own
    # overlap is the identity); one block step lowers the residual norm
    # (R + lambda dR is the residual of psi + lambda PR); the FP32
    # preconditioner against the FP64 one on 8 bands
    overlap(psi, gd.dv, False, S)
    orth_err = float((S - torch.eye(nbands, dtype=S.dtype,
                                    device=S.device)).abs().max())
    apply_hamiltonian(kin, vt, psi, R, phases[-1])
    eps = band_dots(psi, R, gd.dv)
    apply_hamiltonian(kin, vt, psi, R, phases[-1], eps=eps)
    r0 = float(band_dots(R, R, gd.dv).sum())
    rmm(psi, R, eps, phases[-1])
    r1 = float(band_dots(rmm._dR, rmm._dR, gd.dv).sum())
    P64 = Preconditioner(gd, kin, complex, block=8)
    ref = P64(psi[:8], Kpt)
    got = P(psi[:8], Kpt)
    pre_err = float((got - ref).abs().max() / ref.abs().max())
    del P64, ref, got
    # H psi in R, the rotated bands in the block step's (free) PR buffer
    sub = subspace_step(gd, kin, vt, phases[-1], psi, R, rmm._PR)
    X = 128**3
    nk = len(phases)
    herk_flop = nbands * (nbands + 1) / 2 * X * 8
    h_bytes = nbands * X * 32
    per = acc / (steps * nk)
    upd_bytes = nbands * X * 16 * (6 + 3)
    return {'ms_per_step_4k': round(wall, 2),
            'h_residual_ms_per_k': round(per[0], 3),
            'h_residual_hbm_frac': round(h_bytes / per[0] / 1e6 /
                                         HBM_PEAK_GBS, 4),
            'band_dots_ms': round(ms_dots, 3),
            'band_dots_hbm_frac': round(h_bytes / ms_dots / 1e6 /
                                        HBM_PEAK_GBS, 4),
            'rmm_block_step_ms_per_k': round(per[1], 3),
            'precond_ms': round(ms_pc, 3),
            'rmm_step_other_ms': round(per[1] - 2 * ms_pc - per[0] -
                                       ms_dots, 3),
            'rmm_update_bytes_gb': round(upd_bytes / 1e9, 1),
            'rmm_residual_ratio': r1 / r0,
            'overlap_orth_ms_per_k': round(per[2], 3),
            'overlap_ms': round(ms_s, 3),
            'overlap_mfma_tflops': round(herk_flop / ms_s / 1e9, 2),
            'overlap_mfma_frac': round(herk_flop / ms_s / 1e9 / 78.6, 4),
            'orthonormality_err': orth_err,
            'precond_fp32_relerr_vs_fp64': pre_err,
            'subspace': sub,
            'bands': nbands, 'kpoints': nk, 'grid': [128, 128, 128],
            'note': 'RMM-DIIS (diis_steps=1) with the FP32 preconditioner; '
                    'residuals, step lengths, overlap and Cholesky in '
                    'FP64'}


def cpu_spot_check(gd, op_coefs, op_offsets, spot):
    """The oracle as the checker of two full-size bands of the timed GPU
    apply: max relative error per band."""
    import oracle
    out = {}
    for band, (a, got) in spot.items():
        ref = oracle.fd_apply(op_coefs, op_offsets, gd.n_c, gd.neighbor_cd,
                              a)
        out[f'band{band}_relerr'] = float(np.abs(got - ref).max() /
                                          np.abs(ref).max())
    return out


def cpu_baseline(gd, op_coefs, op_offsets, nbands_sample):
    """The reference C extension (oracle/_ref, built from the reference
    sources) on the host cores: Laplace nn=3 over a sample of the bands."""
    threads = len(os.sched_getaffinity(0))
    os.environ['OMP_NUM_THREADS'] = str(threads)
    so = glob.glob(os.path.join(ROOT, 'oracle', '_ref', '_gpaw_ref*.so'))
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 1, (nbands_sample,) + tuple(gd.n_c))
    b = np.empty_like(a)
    n_c = np.ascontiguousarray(gd.n_c)
    if so:
        sys.path.insert(0, os.path.dirname(so[0]))
        import _gpaw_ref
        r = 3
        M = n_c + 2 * r
        off = np.ascontiguousarray(op_offsets @ np.array([M[1] * M[2], M[2],
                                                          1]))
        op = _gpaw_ref.Operator(np.ascontiguousarray(op_coefs), off, n_c, r,
                                np.ascontiguousarray(gd.neighbor_cd), 1, None,
                                1)
        run = lambda: op.apply(a, b)  # noqa: E731
        kind = 'reference'
    else:
        import oracle
        run = lambda: oracle.fd_apply(op_coefs, op_offsets, n_c,  # noqa
                                      gd.neighbor_cd, a, out=b,
                                      nthreads=threads)
        kind = 'port'
    run()
    reps = 0
    t0 = time.perf_counter()
    while True:
        run()
        reps += 1
        el = time.perf_counter() - t0
        if el > 12.0 or reps >= 5000:
            break
    pts = reps * a.size
    return {'value': round(pts / el / 1e9, 4), 'unit': 'GPts/s',
            'cores': threads, 'kind': kind,
            'sample': f'{nbands_sample} of the 256 bands (96^3, f64), '
                      f'{reps} applies, {el:.1f} s, OpenMP over bands'}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; '
                          f'--gpus {args.gpus} is not supported'}))
        sys.exit(2)
    import torch
    import gpaw_amd
    from gpaw_amd import _lib, workloads
    from gpaw_amd.fd_operators import Gradient, Laplace
    _lib.load()
    torch.cuda.init()
    dev = torch.device('cuda', 0)
    if args.stage:
        fn = {'poisson192': lambda: stage_poisson_192(dev),
              'uneven48': lambda: stage_uneven_48(dev),
              'precond128': lambda: stage_precond_128(dev),
              'density96': lambda: stage_density_96(dev),
              'eigen': lambda: stage_eigen_128(dev, args.eigen_bands,
                                               steps=1, warmup=1)}
        print(json.dumps({args.stage: fn[args.stage]()}, default=float),
              flush=True)
        return

    gd = workloads.bands_96_gd()
    nb = args.bands
    lap = Laplace(gd, 1.0, 3)
    log(f'[bench] {nb} bands 96^3 f64, Laplace nn=3 ({lap.npoints} points)')
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    a = torch.rand((nb, 96, 96, 96), dtype=torch.float64, device=dev,
                   generator=g) * 2 - 1
    b = torch.empty_like(a)
    if args.pmc:
        for _ in range(3):
            lap.apply(a, b)
        torch.cuda.synchronize()
        lib = _lib.load()
        nbytes = a.numel() * 8
        for _ in range(3):
            _lib.check(lib.gfd_stream_copy(a.data_ptr(), b.data_ptr(),
                                           nbytes, torch.cuda.current_stream().cuda_stream), "copy")
        torch.cuda.synchronize()
        log(f'[bench] pmc mode: 3 Laplace applies, 3 copies of {nbytes} B')
        return
    ms, wall = timed(lambda: lap.apply(a, b), args.steps, args.warmup)
    pts = nb * 96**3
    value = pts / (ms * 1e-3) / 1e9
    alg_bytes = pts * 16
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    log(f'[bench] laplace: {ms:.4f} ms/step (wall {wall:.4f}), '
        f'{value:.1f} GPts/s, {achieved:.0f} GB/s')

    # a size-free property at full size (the periodic Laplacian of any
    # field sums to zero); two bands are kept on the host for the oracle
    # spot check of the CPU-baseline leg
    parity = {}
    s = b.sum(dim=(1, 2, 3)).abs().max().item()
    parity['max_band_sum_over_l1'] = s / b.abs().sum(dim=(1, 2, 3)).min().item()
    parity['ok'] = bool(parity['max_band_sum_over_l1'] < 1e-12)
    spot = {band: (a[band].cpu().numpy(), b[band].cpu().numpy())
            for band in (0, nb - 1)}
    log('[bench] parity (property)', parity)
    # measured streaming ceiling: a 16-byte copy of the same bytes (after
    # the parity check: it overwrites b)
    lib = _lib.load()
    ms_copy, _ = timed(lambda: _lib.check(lib.gfd_stream_copy(
        a.data_ptr(), b.data_ptr(), a.numel() * 8, torch.cuda.current_stream(
        ).cuda_stream), 'copy'), 10, 2)
    copy_gbs = alg_bytes / (ms_copy * 1e-3) / 1e9
    log(f'[bench] copy ceiling {copy_gbs:.0f} GB/s')

    stages = {}
    if not args.no_stages:
        gx = Gradient(gd, 0, 1.0, 3)
        ms_g, _ = timed(lambda: gx.apply(a, b), max(5, args.steps // 2), 2)
        stages['gradient_x_nn3_bands_96'] = {
            'ms': round(ms_g, 4), 'GPts/s': round(pts / ms_g / 1e6, 1),
            'hbm_frac': round(alg_bytes / ms_g / 1e6 / HBM_PEAK_GBS, 4)}
        del a, b
        torch.cuda.empty_cache()
        from gpaw_amd.transformers import Transformer
        gdd, dens = workloads.density_96()
        coarse = gdd.coarsen()
        R = Transformer(gdd, coarse, 3)
        I = Transformer(coarse, gdd, 3)
        d = torch.from_numpy(dens).to(dev)
        c = torch.empty((4, 48, 48, 48), dtype=torch.float64, device=dev)
        f = torch.empty_like(d)
        ms_r, _ = timed(lambda: R.apply(d, c), 20, 3)
        ms_i, _ = timed(lambda: I.apply(c, f), 20, 3)
        fine_pts = 4 * 96**3
        stages['restrict_nn3_density_96'] = {
            'ms': round(ms_r, 4),
            'GPts/s_fine': round(fine_pts / ms_r / 1e6, 1),
            'note': 'working set 36 MB: Infinity-Cache resident'}
        stages['interpolate_nn3_density_96'] = {
            'ms': round(ms_i, 4),
            'GPts/s_fine': round(fine_pts / ms_i / 1e6, 1)}
        from gpaw_amd.poisson import FDPoissonSolver
        gdp, rho = workloads.poisson_48()
        ps = FDPoissonSolver(nn=3, eps=2e-10)
        ps.set_grid_descriptor(gdp)
        drho = torch.from_numpy(rho).to(dev)
        phi = torch.zeros_like(drho)
        ps.solve(phi, drho)
        torch.cuda.synchronize()
        times = []
        for _ in range(3):
            phi.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            niter = ps.solve(phi, drho)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        stages['poisson_fp64_48'] = {
            'ms_to_eps': round(min(times) * 1e3, 3), 'niter': niter,
            'eps': 2e-10, 'error': ps.last_error}
        from gpaw_amd.preconditioner import Preconditioner
        L = 128 * 0.18 / gpaw_amd.Bohr
        from gpaw_amd.grid_descriptor import GridDescriptor
        gdk = GridDescriptor((128, 128, 128), (L, L, L))
        kin = Laplace(gdk, -0.5, 3, complex)
        P = Preconditioner(gdk, kin, complex, block=32)
        res = torch.rand((32, 128, 128, 128), dtype=torch.complex128,
                         device=dev)
        out = torch.empty_like(res)

        class Kpt:
            phase_cd = np.exp(2j * np.pi * np.array([[0.1], [0.0], [0.25]]) *
                              np.array([1, -1]))
        ms_p, _ = timed(lambda: P(res, Kpt, out=out), 5, 1)
        stages['precond_c128_32x128'] = {
            'ms': round(ms_p, 3),
            'GPts/s': round(32 * 128**3 / ms_p / 1e6, 1),
            'chain': 'FP64 (the reference precision), fused fine level'}
        # weighted operator of the dielectric Poisson solver: eps Laplace +
        # grad(eps) . Gradient (nn=3) on 16 densities of 96^3, weights shared
        from gpaw_amd.fd_operators import Gradient
        from gpaw_amd.wfd_operators import WeightedFDOperator
        Aw = WeightedFDOperator([Laplace(gd, 1.0, 3)] +
                                [Gradient(gd, v, 1.0, 3) for v in range(3)])
        gw = torch.Generator(device=dev)
        gw.manual_seed(77)
        wts = [1.0 + 78.0 * torch.rand((96, 96, 96), dtype=torch.float64,
                                       device=dev, generator=gw)]
        wts += [torch.randn((96, 96, 96), dtype=torch.float64, device=dev,
                            generator=gw) for _ in range(3)]
        Aw.set_weights(wts)
        dw = torch.rand((16, 96, 96, 96), dtype=torch.float64, device=dev,
                        generator=gw)
        ow = torch.empty_like(dw)
        ms_w, _ = timed(lambda: Aw.apply(dw, ow), 5, 1)
        wbytes = 16 * 96**3 * 16 + 4 * 96**3 * 8
        stages['wfd_dielectric_nn3_16x96'] = {
            'ms': round(ms_w, 3),
            'GPts/s': round(16 * 96**3 / ms_w / 1e6, 1),
            'hbm_frac': round(wbytes / ms_w / 1e6 / HBM_PEAK_GBS, 4),
            'bytes': wbytes}
        del Aw, wts, dw, ow
        log('[bench] stages', stages)
        torch.cuda.empty_cache()
        stages['poisson_192_time_to_1e-10'] = stage_poisson_192(dev)
        log('[bench] poisson 192', stages['poisson_192_time_to_1e-10'])
        del res, out, P
        torch.cuda.empty_cache()
        stages['uneven_48_batched_poisson'] = stage_uneven_48(dev)
        log('[bench] uneven 48', stages['uneven_48_batched_poisson'])
        torch.cuda.empty_cache()
        if not args.no_eigen:
            stages['eigen_step_c128_128'] = stage_eigen_128(
                dev, args.eigen_bands)
            log('[bench] eigen', stages['eigen_step_c128_128'])
            torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        log('[bench] CPU baseline ...')
        parity.update(cpu_spot_check(gd, lap.coef_p, lap.offset_pc, spot))
        parity['ok'] = bool(parity['ok'] and max(
            parity['band0_relerr'], parity[f'band{nb - 1}_relerr']) < 1e-13)
        log('[bench] parity', parity)
        cpu = cpu_baseline(gd, lap.coef_p, lap.offset_pc, 16)
        log('[bench] cpu', cpu)

    traffic = traffic_from_profiles('laplace_nn3_bands_96')
    line = {
        'metric': METRIC,
        'value': round(value, 2),
        'unit': 'GPts/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (uniform [-1, 1] bands, seeded)',
        'config': {'workload': 'laplace_nn3_bands_96', 'bands': nb,
                   'grid': [96, 96, 96], 'h_angstrom': 0.2,
                   'stencil': '19-point nn=3 (O(h^6))',
                   'boundary': 'periodic', 'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1),
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': traffic},
        'cpu_baseline': cpu,
        'copy_ceiling': {'GB/s': round(copy_gbs, 1),
                         'laplace_frac_of_copy': round(achieved / copy_gbs,
                                                       4),
                         'note': '16-byte device copy of the same 3.62 GB'},
        'parity': parity,
        'stages': stages,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
