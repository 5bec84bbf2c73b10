"""Multigrid FD Poisson solver on the GPU (drop-in for
gpaw.poisson.FDPoissonSolver, gpaw/poisson.py:114-554).

The equation is -lap(phi) / 4 pi = rho (Gaussian units).  The solver keeps
the reference's public surface -- constructor keywords, ``set_grid_
descriptor``, ``solve`` (charge handling), ``solve_neutral``, ``iterate2``,
``todict``, ``get_description``, ``estimate_memory`` -- and its numerics:
the multigrid hierarchy and V-cycle schedule (below), which fix the
iterates.  The V-cycles run in the native driver (gfd_poisson_*), one call
per solve; the charge corrections are set up on the host once per grid.

Schedule (gpaw/poisson.py:366-405, 451-472, 500-537): levels are halvings
of the grid while every side stays even, at most 8; the fine level uses
the caller's stencil (nn points per side, or the Mehrstellen pair for
nn='M'), coarser ones the 7-point Laplacian; 2 pre- and 1 post-smoothing
sweep on the fine level with Jacobi weight 2/3, 4 and 4 with weight 1
below, 8 post-sweeps on the coarsest; restriction and interpolation with
nn=1 transformers.
"""
from __future__ import annotations

import warnings
from math import erf, pi, sqrt

import numpy as np

from . import _lib
from .arrays import Staging, is_device, ptr, stream_handle, to_device
from .fd_operators import Laplace, LaplaceA, LaplaceB
from .transformers import Transformer


class ConvergenceError(Exception):
    pass


class PoissonConvergenceError(ConvergenceError):
    """Raised when the V-cycles reach maxiter (gpaw/__init__.py:104)."""


_COARSE_WARNING = (
    'The FD multigrid Poisson solver cannot coarsen this grid far enough: '
    'its coarsest level still has more than 36 points along an axis, so '
    'the V-cycles converge slowly or not at all.  Use grid sizes with more '
    'factors of 2, or the FFT solver (PoissonSolver("fast")).')

_RELAX = {'J': (2, 'Jacobi'), 'GS': (1, 'Gauss-Seidel')}


# ------------------------------------------------------ charge corrections

def gaussian_charge(gd, centre=None, width=19.0):
    """(rho_g, phi_g): a unit charge (a / pi)^{3/2} exp(-a r^2) centred at
    `centre` (default: the cell centre) and its potential erf(sqrt(a) r) / r
    (lap phi_g = -4 pi rho_g), on the grid of gd (host arrays).

    The exponent puts exp(-a r^2) at exp(-width) at the nearest cell face:
    a = 4 width max_c |b_c|^2 (b_c the rows of the inverse cell) for the
    centred charge, width / r_min^2 for another centre (r_min its distance
    to the nearest face), as gpaw/utilities/gauss.py:127-135 chooses it.
    Distances below 1e-6 bohr are taken as 1e-6 (the potential is smooth
    there: erf(sqrt(a) r) / r -> 2 sqrt(a / pi))."""
    half = 0.5 * gd.cell_cv.sum(0)
    if centre is None:
        centre = half
        a = 4.0 * width * (gd.icell_cv**2).sum(1).max()
    else:
        centre = np.asarray(centre, float)
        r_min = (0.5 * gd.cell_cv.sum(1) - np.abs(centre - half)).min()
        a = width / r_min**2
    r_vG = gd.get_grid_point_coordinates() - centre[:, None, None, None]
    r = np.sqrt(np.maximum((r_vG**2).sum(0), 1e-12))
    rho = (a / pi)**1.5 * np.exp(-a * r * r)
    verf = np.vectorize(erf, otypes=[float])
    phi = verf(sqrt(a) * r) / r
    return rho, phi


def madelung(cell_cv, tol=1e-15, G=None):
    """Potential at a lattice site of the periodic images (the site itself
    excluded) of a charge -1 plus the neutralising uniform background:
    the constant the reference adds (per unit of net charge) to the
    potential of a charged periodic system (gpaw/utilities/ewald.py:143).

    Ewald summation with splitting parameter G:
      psi(unit charge) = sum_{l != 0} erfc(G |l|) / |l|
                         + (4 pi / V) sum_{g != 0} exp(-g^2 / 4 G^2) / g^2
                         - 2 G / sqrt(pi) - pi / (V G^2),
    summed until the omitted terms are below tol; the result, -psi, does
    not depend on G (default sqrt(pi) / V^(1/3), balancing both sums)."""
    from scipy.special import erfc
    cell = np.asarray(cell_cv, float)
    V = abs(np.linalg.det(cell))
    if G is None:
        G = sqrt(pi) / V**(1.0 / 3.0)
    recip = 2 * pi * np.linalg.inv(cell).T

    def shells(basis, cutoff):
        # all lattice vectors of `basis` shorter than cutoff
        heights = 1.0 / np.linalg.norm(np.linalg.inv(basis).T, axis=1)
        m = np.ceil(cutoff / heights).astype(int) + 1
        idx = np.stack(np.meshgrid(*[np.arange(-k, k + 1) for k in m],
                                   indexing='ij'), -1).reshape(-1, 3)
        v = idx @ basis
        d = np.linalg.norm(v, axis=1)
        return d[(d > 0) & (d < cutoff)]

    # erfc(x) < tol for x > ~5.9 (tol 1e-15); exp(-y) < tol for y > 34.6
    x_max = 6.5
    real = shells(cell, x_max / G)
    g = shells(recip, 2 * G * sqrt(-np.log(tol) + 2))
    psi = (erfc(G * real) / real).sum()
    psi += 4 * pi / V * (np.exp(-g**2 / (4 * G * G)) / g**2).sum()
    psi -= 2 * G / sqrt(pi) + pi / (V * G * G)
    return -psi


# ------------------------------------------------------------- base class

class _PoissonSolver:
    """Interface of every Poisson solver (gpaw/poisson.py:86-111)."""

    def set_grid_descriptor(self, gd):
        raise NotImplementedError

    def solve(self, *args, **kwargs):
        raise NotImplementedError

    def todict(self):
        raise NotImplementedError(type(self).__name__)

    def get_description(self):
        return type(self).__name__

    def estimate_memory(self, mem):
        raise NotImplementedError


class BasePoissonSolver(_PoissonSolver):
    """Charge handling shared by the FD and FFT solvers
    (gpaw/poisson.py:114-319): a neutral density is solved as is; a
    charged one has its charge taken out first -- by a uniform background
    on fully periodic cells (plus, optionally, the Madelung constant
    times the charge), by a Gaussian of the same charge, whose potential
    is added back analytically, on cells with zero boundaries on every
    axis."""

    def __init__(self, *, remove_moment=None, use_charge_center=False,
                 metallic_electrodes=False, eps=None,
                 use_charged_periodic_corrections=False, xp=np):
        if metallic_electrodes not in (False, None, 'single', 'both'):
            raise AssertionError('metallic_electrodes: False, None, '
                                 "'single' or 'both'")
        self.xp = xp
        self.gd = None
        self.eps = eps
        self.remove_moment = remove_moment
        self.use_charge_center = use_charge_center
        self.metallic_electrodes = metallic_electrodes
        self.use_charged_periodic_corrections = \
            use_charged_periodic_corrections
        self.charged_periodic_correction = None
        self._gauss = None

    def todict(self):
        options = {'remove_moment': self.remove_moment,
                   'use_charge_center': self.use_charge_center,
                   'use_charged_periodic_corrections':
                       self.use_charged_periodic_corrections,
                   'metallic_electrodes': self.metallic_electrodes}
        d = {'name': 'basepoisson'}
        d.update({k: v for k, v in options.items() if v})
        return d

    def get_description(self):
        notes = []
        if self.remove_moment is not None:
            notes.append(f'    Multipole moments removed up to '
                         f'L={self.remove_moment}')
        if self.use_charge_center:
            notes.append('    Net charge compensated at the centre of the '
                         'majority charge')
        if self.use_charged_periodic_corrections:
            notes.append('    Madelung correction for the neutralising '
                         'background')
        return '\n'.join(notes)

    def _gaussian(self, rho):
        """Device (rho_g, phi_g) of the compensating Gaussian (cached per
        grid; recomputed for a charge-centred one)."""
        centre = None
        if self.use_charge_center:
            centre = self._charge_centre(rho)
        elif self._gauss is not None:
            return self._gauss
        r, p = gaussian_charge(self.gd, centre)
        g = (to_device(r), to_device(p)) if is_device(rho) else (r, p)
        if centre is None:
            self._gauss = g
        return g

    def _charge_centre(self, rho):
        """Centre of the majority-sign charge, at least 7 grid spacings
        from every face (gpaw/poisson.py:248-264)."""
        gd = self.gd
        a = rho.detach().cpu().numpy() if is_device(rho) else \
            np.asarray(rho)
        q = gd.integrate(a)
        major = np.where(a * np.sign(q) > 0, a * np.sign(q), 0.0)
        r_vG = gd.get_grid_point_coordinates()
        centre = np.array([(major * r_vG[v]).sum() for v in range(3)]) / \
            major.sum()
        margin = gd.h_cv.T @ np.full(3, 7.0)
        upper = gd.h_cv.T @ gd.N_c - margin
        if (centre > upper).any() or (centre < margin).any():
            raise RuntimeError('Poisson solver: center of charge outside '
                               'borders - please increase box')
        return centre

    def solve(self, phi, rho, charge=None, maxcharge=1e-6,
              zero_initial_phi=False, timer=None):
        """Solve for phi (in place; initial guess unless zero_initial_phi)
        and return the number of iterations (gpaw/poisson.py:182-312)."""
        self._init()
        gd = self.gd
        if tuple(phi.shape) != tuple(gd.n_c) or \
                tuple(rho.shape) != tuple(gd.n_c):
            raise AssertionError('phi and rho must have the grid shape')
        q = gd.integrate(rho)
        if self.remove_moment:
            raise NotImplementedError(
                'remove_moment (deprecated in the reference, which points '
                'to its moment-correction wrapper) is not supported')
        if charge is None:
            charge = q
        # uniform density of the actual charge
        background = q / (gd.dv * np.prod(gd.get_size_of_global_array()))
        if abs(charge) <= maxcharge:
            return self.solve_neutral(phi, rho - background, timer=timer)
        if gd.pbc_c.all():
            if zero_initial_phi:
                phi[...] = 0.0
            niter = self.solve_neutral(phi, rho - background, timer=timer)
            if self.use_charged_periodic_corrections:
                if self.charged_periodic_correction is None:
                    self.charged_periodic_correction = madelung(gd.cell_cv)
                phi += q * self.charged_periodic_correction
            return niter
        if not gd.pbc_c.any():
            rho_g, phi_g = self._gaussian(rho)
            if zero_initial_phi:
                phi[...] = 0.0
            else:
                phi -= q * phi_g
            niter = self.solve_neutral(phi, rho - q * rho_g, timer=timer)
            phi += q * phi_g
            return niter
        if self.metallic_electrodes == 'both':
            return self.solve_neutral(phi, rho, timer=timer)
        if self.metallic_electrodes == 'single':
            raise NotImplementedError(
                'metallic_electrodes="single" needs the dipole-layer '
                'correction of gpaw/dipole_correction.py, outside the FD '
                'grid path')
        raise NotImplementedError('Charged systems with mixed periodic/zero '
                                  'boundary conditions')


# ------------------------------------------------------------- FD solver

class FDPoissonSolver(BasePoissonSolver):
    def __init__(self, nn=3, relax='J', eps=2e-10, maxiter=1000,
                 remove_moment=None, use_charge_center=False,
                 metallic_electrodes=False,
                 use_charged_periodic_corrections=False, **kwargs):
        if relax not in _RELAX:
            raise NotImplementedError(f'Relaxation method {relax}')
        super().__init__(
            remove_moment=remove_moment, use_charge_center=use_charge_center,
            metallic_electrodes=metallic_electrodes,
            use_charged_periodic_corrections=use_charged_periodic_corrections,
            **kwargs)
        self.nn = nn
        self.relax = relax
        self.relax_method = _RELAX[relax][0]
        self.eps = eps
        self.maxiter = maxiter
        self.description = None
        self.last_error = None
        self._initialized = False
        self._native = None
        self._rho_b = None

    def todict(self):
        d = super().todict()
        d.update(name='fd', nn=self.nn, relax=self.relax, eps=self.eps)
        return d

    def get_stencil(self):
        return self.nn

    def create_laplace(self, gd, scale=1.0, n=1, dtype=float):
        """The Laplacian of one level (a hook for subclasses, as in the
        reference)."""
        return Laplace(gd, scale, n, dtype, xp=self.xp)

    def set_grid_descriptor(self, gd):
        """Build the multigrid hierarchy for gd (schedule in the module
        docstring)."""
        self._free_native()
        self.gd = gd
        self._gauss = None
        self._rho_b = None
        fine = -0.25 / pi                       # -lap / 4 pi
        if self.nn == 'M':
            if not gd.orthogonal:
                raise RuntimeError('Cannot use Mehrstellen stencil with '
                                   'non orthogonal cell.')
            self.operators = [LaplaceA(gd, -fine, xp=self.xp)]
            self.B = LaplaceB(gd)
        else:
            self.operators = [self.create_laplace(gd, fine, self.nn)]
            self.B = None
        self.restrictors, self.interpolators = [], []
        self.presmooths, self.postsmooths = [2], [1]
        self.weights = [2.0 / 3.0]
        level_gd = gd
        for _ in range(8):
            try:
                coarse = level_gd.coarsen()
            except ValueError:
                break
            self.operators.append(self.create_laplace(coarse, fine, 1))
            self.restrictors.append(Transformer(level_gd, coarse,
                                                xp=self.xp))
            self.interpolators.append(Transformer(coarse, level_gd,
                                                  xp=self.xp))
            self.presmooths.append(4)
            self.postsmooths.append(4)
            self.weights.append(1.0)
            level_gd = coarse
        self.levels = len(self.restrictors)
        self._initialized = False
        if self.operators[-1].gd.N_c.max() > 36:
            warnings.warn(_COARSE_WARNING + '\n' + self.get_description(),
                          stacklevel=2)

    def get_description(self):
        method = _RELAX[self.relax][1]
        coarsest = self.operators[-1].gd.N_c
        text = [f'{method} multigrid solver, {self.levels + 1} levels',
                f'    Coarsest grid: {" x ".join(str(n) for n in coarsest)} '
                'points']
        if coarsest.max() > 24:
            text.append('    (more than 24 points per side on the coarsest '
                        'level: add levels)')
        text += [f'    Stencil: {self.operators[0].description}',
                 f'    Max iterations: {self.maxiter}',
                 f'    Tolerance: {self.eps:e}']
        base = super().get_description()
        if base:
            text.append(base)
        return '\n'.join(text)

    def _init(self):
        if self._initialized:
            return
        self.step = 0.66666666 / self.operators[0].get_diagonal_element()
        self.presmooths[self.levels] = 8
        self.postsmooths[self.levels] = 8
        self._initialized = True

    def _free_native(self):
        if getattr(self, '_native', None) is not None:
            _lib.load().gfd_poisson_free(self._native)
            self._native = None

    def __del__(self):
        try:
            self._free_native()
        except Exception:
            pass

    def _handle(self):
        self._init()
        if self._native is None:
            lib = _lib.load()
            L = self.levels
            V = _lib.ctypes.c_void_p

            def handles(objs):
                return (V * max(len(objs), 1))(
                    *([o._handle().value for o in objs] or [None]))
            h = _lib.ctypes.c_void_p()
            _lib.check(lib.gfd_poisson_new(
                L, handles(self.operators), handles(self.restrictors),
                handles(self.interpolators),
                (_lib.ctypes.c_int * (L + 1))(*self.presmooths),
                (_lib.ctypes.c_int * (L + 1))(*self.postsmooths),
                (_lib.ctypes.c_double * (L + 1))(*self.weights),
                self.relax_method, float(self.gd.dv), _lib.ctypes.byref(h)),
                'FDPoissonSolver')
            self._native = h
        return self._native

    def _rhs(self, rho_dev):
        """The right-hand side the V-cycles solve for: rho, or B rho for the
        Mehrstellen pair (gpaw/poisson.py:481-484)."""
        if self.B is None:
            return rho_dev
        if self._rho_b is None or self._rho_b.shape != rho_dev.shape:
            self._rho_b = rho_dev.new_empty(rho_dev.shape)
        self.B.apply(rho_dev, self._rho_b)
        return self._rho_b

    def solve_neutral(self, phi, rho, timer=None):
        """V-cycles until dv sum r^2 <= eps; the mean of phi is removed on
        fully periodic grids (gpaw/poisson.py:474-498)."""
        h = self._handle()
        st = Staging(phi, rho)
        niter = _lib.ctypes.c_int(0)
        err = _lib.ctypes.c_double(0.0)
        status = _lib.load().gfd_poisson_solve(
            h, ptr(st.dev[0]), ptr(self._rhs(st.dev[1])), float(self.eps),
            int(self.maxiter), int(bool(self.gd.pbc_c.all())),
            _lib.ctypes.byref(niter), _lib.ctypes.byref(err),
            stream_handle())
        st.download(0)
        self.last_error = err.value
        if status == _lib.GFD_ECONVERGE:
            raise PoissonConvergenceError(
                f'Poisson solver did not converge in {self.maxiter} '
                'iterations!')
        _lib.check(status, 'FDPoissonSolver.solve_neutral')
        return niter.value

    def solve_mixed(self, phi, rho, tol=1e-10, levels=None, pre=2, post=2,
                    coarse_sweeps=8, w0=-1.0, wc=-1.0, inner=1,
                    max_outer=100, fp64_inner=False):
        """FP32 V-cycles inside FP64 iterative refinement until
        |L phi - rho| / |rho| <= tol (phi is the initial guess and the
        result).  Returns (refinement steps, final relative residual).
        fp64_inner: the same cycles in FP64 (isolates what FP32 gains).

        levels: depth of the FP32 hierarchy (None: all levels of the FP64
        one).  The coarsest FP32 level gets coarse_sweeps Jacobi sweeps
        (Chebyshev steps when it is part of the LDS-resident coarse solve),
        so a shallow hierarchy leaves smooth error behind and needs more
        refinement steps.  Not for the Mehrstellen pair (its right-hand
        side is B rho)."""
        if self.B is not None:
            raise NotImplementedError('solve_mixed: standard stencils only')
        h = self._handle()
        if levels is None:
            levels = self.levels + 1
        st = Staging(phi, rho)
        n = _lib.ctypes.c_int(0)
        rel = _lib.ctypes.c_double(0.0)
        status = _lib.load().gfd_poisson_solve_mixed(
            h, ptr(st.dev[0]), ptr(st.dev[1]), float(tol), int(levels),
            int(pre), int(post), int(coarse_sweeps), float(w0), float(wc),
            int(inner), int(max_outer), int(bool(fp64_inner)),
            _lib.ctypes.byref(n),
            _lib.ctypes.byref(rel), stream_handle())
        st.download(0)
        if status == _lib.GFD_ECONVERGE:
            raise PoissonConvergenceError(
                f'mixed-precision solve: relative residual {rel.value:.3g} '
                f'after {n.value} steps')
        _lib.check(status, 'FDPoissonSolver.solve_mixed')
        return n.value, rel.value

    def iterate2(self, step, level=0, phi=None, rho=None):
        """One V-cycle from level 0; returns dv sum r^2
        (gpaw/poisson.py:500-537).  rho is the right-hand side the cycles
        solve for (already B rho for the Mehrstellen pair)."""
        if level != 0:
            raise NotImplementedError('iterate2 is driven from level 0')
        h = self._handle()
        st = Staging(phi, rho)
        err = _lib.ctypes.c_double(0.0)
        _lib.check(_lib.load().gfd_poisson_iterate(
            h, ptr(st.dev[0]), ptr(st.dev[1]), _lib.ctypes.byref(err),
            stream_handle()), 'FDPoissonSolver.iterate2')
        st.download(0)
        return err.value

    def estimate_memory(self, mem):
        """rho, phi and residual on every level but phi on the finest (the
        caller's array)."""
        total = 0
        for l, op in enumerate(self.operators[:-1]):
            arrays = 2 if l == 0 else 3
            total += arrays * op.gd.bytecount()
        mem.subnode(f'rho, phi, residual [{self.levels} levels]', total)

    def __repr__(self):
        return (f'FDPoissonSolver(relax={self.relax!r}, nn={self.nn!r}, '
                f'eps={self.eps:e})')


def PoissonSolver(name='fd', **kwargs):
    """Factory of the solvers this library provides (the FD multigrid and
    the FFT solver)."""
    if name == 'fast':
        from .fft_poisson import FastPoissonSolver
        return FastPoissonSolver(**kwargs)
    if name != 'fd':
        raise NotImplementedError(f'Poisson solver {name!r}: only the FD '
                                  'multigrid and FFT solvers are part of '
                                  'this library')
    return FDPoissonSolver(**kwargs)
