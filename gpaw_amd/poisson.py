"""Multigrid FD Poisson solver on the GPU (drop-in for
gpaw.poisson.FDPoissonSolver, gpaw/poisson.py:114-554).

The hierarchy, smoothing schedule, convergence test and charge handling are
the reference's; the V-cycles run in the native driver (gfd_poisson_*),
one call per solve.
"""
from __future__ import annotations

import warnings
from math import pi

import numpy as np

from . import _lib
from .arrays import Staging, is_device, ptr, stream_handle
from .fd_operators import Laplace, LaplaceA, LaplaceB
from .transformers import Transformer


class ConvergenceError(Exception):
    pass


class PoissonConvergenceError(ConvergenceError):
    """gpaw/__init__.py:104."""


POISSON_GRID_WARNING = """Grid unsuitable for FDPoissonSolver!

Consider using FastPoissonSolver instead.

The FDPoissonSolver does not have sufficient multigrid levels for good
performance and will converge inefficiently if at all, or yield wrong
results."""


class _PoissonSolver:
    def set_grid_descriptor(self, gd):
        raise NotImplementedError()

    def solve(self):
        raise NotImplementedError()

    def todict(self):
        raise NotImplementedError(self.__class__.__name__)

    def get_description(self):
        return self.__class__.__name__

    def estimate_memory(self, mem):
        raise NotImplementedError()


class BasePoissonSolver(_PoissonSolver):
    """Charge / background handling of gpaw/poisson.py:114-312 for the
    neutral and the charged fully periodic cases."""

    def __init__(self, *, remove_moment=None, use_charge_center=False,
                 metallic_electrodes=False, eps=None,
                 use_charged_periodic_corrections=False, xp=np):
        self.xp = xp
        self.gd = None
        self.remove_moment = remove_moment
        self.use_charge_center = use_charge_center
        self.use_charged_periodic_corrections = \
            use_charged_periodic_corrections
        self.charged_periodic_correction = None
        self.eps = eps
        self.metallic_electrodes = metallic_electrodes
        assert self.metallic_electrodes in [False, None, 'single', 'both']

    def todict(self):
        d = {'name': 'basepoisson'}
        if self.remove_moment:
            d['remove_moment'] = self.remove_moment
        if self.use_charge_center:
            d['use_charge_center'] = self.use_charge_center
        if self.use_charged_periodic_corrections:
            d['use_charged_periodic_corrections'] = \
                self.use_charged_periodic_corrections
        if self.metallic_electrodes:
            d['metallic_electrodes'] = self.metallic_electrodes
        return d

    def get_description(self):
        lines = []
        if self.remove_moment is not None:
            lines.append('    Remove moments up to L=%d' % self.remove_moment)
        if self.use_charge_center:
            lines.append('    Compensate for charged system using center of '
                         'majority charge')
        if self.use_charged_periodic_corrections:
            lines.append('    Subtract potential of homogeneous background')
        return '\n'.join(lines)

    def solve(self, phi, rho, charge=None, maxcharge=1e-6,
              zero_initial_phi=False, timer=None):
        self._init()
        assert tuple(phi.shape) == tuple(self.gd.n_c)
        assert tuple(rho.shape) == tuple(self.gd.n_c)
        actual_charge = self.gd.integrate(rho)
        background = (actual_charge / self.gd.dv /
                      self.gd.get_size_of_global_array().prod())
        if self.remove_moment:
            raise NotImplementedError('remove_moment is deprecated in the '
                                      'reference; use a moment-correction '
                                      'wrapper')
        if charge is None:
            charge = actual_charge
        if abs(charge) <= maxcharge:
            return self.solve_neutral(phi, rho - background, timer=timer)
        elif abs(charge) > maxcharge and self.gd.pbc_c.all():
            if zero_initial_phi:
                phi[:] = 0.0
            iters = self.solve_neutral(phi, rho - background, timer=timer)
            if self.use_charged_periodic_corrections:
                raise NotImplementedError('Madelung correction (host-side '
                                          'Ewald sum) is not part of the FD '
                                          'grid path')
            return iters
        raise NotImplementedError('charged systems with zero boundaries need '
                                  'the Gaussian monopole removal of '
                                  'gpaw/poisson.py:233-283')


class FDPoissonSolver(BasePoissonSolver):
    def __init__(self, nn=3, relax='J', eps=2e-10, maxiter=1000,
                 remove_moment=None, use_charge_center=False,
                 metallic_electrodes=False,
                 use_charged_periodic_corrections=False, **kwargs):
        super().__init__(
            remove_moment=remove_moment,
            use_charge_center=use_charge_center,
            metallic_electrodes=metallic_electrodes,
            use_charged_periodic_corrections=use_charged_periodic_corrections,
            **kwargs)
        self.eps = eps
        self.relax = relax
        self.nn = nn
        self.maxiter = maxiter
        if relax == 'GS':
            self.relax_method = 1
        elif relax == 'J':
            self.relax_method = 2
        else:
            raise NotImplementedError('Relaxation method %s' % relax)
        self.description = None
        self._initialized = False
        self._native = None

    def todict(self):
        d = super().todict()
        d.update({'name': 'fd', 'nn': self.nn, 'relax': self.relax,
                  'eps': self.eps})
        return d

    def get_stencil(self):
        return self.nn

    def create_laplace(self, gd, scale=1.0, n=1, dtype=float):
        return Laplace(gd, scale, n, dtype, xp=self.xp)

    def set_grid_descriptor(self, gd):
        """Multigrid hierarchy of gpaw/poisson.py:366-430."""
        self.gd = gd
        scale = -0.25 / pi
        if self.nn == 'M':
            if not gd.orthogonal:
                raise RuntimeError('Cannot use Mehrstellen stencil with '
                                   'non orthogonal cell.')
            self.operators = [LaplaceA(gd, -scale, xp=self.xp)]
            self.B = LaplaceB(gd)
        else:
            self.operators = [self.create_laplace(gd, scale, self.nn)]
            self.B = None
        self.interpolators = []
        self.restrictors = []
        level = 0
        self.presmooths = [2]
        self.postsmooths = [1]
        self.weights = [2.0 / 3.0]
        while level < 8:
            try:
                gd2 = gd.coarsen()
            except ValueError:
                break
            self.operators.append(self.create_laplace(gd2, scale, 1))
            self.interpolators.append(Transformer(gd2, gd, xp=self.xp))
            self.restrictors.append(Transformer(gd, gd2, xp=self.xp))
            self.presmooths.append(4)
            self.postsmooths.append(4)
            self.weights.append(1.0)
            level += 1
            gd = gd2
        self.levels = level
        if self.operators[-1].gd.N_c.max() > 36:
            warnings.warn('\n'.join([POISSON_GRID_WARNING, '',
                                     self.get_description()]), stacklevel=2)
        self._initialized = False
        self._free_native()

    def get_description(self):
        name = {1: 'Gauss-Seidel', 2: 'Jacobi'}[self.relax_method]
        coarsest_grid = self.operators[-1].gd.N_c
        coarsest_grid_string = ' x '.join([str(N) for N in coarsest_grid])
        assert self.levels + 1 == len(self.operators)
        lines = ['%s solver with %d multi-grid levels'
                 % (name, self.levels + 1),
                 '    Coarsest grid: %s points' % coarsest_grid_string]
        if coarsest_grid.max() > 24:
            lines.extend(['    Warning: Coarse grid has more than 24 points.',
                          '             More multi-grid levels recommended.'])
        lines.extend(['    Stencil: %s' % self.operators[0].description,
                      '    Max iterations: %d' % self.maxiter])
        lines.extend(['    Tolerance: %e' % self.eps])
        lines.append(super().get_description())
        return '\n'.join(lines)

    def _init(self):
        if self._initialized:
            return
        self.step = 0.66666666 / self.operators[0].get_diagonal_element()
        self.presmooths[self.levels] = 8
        self.postsmooths[self.levels] = 8
        self._initialized = True

    def _free_native(self):
        if self._native is not None:
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
            if self.B is not None:
                raise NotImplementedError('Mehrstellen (nn="M") V-cycle')
            lib = _lib.load()
            L = self.levels
            V = _lib.ctypes.c_void_p
            ops = (V * (L + 1))(*[op._handle().value for op in
                                  self.operators])
            rst = (V * max(L, 1))(*[t._handle().value for t in
                                    self.restrictors] or [None])
            itp = (V * max(L, 1))(*[t._handle().value for t in
                                    self.interpolators] or [None])
            pre = (_lib.ctypes.c_int * (L + 1))(*self.presmooths)
            post = (_lib.ctypes.c_int * (L + 1))(*self.postsmooths)
            w = (_lib.ctypes.c_double * (L + 1))(*self.weights)
            h = _lib.ctypes.c_void_p()
            _lib.check(lib.gfd_poisson_new(L, ops, rst, itp, pre, post, w,
                                           self.relax_method,
                                           float(self.gd.dv),
                                           _lib.ctypes.byref(h)),
                       'FDPoissonSolver')
            self._native = h
        return self._native

    def solve_neutral(self, phi, rho, timer=None):
        """V-cycles until dv sum r^2 <= eps (gpaw/poisson.py:474-498)."""
        h = self._handle()
        st = Staging(phi, rho)
        niter = _lib.ctypes.c_int(0)
        err = _lib.ctypes.c_double(0.0)
        status = _lib.load().gfd_poisson_solve(
            h, ptr(st.dev[0]), ptr(st.dev[1]), float(self.eps),
            int(self.maxiter), int(bool(self.gd.pbc_c.all())),
            _lib.ctypes.byref(niter), _lib.ctypes.byref(err),
            stream_handle())
        st.download(0)
        self.last_error = err.value
        if status == _lib.GFD_ECONVERGE:
            raise PoissonConvergenceError(
                'Poisson solver did not converge in %d iterations!'
                % self.maxiter)
        _lib.check(status, 'FDPoissonSolver.solve_neutral')
        return niter.value

    def solve_mixed(self, phi, rho, tol=1e-10, levels=None, pre=2, post=2,
                    coarse_sweeps=8, w0=-1.0, wc=-1.0, inner=1,
                    max_outer=100):
        """FP32 V-cycles inside FP64 iterative refinement until
        |L phi - rho| / |rho| <= tol (phi is the initial guess and the
        result).  Returns (refinement steps, final relative residual).

        levels: depth of the FP32 hierarchy (None: all levels of the FP64
        one).  The coarsest FP32 level gets coarse_sweeps Jacobi sweeps, so
        a shallow hierarchy leaves smooth error behind and needs more
        refinement steps."""
        h = self._handle()
        if levels is None:
            levels = self.levels + 1
        st = Staging(phi, rho)
        n = _lib.ctypes.c_int(0)
        rel = _lib.ctypes.c_double(0.0)
        status = _lib.load().gfd_poisson_solve_mixed(
            h, ptr(st.dev[0]), ptr(st.dev[1]), float(tol), int(levels),
            int(pre), int(post), int(coarse_sweeps), float(w0), float(wc),
            int(inner), int(max_outer), _lib.ctypes.byref(n),
            _lib.ctypes.byref(rel), stream_handle())
        st.download(0)
        if status == _lib.GFD_ECONVERGE:
            raise PoissonConvergenceError(
                'mixed-precision solve: relative residual %.3g after %d '
                'steps' % (rel.value, n.value))
        _lib.check(status, 'FDPoissonSolver.solve_mixed')
        return n.value, rel.value

    def iterate2(self, step, level=0, phi=None, rho=None):
        """One V-cycle at level 0; returns dv sum r^2
        (gpaw/poisson.py:500-537).  phi/rho default to the arrays of the
        last solve."""
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
        gdbytes = self.gd.bytecount()
        nbytes = -gdbytes
        for level in range(self.levels):
            nbytes += 3 * gdbytes
            gdbytes //= 8
        mem.subnode('rho, phi, residual [%d levels]' % self.levels, nbytes)

    def __repr__(self):
        template = 'FDPoissonSolver(relax=\'%s\', nn=%s, eps=%e)'
        return template % (self.relax, repr(self.nn), self.eps)


def PoissonSolver(name='fd', **kwargs):
    if name == 'fast':
        from .fft_poisson import FastPoissonSolver
        return FastPoissonSolver(**kwargs)
    if name != 'fd':
        raise NotImplementedError(f'Poisson solver {name!r}: only the FD '
                                  'multigrid and FFT solvers are part of '
                                  'this library')
    return FDPoissonSolver(**kwargs)
