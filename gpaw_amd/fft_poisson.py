"""FFT / DST Poisson solver on the GPU (drop-in for
gpaw.poisson.FastPoissonSolver, gpaw/poisson.py:846-1028).

The FD Laplacian (`Laplace(gd, -1/4pi, nn)`) is diagonal in the Fourier
basis of every periodic axis and in the DST-I basis of every zero-boundary
axis.  Its eigenvalues are computed here from the stencil, as the reference
does (gpaw/poisson.py:929-961); the solve is one native call: odd extension
of the zero-boundary axes, a real 3-D FFT (hipFFT), division by the
eigenvalues (the zero mode dropped), the inverse FFT and the interior
extraction.  It is exact for the FD operator: the result equals the
converged multigrid solution.
"""
from __future__ import annotations

from math import pi

import numpy as np

from . import _lib
from .arrays import Staging, ptr, stream_handle
from .fd_operators import Laplace
from .poisson import BasePoissonSolver


class BadAxesError(ValueError):
    pass


class FastPoissonSolver(BasePoissonSolver):
    def __init__(self, nn=3, **kwargs):
        super().__init__(**kwargs)
        self.nn = nn
        self.use_cholesky = False
        self._native = None

    def _init(self):
        pass

    def set_grid_descriptor(self, gd):
        self._free()
        self.gd = gd
        pbc_c = np.array(gd.pbc_c, dtype=bool)
        dotprods = gd.cell_cv @ gd.cell_cv.T
        orthogonal_c = (np.abs(dotprods) > 1e-10).sum(axis=0) == 1
        if not all(pbc_c | orthogonal_c):
            raise BadAxesError('Each axis must be periodic or orthogonal '
                               'to other axes.  But we have pbc={} '
                               'and orthogonal={}'
                               .format(pbc_c.astype(int),
                                       orthogonal_c.astype(int)))
        axes = np.arange(3)
        self.fft_axes = list(axes[pbc_c])
        self.fst_axes = sorted(axes[~pbc_c], key=lambda c: gd.N_c[c])
        self.cholesky_axes = []
        laplace = Laplace(gd, -0.25 / pi, self.nn)
        self.stencil_description = laplace.description
        n_c = np.array(gd.n_c, dtype=np.int64)
        zero = (~pbc_c).astype(np.int32)
        m_c = np.where(pbc_c, n_c, 2 * (n_c + 1))
        # eigenvalues on the extended grid's half spectrum (C order)
        k = [np.arange(m_c[0]) / m_c[0], np.arange(m_c[1]) / m_c[1],
             np.arange(m_c[2] // 2 + 1) / m_c[2]]
        lam = np.zeros((m_c[0], m_c[1], m_c[2] // 2 + 1))
        for coef, off in zip(laplace.coef_p, laplace.offset_pc):
            if not any(off):
                continue
            ph = (2 * pi * off[0] * k[0][:, None, None] +
                  2 * pi * off[1] * k[1][None, :, None] +
                  2 * pi * off[2] * k[2][None, None, :])
            lam += coef * (np.cos(ph) - 1.0)
        with np.errstate(divide='ignore'):
            self.inv_lambda = np.ascontiguousarray(
                np.where(np.abs(lam) > 1e-10, 1.0 / lam, 0.0))
        self._n_c = n_c
        self._zero = zero

    def _handle(self):
        if self._native is None:
            h = _lib.ctypes.c_void_p()
            n = np.ascontiguousarray(self._n_c, dtype=np.int64)
            z = np.ascontiguousarray(self._zero, dtype=np.int32)
            _lib.check(_lib.load().gfd_fft_poisson_new(
                n.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_long)),
                z.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_int)),
                self.inv_lambda.ctypes.data_as(
                    _lib.ctypes.POINTER(_lib.ctypes.c_double)),
                _lib.ctypes.byref(h)), 'FastPoissonSolver')
            self._native = h
        return self._native

    def _free(self):
        if getattr(self, '_native', None) is not None:
            _lib.load().gfd_fft_poisson_free(self._native)
            self._native = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass

    def solve_neutral(self, phi_g, rho_g, timer=None):
        """phi = L^-1 rho (zero mode dropped); returns 1 (non-iterative),
        gpaw/poisson.py:963-1005."""
        h = self._handle()
        st = Staging(phi_g, rho_g)
        _lib.check(_lib.load().gfd_fft_poisson_solve(
            h, ptr(st.dev[1]), ptr(st.dev[0]), stream_handle()),
            'FastPoissonSolver.solve_neutral')
        st.download(0)
        return 1

    def todict(self):
        d = super().todict()
        d.update({'name': 'fast', 'nn': self.nn})
        return d

    def estimate_memory(self, mem):
        pass

    def get_description(self):
        lines = [f'{self.__class__.__name__} using',
                 f'    Stencil: {self.stencil_description}',
                 f'    FFT axes: {self.fft_axes}',
                 f'    FST axes: {self.fst_axes}']
        lines.append(BasePoissonSolver.get_description(self))
        return '\n'.join(lines)
