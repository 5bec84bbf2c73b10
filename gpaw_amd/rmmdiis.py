"""RMM-DIIS block step of the FD eigensolver on the GPU, grid part.

``RMMDIIS.block_step`` (gpaw/new/pwfd/rmmdiis.py:123-257) with its default
``diis_steps=1`` for an FD Hamiltonian without PAW projections:

1. ``PR = P R``                          the kinetic preconditioner
2. ``dR = (H - eps_n) PR``               one fused stencil pass
3. ``lambda_n = -Re<dR|R> / <dR|dR>``    one pass over dR and R
4. ``psi += lambda_n PR``, ``dR <- R + lambda_n dR``   one pass
5. ``PR = P dR``, ``psi += lambda_n PR`` (or ``trial_step``)

Each line is one or two kernel launches on the whole band block; the
projector terms of ``calculate_residuals`` are outside the grid path.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from . import _lib
from .arrays import ptr, stream_handle, torch
from .bands import apply_hamiltonian


def _check(psi, *others):
    if not psi.is_contiguous() or psi.dtype not in (torch.float64,
                                                    torch.complex128):
        raise TypeError('band arrays: contiguous float64 or complex128')
    for o in others:
        if o.shape != psi.shape or o.dtype != psi.dtype or \
                not o.is_contiguous():
            raise ValueError('band arrays: equal shapes and dtypes')


def rmm_lambda(dR, R, out=None):
    """lambda_n = -Re<dR_n|R_n> / <dR_n|dR_n> (rmmdiis.py:150-155)."""
    _check(dR, R)
    nb = dR.shape[0]
    npts = int(np.prod(dR.shape[1:]))
    if out is None:
        out = torch.empty(nb, dtype=torch.float64, device=dR.device)
    _lib.check(_lib.load().gfd_rmm_lambda(
        int(dR.is_complex()), ptr(dR), ptr(R), nb, npts, npts, ptr(out),
        stream_handle()), 'rmm_lambda')
    return out


def band_axpy(alpha, x, y, scale=1.0):
    """y_n += scale alpha_n x_n (alpha: float64 device tensor)."""
    _check(y, x)
    nb = y.shape[0]
    npts = int(np.prod(y.shape[1:]))
    _lib.check(_lib.load().gfd_band_update(
        int(y.is_complex()), 0, ptr(alpha), float(scale), ptr(x), ptr(y),
        None, None, nb, npts, npts, stream_handle()), 'band_axpy')
    return y


def rmm_update(lam, psi, PR, R, dR):
    """psi += lambda PR and dR <- R + lambda dR in one pass
    (rmmdiis.py:157-162)."""
    _check(psi, PR, R, dR)
    nb = psi.shape[0]
    npts = int(np.prod(psi.shape[1:]))
    _lib.check(_lib.load().gfd_band_update(
        int(psi.is_complex()), 1, ptr(lam), 1.0, ptr(PR), ptr(psi), ptr(R),
        ptr(dR), nb, npts, npts, stream_handle()), 'rmm_update')


class BlockStep:
    """One RMM-DIIS block step (diis_steps=1) over a band block.

    kin: the Laplace(-1/2) operator, vt: the real local potential (float64
    device tensor), preconditioner: gpaw_amd.preconditioner.Preconditioner
    for the same grid.  Work arrays for nbands bands are kept between
    calls.
    """

    def __init__(self, kin, vt, preconditioner, trial_step=None):
        self.kin = kin
        self.vt = vt
        self.pc = preconditioner
        self.trial_step = trial_step
        self._PR = self._dR = None
        self._lam = self._trial = None

    def _work(self, psi):
        if self._PR is None or self._PR.shape != psi.shape or \
                self._PR.dtype != psi.dtype:
            self._PR = torch.empty_like(psi)
            self._dR = torch.empty_like(psi)
            self._lam = torch.empty(psi.shape[0], dtype=torch.float64,
                                    device=psi.device)
            if self.trial_step is not None:
                self._trial = torch.full((psi.shape[0],),
                                         float(self.trial_step),
                                         dtype=torch.float64,
                                         device=psi.device)
        return self._PR, self._dR, self._lam

    def __call__(self, psi, R, eig, phase_cd=None):
        """Update psi (nbands, n0, n1, n2) in place from its residuals R
        = (H - eig) psi; eig is a float64 device tensor of band
        energies.  Returns the step lengths lambda_n."""
        _check(psi, R)
        PR, dR, lam = self._work(psi)
        kpt = SimpleNamespace(phase_cd=phase_cd)
        self.pc(R, kpt, out=PR)
        apply_hamiltonian(self.kin, self.vt, PR, dR, phase_cd, eps=eig)
        rmm_lambda(dR, R, out=lam)
        rmm_update(lam, psi, PR, R, dR)
        self.pc(dR, kpt, out=PR)
        band_axpy(lam if self._trial is None else self._trial, PR, psi)
        return lam
