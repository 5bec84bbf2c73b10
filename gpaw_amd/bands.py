"""FD eigensolver step on the GPU: H apply, band overlap / subspace
matrices (FP64 MFMA) and Cholesky orthonormalisation.

- ``apply_hamiltonian``: FDHamiltonian.apply_local_potential
  (gpaw/new/fd/hamiltonian.py:23-32): H psi = kin(psi) + vt psi
- ``overlap``: matrix_elements(psit, psit, cc=True)
  (gpaw/core/arrays.py:183-262): S[i, j] = dv sum psi_i conj(psi_j)
- ``subspace_matrix``: matrix_elements(psit, Htpsit):
  H[i, j] = dv sum conj(psi_i) (H psi)_j
- ``apply_hamiltonian(..., eps=)``: the residuals (H - eps_n) psi_n of
  calculate_residuals (gpaw/new/pwfd/eigensolver.py:197-227), grid part
- ``band_dots``: dv Re <a_n|b_n> per band (band energies)
- ``orthonormalize``: gpaw/new/pwfd/wave_functions.py:191-246 for the grid
  part (no PAW projections): S -> L = chol(S), C = L^-1, psi <- C psi
  (Matrix.invcholesky, gpaw/core/matrix.py:382-417)

Arrays are torch complex128 (or float64 for a real H apply) device tensors
of shape (nbands, n0, n1, n2).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .arrays import ptr, stream_handle, torch


def _npts(psi):
    return int(np.prod(psi.shape[1:]))


def apply_hamiltonian(kin, vt, psi, out, phase_cd=None, eps=None):
    """out = kin(psi) + vt psi (vt: real grid, float64 device tensor); with
    eps (float64 device tensor, one value per band) the residuals
    out_n = (kin + vt - eps_n) psi_n in the same pass
    (gpaw/new/pwfd/eigensolver.py:197-227)."""
    if tuple(psi.shape[-3:]) != kin.shape or tuple(vt.shape) != kin.shape:
        raise ValueError('grid shape mismatch')
    if psi.shape != out.shape or not psi.is_contiguous() or \
            not out.is_contiguous():
        raise ValueError('psi / out: equal, contiguous shapes')
    if vt.dtype != torch.float64 or not vt.is_contiguous():
        raise TypeError('vt must be a contiguous float64 tensor')
    keep, ph = (None, None)
    if kin.dtype == complex:
        keep, ph = _lib.phases_arg(np.ones((3, 2), complex)
                                   if phase_cd is None else phase_cd)
    nb = int(np.prod(psi.shape[:-3]))
    if eps is not None:
        if eps.dtype != torch.float64 or eps.numel() != nb or \
                not eps.is_contiguous():
            raise ValueError('eps: contiguous float64, one per band')
        _lib.check(_lib.load().gfd_h_residual(
            kin._handle(), ptr(vt), ptr(eps), ptr(psi), ptr(out), nb, ph,
            stream_handle()), 'apply_hamiltonian(eps)')
        return out
    _lib.check(_lib.load().gfd_h_apply(kin._handle(), ptr(vt), ptr(psi),
                                       ptr(out), nb, ph, stream_handle()),
               'apply_hamiltonian')
    return out


def band_dots(a, b, dv, out=None):
    """out[n] = dv Re <a_n|b_n> for (nbands, ...) float64 / complex128
    device tensors."""
    nb = a.shape[0]
    npts = _npts(a)
    if b.shape != a.shape or b.dtype != a.dtype:
        raise ValueError('a and b: same shape and dtype')
    if out is None:
        out = torch.empty(nb, dtype=torch.float64, device=a.device)
    _lib.check(_lib.load().gfd_band_dots(
        int(a.dtype == torch.complex128), ptr(a), ptr(b), nb, npts, npts,
        float(dv), ptr(out), stream_handle()), 'band_dots')
    return out


def overlap(psi, dv, lower_only=True, out=None):
    """S[i, j] = dv sum_x psi_i(x) conj(psi_j(x)) (lower tiles only when
    lower_only: the part invcholesky reads)."""
    nb = psi.shape[0]
    if out is None:
        out = torch.zeros((nb, nb), dtype=torch.complex128,
                          device=psi.device)
    _lib.check(_lib.load().gfd_band_overlap(
        ptr(psi), nb, _npts(psi), _npts(psi), float(dv), ptr(out),
        int(lower_only), stream_handle()), 'overlap')
    return out


def subspace_matrix(psi, hpsi, dv, out=None):
    """H[i, j] = dv sum_x conj(psi_i(x)) hpsi_j(x)."""
    nb = psi.shape[0]
    if out is None:
        out = torch.empty((nb, nb), dtype=torch.complex128,
                          device=psi.device)
    _lib.check(_lib.load().gfd_band_matrix(
        ptr(psi), ptr(hpsi), nb, _npts(psi), _npts(psi), float(dv),
        ptr(out), stream_handle()), 'subspace_matrix')
    return out


def orthonormalize(psi, dv, out=None, S=None):
    """psi <- C psi with C = chol(S)^-1, S = overlap(psi) (lower);
    returns (out, C)."""
    nb = psi.shape[0]
    S = overlap(psi, dv, True, S)
    if out is None:
        out = torch.empty_like(psi)
    _lib.check(_lib.load().gfd_band_cholesky_orth(
        ptr(S), nb, ptr(psi), _npts(psi), _npts(psi), ptr(out),
        stream_handle()), 'orthonormalize')
    return out, S
