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
- ``subspace_diagonalize``: subspace_eigenvalues + canonical_transformation
  (gpaw/new/pwfd/wave_functions.py:289-368), grid part: eigenvalues and
  eigenvectors of H[i, j] = <psi_i|H|psi_j>, psi_m <- sum_k v_m[k] psi_k

Arrays are torch complex128 (or float64 for a real H apply) device tensors
of shape (nbands, n0, n1, n2).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .arrays import ptr, stream_handle, torch


def _npts(psi):
    return int(np.prod(psi.shape[1:]))


def _bands(x, name, dtype=None):
    """Check a (nbands, ...) device tensor before its pointer reaches the
    library: the kernels read the dtype's element size over the whole
    shape, so a wrong dtype or a strided view would read out of bounds."""
    if not (torch is not None and isinstance(x, torch.Tensor)) or \
            x.device.type != 'cuda':
        raise TypeError(f'{name} must be a device (cuda) tensor')
    want = torch.complex128 if dtype is None else dtype
    if x.dtype != want:
        raise TypeError(f'{name} must be {want}, not {x.dtype}')
    if not x.is_contiguous():
        raise ValueError(f'{name} must be contiguous')
    if x.ndim < 2:
        raise ValueError(f'{name} must have a band axis and grid axes')


def apply_hamiltonian(kin, vt, psi, out, phase_cd=None, eps=None):
    """out = kin(psi) + vt psi (vt: real grid, float64 device tensor); with
    eps (float64 device tensor, one value per band) the residuals
    out_n = (kin + vt - eps_n) psi_n in the same pass
    (gpaw/new/pwfd/eigensolver.py:197-227)."""
    want = torch.complex128 if kin.dtype == complex else torch.float64
    _bands(psi, 'psi', want)
    _bands(out, 'out', want)
    if tuple(psi.shape[-3:]) != kin.shape or tuple(vt.shape) != kin.shape:
        raise ValueError('grid shape mismatch')
    if psi.shape != out.shape:
        raise ValueError('psi and out must have equal shapes')
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


def _square(m, nb):
    if m.dtype != torch.complex128 or not m.is_contiguous() or \
            tuple(m.shape) != (nb, nb) or m.device.type != 'cuda':
        raise ValueError(f'matrix: contiguous complex128 ({nb}, {nb}) '
                         'device tensor')


def _bands_vector(v, nb):
    if v.dtype != torch.float64 or not v.is_contiguous() or \
            v.numel() != nb or v.device.type != 'cuda':
        raise ValueError(f'out: contiguous float64 device tensor of {nb}')


def band_dots(a, b, dv, out=None):
    """out[n] = dv Re <a_n|b_n> for (nbands, ...) float64 / complex128
    device tensors."""
    if a.dtype not in (torch.float64, torch.complex128):
        raise TypeError('band_dots: float64 or complex128 bands')
    _bands(a, 'a', a.dtype)
    _bands(b, 'b', a.dtype)
    nb = a.shape[0]
    npts = _npts(a)
    if b.shape != a.shape:
        raise ValueError('a and b: same shape')
    if out is not None:
        _bands_vector(out, nb)
    if out is None:
        out = torch.empty(nb, dtype=torch.float64, device=a.device)
    _lib.check(_lib.load().gfd_band_dots(
        int(a.dtype == torch.complex128), ptr(a), ptr(b), nb, npts, npts,
        float(dv), ptr(out), stream_handle()), 'band_dots')
    return out


def overlap(psi, dv, lower_only=True, out=None):
    """S[i, j] = dv sum_x psi_i(x) conj(psi_j(x)) (lower tiles only when
    lower_only: the part invcholesky reads)."""
    _bands(psi, 'psi')
    nb = psi.shape[0]
    if out is not None:
        _square(out, nb)
    if out is None:
        out = torch.zeros((nb, nb), dtype=torch.complex128,
                          device=psi.device)
    _lib.check(_lib.load().gfd_band_overlap(
        ptr(psi), nb, _npts(psi), _npts(psi), float(dv), ptr(out),
        int(lower_only), stream_handle()), 'overlap')
    return out


def subspace_matrix(psi, hpsi, dv, out=None):
    """H[i, j] = dv sum_x conj(psi_i(x)) hpsi_j(x)."""
    _bands(psi, 'psi')
    _bands(hpsi, 'hpsi')
    if hpsi.shape != psi.shape:
        raise ValueError('psi and hpsi must have equal shapes')
    nb = psi.shape[0]
    if out is not None:
        _square(out, nb)
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
    _bands(psi, 'psi')
    nb = psi.shape[0]
    S = overlap(psi, dv, True, S)
    if out is None:
        out = torch.empty_like(psi)
    _bands(out, 'out')
    if out.shape != psi.shape or out.data_ptr() == psi.data_ptr():
        raise ValueError('out: the shape of psi, not psi itself')
    _lib.check(_lib.load().gfd_band_cholesky_orth(
        ptr(S), nb, ptr(psi), _npts(psi), _npts(psi), ptr(out),
        stream_handle()), 'orthonormalize')
    return out, S


def subspace_diagonalize(H, psi, out=None, eps=None):
    """Eigenvalues of the Hermitian subspace matrix H (complex128 device
    tensor (nbands, nbands), e.g. subspace_matrix(psi, H psi, dv); its lower
    part is read) and the canonical rotation out_m = sum_k v_m[k] psi_k.
    On return H[m, k] = v_m[k] (eigenvectors as rows, as the reference's
    H_nn after subspace_eigenvalues).  Returns (eps, out); eps ascending,
    float64 device tensor."""
    _bands(psi, 'psi')
    nb = psi.shape[0]
    _square(H, nb)
    if out is None:
        out = torch.empty_like(psi)
    _bands(out, 'out')
    if out.shape != psi.shape or out.data_ptr() == psi.data_ptr():
        raise ValueError('out: the shape of psi, not psi itself')
    if eps is None:
        eps = torch.empty(nb, dtype=torch.float64, device=psi.device)
    _bands_vector(eps, nb)
    _lib.check(_lib.load().gfd_band_subspace_diag(
        ptr(H), nb, ptr(psi), _npts(psi), _npts(psi), ptr(out), ptr(eps),
        stream_handle()), 'subspace_diagonalize')
    return eps, out
