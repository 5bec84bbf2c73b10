"""Band contractions (FP64 MFMA), Cholesky orthonormalisation and the H
apply, against numpy FP64 (matmul / cholesky / inv) and the oracle stencil.

These rows have no golden vectors in the reference's tests (the reference
calls BLAS / scipy): they are pinned by numpy's FP64 results and by the
invariants of the algebra (orthonormality after the rotation)."""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle
from gpaw_amd.bands import (apply_hamiltonian, band_dots, orthonormalize,
                            overlap, subspace_diagonalize, subspace_matrix)
from gpaw_amd.fd_operators import Laplace
from gpaw_amd.grid_descriptor import GridDescriptor
from gpu_util import dev, host, say
from helpers import bloch_phases, relerr

pytestmark = pytest.mark.gpu


def psi_of(rng, nb, n):
    return (rng.standard_normal((nb,) + n) +
            1j * rng.standard_normal((nb,) + n))


@pytest.mark.parametrize('nb,n', [(64, (8, 8, 8)), (200, (12, 10, 16)),
                                  (256, (16, 16, 16))])
def test_overlap_and_subspace(nb, n):
    rng = np.random.default_rng(nb)
    psi = psi_of(rng, nb, n)
    hpsi = psi_of(rng, nb, n)
    dv = 0.37
    A = psi.reshape(nb, -1)
    B = hpsi.reshape(nb, -1)
    say('overlap', nb, n)
    S = host(overlap(dev(psi), dv, lower_only=False))
    S_ref = dv * A @ A.conj().T
    assert relerr(S, S_ref) < 1e-13
    H = host(subspace_matrix(dev(psi), dev(hpsi), dv))
    H_ref = dv * A.conj() @ B.T
    assert relerr(H, H_ref) < 1e-13
    Sl = host(overlap(dev(psi), dv, lower_only=True))
    low = np.tril_indices(nb)
    assert relerr(Sl[low], S_ref[low]) < 1e-13


def test_orthonormalize():
    rng = np.random.default_rng(3)
    nb, n = 192, (12, 12, 16)
    dv = 0.05
    psi = psi_of(rng, nb, n)
    A = psi.reshape(nb, -1)
    S_ref = dv * A @ A.conj().T
    L = sla.cholesky(S_ref, lower=True)
    C = sla.inv(L)
    ref = (C @ A).reshape(psi.shape)
    say('orthonormalize', nb, n)
    out, Cdev = orthonormalize(dev(psi), dv)
    got = host(out)
    assert relerr(got, ref) < 1e-12
    G = got.reshape(nb, -1)
    assert np.abs(dv * G @ G.conj().T - np.eye(nb)).max() < 1e-12


def test_h_apply():
    rng = np.random.default_rng(4)
    gd = GridDescriptor((16, 12, 20), (3.0, 2.5, 4.0))
    for dtype in (float, complex):
        kin = Laplace(gd, -0.5, 3, dtype)
        psi = rng.standard_normal((3,) + tuple(gd.n_c))
        if dtype == complex:
            psi = psi + 1j * rng.standard_normal(psi.shape)
        vt = rng.standard_normal(tuple(gd.n_c))
        ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
        ref = oracle.fd_apply(kin.coef_p, kin.offset_pc, gd.n_c,
                              gd.neighbor_cd, psi, ph) + psi * vt
        d = dev(psi)
        out = d.new_empty(d.shape)
        say('h apply', dtype.__name__)
        apply_hamiltonian(kin, dev(vt), d, out, ph)
        assert relerr(host(out), ref) < 1e-13


@pytest.mark.parametrize('dtype', [float, complex])
def test_h_residual_and_band_dots(dtype):
    """The fused residual (H - eps_n) psi_n against the oracle stencil plus
    numpy, and the per-band energies against numpy's vdot."""
    import torch
    rng = np.random.default_rng(5)
    gd = GridDescriptor((12, 16, 10), (2.0, 3.0, 2.5),
                        pbc_c=(True, False, True))
    nb = 5
    kin = Laplace(gd, -0.5, 3, dtype)
    psi = rng.standard_normal((nb,) + tuple(gd.n_c))
    if dtype == complex:
        psi = psi + 1j * rng.standard_normal(psi.shape)
    vt = rng.standard_normal(tuple(gd.n_c))
    eps = rng.standard_normal(nb)
    ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
    hpsi = oracle.fd_apply(kin.coef_p, kin.offset_pc, gd.n_c,
                           gd.neighbor_cd, psi, ph) + psi * vt
    ref = hpsi - eps[:, None, None, None] * psi
    d = dev(psi)
    out = d.new_empty(d.shape)
    say('h residual', dtype.__name__)
    apply_hamiltonian(kin, dev(vt), d, out, ph,
                      eps=torch.tensor(eps, dtype=torch.float64,
                                       device=d.device))
    assert relerr(host(out), ref) < 1e-13
    # band energies: dv Re <psi_n|H psi_n>
    dv = gd.dv
    e_ref = np.array([dv * np.vdot(psi[n], hpsi[n]).real for n in range(nb)])
    e = host(band_dots(d, dev(hpsi), dv))
    assert np.abs(e - e_ref).max() < 1e-12 * np.abs(e_ref).max()
    # residual orthogonal to psi_n when eps_n is its Rayleigh quotient
    norms = np.array([dv * np.vdot(psi[n], psi[n]).real for n in range(nb)])
    rq = torch.tensor(e_ref / norms, dtype=torch.float64, device=d.device)
    apply_hamiltonian(kin, dev(vt), d, out, ph, eps=rq)
    ov = host(band_dots(d, out, dv))
    assert np.abs(ov).max() < 1e-11 * np.abs(e_ref).max()


@pytest.mark.parametrize('nb', [96, 200])
def test_subspace_diagonalize(nb):
    """Eigenvalues against numpy's eigh of the same (lower-half) matrix,
    eigenvectors by H v = eps v, the rotated bands against numpy's rotation
    with the returned vectors, and the subspace matrix of the rotated bands
    diagonal with the eigenvalues on it (the canonical form)."""
    import torch
    rng = np.random.default_rng(900 + nb)
    gd = GridDescriptor((12, 16, 10), (2.4, 3.2, 2.0))
    kin = Laplace(gd, -0.5, 3, complex)
    ph = bloch_phases(rng, gd.pbc_c)
    vt = dev(rng.uniform(-2, 0, tuple(gd.n_c)))
    psi0 = psi_of(rng, nb, tuple(gd.n_c))
    psi, _ = orthonormalize(dev(psi0), gd.dv)
    hpsi = torch.empty_like(psi)
    apply_hamiltonian(kin, vt, psi, hpsi, ph)
    H = subspace_matrix(psi, hpsi, gd.dv)
    Hh = host(H)
    low = np.tril(Hh)
    Hfull = low + np.tril(Hh, -1).conj().T  # the lower half, as read
    eps_ref = np.linalg.eigvalsh(Hfull)
    say('subspace diagonalize', nb)
    eps, out = subspace_diagonalize(H, psi)
    e = host(eps)
    assert np.abs(e - eps_ref).max() < 1e-12 * np.abs(eps_ref).max()
    assert (np.diff(e) >= 0).all()
    V = host(H)  # rows v_m
    r = Hfull @ V.T - V.T * e[None, :]
    assert np.abs(r).max() < 1e-11 * np.abs(eps_ref).max()
    assert np.abs(V.conj() @ V.T - np.eye(nb)).max() < 1e-12
    A = host(psi).reshape(nb, -1)
    ref = (V @ A).reshape(psi.shape)
    assert relerr(host(out), ref) < 1e-12
    # canonical form: <out_m|H|out_k> = eps_m delta_mk
    hout = torch.empty_like(out)
    apply_hamiltonian(kin, vt, out, hout, ph)
    H2 = host(subspace_matrix(out, hout, gd.dv))
    assert np.abs(H2 - np.diag(e)).max() < 1e-10 * np.abs(e).max()


def test_dense_rotation_large_k():
    """Bands beyond one 128-row tile on both sides of the rotation: 300
    bands (3 tiles, the last partial) over 7000 points (partial point
    tile), against numpy."""
    import torch
    rng = np.random.default_rng(77)
    nb = 300
    H0 = rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))
    H0 = H0 + H0.conj().T
    psi = psi_of(rng, nb, (7, 10, 100))
    eps, out = subspace_diagonalize(dev(H0.copy()), dev(psi))
    w, U = np.linalg.eigh(H0)  # columns: eigenvectors
    assert np.abs(host(eps) - w).max() < 1e-12 * np.abs(w).max()
    # v_m = U[:, m] up to a phase: compare phase-invariant products
    got = host(out).reshape(nb, -1)
    ref = U.T @ psi.reshape(nb, -1)
    ov = np.abs(np.einsum('mx,mx->m', got.conj(), ref))
    nrm = np.linalg.norm(got, axis=1) * np.linalg.norm(ref, axis=1)
    assert np.abs(ov / nrm - 1).max() < 1e-12
