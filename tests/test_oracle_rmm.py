"""The oracle's RMM-DIIS block step (gpaw/new/pwfd/rmmdiis.py:123-257,
diis_steps=1) pinned by its algebra: the step length minimises the new
residual norm (checked by brute force over the step length and by the
orthogonality it implies), R + lambda dR is the residual of
psi + lambda PR, and an exact eigenpair is a fixed point."""
import numpy as np
import pytest

import oracle
from gpaw_amd.fd_operators import Laplace
from gpaw_amd.grid_descriptor import GridDescriptor
from gpaw_amd.preconditioner import Preconditioner
from helpers import bloch_phases, rand, relerr, spec, tspec


def setup(dtype, seed, N=(8, 12, 16), pbc=True, nb=3):
    rng = np.random.default_rng(seed)
    gd = GridDescriptor(N, (2.0, 3.0, 2.5), pbc)
    kin = Laplace(gd, -0.5, 3, dtype)
    P = Preconditioner(gd, kin, dtype, block=1)
    ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None

    def precond(x):
        return oracle.fd_precond(spec(P.kin0), spec(P.kin1), spec(P.kin2),
                                 tspec(P.restrictor_object0),
                                 tspec(P.restrictor_object1),
                                 tspec(P.interpolator_object1),
                                 tspec(P.interpolator_object2), x, P.step,
                                 ph)

    ks = spec(kin)
    vt = rng.uniform(-1, 1, tuple(gd.n_c))

    def resid(x, eig):
        hx = oracle.fd_apply(ks.coef_p, ks.offset_pc, ks.n_c,
                             ks.neighbor_cd, x, ph) + x * vt
        return hx - eig.reshape((-1, 1, 1, 1)) * x

    psi = rand(rng, (nb,) + tuple(gd.n_c), dtype)
    # Rayleigh quotients as band energies
    hpsi = resid(psi, np.zeros(nb))
    eig = np.array([np.vdot(psi[n], hpsi[n]).real /
                    np.vdot(psi[n], psi[n]).real for n in range(nb)])
    R = resid(psi, eig)
    return ks, vt, precond, resid, psi, R, eig, ph


@pytest.mark.parametrize('dtype', [float, complex])
def test_step_length_is_line_minimum(dtype):
    ks, vt, precond, resid, psi, R, eig, ph = setup(dtype, 1)
    new, it = oracle.rmm_block_step(ks, vt, precond, psi, R, eig, ph)
    lam = it['lam']
    assert (lam != 0).all()
    for n in range(psi.shape[0]):
        # brute force: |resid(psi + l PR)|^2 around l = lambda, vertex of
        # the parabola through three samples
        def f(l):
            x = psi[n:n + 1] + l * it['PR'][n:n + 1]
            return float(np.vdot(r := resid(x, eig[n:n + 1]), r).real)
        h = 0.05 * abs(lam[n])
        fm, f0, fp = f(lam[n] - h), f(lam[n]), f(lam[n] + h)
        assert f0 < fm and f0 < fp
        vertex = lam[n] - h * (fp - fm) / (2 * (fp - 2 * f0 + fm))
        assert abs(vertex - lam[n]) < 1e-7 * abs(lam[n])
        # orthogonality at the minimum: Re<dR|R + lambda dR> = 0
        g = np.vdot(it['dR'][n], it['R1'][n]).real
        assert abs(g) < 1e-12 * np.vdot(R[n], R[n]).real ** 0.5 * \
            np.vdot(it['dR'][n], it['dR'][n]).real ** 0.5
    # linearity: R + lambda dR is the residual of psi + lambda PR
    assert relerr(it['R1'], resid(it['psi1'], eig)) < 1e-12
    # the step lowers the residual norm of every band
    for n in range(psi.shape[0]):
        assert np.linalg.norm(it['R1'][n]) < np.linalg.norm(R[n])


def test_trial_step():
    ks, vt, precond, resid, psi, R, eig, ph = setup(complex, 2)
    new, it = oracle.rmm_block_step(ks, vt, precond, psi, R, eig, ph,
                                    trial_step=0.1)
    ref, it2 = oracle.rmm_block_step(ks, vt, precond, psi, R, eig, ph)
    # the first update is the same; the last one uses the fixed step
    assert relerr(it['psi1'], it2['psi1']) == 0
    assert relerr(new - it['psi1'], 0.1 * precond(it['R1'])) < 1e-14


def test_eigenpair_is_fixed_point():
    """A plane wave is an eigenvector of the periodic FD Hamiltonian with a
    constant potential; its residual is rounding noise and the step leaves
    it unchanged."""
    N = (8, 12, 16)
    gd = GridDescriptor(N, (2.0, 3.0, 2.5), True)
    kin = Laplace(gd, -0.5, 3, complex)
    ks = spec(kin)
    P = Preconditioner(gd, kin, complex, block=1)

    def precond(x):
        return oracle.fd_precond(spec(P.kin0), spec(P.kin1), spec(P.kin2),
                                 tspec(P.restrictor_object0),
                                 tspec(P.restrictor_object1),
                                 tspec(P.interpolator_object1),
                                 tspec(P.interpolator_object2), x, P.step)

    m = np.array([1, 2, 3])
    idx = np.indices(N).reshape(3, -1).T
    psi = np.exp(2j * np.pi * (idx @ (m / np.array(N)))).reshape((1,) + N)
    theta = 2 * np.pi * m / np.array(N)
    e_kin = float(np.sum(ks.coef_p * np.exp(1j * ks.offset_pc @ theta)).real)
    v0 = 0.3
    vt = np.full(N, v0)
    eig = np.array([e_kin + v0])
    hx = oracle.fd_apply(ks.coef_p, ks.offset_pc, ks.n_c, ks.neighbor_cd,
                         psi) + psi * vt
    R = hx - eig[0] * psi
    assert np.abs(R).max() < 1e-12
    new, _ = oracle.rmm_block_step(ks, vt, precond, psi, R, eig)
    assert relerr(new, psi) < 1e-12
