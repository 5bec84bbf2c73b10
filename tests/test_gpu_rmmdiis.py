"""RMM-DIIS block step on the GPU (gpaw/new/pwfd/rmmdiis.py:123-257,
diis_steps=1, grid part) against the oracle's restatement, which is pinned
by the step's algebra in test_oracle_rmm.py."""
import numpy as np
import pytest

import oracle
from gpaw_amd.fd_operators import Laplace
from gpaw_amd.grid_descriptor import GridDescriptor
from gpaw_amd.preconditioner import Preconditioner
from gpaw_amd.rmmdiis import BlockStep, band_axpy, rmm_lambda
from gpu_util import dev, host, say
from helpers import bloch_phases, rand, relerr, spec, tspec

pytestmark = pytest.mark.gpu


def problem(dtype, seed, N, pbc, nb):
    rng = np.random.default_rng(seed)
    gd = GridDescriptor(N, (3.0, 2.5, 4.0), pbc)
    kin = Laplace(gd, -0.5, 3, dtype)
    ph = bloch_phases(rng, gd.pbc_c) if dtype == complex else None
    ks = spec(kin)
    vt = rng.uniform(-1, 1, tuple(gd.n_c))
    psi = rand(rng, (nb,) + tuple(gd.n_c), dtype)
    hpsi = oracle.fd_apply(ks.coef_p, ks.offset_pc, ks.n_c, ks.neighbor_cd,
                           psi, ph) + psi * vt
    eig = np.array([np.vdot(psi[n], hpsi[n]).real /
                    np.vdot(psi[n], psi[n]).real for n in range(nb)])
    R = hpsi - eig[:, None, None, None] * psi
    return gd, kin, ks, vt, ph, psi, R, eig


@pytest.mark.parametrize('dtype,pbc,trial', [
    (complex, True, None), (float, True, None), (float, False, None),
    (complex, (True, True, False), 0.1)])
def test_block_step(dtype, pbc, trial):
    import torch
    gd, kin, ks, vt, ph, psi, R, eig = problem(dtype, 7, (16, 12, 20), pbc,
                                               5)
    P = Preconditioner(gd, kin, dtype, block=2)

    def precond(x):
        return oracle.fd_precond(spec(P.kin0), spec(P.kin1), spec(P.kin2),
                                 tspec(P.restrictor_object0),
                                 tspec(P.restrictor_object1),
                                 tspec(P.interpolator_object1),
                                 tspec(P.interpolator_object2), x, P.step,
                                 ph)

    ref, it = oracle.rmm_block_step(ks, vt, precond, psi, R, eig, ph,
                                    trial_step=trial)
    say('rmm block step', dtype.__name__, pbc, trial)
    step = BlockStep(kin, dev(vt), P, trial_step=trial)
    d = dev(psi)
    lam = step(d, dev(R), torch.tensor(eig, device=d.device), ph)
    assert relerr(host(lam), it['lam']) < 1e-12
    assert relerr(host(d), ref) < 1e-12
    # second call on the same work arrays
    d2 = dev(psi)
    step(d2, dev(R), torch.tensor(eig, device=d.device), ph)
    assert relerr(host(d2), ref) < 1e-12


def test_lambda_and_axpy_kernels():
    import torch
    rng = np.random.default_rng(8)
    nb, n = 7, (9, 10, 11)
    for dtype in (float, complex):
        a = rand(rng, (nb,) + n, dtype)
        b = rand(rng, (nb,) + n, dtype)
        a[3] = 0          # zero dR: lambda 0, no division
        lam_ref = np.array([-np.vdot(a[i], b[i]).real /
                            np.vdot(a[i], a[i]).real if i != 3 else 0.0
                            for i in range(nb)])
        say('rmm lambda', dtype.__name__)
        lam = rmm_lambda(dev(a), dev(b))
        assert relerr(host(lam), lam_ref) < 1e-13
        y = dev(b)
        band_axpy(lam, dev(a), y, scale=0.5)
        assert relerr(host(y), b + 0.5 * lam_ref[:, None, None, None] * a) \
            < 1e-15
    with pytest.raises(ValueError):
        band_axpy(lam, dev(a), dev(b[:2]))
    del torch


def test_steps_reduce_residual():
    """Repeated steps with re-evaluated Rayleigh quotients: within each step
    the line search lowers every band's residual (|R + lambda dR| < |R|),
    and the residual norms fall over the steps, at 32^3 with 16 bands."""
    import torch
    from gpaw_amd.bands import apply_hamiltonian, band_dots
    rng = np.random.default_rng(9)
    gd = GridDescriptor((32, 32, 32), (6.0, 6.0, 6.0), True)
    kin = Laplace(gd, -0.5, 3, complex)
    ph = bloch_phases(rng, gd.pbc_c)
    vt = dev(rng.uniform(-1, 1, tuple(gd.n_c)))
    psi = dev(rand(rng, (16,) + tuple(gd.n_c), complex))
    P = Preconditioner(gd, kin, complex, block=16)
    step = BlockStep(kin, vt, P)
    R = torch.empty_like(psi)
    norms = []
    say('rmm convergence')
    for _ in range(6):
        apply_hamiltonian(kin, vt, psi, R, ph)
        eig = band_dots(psi, R, gd.dv) / band_dots(psi, psi, gd.dv)
        apply_hamiltonian(kin, vt, psi, R, ph, eps=eig)
        n2 = band_dots(psi, psi, gd.dv)
        r0 = host(band_dots(R, R, gd.dv) / n2)
        norms.append(r0)
        step(psi, R, eig, ph)
        r1 = host(band_dots(step._dR, step._dR, gd.dv) / n2)
        assert (r1 < r0).all()
    norms = np.array(norms)
    assert (norms[-1] < 0.95 * norms[0]).all()
