"""CPU oracle for the GPAW FD grid path (TEST INFRASTRUCTURE ONLY).

This package is the checker the HIP path is compared against.  Only
``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import it; the product (``gpaw_amd``) never does.

It wraps ``oracle/_build/liboracle.so`` (plain C, ``bmgs_oracle.c``), a
restatement of the reference CPU kernels in their operation order, and adds
Python restatements of the reference glue that chains them:

- ``fd_precond``      gpaw/preconditioner.py:35-93 and
                      c/fd_preconditioner.cpp:221-345
- ``PoissonVCycle``   gpaw/poisson.py:474-537 (solve_neutral / iterate2)
- ``rmm_block_step``  gpaw/new/pwfd/rmmdiis.py:123-257 (diis_steps=1, grid
                      part: no projector terms)
- ``wfd_apply``, ``wfd_relax``  c/bmgs/wfd.cpp, c/bmgs/wrelax.cpp (weighted
                      operators of gpaw/wfd_operators.py), in the C restatement

Pinning (what checks each function; see tests/test_oracle_*.py):

- ``fd_apply``, ``fd_relax``, ``transform``: bit-exact against the reference
  C extension built from the reference sources (``oracle/_ref``) and against
  the committed fixtures in ``tests/golden``; plus closed forms (polynomial
  exactness, exact gradient vectors, Bloch plane-wave eigenvalue, integral
  conservation, brute-force dense operator on tiny grids).
- ``fd_precond``: bit-exact against the reference ``fd_precond`` (fixtures).
- ``PoissonVCycle``: against the same V-cycle driving the reference C
  operators (fixtures: niter and phi), and against the spectral closed form
  of the converged periodic solution.
- ``wfd_apply``, ``wfd_relax``: bit-exact against the reference WOperator
  built from the reference sources (``oracle/_ref``), and with one unit
  weight against ``fd_apply`` / ``fd_relax``.
- ``rmm_block_step``: by what the step's algebra fixes -- lambda minimises
  |R + lambda dR| (Re<dR|R + lambda dR> = 0), R + lambda dR is the residual
  of psi + lambda PR (linearity of H), an exact eigenpair is a fixed point,
  and on a one-band quadratic model the step equals the closed-form line
  minimum.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, '_build', 'liboracle.so')
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


def build():
    """Compile the C restatement (and the reference build when present)."""
    import subprocess
    subprocess.run(['make', '-s', '-C', _HERE, 'oracle'], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_fd_apply.argtypes = [
            ctypes.c_int, _dp, _lp, _lp, ctypes.c_int, _lp, ctypes.c_int,
            ctypes.c_int, _dp, _dp, _dp, ctypes.c_long, ctypes.c_int]
        L.oracle_fd_relax.argtypes = [
            ctypes.c_int, ctypes.c_int, _dp, _lp, _lp, ctypes.c_int, _lp,
            ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.oracle_transform.argtypes = [
            ctypes.c_int, ctypes.c_int, _lp, _lp, _lp, _lp, _lp, _lp,
            ctypes.c_int, _dp, _dp, _dp, ctypes.c_long, ctypes.c_int]
        _ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_wfd_apply.argtypes = [
            ctypes.c_int, _ip, _dp, _lp, _dp, _lp, ctypes.c_int, _lp,
            ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_long]
        L.oracle_wfd_relax.argtypes = [
            ctypes.c_int, ctypes.c_int, _ip, _dp, _lp, _dp, _lp,
            ctypes.c_int, _lp, ctypes.c_int, _dp, _dp, ctypes.c_int,
            ctypes.c_double]
        for f in (L.oracle_fd_apply, L.oracle_fd_relax, L.oracle_transform,
                  L.oracle_wfd_apply, L.oracle_wfd_relax):
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _l(a):
    return a.ctypes.data_as(_lp)


def _longs(x, shape=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int64))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _phases(phase_cd, real):
    if real:
        return None, None
    if phase_cd is None:
        phase_cd = np.ones((3, 2), complex)
    ph = np.ascontiguousarray(np.asarray(phase_cd, complex).reshape(3, 2))
    return ph, _d(ph.view(float))


def _threads(nthreads):
    if nthreads is None:
        nthreads = len(os.sched_getaffinity(0))
    return int(nthreads)


def is_cfd(offset_pc):
    """Axis-only stencil test of gpaw/fd_operators.py:48-52."""
    return all(sum(o != 0 for o in off) < 2 for off in np.asarray(offset_pc))


def fd_apply(coef_p, offset_pc, n_c, neighbor_cd, a, phase_cd=None,
             out=None, nthreads=None):
    """Reference FDOperator.apply (c/operators.cpp:245-316) on host arrays."""
    a = np.ascontiguousarray(a)
    real = a.dtype == np.float64
    assert real or a.dtype == np.complex128
    n_c = _longs(n_c)
    assert a.shape[-3:] == tuple(n_c)
    coef = np.ascontiguousarray(np.asarray(coef_p, float))
    off = _longs(offset_pc, (-1, 3))
    rng = int(np.abs(off).max())
    nb = _longs(neighbor_cd)
    if out is None:
        out = np.empty_like(a)
    nin = int(np.prod(a.shape[:-3], dtype=np.int64))
    ph, php = _phases(phase_cd, real)
    err = lib().oracle_fd_apply(len(coef), _d(coef), _l(off), _l(n_c), rng,
                                _l(nb), int(real), int(is_cfd(off)), php,
                                _d(a.view(float)), _d(out.view(float)), nin,
                                _threads(nthreads))
    if err:
        raise ValueError(f'oracle_fd_apply: error {err}')
    return out


def fd_relax(method, coef_p, offset_pc, n_c, neighbor_cd, f, src, nrelax,
             w=1.0):
    """Reference FDOperator.relax (c/operators.cpp:50-92); f updated."""
    assert f.dtype == np.float64 and f.flags.c_contiguous
    src = np.ascontiguousarray(src, float)
    coef = np.ascontiguousarray(np.asarray(coef_p, float))
    off = _longs(offset_pc, (-1, 3))
    rng = int(np.abs(off).max())
    n_c = _longs(n_c)
    nb = _longs(neighbor_cd)
    err = lib().oracle_fd_relax(int(method), len(coef), _d(coef), _l(off),
                                _l(n_c), rng, _l(nb), int(is_cfd(off)),
                                _d(f), _d(src), int(nrelax), float(w))
    if err:
        raise ValueError(f'oracle_fd_relax: error {err}')
    return f


def _wstencils(coef_ps, offset_ps):
    nc = np.ascontiguousarray([len(c) for c in coef_ps], dtype=np.int32)
    coef = np.ascontiguousarray(np.concatenate(
        [np.asarray(c, float) for c in coef_ps]))
    off = _longs(np.concatenate([np.asarray(o).reshape(-1, 3)
                                 for o in offset_ps]), (-1, 3))
    return nc, coef, off


def wfd_apply(coef_ps, offset_ps, weights, n_c, neighbor_cd, a,
              phase_cd=None):
    """Reference WeightedFDOperator.apply (c/woperators.cpp, c/bmgs/wfd.cpp):
    out = sum_w weights[w] * (stencil w applied to a); every stencil's
    first point is the centre (gpaw/wfd_operators.py:24-44)."""
    a = np.ascontiguousarray(a)
    real = a.dtype == np.float64
    n_c = _longs(n_c)
    nc, coef, off = _wstencils(coef_ps, offset_ps)
    wts = np.ascontiguousarray(np.asarray(weights, float))
    assert wts.shape == (len(coef_ps),) + tuple(n_c)
    rng = int(np.abs(off).max())
    nb = _longs(neighbor_cd)
    out = np.empty_like(a)
    nin = int(np.prod(a.shape[:-3], dtype=np.int64))
    ph, php = _phases(phase_cd, real)
    err = lib().oracle_wfd_apply(
        len(nc), nc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(coef),
        _l(off), _d(wts), _l(n_c), rng, _l(nb), int(real),
        int(is_cfd(off)), php, _d(a.view(float)), _d(out.view(float)), nin)
    if err:
        raise ValueError(f'oracle_wfd_apply: error {err}')
    return out


def wfd_relax(method, coef_ps, offset_ps, weights, n_c, neighbor_cd, f, src,
              nrelax, w=1.0):
    """Reference WeightedFDOperator.relax (c/woperators.cpp:55-94,
    c/bmgs/wrelax.cpp); f updated in place."""
    assert f.dtype == np.float64 and f.flags.c_contiguous
    src = np.ascontiguousarray(src, float)
    n_c = _longs(n_c)
    nc, coef, off = _wstencils(coef_ps, offset_ps)
    wts = np.ascontiguousarray(np.asarray(weights, float))
    rng = int(np.abs(off).max())
    nb = _longs(neighbor_cd)
    err = lib().oracle_wfd_relax(
        int(method), len(nc), nc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
        _d(coef), _l(off), _d(wts), _l(n_c), rng, _l(nb), int(is_cfd(off)),
        _d(f), _d(src), int(nrelax), float(w))
    if err:
        raise ValueError(f'oracle_wfd_relax: error {err}')
    return f


def transform(interpolate, k, size_in, size_out, pad_cd, npad_cd, skip_cd,
              neighbor_cd, a, phases=None, out=None, nthreads=None):
    """Reference Transformer.apply (c/transformers.cpp:112-158)."""
    a = np.ascontiguousarray(a)
    real = a.dtype == np.float64
    size_in = _longs(size_in)
    size_out = _longs(size_out)
    assert a.shape[-3:] == tuple(size_in)
    if out is None:
        out = np.empty(a.shape[:-3] + tuple(size_out), a.dtype)
    nin = int(np.prod(a.shape[:-3], dtype=np.int64))
    ph, php = _phases(phases, real)
    err = lib().oracle_transform(
        int(bool(interpolate)), int(k), _l(size_in), _l(size_out),
        _l(_longs(pad_cd)), _l(_longs(npad_cd)), _l(_longs(skip_cd)),
        _l(_longs(neighbor_cd)), int(real), php, _d(a.view(float)),
        _d(out.view(float)), nin, _threads(nthreads))
    if err:
        raise ValueError(f'oracle_transform: error {err}')
    return out


# --------------------------------------------------------------------------
# Glue restatements over a kernel backend.  A backend provides
#   apply(op, a, out, phases)      op = StencilSpec
#   relax(op, f, src, n, w)        (in place on f)
#   transform(t, a, out, phases)   t = TransferSpec
# ``OracleBackend`` uses the C restatement above; tests/golden/gen_golden.py
# provides one backed by the reference extension.
# --------------------------------------------------------------------------

class StencilSpec:
    def __init__(self, coef_p, offset_pc, n_c, neighbor_cd):
        self.coef_p = np.ascontiguousarray(np.asarray(coef_p, float))
        self.offset_pc = np.asarray(offset_pc, int).reshape((-1, 3))
        self.n_c = np.asarray(n_c, int)
        self.neighbor_cd = np.asarray(neighbor_cd, int).reshape((3, 2))


class TransferSpec:
    def __init__(self, interpolate, k, size_in, size_out, pad_cd, npad_cd,
                 skip_cd, neighbor_cd):
        self.interpolate = bool(interpolate)
        self.k = int(k)
        self.size_in = np.asarray(size_in, int)
        self.size_out = np.asarray(size_out, int)
        self.pad_cd = np.asarray(pad_cd, int).reshape((3, 2))
        self.npad_cd = np.asarray(npad_cd, int).reshape((3, 2))
        self.skip_cd = np.asarray(skip_cd, int).reshape((3, 2))
        self.neighbor_cd = np.asarray(neighbor_cd, int).reshape((3, 2))


class OracleBackend:
    def __init__(self, nthreads=None):
        self.nthreads = nthreads

    def apply(self, op, a, out, phases=None):
        fd_apply(op.coef_p, op.offset_pc, op.n_c, op.neighbor_cd, a,
                 phases, out=out, nthreads=self.nthreads)

    def relax(self, op, f, src, n, w, method=2):
        fd_relax(method, op.coef_p, op.offset_pc, op.n_c, op.neighbor_cd, f,
                 src, n, w)

    def transform(self, t, a, out, phases=None):
        transform(t.interpolate, t.k, t.size_in, t.size_out, t.pad_cd,
                  t.npad_cd, t.skip_cd, t.neighbor_cd, a, phases, out=out,
                  nthreads=self.nthreads)


def fd_precond(kin0, kin1, kin2, rest0, rest1, intp1, intp2, residuals,
               step, phases=None, backend=None):
    """Kinetic preconditioner chain (c/fd_preconditioner.cpp:221-345, the
    same sequence as gpaw/preconditioner.py:74-93); returns d0."""
    be = backend or OracleBackend()
    res = np.ascontiguousarray(residuals)
    nb = res.shape[:-3]
    d0 = np.empty_like(res)
    q0 = np.empty_like(res)
    r1 = np.empty(nb + tuple(kin1.n_c), res.dtype)
    q1 = np.empty_like(r1)
    r2 = np.empty(nb + tuple(kin2.n_c), res.dtype)
    q2 = np.empty_like(r2)
    be.transform(rest0, -res, r1, phases)
    d1 = (4 * step) * r1
    be.apply(kin1, d1, q1, phases)
    q1 -= r1
    be.transform(rest1, q1, r2, phases)
    d2 = (16 * step) * r2
    be.apply(kin2, d2, q2, phases)
    q2 -= r2
    d2 -= (16 * step) * q2
    be.transform(intp2, d2, q1, phases)
    d1 -= q1
    be.apply(kin1, d1, q1, phases)
    q1 -= r1
    d1 -= (4 * step) * q1
    be.transform(intp1, -d1, d0, phases)
    be.apply(kin0, d0, q0, phases)
    q0 -= res
    d0 -= step * q0
    d0 *= -1.0
    return d0


class PoissonVCycle:
    """FDPoissonSolver V-cycles (gpaw/poisson.py:451-537) over a backend.

    ops[l] (StencilSpec), restrictors[l] / interpolators[l] (TransferSpec)
    for l = 0..L; smoothing counts and Jacobi weights as the reference's
    set_grid_descriptor / _init produce them; relax_method 2 (Jacobi) or
    1 (Gauss-Seidel), passed to every relax call as the reference's
    iterate2 does.
    """

    def __init__(self, ops, restrictors, interpolators, presmooths,
                 postsmooths, weights, dv, backend=None, relax_method=2):
        self.ops = ops
        self.restrictors = restrictors
        self.interpolators = interpolators
        self.presmooths = list(presmooths)
        self.postsmooths = list(postsmooths)
        self.weights = list(weights)
        self.levels = len(ops) - 1
        self.dv = dv
        self.be = backend or OracleBackend()
        self.method = relax_method
        self.phis = [None] + [np.empty(tuple(o.n_c)) for o in ops[1:]]
        self.rhos = [np.empty(tuple(o.n_c)) for o in ops]
        self.residuals = [np.empty(tuple(o.n_c)) for o in ops]

    def _relax(self, op, level, n):
        if self.method == 2:
            self.be.relax(op, self.phis[level], self.rhos[level], n,
                          self.weights[level])
        else:
            self.be.relax(op, self.phis[level], self.rhos[level], n,
                          self.weights[level], self.method)

    def iterate2(self, level=0):
        residual = self.residuals[level]
        op = self.ops[level]
        if level < self.levels:
            self._relax(op, level, self.presmooths[level])
            self.be.apply(op, self.phis[level], residual)
            residual -= self.rhos[level]
            self.be.transform(self.restrictors[level], residual,
                              self.rhos[level + 1])
            self.phis[level + 1][:] = 0.0
            self.iterate2(level + 1)
            self.be.transform(self.interpolators[level],
                              self.phis[level + 1], residual)
            self.phis[level] -= residual
        self._relax(op, level, self.postsmooths[level])
        if level == 0:
            self.be.apply(op, self.phis[level], residual)
            residual -= self.rhos[level]
            return float(np.dot(residual.ravel(), residual.ravel())) * self.dv

    def solve_neutral(self, phi, rho, eps, maxiter=1000, pbc_all=True,
                      fixed_iterations=None):
        """Returns niter; with fixed_iterations, runs exactly that many
        V-cycles (for comparisons at equal niter) and skips the test."""
        self.phis[0] = phi
        self.rhos[0][:] = rho
        if fixed_iterations is not None:
            for _ in range(fixed_iterations):
                self.last_error = self.iterate2()
            niter = fixed_iterations
        else:
            niter = 1
            while True:
                self.last_error = self.iterate2()
                if not (self.last_error > eps and niter < maxiter):
                    break
                niter += 1
            if niter == maxiter:
                raise RuntimeError('Poisson solver did not converge in %d '
                                   'iterations!' % maxiter)
        if pbc_all:
            phi_ave = float(np.sum(phi.ravel())) / phi.size
            phi -= phi_ave
        return niter


def rmm_block_step(kin, vt, precond, psi, R, eig, phases=None,
                   trial_step=None):
    """RMM-DIIS block step with diis_steps=1 on the grid part
    (gpaw/new/pwfd/rmmdiis.py:123-257): precond(x) is the kinetic
    preconditioner, kin a StencilSpec, vt the local potential, eig the
    band energies.  Returns (psi_new, intermediates)."""
    def resid(x):
        # calculate_residuals without projections
        # (gpaw/new/pwfd/eigensolver.py:197-227)
        hx = fd_apply(kin.coef_p, kin.offset_pc, kin.n_c, kin.neighbor_cd,
                      x, phases) + x * vt
        return hx - eig.reshape((-1, 1, 1, 1)) * x

    nb = psi.shape[0]
    PR = precond(R)
    dR = resid(PR)
    a = np.array([-np.vdot(dR[n], R[n]).real for n in range(nb)])
    b = np.array([np.vdot(dR[n], dR[n]).real for n in range(nb)])
    lam = np.where(b > 0, a / np.where(b > 0, b, 1.0), 0.0)
    lr = lam.reshape((-1, 1, 1, 1))
    R1 = R + lr * dR
    psi1 = psi + lr * PR
    PR2 = precond(R1)
    step = lr if trial_step is None else trial_step
    return psi1 + step * PR2, {'PR': PR, 'dR': dR, 'lam': lam, 'R1': R1,
                               'psi1': psi1}
