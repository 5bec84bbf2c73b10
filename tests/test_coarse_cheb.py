"""Eigenvalue bounds of the coarse solve's Chebyshev iteration, on the
CPU: the library's bounds (host code, no device calls) against the
eigenvalues of the dense matrix D^-1 A of the same radius-1 stencil,
built point by point with its periodic wrap / zero boundary."""
import ctypes

import numpy as np
import pytest

from gpaw_amd import _lib


def lib_bounds(n, periodic, cp, c0):
    ni = (ctypes.c_int * 3)(*n)
    pi = (ctypes.c_int * 3)(*[int(p) for p in periodic])
    cd = (ctypes.c_double * 3)(*cp)
    lo, hi = ctypes.c_double(), ctypes.c_double()
    st = _lib.load().gfd_coarse_cheb_bounds(ni, pi, cd, c0, ctypes.byref(lo),
                                            ctypes.byref(hi))
    _lib.check(st, 'gfd_coarse_cheb_bounds')
    return lo.value, hi.value


def dense(n, periodic, cp, c0):
    idx = np.arange(int(np.prod(n))).reshape(n)
    A = np.zeros((idx.size, idx.size))
    for i in np.ndindex(*n):
        p = idx[i]
        A[p, p] += c0
        for c in range(3):
            for s in (1, -1):
                j = list(i)
                j[c] += s
                if not 0 <= j[c] < n[c]:
                    if not periodic[c]:
                        continue
                    j[c] %= n[c]
                A[p, idx[tuple(j)]] += cp[c]
    return np.linalg.eigvalsh(A / c0)


CASES = [((5, 7, 4), (True, True, True), (0.3, 0.35, 0.4)),
         ((5, 6, 7), (False, False, False), (0.3, 0.3, 0.3)),
         ((4, 6, 5), (True, False, True), (0.5, 0.3, 0.4)),
         ((3, 9, 5), (False, True, False), (0.2, 0.6, 0.3)),
         ((2, 8, 6), (True, True, False), (0.25, 0.25, 0.5))]


@pytest.mark.parametrize('n,periodic,h', CASES)
def test_bounds_bracket_dense_spectrum(n, periodic, h):
    cp = [1.0 / x**2 for x in h]
    c0 = -2.0 * sum(cp)
    lo, hi = lib_bounds(n, periodic, cp, c0)
    ev = dense(n, periodic, cp, c0)
    nonzero = ev[ev > 1e-10]
    # the exact extremes, widened by 0.9 and 1.05
    assert lo == pytest.approx(0.9 * nonzero[0], rel=1e-12)
    assert hi == pytest.approx(1.05 * ev[-1], rel=1e-12)
    # a singular (all-periodic) operator keeps its constant mode at 0
    assert (len(nonzero) < len(ev)) == all(periodic)


def test_chebyshev_steps_reduce_error():
    """The iteration those bounds drive, in numpy: 16 steps reduce every
    error component of a 6 x 7 x 5 zero-boundary problem at least as
    |T_16|^-1 predicts, far below 16 weighted-Jacobi sweeps."""
    n, periodic = (6, 7, 5), (False, False, False)
    cp = [1.0, 1.2, 0.8]
    c0 = -2.0 * sum(cp)
    lo, hi = lib_bounds(n, periodic, cp, c0)
    ev = dense(n, periodic, cp, c0)
    theta, delta = (hi + lo) / 2, (hi - lo) / 2
    sigma = theta / delta
    rho = 1 / sigma
    lam = ev                      # error components in the eigenbasis
    e = np.ones_like(lam)
    d = e * 0
    for s in range(16):
        z = -lam * e              # D^-1 (r - A x) for r = A x_exact
        if s == 0:
            d = z / theta
        else:
            rn = 1 / (2 * sigma - rho)
            d = rn * rho * d + 2 * rn / delta * z
            rho = rn
        e = e + d
    kappa = hi / lo
    bound = 2 * ((kappa**0.5 - 1) / (kappa**0.5 + 1))**16
    assert np.abs(e).max() <= bound * 1.0001
    jac = np.abs((1 - 0.8 * lam)**16).max()
    assert np.abs(e).max() < 0.05 * jac
