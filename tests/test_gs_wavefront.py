"""The Gauss-Seidel wavefront order, on the CPU.

The GPU sweep updates point (i0, i1, i2) at step a0 i0 + a1 i1 + i2 with
the weights gfd_gs_order returns (host code of the library, no device
calls), reading in-grid neighbours from the array being updated and
wrapped ones from a start-of-sweep copy.  Here that schedule is replayed
in numpy with the library's weights and compared bit for bit with the
oracle's lexicographic sweep (c/bmgs/relax.cpp:12-40), which is pinned to
the reference build in test_oracle_ref.py."""
import ctypes

import numpy as np
import pytest

import oracle
from gpaw_amd import _lib
from gpaw_amd.fd_operators import GUCLaplace, Laplace, LaplaceA
from gpaw_amd.grid_descriptor import GridDescriptor


def gs_weights(off_pc):
    off = np.ascontiguousarray(np.asarray(off_pc, dtype=np.int32))
    a0, a1 = ctypes.c_int(0), ctypes.c_int(0)
    ip = ctypes.POINTER(ctypes.c_int)
    st = _lib.load().gfd_gs_order(off.ctypes.data_as(ip), len(off),
                                  ctypes.byref(a0), ctypes.byref(a1))
    _lib.check(st, 'gfd_gs_order')
    return a0.value, a1.value


def wavefront_sweeps(coef, off, n, pbc, f, src, nsweeps, a0, a1):
    """The GPU schedule in numpy: every point of one step at once."""
    f = f.copy()
    n = tuple(int(x) for x in n)
    i0, i1 = np.meshgrid(np.arange(n[0]), np.arange(n[1]), indexing='ij')
    inv_c0 = 1.0 / coef[0]
    for _ in range(nsweeps):
        snap = f.copy()
        for t in range(a0 * (n[0] - 1) + a1 * (n[1] - 1) + n[2]):
            i2 = t - a0 * i0 - a1 * i1
            m = (i2 >= 0) & (i2 < n[2])
            if not m.any():
                continue
            p = (i0[m], i1[m], i2[m])
            x = np.zeros(len(p[0]))
            for c in range(1, len(coef)):
                j = []
                wrapped = np.zeros(len(p[0]), bool)
                inside = np.ones(len(p[0]), bool)
                for d in range(3):
                    jd = p[d] + off[c][d]
                    out = (jd < 0) | (jd >= n[d])
                    if pbc[d]:
                        wrapped |= out
                        jd = jd % n[d]
                    else:
                        inside &= ~out
                        jd = np.clip(jd, 0, n[d] - 1)
                    j.append(jd)
                j = tuple(j)
                v = np.where(wrapped, snap[j], f[j])
                v = np.where(inside, v, 0.0)
                x = x + v * coef[c]
            f[p] = (src[p] - x) * inv_c0
    return f


def _op(kind, gd):
    if kind == 'mehrstellen':
        return LaplaceA(gd, 0.25 / np.pi)
    if kind == 'hexagonal':
        return GUCLaplace(gd, -0.25 / np.pi, 2)
    return Laplace(gd, -0.25 / np.pi, int(kind[2]))


CASES = [('nn1', (6, 5, 7), (True, True, True)),
         ('nn2', (7, 6, 5), (True, False, True)),
         ('nn3', (8, 7, 6), (False, True, True)),
         ('mehrstellen', (6, 7, 5), (False, False, False)),
         ('mehrstellen', (6, 6, 6), (True, True, True)),
         ('hexagonal', (6, 6, 5), (True, True, True))]


@pytest.mark.parametrize('kind,N,pbc', CASES)
def test_wavefront_order_is_lexicographic_sweep(kind, N, pbc):
    rng = np.random.default_rng(90 + CASES.index((kind, N, pbc)))
    cell = np.array(N) * 0.3
    if kind == 'hexagonal':
        cell = np.array([(1, 0, 0), (0.5, 3**0.5 / 2, 0), (0, 0, 1.1)]) * \
            np.array(N)[:, None] * 0.3
    gd = GridDescriptor(N, cell, pbc)
    op = _op(kind, gd)
    assert op.offset_pc[0].tolist() == [0, 0, 0]
    a0, a1 = gs_weights(op.offset_pc)
    # the defining property: C-order predecessors lower the step
    for o in op.offset_pc[1:]:
        t = a0 * o[0] + a1 * o[1] + o[2]
        before = tuple(o) < (0, 0, 0)
        assert (t < 0) if before else (t > 0), (tuple(o), a0, a1)
    n = tuple(int(x) for x in gd.n_c)
    f0 = rng.uniform(-1, 1, n)
    src = rng.uniform(-1, 1, n)
    ref = f0.copy()
    oracle.fd_relax(1, op.coef_p, op.offset_pc, gd.n_c, gd.neighbor_cd, ref,
                    src, 2, 1.0)
    got = wavefront_sweeps(op.coef_p, op.offset_pc, n, gd.pbc_c, f0, src, 2,
                           a0, a1)
    assert (got == ref).all(), np.abs(got - ref).max()


def test_order_weights_axis_stencils():
    """Axis-only stencils need no skew: (1, 1) for every radius."""
    for r in (1, 2, 3, 4):
        off = [(0, 0, 0)]
        for d in range(3):
            for k in range(1, r + 1):
                for s in (k, -k):
                    o = [0, 0, 0]
                    o[d] = s
                    off.append(tuple(o))
        assert gs_weights(off) == (1, 1)


def test_order_weights_mehrstellen_skew():
    """(0, 1, -1) follows the centre in C order but has o2 < 0: a1 = 2."""
    off = [(0, 0, 0), (0, 1, -1), (0, -1, 1), (1, 0, 0), (-1, 0, 0)]
    assert gs_weights(off) == (1, 2)
    off += [(1, -2, 0), (-1, 2, 0)]
    a0, a1 = gs_weights(off)
    assert (a0, a1) == (5, 2)
