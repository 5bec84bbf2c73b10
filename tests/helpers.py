"""Shared test helpers: oracle specs from drop-in objects, error norms."""
import numpy as np

import oracle


def spec(op):
    return oracle.StencilSpec(op.coef_p, op.offset_pc, op.gd.n_c,
                              op.gd.neighbor_cd)


def tspec(t):
    return oracle.TransferSpec(t.interpolate, 2 * t.nn, t.gdin.n_c,
                               t.gdout.n_c, t.pad_cd, t.neighborpad_cd,
                               t.skip_cd, t.gdin.neighbor_cd)


def vcycle(solver, backend=None):
    """oracle.PoissonVCycle with the hierarchy of an FDPoissonSolver."""
    solver._init()
    return oracle.PoissonVCycle([spec(o) for o in solver.operators],
                                [tspec(t) for t in solver.restrictors],
                                [tspec(t) for t in solver.interpolators],
                                solver.presmooths, solver.postsmooths,
                                solver.weights, solver.gd.dv, backend,
                                getattr(solver, 'relax_method', 2))


def relerr(x, ref):
    """max |x - ref| / max |ref| (the parity metric)."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.abs(ref).max()
    if den == 0:
        return float(np.abs(x).max())
    return float(np.abs(x - ref).max() / den)


def rand(rng, shape, dtype=float):
    a = rng.uniform(-1, 1, shape)
    if dtype == complex:
        a = a + 1j * rng.uniform(-1, 1, shape)
    return a


def bloch_phases(rng, pbc):
    k = rng.uniform(-0.5, 0.5, 3) * np.asarray(pbc)
    return np.exp(2j * np.pi * k[:, None] * np.array([1, -1]))


def case_names(npz, group):
    return sorted({k.split('/')[1] for k in npz.files
                   if k.startswith(group + '/')})


def case(npz, group, name):
    pre = f'{group}/{name}/'
    return {k[len(pre):]: npz[k] for k in npz.files if k.startswith(pre)}
