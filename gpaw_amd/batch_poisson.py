"""Many independent FD Poisson problems of uneven shape, solved together.

Each problem has the hierarchy of FDPoissonSolver (gpaw/poisson.py:366-430:
Laplace(scale=-1/4pi, nn) on the fine level, nn=1 below, nn=1 K=2
transfers, Jacobi weights 2/3 and 1).  The batched solve runs FP32 V-cycles
over every grid at once -- one persistent launch per stencil or transfer
stage of a level -- inside FP64 residual refinement, with one refinement
step captured as a hipGraph.  Tile shapes are timed per level on the actual
batch the first time the batch runs (``tune``).  With ``coarse=True``
each grid's levels from the coarsest batched one down to its own coarsest
(GPAW coarsens while every side is even, so depths differ between grids)
run in one LDS-resident workgroup per grid, a single launch per V-cycle.

There is no counterpart in the reference, which solves one grid at a time.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .arrays import ptr, stream_handle, torch
from .poisson import FDPoissonSolver, PoissonConvergenceError


class BatchedPoissonSolver:
    def __init__(self, nn=3, levels=4, coarse=True, w0=None):
        """levels: hierarchy levels run as batched launches; w0: Jacobi
        weight of the fine level (None: the reference's 2/3)."""
        self.nn = nn
        self.levels = levels
        self.coarse = coarse
        self.w0 = w0
        self.solvers = []
        self._native = None

    def set_grid_descriptors(self, gds):
        self._free()
        self.gds = list(gds)
        self.solvers = []
        for gd in self.gds:
            s = FDPoissonSolver(nn=self.nn)
            s.set_grid_descriptor(gd)
            self.solvers.append(s)
        self.L = min([self.levels] + [s.levels + 1 for s in self.solvers])
        self.weights = list(self.solvers[0].weights[:self.L])
        if self.w0 is not None:
            self.weights[0] = float(self.w0)

    def _handle(self):
        if self._native is None:
            L, G = self.L, len(self.solvers)
            V = _lib.ctypes.c_void_p
            ops = (V * (G * L))(*[s.operators[l]._handle().value
                                  for s in self.solvers for l in range(L)])
            nt = max(G * (L - 1), 1)
            rst = (V * nt)(*([s.restrictors[l]._handle().value
                              for s in self.solvers for l in range(L - 1)]
                             or [None]))
            itp = (V * nt)(*([s.interpolators[l]._handle().value
                              for s in self.solvers for l in range(L - 1)]
                             or [None]))
            w = (_lib.ctypes.c_double * L)(*self.weights)
            h = V()
            _lib.check(_lib.load().gfd_batch_poisson_new(
                G, L, ops, rst, itp, w, _lib.ctypes.byref(h)),
                'BatchedPoissonSolver')
            self._native = h
            if self.coarse:
                self._set_coarse(True)
        return self._native

    def _set_coarse(self, on):
        lib = _lib.load()
        V = _lib.ctypes.c_void_p
        if not on:
            _lib.check(lib.gfd_batch_poisson_set_coarse(
                self._native, None, None, None, None, None), 'set_coarse')
            return
        nlev = [s.levels + 1 for s in self.solvers]
        ops = [op._handle().value for s in self.solvers
               for op in s.operators]
        rst = [t._handle().value for s in self.solvers
               for t in s.restrictors] or [None]
        itp = [t._handle().value for s in self.solvers
               for t in s.interpolators] or [None]
        w = list(max(self.solvers, key=lambda s: s.levels).weights)
        if self.w0 is not None:
            w[0] = float(self.w0)
        _lib.check(lib.gfd_batch_poisson_set_coarse(
            self._native, (_lib.ctypes.c_int * len(nlev))(*nlev),
            (V * len(ops))(*ops), (V * len(rst))(*rst),
            (V * len(itp))(*itp), (_lib.ctypes.c_double * len(w))(*w)),
            'set_coarse')

    def _free(self):
        if self._native is not None:
            _lib.load().gfd_batch_poisson_free(self._native)
            self._native = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass

    def tune(self):
        """Time the candidate tile shapes of both stencil kernels on this
        batch (also done on the first solve); returns tiles()."""
        lib = _lib.load()
        _lib.check(lib.gfd_batch_poisson_tune(self._handle(),
                                              stream_handle()), 'tune')
        return self.tiles()

    def tiles(self):
        """{level: dict} of the chosen kernel and tile per level, level -1
        being the FP64 fine-level residual: 'kernel' ('regs' register
        tiles or 'march' x-march tiles), 'xc' x planes per tile, 'tiles',
        and bz / by / pz / blocks (register tiles) or nt / py / zmax / pf
        (x-march tiles, pf planes of loads in flight)."""
        lib = _lib.load()
        out = {}
        c = _lib.ctypes
        for level in range(-1, self.L):
            v = (c.c_long * 11)()
            _lib.check(lib.gfd_batch_poisson_tile_info(self._handle(), level,
                                                       v, 11), 'tiles')
            d = {'kernel': 'march' if v[0] == 1 else 'regs', 'xc': v[1],
                 'tiles': v[2]}
            if v[0] == 1:
                d.update(nt=v[7], py=v[8], zmax=v[9], pf=v[10])
            else:
                d.update(bz=v[3], by=v[4], pz=v[5], blocks=v[6])
            out[level] = d
        return out

    def level_points(self, level):
        return _lib.load().gfd_batch_poisson_level_points(self._handle(),
                                                          level)

    def pack(self, arrays, level=0, dtype=torch.float64):
        """Concatenate one array per grid (level-`level` shapes) into the
        batch's level layout."""
        return torch.cat([a.reshape(-1).to(dtype) for a in arrays])

    def unpack(self, flat, level=0):
        out, i = [], 0
        for s in self.solvers:
            n = tuple(s.operators[level].gd.n_c)
            k = int(np.prod(n))
            out.append(flat[i:i + k].reshape(n))
            i += k
        return out

    def apply(self, level, flat_in, flat_out):
        """flat_out = L_level flat_in on every grid (FP64 or FP32 by the
        arrays' dtype)."""
        fp32 = flat_in.dtype == torch.float32
        _lib.check(_lib.load().gfd_batch_poisson_apply(
            self._handle(), level, int(fp32), ptr(flat_in), ptr(flat_out),
            stream_handle()), 'batched apply')
        return flat_out

    def transfer(self, level, interpolate, flat_in, flat_out, base=None):
        """Restriction level -> level + 1, or flat_out = base - I(flat_in)
        from level + 1 to level."""
        fp32 = flat_in.dtype == torch.float32
        _lib.check(_lib.load().gfd_batch_poisson_transfer(
            self._handle(), level, int(interpolate), int(fp32), ptr(flat_in),
            ptr(flat_out), None if base is None else ptr(base),
            stream_handle()), 'batched transfer')
        return flat_out

    def solve_mixed(self, phis, rhos, tol=1e-10, pre=3, post=3,
                    coarse_sweeps=16, max_outer=100, graph=True):
        """Refine every phi_g (float64 device tensors, in/out) until
        |L phi_g - rho_g| / |rho_g| <= tol for all g.  Returns
        (refinement steps, relative residuals)."""
        G = len(self.solvers)
        if len(phis) != G or len(rhos) != G:
            raise ValueError('one phi and one rho per grid')
        for s, p, r in zip(self.solvers, phis, rhos):
            n = tuple(s.gd.n_c)
            for a in (p, r):
                if tuple(a.shape) != n or a.dtype != torch.float64 or \
                        not a.is_contiguous():
                    raise ValueError('phi / rho: contiguous float64 of the '
                                     'grid shape %s' % (n,))
        P = _lib.ctypes.c_void_p * G
        pp = P(*[ptr(p) for p in phis])
        rr = P(*[ptr(r) for r in rhos])
        n = _lib.ctypes.c_int(0)
        rel = (_lib.ctypes.c_double * G)()
        status = _lib.load().gfd_batch_poisson_solve_mixed(
            self._handle(), pp, rr, float(tol), int(pre), int(post),
            int(coarse_sweeps), int(max_outer), int(bool(graph)),
            _lib.ctypes.byref(n), rel, stream_handle())
        relres = np.array(rel[:])
        if status == _lib.GFD_ENOTSUP and self.coarse:
            # some grid's coarse levels exceed one workgroup's LDS
            self.coarse = False
            self._set_coarse(False)
            return self.solve_mixed(phis, rhos, tol, pre, post,
                                    coarse_sweeps, max_outer, graph)
        if status == _lib.GFD_ECONVERGE:
            raise PoissonConvergenceError(
                'batched mixed solve: max relative residual %.3g after %d '
                'steps' % (relres.max(), n.value))
        _lib.check(status, 'BatchedPoissonSolver.solve_mixed')
        return n.value, relres
