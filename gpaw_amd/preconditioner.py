"""Multigrid kinetic preconditioner on the GPU (drop-in for
gpaw.preconditioner.Preconditioner, gpaw/preconditioner.py:9-93).

The whole 3-level chain of c/fd_preconditioner.cpp:221-345 runs as one
native call (gfd_fd_precond) over a block of bands: four transfers and
three stencil passes with the element-wise steps fused into them.

``fp32=True`` (a keyword the reference does not have) runs the chain in
FP32 (gfd_fd_precond32): the first restriction reads the FP64 residuals
and narrows them, the work arrays are float32 / complex64, and the last
stencil pass writes the FP64 result.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .arrays import Staging, device, is_device, ptr, stream_handle, torch
from .fd_operators import Laplace
from .transformers import Transformer


class Preconditioner:
    def __init__(self, gd0, kin0, dtype=float, block=1, xp=np, fp32=False):
        gd1 = gd0.coarsen()
        gd2 = gd1.coarsen()
        self.gd0, self.gd1, self.gd2 = gd0, gd1, gd2
        self.kin0 = kin0
        self.kin1 = Laplace(gd1, -0.5, 1, dtype, xp=xp)
        self.kin2 = Laplace(gd2, -0.5, 1, dtype, xp=xp)
        self.dtype = dtype
        self.block = block
        self.step = 0.66666666 / kin0.get_diagonal_element()
        self.restrictor_object0 = Transformer(gd0, gd1, 1, dtype, xp=xp)
        self.restrictor_object1 = Transformer(gd1, gd2, 1, dtype, xp=xp)
        self.interpolator_object2 = Transformer(gd2, gd1, 1, dtype, xp=xp)
        self.interpolator_object1 = Transformer(gd1, gd0, 1, dtype, xp=xp)
        self.restrictor0 = self.restrictor_object0.apply
        self.restrictor1 = self.restrictor_object1.apply
        self.interpolator2 = self.interpolator_object2.apply
        self.interpolator1 = self.interpolator_object1.apply
        self.use_c_precond = True
        self.xp = xp
        self.fp32 = bool(fp32)
        self._scratch = None

    def calculate_kinetic_energy(self, psit_xG, kpt):
        return None

    def _work(self, nb):
        """(fine q0, level-1 r1 d1 q1, level-2 r2 d2 q2) of at least nb
        bands."""
        if self._scratch is None or self._scratch[1].shape[1] < nb:
            cplx = self.dtype == complex
            f64 = torch.complex128 if cplx else torch.float64
            wdt = (torch.complex64 if cplx else torch.float32) \
                if self.fp32 else f64
            dev = device()
            n = max(nb, self.block)
            self._scratch = (
                torch.zeros((n,) + tuple(self.gd0.n_c), dtype=wdt,
                            device=dev),
                torch.zeros((3, n) + tuple(self.gd1.n_c), dtype=wdt,
                            device=dev),
                torch.zeros((3, n) + tuple(self.gd2.n_c), dtype=wdt,
                            device=dev))
        return self._scratch

    def __call__(self, residuals, kpt, ekin=None, out=None):
        if residuals.ndim == 3:
            if out is None:
                return self.__call__(residuals[None], kpt)[0]
            return self.__call__(residuals[None], kpt, out=out[None])[0]
        nb = len(residuals)
        if nb > self.block:
            if out is None:
                out = (torch.empty_like(residuals) if is_device(residuals)
                       else np.empty_like(residuals))
            for n1 in range(0, nb, self.block):
                self(residuals[n1:n1 + self.block], kpt,
                     out=out[n1:n1 + self.block])
            return out
        if tuple(residuals.shape[-3:]) != tuple(self.gd0.n_c):
            raise TypeError('Input grid shape does not match with arrays d0 '
                            'and q0.')
        phases = getattr(kpt, 'phase_cd', None) if kpt is not None else None
        s0, s1, s2 = self._work(nb)
        q0 = s0[:nb]
        r1, d1, q1 = s1[0, :nb], s1[1, :nb], s1[2, :nb]
        r2, d2, q2 = s2[0, :nb], s2[1, :nb], s2[2, :nb]
        st = Staging(residuals)
        res = st.dev[0].contiguous()
        if out is None:
            d0 = torch.empty_like(res)
        else:
            d0 = out if is_device(out) else torch.empty_like(res)
        keep, ph = (None, None)
        if self.dtype == complex:
            keep, ph = _lib.phases_arg(np.ones((3, 2), complex)
                                       if phases is None else phases)
        lib = _lib.load()
        fn = lib.gfd_fd_precond32 if self.fp32 else lib.gfd_fd_precond
        _lib.check(fn(
            self.restrictor_object0._handle(),
            self.restrictor_object1._handle(),
            self.interpolator_object1._handle(),
            self.interpolator_object2._handle(),
            self.kin0._handle(), self.kin1._handle(), self.kin2._handle(),
            ptr(d0), ptr(q0), ptr(r1), ptr(d1), ptr(q1), ptr(r2), ptr(d2),
            ptr(q2), ptr(res), nb, float(self.step), ph, stream_handle()),
            'fd_precond')
        if out is not None and not is_device(out):
            out[...] = d0.cpu().numpy()
            return out
        if not is_device(residuals):
            return d0.cpu().numpy()
        return d0


class BatchedPreconditioner:
    """The FP32 kinetic preconditioner on many independent real grids of
    uneven shape at once (no counterpart in the reference, which handles
    one grid descriptor per Preconditioner).

    Each grid gets the hierarchy of Preconditioner: K0 = the caller's
    kinetic stencil order (nn, scale -1/2) on the grid, nn=1 kinetic
    operators on its two coarsenings, nn=1 transfers, and its own step
    0.66666666 / diag(K0) (gpaw/preconditioner.py:17-33).  One call runs
    the chain of c/fd_preconditioner.cpp:221-345 on every grid in eight
    launches (gfd_batch_precond32), FP32 work arrays, FP64 residuals and
    results."""

    def __init__(self, gds, nn=3):
        self.gds = list(gds)
        self.levels = []
        for gd in self.gds:
            gd1 = gd.coarsen()
            gd2 = gd1.coarsen()
            kin = [Laplace(gd, -0.5, nn), Laplace(gd1, -0.5, 1),
                   Laplace(gd2, -0.5, 1)]
            rst = [Transformer(gd, gd1, 1), Transformer(gd1, gd2, 1)]
            itp = [Transformer(gd1, gd, 1), Transformer(gd2, gd1, 1)]
            self.levels.append((kin, rst, itp))
        self.steps = np.array([0.66666666 / k[0].get_diagonal_element()
                               for k, _, _ in self.levels])
        self._native = None
        self._res = None

    def _handle(self):
        if self._native is None:
            G = len(self.gds)
            V = _lib.ctypes.c_void_p
            ops = (V * (3 * G))(*[k[l]._handle().value
                                  for k, _, _ in self.levels
                                  for l in range(3)])
            rst = (V * (2 * G))(*[r[l]._handle().value
                                  for _, r, _ in self.levels
                                  for l in range(2)])
            # level pair l: interpolation from l + 1 back to l
            itp = (V * (2 * G))(*[i[l]._handle().value
                                  for _, _, i in self.levels
                                  for l in range(2)])
            w = (_lib.ctypes.c_double * 3)(1.0, 1.0, 1.0)
            h = V()
            _lib.check(_lib.load().gfd_batch_poisson_new(
                G, 3, ops, rst, itp, w, _lib.ctypes.byref(h)),
                'BatchedPreconditioner')
            self._native = h
        return self._native

    def __del__(self):
        try:
            if self._native is not None:
                _lib.load().gfd_batch_poisson_free(self._native)
                self._native = None
        except Exception:
            pass

    @property
    def npoints(self):
        """Points of all grids: the length of the packed layout."""
        return int(sum(int(np.prod(gd.n_c)) for gd in self.gds))

    def pack(self, arrays):
        """The per-grid arrays back to back (grid g after grids 0..g-1),
        the layout apply_packed works on."""
        return torch.cat([a.reshape(-1) for a in arrays])

    def apply_packed(self, res, out):
        """The chain on residuals already in the packed layout (float64,
        npoints long) into ``out`` (same layout, not aliasing res): no
        gathering or scattering of per-grid arrays."""
        n = self.npoints
        for name, t in (('res', res), ('out', out)):
            if (t.dtype != torch.float64 or t.numel() != n or
                    not t.is_contiguous()):
                raise TypeError('%s: contiguous float64 array of %d points'
                                % (name, n))
        steps = (_lib.ctypes.c_double * len(self.steps))(*self.steps)
        _lib.check(_lib.load().gfd_batch_precond32(
            self._handle(), steps, ptr(res), ptr(out), stream_handle()),
            'BatchedPreconditioner')
        return out

    def __call__(self, residuals, out=None):
        """residuals: one float64 device array per grid (its shape);
        returns (or fills ``out`` with) the preconditioned arrays."""
        if len(residuals) != len(self.gds):
            raise ValueError('one residual array per grid')
        for gd, r in zip(self.gds, residuals):
            if tuple(r.shape) != tuple(gd.n_c) or r.dtype != torch.float64:
                raise TypeError('residuals: float64 arrays of the grid '
                                'shapes')
        flat = self.pack(residuals)
        res_out = torch.empty_like(flat)
        self.apply_packed(flat, res_out)
        if out is None:
            out = [None] * len(self.gds)
        i = 0
        for g, gd in enumerate(self.gds):
            n = tuple(int(x) for x in gd.n_c)
            k = int(np.prod(n))
            v = res_out[i:i + k].reshape(n)
            if out[g] is None:
                out[g] = v
            else:
                out[g].copy_(v)
            i += k
        return out
