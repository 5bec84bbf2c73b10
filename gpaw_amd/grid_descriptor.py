"""Minimal, API-compatible GridDescriptor (single domain).

The reference descriptor (gpaw/old/grid_descriptor.py:38-160, with
gpaw/old/domain.py) needs ASE and MPI; the FD hot path only uses its layout
contract, restated here:

- ``N_c`` grid points per axis; a non-periodic axis stores ``N - 1`` points
  starting at index 1 (``beg_c = 1``), a periodic one all ``N``
  (grid_descriptor.py:119-140);
- ``h_cv = cell_cv / N_c``, ``dv = |det cell| / prod(N_c)`` (:142-144);
- ``neighbor_cd`` is the neighbouring rank, ``-1`` at a zero boundary
  (domain.py:110-138) -- on one domain it is 0 for periodic axes;
- ``coarsen`` / ``refine`` halve / double ``N_c`` (:327-341).

Arrays come from ``xp``: numpy (host) or the device namespace
``gpaw_amd.arrays`` (torch tensors in HBM).
"""
from __future__ import annotations

import numbers

import numpy as np


class SerialCommunicator:
    """Stand-in for the one-rank MPI communicator (gpaw/mpi.py:728-840)."""
    size = 1
    rank = 0
    parent = None

    def sum_scalar(self, x):
        return x

    def sum(self, a, root=-1):
        return a

    def max(self, x):
        return x

    def min(self, x):
        return x

    def barrier(self):
        pass

    def get_c_object(self):
        return None


serial_comm = SerialCommunicator()


class BadGridError(ValueError):
    pass


class GridDescriptor:
    ndim = 3

    def __init__(self, N_c, cell_cv=(1, 1, 1), pbc_c=True, comm=None,
                 parsize_c=None, allow_empty_domains=False):
        if isinstance(pbc_c, (bool, int, np.bool_)):
            pbc_c = (bool(pbc_c),) * 3
        if comm is None:
            comm = serial_comm
        if getattr(comm, 'size', 1) != 1:
            raise NotImplementedError('domain decomposition: one GPU, one '
                                      'domain')
        self.comm = comm
        self.N_c = np.array(N_c, int)
        if (self.N_c != np.asarray(N_c)).any():
            raise ValueError('Non-int number of grid points %s' % (N_c,))
        cell = np.array(cell_cv, float)
        if cell.ndim == 1:
            cell = np.diag(cell)
        self.cell_cv = cell
        self.pbc_c = np.array(pbc_c, bool)
        self.parsize_c = np.ones(3, int)
        self.parpos_c = np.zeros(3, int)
        self.stride_c = np.array([1, 1, 1])
        self.rank = 0
        self.beg_c = np.where(self.pbc_c, 0, 1).astype(int)
        self.end_c = self.N_c.copy()
        self.n_c = self.end_c - self.beg_c
        if (self.n_c <= 0).any():
            raise BadGridError('Grid %s too small' % self.N_c)
        self.n_cp = [np.array([b, e]) for b, e in zip(self.beg_c, self.end_c)]
        self.h_cv = self.cell_cv / self.N_c[:, np.newaxis]
        self.volume = abs(np.linalg.det(self.cell_cv))
        self.dv = self.volume / self.N_c.prod()
        self.orthogonal = not (self.cell_cv -
                               np.diag(self.cell_cv.diagonal())).any()
        self.icell_cv = np.linalg.inv(self.cell_cv).T
        cell_c = (self.cell_cv**2).sum(1)**0.5
        self.xxxucell_cv = np.array([self.cell_cv[x] / cell_c[x]
                                     for x in range(3)])
        self.xxxiucell_cv = np.linalg.inv(self.xxxucell_cv.T)
        self.neighbor_cd = np.where(self.pbc_c[:, None], 0, -1) * \
            np.ones((3, 2), int)
        self.neighbor_cd = np.ascontiguousarray(self.neighbor_cd, dtype=int)
        self.sdisp_cd = np.zeros((3, 2), int)
        for c in range(3):
            if self.pbc_c[c]:
                self.sdisp_cd[c] = [1, -1]

    # -- descriptors -----------------------------------------------------
    def new_descriptor(self, N_c=None, cell_cv=None, pbc_c=None, comm=None,
                       parsize_c=None, allow_empty_domains=False):
        return self.__class__(self.N_c if N_c is None else N_c,
                              self.cell_cv if cell_cv is None else cell_cv,
                              self.pbc_c if pbc_c is None else pbc_c,
                              self.comm if comm is None else comm)

    def coarsen(self):
        if (self.N_c % 2).any():
            raise ValueError('Grid %s not divisible by 2!' % self.N_c)
        return self.new_descriptor(self.N_c // 2)

    def refine(self):
        return self.new_descriptor(self.N_c * 2)

    def get_size_of_global_array(self, pad=False):
        if pad:
            return self.N_c
        return self.N_c - 1 + self.pbc_c

    def bytecount(self, dtype=float):
        return int(np.prod(self.n_c)) * np.array(1, dtype).itemsize

    def __repr__(self):
        return ('GridDescriptor(%s, cell_cv=%s, pbc_c=%s)' %
                (self.N_c.tolist(), self.cell_cv.tolist(),
                 self.pbc_c.tolist()))

    # -- arrays ----------------------------------------------------------
    def _new_array(self, n, dtype, zero, global_array, pad, xp):
        shape = (self.get_size_of_global_array(pad) if global_array
                 else self.n_c)
        if isinstance(n, numbers.Integral):
            n = (n,)
        shape = tuple(n) + tuple(int(s) for s in shape)
        if xp is None:
            xp = np
        if zero:
            return xp.zeros(shape, dtype)
        return xp.empty(shape, dtype)

    def zeros(self, n=(), dtype=float, global_array=False, pad=False, xp=np):
        return self._new_array(n, dtype, True, global_array, pad, xp)

    def empty(self, n=(), dtype=float, global_array=False, pad=False, xp=np):
        return self._new_array(n, dtype, False, global_array, pad, xp)

    def collect(self, a_xg, out=None, broadcast=False):
        if out is None:
            return a_xg
        out[:] = a_xg
        return out

    def distribute(self, B_xg, out=None):
        if out is None:
            return B_xg.copy()
        out[:] = B_xg
        return out

    def integrate(self, a_xg, b_yg=None, global_integral=True,
                  hermitian=False):
        """sum(a) dv, or <a|b> dv (gpaw/old/grid_descriptor.py:274-325)."""
        from .arrays import is_device
        if (is_device(a_xg) and b_yg is None and
                str(a_xg.dtype) == 'torch.float64'):
            from . import _lib
            from .arrays import ptr, stream_handle
            lib = _lib.load()
            a = a_xg.contiguous()
            ng = int(np.prod(a.shape[-3:]))
            flat = a.reshape(-1, ng)
            out = np.empty(len(flat))
            r = _lib.ctypes.c_double()
            for i in range(len(flat)):
                _lib.check(lib.gfd_dot(ptr(flat[i]), None, ng,
                                       _lib.ctypes.byref(r), stream_handle()),
                           'integrate')
                out[i] = r.value * self.dv
            if a.dim() == 3:
                return float(out[0])
            return out.reshape(tuple(a.shape[:-3]))
        if is_device(a_xg):
            a = a_xg.detach().cpu().numpy()
            b = None if b_yg is None else b_yg.detach().cpu().numpy()
        else:
            a, b = np.asarray(a_xg), (None if b_yg is None
                                     else np.asarray(b_yg))
        xshape = a.shape[:-3]
        if b is None:
            result = a.reshape(xshape + (-1,)).sum(axis=-1) * self.dv
            if result.ndim == 0:
                return result.item()
            return result
        A = a.reshape((-1, int(np.prod(a.shape[-3:]))))
        B = b.reshape((-1, int(np.prod(b.shape[-3:]))))
        result_yx = (B @ A.conj().T) * self.dv
        yshape = b.shape[:-3]
        result = result_yx.T.reshape(xshape + yshape)
        if result.ndim == 0:
            return result.item()
        return result

    def get_grid_point_coordinates(self, dtype=float, global_array=False):
        r_vG = np.dot(np.indices(self.n_c, dtype).T + self.beg_c,
                      self.h_cv).T.copy()
        return r_vG

    def flat_index(self, G_c):
        g1, g2, g3 = np.asarray(G_c) - self.beg_c
        return g3 + self.n_c[2] * (g2 + g1 * self.n_c[1])
