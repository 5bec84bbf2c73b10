"""ctypes binding of libgpaw_fd.so (the C ABI declared in include/gpaw_fd.h).

The library is the only compute path of this package: if it is missing, or
no GPU is visible when a kernel is called, the call raises -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libgpaw_fd.so')

GFD_OK = 0
GFD_EINVAL = 1
GFD_EHIP = 2
GFD_ENOMEM = 3
GFD_ENOTSUP = 4
GFD_ECONVERGE = 5

# every symbol of include/gpaw_fd.h with (restype, argtypes)
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)
_ip = ctypes.POINTER(ctypes.c_int)
_i = ctypes.c_int
_l = ctypes.c_long
_d = ctypes.c_double

SIGNATURES = {
    'gfd_last_error': (ctypes.c_char_p, []),
    'gfd_abi_version': (_i, []),
    'gfd_operator_new': (_i, [_i, _dp, _lp, _lp, _i, _lp, _i, _i,
                              ctypes.POINTER(_vp)]),
    'gfd_operator_free': (None, [_vp]),
    'gfd_operator_apply': (_i, [_vp, _vp, _vp, _l, _dp, _vp]),
    'gfd_operator_relax': (_i, [_vp, _i, _vp, _vp, _i, _d, _vp]),
    'gfd_operator_get_diagonal_element': (_d, [_vp]),
    'gfd_operator_residual': (_i, [_vp, _vp, _vp, _vp, _l, _dp, _dp, _vp]),
    'gfd_operator_residual_noalloc': (_i, [_vp, _vp, _vp, _vp, _l, _dp, _dp,
                                           _vp]),
    'gfd_operator_residual_noalloc_max_blocks': (_l, []),
    'gfd_gs_order': (_i, [_ip, _i, _ip, _ip]),
    'gfd_transformer_new': (_i, [_lp, _lp, _i, _lp, _lp, _lp, _lp, _i, _i,
                                 ctypes.POINTER(_vp)]),
    'gfd_transformer_free': (None, [_vp]),
    'gfd_transformer_apply': (_i, [_vp, _vp, _vp, _l, _dp, _vp]),
    'gfd_fd_precond': (_i, [_vp] * 7 + [_vp] * 8 + [_vp, _l, _d, _dp, _vp]),
    'gfd_fd_precond32': (_i, [_vp] * 7 + [_vp] * 8 + [_vp, _l, _d, _dp, _vp]),
    'gfd_poisson_new': (_i, [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                             ctypes.POINTER(_vp), _ip, _ip, _dp, _i, _d,
                             ctypes.POINTER(_vp)]),
    'gfd_poisson_free': (None, [_vp]),
    'gfd_poisson_iterate': (_i, [_vp, _vp, _vp, _dp, _vp]),
    'gfd_poisson_solve': (_i, [_vp, _vp, _vp, _d, _i, _i, _ip, _dp, _vp]),
    'gfd_poisson_set_graph': (_i, [_vp, _i]),
    'gfd_poisson_solve_mixed': (_i, [_vp, _vp, _vp, _d, _i, _i, _i, _i, _d,
                                     _d, _i, _i, _i, _ip, _dp, _vp]),
    'gfd_batch_poisson_new': (_i, [_i, _i, ctypes.POINTER(_vp),
                                   ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                   _dp, ctypes.POINTER(_vp)]),
    'gfd_batch_poisson_free': (None, [_vp]),
    'gfd_batch_poisson_set_coarse': (_i, [_vp, _ip, ctypes.POINTER(_vp),
                                          ctypes.POINTER(_vp),
                                          ctypes.POINTER(_vp), _dp]),
    'gfd_batch_poisson_level_points': (_l, [_vp, _i]),
    'gfd_batch_poisson_tune': (_i, [_vp, _vp]),
    'gfd_batch_poisson_tiles': (_i, [_vp, _i, _ip, _ip, _ip,
                                     ctypes.POINTER(_l)]),
    'gfd_batch_poisson_tile_info': (_i, [_vp, _i, ctypes.POINTER(_l), _i]),
    'gfd_coarse_cheb_bounds': (_i, [_ip, _ip, ctypes.POINTER(ctypes.c_double),
                                    ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]),
    'gfd_batch_poisson_apply': (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    'gfd_batch_poisson_transfer': (_i, [_vp, _i, _i, _i, _vp, _vp, _vp,
                                        _vp]),
    'gfd_batch_poisson_solve_mixed': (_i, [_vp, ctypes.POINTER(_vp),
                                           ctypes.POINTER(_vp), _d, _i, _i,
                                           _i, _i, _i, _ip, _dp, _vp]),
    'gfd_batch_precond32': (_i, [_vp, _dp, _vp, _vp, _vp]),
    'gfd_fft_poisson_spectrum_shape': (_l, [ctypes.POINTER(_l), _ip,
                                            ctypes.POINTER(_l)]),
    'gfd_fft_poisson_new': (_i, [ctypes.POINTER(_l), _ip, _dp,
                                 ctypes.POINTER(_vp)]),
    'gfd_fft_poisson_free': (None, [_vp]),
    'gfd_fft_poisson_solve': (_i, [_vp, _vp, _vp, _vp]),
    'gfd_woperator_new': (_i, [_i, _vp, _dp, _lp, _lp, _i, _lp, _i, _vp,
                               ctypes.POINTER(_vp)]),
    'gfd_woperator_free': (None, [_vp]),
    'gfd_woperator_apply': (_i, [_vp, _vp, _vp, _l, _dp, _vp]),
    'gfd_woperator_relax': (_i, [_vp, _i, _vp, _vp, _i, _d, _vp]),
    'gfd_woperator_get_diagonal_element': (_i, [_vp, _dp, _vp]),
    'gfd_h_residual': (_i, [_vp, _vp, _vp, _vp, _vp, _l, _dp, _vp]),
    'gfd_band_dots': (_i, [_i, _vp, _vp, _i, _l, _l, _d, _vp, _vp]),
    'gfd_rmm_lambda': (_i, [_i, _vp, _vp, _i, _l, _l, _vp, _vp]),
    'gfd_band_update': (_i, [_i, _i, _vp, _d, _vp, _vp, _vp, _vp, _i, _l, _l,
                             _vp]),
    'gfd_h_apply': (_i, [_vp, _vp, _vp, _vp, _l, _dp, _vp]),
    'gfd_band_overlap': (_i, [_vp, _i, _l, _l, _d, _vp, _i, _vp]),
    'gfd_band_matrix': (_i, [_vp, _vp, _i, _l, _l, _d, _vp, _vp]),
    'gfd_band_subspace_diag': (_i, [_vp, _i, _vp, _l, _l, _vp, _vp, _vp]),
    'gfd_band_cholesky_orth': (_i, [_vp, _i, _vp, _l, _l, _vp, _vp]),
    'gfd_dot': (_i, [_vp, _vp, _l, _dp, _vp]),
    'gfd_add_scalar': (_i, [_vp, _l, _d, _vp]),
    'gfd_stream_copy': (_i, [_vp, _vp, _l, _vp]),
    'gfd_set_tuning': (_i, [ctypes.c_char_p, _l]),
    'gfd_row_tile_candidates': (_i, [_i, _i, _lp, _i, _lp, _i]),
}

_lib = None


class GPUError(RuntimeError):
    """A call into libgpaw_fd.so failed."""

    def __init__(self, status, message):
        super().__init__(f'[gfd status {status}] {message}')
        self.status = status


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f'{LIB_PATH} is missing: build it with '
                '`python -m gpaw_amd._build` (hipcc, gfx950)')
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(status, what=''):
    if status != GFD_OK:
        msg = load().gfd_last_error().decode(errors='replace')
        if status == GFD_EINVAL:
            raise ValueError(f'{what}: {msg}')
        if status == GFD_ENOTSUP:
            raise NotImplementedError(f'{what}: {msg}')
        raise GPUError(status, f'{what}: {msg}')


def longs(x):
    """Contiguous C long array (keeps a reference via the returned tuple)."""
    import numpy as np
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int64).ravel())
    return a, a.ctypes.data_as(_lp)


def doubles(x):
    import numpy as np
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
    return a, a.ctypes.data_as(_dp)


def phases_arg(phase_cd):
    """phase_cd (3, 2) complex -> 12 doubles, or NULL."""
    import numpy as np
    if phase_cd is None:
        return None, None
    ph = np.ascontiguousarray(np.asarray(phase_cd, dtype=complex).reshape(3, 2))
    return ph, ph.view(np.float64).ctypes.data_as(_dp)
