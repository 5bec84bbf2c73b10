"""The LDS-DMA row kernel (k_xdma: plain passes on real data) against the
oracle, on tiles forced through the tile override: every radius, rows per
thread 1-4, 1-4 planes in flight, x chunks that do and do not divide the
grid, periodic and zero boundaries.  Tolerance 1e-13 relative max-norm."""
import numpy as np
import pytest

import oracle
from gpaw_amd import _lib
from gpaw_amd.fd_operators import Gradient, Laplace
from gpaw_amd.grid_descriptor import GridDescriptor
from gpu_util import dev, host, kernel_path, say
from helpers import rand, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-13


def tile_ok(n, R, nseg, py, A):
    """Mirror of the kernel's tile check for 8-byte data."""
    PZ = 2
    if n[2] % PZ or n[1] % (nseg * py) or R > n[0] or R > n[1]:
        return False
    if nseg < R or A > 4:
        return False
    CPR = n[2] // PZ
    NT = CPR * nseg
    if NT < 64 or NT > 512:
        return False
    TH = nseg * py + 2 * R
    kd = -(-TH * CPR // NT)
    return kd <= 8 and (A + 1) * kd * NT * 16 <= 160 * 1024


def tile_code(n2, nseg, py, xchunk, A):
    return n2 | (nseg << 12) | (py << 24) | (xchunk << 32) | (A << 48)


def ref_apply(op, a):
    return oracle.fd_apply(op.coef_p, op.offset_pc, op.gd.n_c,
                           op.gd.neighbor_cd, a)


# the (radius, rows per thread, halo chunks per thread) instantiations of
# k_xdma (rows_dma.h); halo chunks = ceil(2 R / row segments)
XD_CASES = [(R, PY, H) for R in (1, 2, 3, 4) for PY in (1, 2, 3, 4)
            for H in (1, 2)]
BCS = [(True, True, True), (False, False, False), (True, False, True),
       (False, True, False)]


def case_grid(R, PY, H):
    """A grid shape and row-segment count that the (R, PY, H) instance
    takes: 2 tiles along y, 64 or 128 points along z."""
    for n2 in (64, 128):
        for nseg in range(1, 17):
            if -(-2 * R // nseg) != H:
                continue
            n = (R + 7, 2 * nseg * PY, n2)
            if tile_ok(n, R, nseg, PY, 1):
                return n, nseg
    raise AssertionError(f'no grid for {(R, PY, H)}')


@pytest.mark.parametrize('R,PY,H', XD_CASES)
def test_xdma_every_instance(R, PY, H):
    """Each k_xdma instantiation on a grid and tile it takes, 1 to 4
    planes in flight (where the ring fits LDS), whole and 5-plane x chunks,
    the boundary kinds in turn (one in four cases all-zero)."""
    lib = _lib.load()
    i = XD_CASES.index((R, PY, H))
    rng = np.random.default_rng(700 + i)
    n, nseg = case_grid(R, PY, H)
    pbc = BCS[i % len(BCS)]
    N = tuple(x if p else x + 1 for x, p in zip(n, pbc))
    gd = GridDescriptor(N, (3.0, 2.0, 4.0), pbc)
    assert tuple(gd.n_c) == n
    op = Laplace(gd, -0.5, R)
    a = rand(rng, (3,) + n)
    ref = ref_apply(op, a)
    for A in (1, 2, 3, 4):
        if not tile_ok(n, R, nseg, PY, A):
            assert A > 2
            continue
        for xc in (n[0], 5):
            say('xdma', (R, PY, H), n, pbc, nseg, A, xc)
            try:
                _lib.check(lib.gfd_set_tuning(
                    b'xr_tile', tile_code(n[2], nseg, PY, xc, A)), 'tuning')
                with kernel_path('row'):
                    d = dev(a)
                    out = d.new_empty(d.shape)
                    op.apply(d, out)
                    got = host(out)
            finally:
                lib.gfd_set_tuning(b'xr_tile', 0)
            assert relerr(got, ref) < TOL, (n, pbc, nseg, PY, A, xc)


CASES = [((16, 24, 32), (True, True, True)),
         ((21, 13, 64), (False, False, False)),
         ((13, 8, 96), (True, False, True)),
         ((9, 16, 49), (False, True, False))]


@pytest.mark.parametrize('nn', [1, 2, 3, 4])
def test_xdma_laplace(nn):
    """Every tile that fits each grid (rows per thread 1-4, 1-2 planes in
    flight), whole-grid and 5-plane x chunks."""
    lib = _lib.load()
    rng = np.random.default_rng(70 + nn)
    for N, pbc in CASES:
        gd = GridDescriptor(N, (3.0, 2.0, 4.0), pbc)
        n = tuple(int(x) for x in gd.n_c)
        if n[2] % 2:
            continue
        op = Laplace(gd, -0.5, nn)
        a = rand(rng, (3,) + n)
        ref = ref_apply(op, a)
        tiles = [(s, p, A) for s in range(1, 9) for p in (1, 2, 3, 4)
                 for A in (1, 2, 3) if tile_ok(n, nn, s, p, A)]
        assert tiles, (n, nn)
        for nseg, py, A in tiles:
            for xc in (n[0], 5):
                xc = min(xc, n[0])
                say('xdma', n, pbc, nn, nseg, py, A, xc)
                try:
                    _lib.check(lib.gfd_set_tuning(
                        b'xr_tile', tile_code(n[2], nseg, py, xc, A)),
                        'tuning')
                    with kernel_path('row'):
                        d = dev(a)
                        out = d.new_empty(d.shape)
                        op.apply(d, out)
                        got = host(out)
                finally:
                    lib.gfd_set_tuning(b'xr_tile', 0)
                assert relerr(got, ref) < TOL, (n, pbc, nseg, py, A, xc)


def test_xdma_gradient_and_tuner():
    """A one-axis stencil through the DMA kernel, and the tuner choosing
    between both row kernels (DMA candidates on) on a 64-band batch."""
    lib = _lib.load()
    rng = np.random.default_rng(79)
    gd = GridDescriptor((32, 32, 32), (5.0, 5.0, 5.0))
    n = tuple(int(x) for x in gd.n_c)
    a = rand(rng, (4,) + n)
    for v in range(3):
        op = Gradient(gd, v, 1.0, 3)
        ref = ref_apply(op, a)
        try:
            _lib.check(lib.gfd_set_tuning(
                b'xr_tile', tile_code(n[2], 4, 2, 16, 2)), 'tuning')
            with kernel_path('row'):
                d = dev(a)
                out = d.new_empty(d.shape)
                op.apply(d, out)
        finally:
            lib.gfd_set_tuning(b'xr_tile', 0)
        assert relerr(host(out), ref) < TOL
    import torch
    op = Laplace(gd, -0.5, 3)
    b = torch.rand((64,) + n, dtype=torch.float64, device='cuda')
    out = torch.empty_like(b)
    try:
        _lib.check(lib.gfd_set_tuning(b'xdma', 1), 'tuning')
        say('xdma tuner')
        op.apply(b, out)
        op.apply(b, out)
    finally:
        lib.gfd_set_tuning(b'xdma', 0)
    for band in (0, 63):
        assert relerr(host(out[band]), ref_apply(op, host(b[band]))) < TOL


@pytest.mark.parametrize('order', [0, 1, 2])
def test_xdma_block_orders(order):
    """The three block orders of the DMA kernel (XCD-contiguous tiles,
    linear, bands interleaved over XCDs; the last needs a multiple of 8
    grids) on 8 grids with short x chunks."""
    lib = _lib.load()
    rng = np.random.default_rng(80 + order)
    gd = GridDescriptor((16, 25, 32), (3.0, 2.0, 4.0), (True, False, True))
    n = tuple(int(x) for x in gd.n_c)
    op = Laplace(gd, -0.5, 3)
    a = rand(rng, (8,) + n)
    ref = ref_apply(op, a)
    nseg, py = next((s, p) for s in range(3, 9) for p in (1, 2)
                    if tile_ok(n, 3, s, p, 2))
    try:
        _lib.check(lib.gfd_set_tuning(
            b'xr_tile', tile_code(n[2], nseg, py, 5, 2) | (order << 52)),
            'tuning')
        with kernel_path('row'):
            say('xdma order', order, nseg, py)
            d = dev(a)
            out = d.new_empty(d.shape)
            op.apply(d, out)
            got = host(out)
    finally:
        lib.gfd_set_tuning(b'xr_tile', 0)
    assert relerr(got, ref) < TOL
