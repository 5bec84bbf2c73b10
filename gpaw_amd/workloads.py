"""Seeded synthetic inputs for the FD grid-path workloads.

Lengths are in bohr (1 bohr = 0.52917721067 Angstrom).  Densities are sums
of normalised Gaussians q (pi s^2)^(-3/2) exp(-r^2 / s^2) with
minimum-image distances on periodic axes.

Workloads (sizes as named):
  laplace_48     8 real grids 48^3, cell 9.6 A cubic, periodic, U[-1, 1]
  poisson_48     48^3 Poisson, 4 Gaussians, widths U[0.6, 1.2], q = +-1
  bands_96       256 real grids 96^3 (h ~ 0.2 A), U[-1, 1]
  density_96     4 densities on 96^3, 32 positive Gaussians, U[0.5, 1.5]
  poisson_192    192^3, cell 38.4 A, 64 Gaussians q = +-U[0.5, 2], widths
                 U[0.5, 1.5], uniform neutralising background
  uneven_48      48 independent grids, sides from {40, 48, ..., 160},
                 alternately periodic / zero boundary, 8-64 Gaussians each
  eigen_128      1024 complex bands on 128^3 (h = 0.18 A), 4 Bloch
                 k-points, V = 128 Gaussian wells (depth U[0.5, 3] Ha,
                 widths U[0.8, 1.6]), bands N(0, 1)

Large grids use gaussian_sum, which builds each Gaussian as the outer
product of three 1-D factors (equal to gaussian_density up to rounding).
"""
from __future__ import annotations

import numpy as np

from . import Bohr
from .grid_descriptor import GridDescriptor


def gaussian_density(gd, centres_v, widths, charges):
    """Sum of normalised Gaussians on the grid of gd (host array)."""
    r_vG = gd.get_grid_point_coordinates()
    cell = gd.cell_cv.diagonal()
    rho = np.zeros(tuple(gd.n_c))
    for c_v, s, q in zip(centres_v, widths, charges):
        d2 = np.zeros(tuple(gd.n_c))
        for v in range(3):
            d = r_vG[v] - c_v[v]
            if gd.pbc_c[v]:
                d -= cell[v] * np.round(d / cell[v])
            d2 += d * d
        rho += q * (np.pi * s * s)**-1.5 * np.exp(-d2 / (s * s))
    return rho


def gaussian_sum(gd, centres_v, widths, amplitudes):
    """sum_i a_i exp(-|r - c_i|^2 / s_i^2) on the grid of gd (orthorhombic
    cell), as outer products of per-axis factors."""
    n = tuple(gd.n_c)
    out = np.zeros(n)
    h = gd.h_cv.diagonal()
    cell = gd.cell_cv.diagonal()
    x_v = [(np.arange(n[v]) + gd.beg_c[v]) * h[v] for v in range(3)]
    for c_v, s, a in zip(centres_v, widths, amplitudes):
        f = []
        for v in range(3):
            d = x_v[v] - c_v[v]
            if gd.pbc_c[v]:
                d -= cell[v] * np.round(d / cell[v])
            f.append(np.exp(-d * d / (s * s)))
        out += a * (f[0][:, None, None] * f[1][None, :, None]) * \
            f[2][None, None, :]
    return out


def uniform_bands(nbands, N, seed, dtype=float):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, (nbands, N, N, N))
    if dtype == complex:
        a = a + 1j * rng.uniform(-1.0, 1.0, (nbands, N, N, N))
    return a


def laplace_48(seed=1001):
    L = 9.6 / Bohr
    gd = GridDescriptor((48, 48, 48), (L, L, L), True)
    return gd, uniform_bands(8, 48, seed)


def poisson_48(seed=1002):
    L = 9.6 / Bohr
    gd = GridDescriptor((48, 48, 48), (L, L, L), True)
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, L, (4, 3))
    widths = rng.uniform(0.6, 1.2, 4)
    charges = np.array([1.0, -1.0, 1.0, -1.0])
    return gd, gaussian_density(gd, centres, widths, charges)


def bands_96_gd():
    L = 96 * 0.2 / Bohr
    return GridDescriptor((96, 96, 96), (L, L, L), True)


def density_96(seed=1004, ndens=4):
    gd = bands_96_gd()
    rng = np.random.default_rng(seed)
    L = gd.cell_cv[0, 0]
    out = np.empty((ndens, 96, 96, 96))
    for i in range(ndens):
        centres = rng.uniform(0.0, L, (32, 3))
        widths = rng.uniform(0.5, 1.5, 32)
        out[i] = gaussian_density(gd, centres, widths, np.ones(32))
    return gd, out


def poisson_192(seed=1005):
    L = 38.4 / Bohr
    gd = GridDescriptor((192, 192, 192), (L, L, L), True)
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, L, (64, 3))
    widths = rng.uniform(0.5, 1.5, 64)
    charges = rng.uniform(0.5, 2.0, 64) * rng.choice([-1.0, 1.0], 64)
    rho = gaussian_sum(gd, centres, widths,
                       charges * (np.pi * widths**2)**-1.5)
    rho -= rho.mean()
    return gd, rho


def uneven_48(seed=1007, nproblems=48):
    """[(gd, rho)]: sides from {40, 48, ..., 160}; even problems periodic
    (rho neutralised), odd ones zero boundary."""
    rng = np.random.default_rng(seed)
    sides = np.arange(40, 161, 8)
    out = []
    for i in range(nproblems):
        N = tuple(int(x) for x in rng.choice(sides, 3))
        h = 0.2 / Bohr
        cell = tuple(h * x for x in N)
        pbc = (i % 2 == 0)
        gd = GridDescriptor(N, cell, pbc)
        nc = int(rng.integers(8, 65))
        centres = rng.uniform(0.0, 1.0, (nc, 3)) * np.array(cell)
        widths = rng.uniform(0.5, 1.5, nc)
        charges = rng.uniform(0.5, 2.0, nc) * rng.choice([-1.0, 1.0], nc)
        rho = gaussian_sum(gd, centres, widths,
                           charges * (np.pi * widths**2)**-1.5)
        if pbc:
            rho -= rho.mean()
        out.append((gd, rho))
    return out


def eigen_128_gd():
    L = 128 * 0.18 / Bohr
    return GridDescriptor((128, 128, 128), (L, L, L), True)


def eigen_128_potential(gd, seed=1006, nwells=128):
    """V(r) = -sum_i d_i exp(-|r - c_i|^2 / s_i^2), d ~ U[0.5, 3] Ha,
    s ~ U[0.8, 1.6] bohr."""
    rng = np.random.default_rng(seed)
    L = gd.cell_cv.diagonal()
    centres = rng.uniform(0.0, 1.0, (nwells, 3)) * L
    depth = rng.uniform(0.5, 3.0, nwells)
    widths = rng.uniform(0.8, 1.6, nwells)
    return gaussian_sum(gd, centres, widths, -depth)


def eigen_128_kpoints(seed=1006, nk=4):
    """Bloch phases phase_cd[c] = [exp(2 pi i k_c), exp(-2 pi i k_c)] of nk
    k-points drawn uniformly in the Brillouin zone, shape (nk, 3, 2)."""
    rng = np.random.default_rng(seed + 1)
    k = rng.uniform(-0.5, 0.5, (nk, 3))
    return np.exp(2j * np.pi * k[:, :, None] * np.array([1.0, -1.0]))
