"""gpaw_amd: GPAW's finite-difference grid hot path on AMD MI355X (gfx950).

Drop-in classes with the reference's call signatures, computing in
hand-written HIP kernels (libgpaw_fd.so, C ABI in include/gpaw_fd.h):

- ``fd_operators``: FDOperator, Laplace/GUCLaplace, Gradient, LaplaceA/B
- ``wfd_operators``: WeightedFDOperator
- ``transformers``: Transformer (restrict / interpolate, nn = 1..4)
- ``poisson``: FDPoissonSolver (Jacobi multigrid V-cycles, FP64 or FP32
  inside FP64 refinement)
- ``fft_poisson``: FastPoissonSolver (hipFFT)
- ``batch_poisson``: BatchedPoissonSolver (many uneven grids at once)
- ``preconditioner``: Preconditioner (3-level kinetic multigrid)
- ``bands``: H apply / residuals, band overlap and subspace matrices,
  Cholesky orthonormalisation
- ``rmmdiis``: the RMM-DIIS block step
- ``grid_descriptor``: GridDescriptor (single-domain layout contract)
- ``workloads``: the synthetic inputs of the measured configurations

Device arrays are torch tensors (ROCm) in HBM; numpy arrays are staged.
"""
__version__ = '0.1.0'

Bohr = 0.52917721067  # Angstrom (ASE's value, which the reference uses)

from .grid_descriptor import GridDescriptor  # noqa: E402

__all__ = ['GridDescriptor', 'Bohr', '__version__']
