/* C ABI of the gfx950 finite-difference grid-path library (libgpaw_fd.so).
 *
 * Drop-in boundary for the GPAW FD grid hot path.  Every entry point below
 * replaces one binding of the reference's C extension `_gpaw` (method table
 * c/_gpaw.h:182-194 and the GPU methods in c/gpu/operators.cpp,
 * c/gpu/transformers.cpp); each declaration cites the function it replaces.
 *
 * Conventions
 *   - Plain C types only.  Array arguments are device pointers (HBM) to
 *     C-order grids (n0, n1, n2), n2 contiguous, `nin` grids back to back;
 *     complex data is interleaved (re, im) double.
 *   - `phases` is the reference's phase_cd, a (3, 2) complex array given as
 *     12 doubles [axis][side][re, im]; NULL for real data.
 *   - `stream` is a hipStream_t (NULL: the null stream).  Calls are
 *     asynchronous on that stream unless noted; the caller synchronises.
 *   - Every function returns a status (GFD_OK == 0); on failure
 *     gfd_last_error() describes it.  The library never aborts the process.
 *   - Handles own device memory for their scratch; one handle must not be
 *     used from two streams at once.
 */
#ifndef GPAW_FD_H
#define GPAW_FD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GFD_OK = 0,
    GFD_EINVAL = 1,    /* bad argument (reference: TypeError / assert) */
    GFD_EHIP = 2,      /* HIP runtime error */
    GFD_ENOMEM = 3,    /* device allocation failed */
    GFD_ENOTSUP = 4,   /* combination not supported on the GPU */
    GFD_ECONVERGE = 5  /* reference: PoissonConvergenceError */
};

/* Message for the last failing call on this thread. */
const char* gfd_last_error(void);

/* ABI version of this header. */
int gfd_abi_version(void);

typedef struct gfd_operator gfd_operator;
typedef struct gfd_transformer gfd_transformer;
typedef struct gfd_poisson gfd_poisson;

/* ---------------------------------------------------------------- stencils
 * Replaces _gpaw.Operator(coefs, offsets, size, range, neighbors, real,
 * comm, cfd[, use_gpu]) -- c/operators.cpp:379-420, bound from
 * gpaw/fd_operators.py:91-93.  `offsets` are the flattened offsets with the
 * padded strides ((n1+2r)(n2+2r), n2+2r, 1) of gpaw/fd_operators.py:54-63;
 * neighbors is neighbor_cd (3, 2): >= 0 periodic (single domain), -1 zero.
 * comm must be serial (single GPU); cfd is recomputed from the offsets. */
int gfd_operator_new(int ncoefs, const double* coefs, const long* offsets,
                     const long size[3], int range, const long neighbors[6],
                     int real, int cfd, gfd_operator** out);
void gfd_operator_free(gfd_operator* op);

/* Replaces Operator.apply_gpu(in, out, shape, dtype[, phases]) --
 * c/gpu/operators.cpp:344-385 (CPU: Operator.apply, c/operators.cpp:245-316).
 * nin = number of grids.  in and out must not overlap. */
int gfd_operator_apply(gfd_operator* op, const void* in, void* out, long nin,
                       const double* phases, void* stream);

/* Replaces Operator.relax_gpu(method, f, src, nrelax[, w]) --
 * c/gpu/operators.cpp:200-221 (CPU: Operator.relax, c/operators.cpp:50-92).
 * method 2 = weighted Jacobi (c/bmgs/relax.cpp:41-65); method 1 =
 * lexicographic Gauss-Seidel (c/bmgs/relax.cpp:12-40), bit for bit with the
 * CPU sweep (the reference GPU kernel has no method 1,
 * c/gpu/kernels/relax.cpp:72-80).  Real data only. */
int gfd_operator_relax(gfd_operator* op, int method, double* f,
                       const double* src, int nrelax, double w, void* stream);

/* No counterpart in the reference (host only, no device calls): the step
 * weights of the Gauss-Seidel sweep of gfd_operator_relax(method 1).  The
 * sweep updates point (i0, i1, i2) at step a0 i0 + a1 i1 + i2; the
 * weights make every stencil offset (o0, o1, o2 triples, npoints of them)
 * that precedes the centre in C order lower the step, and every one that
 * follows it raise the step, so each step reads exactly the values the
 * lexicographic sweep of c/bmgs/relax.cpp:12-40 reads.  GFD_ENOTSUP when
 * no such weights exist. */
int gfd_gs_order(const int* offsets, int npoints, int* a0, int* a1);

/* Replaces Operator.get_diagonal_element() -- c/operators.cpp:319-332. */
double gfd_operator_get_diagonal_element(const gfd_operator* op);

/* Fused residual r = op(in) - b with sum |r|^2 written to *norm2 (host,
 * synchronous when norm2 != NULL).  Replaces the pair
 * `op.apply(phi, r); r -= rho; dot(r, r)` of gpaw/poisson.py:527-532. */
int gfd_operator_residual(gfd_operator* op, const void* in, const void* b,
                          void* r, long nin, const double* phases,
                          double* norm2, void* stream);

/* No counterpart in the reference.  gfd_operator_residual for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): same arguments and results, but
 * the norm's partial sums live in a fixed per-GPU device array of 1 MiB
 * instead of buffers the operator grows, and the tile tuner is not run (a
 * grid it has not timed yet gets the heuristic tile).  Returns GFD_EINVAL
 * when the pass would launch more blocks than
 * gfd_operator_residual_noalloc_max_blocks() (split the batch).  Calls
 * that overlap in time on one GPU share that array and race. */
int gfd_operator_residual_noalloc(gfd_operator* op, const void* in,
                                  const void* b, void* r, long nin,
                                  const double* phases, double* norm2,
                                  void* stream);
long gfd_operator_residual_noalloc_max_blocks(void);

/* ---------------------------------------------------------------- transfers
 * Replaces _gpaw.Transformer(size_in, size_out, k, pad, npad, skip,
 * neighbors, real, comm, interpolate[, use_gpu]) --
 * c/transformers.cpp:196-246, bound from gpaw/transformers.py:71-77. */
int gfd_transformer_new(const long size_in[3], const long size_out[3], int k,
                        const long pad[6], const long npad[6],
                        const long skip[6], const long neighbors[6], int real,
                        int interpolate, gfd_transformer** out);
void gfd_transformer_free(gfd_transformer* t);

/* Replaces Transformer.apply_gpu(in, out, shape, dtype[, phases]) --
 * c/gpu/transformers.cpp:189-220 (CPU: c/transformers.cpp:112-158). */
int gfd_transformer_apply(gfd_transformer* t, const void* in, void* out,
                          long nin, const double* phases, void* stream);

/* ---------------------------------------------------------------- precond
 * Replaces _gpaw.fd_precond(rest0, rest1, intp1, intp2, kin0, kin1, kin2,
 * d0, q0, r1, d1, q1, r2, d2, q2, residuals, nresiduals, step[, phases]) --
 * c/fd_preconditioner.cpp:106-351, bound from gpaw/preconditioner.py:68-72.
 * The negated residuals are not needed (the sign is folded into the first
 * restriction).  Result in d0; q0, r1, d1, q1, r2, d2, q2 are work arrays of
 * nin grids each on the fine, first and second coarse levels. */
int gfd_fd_precond(gfd_transformer* rest0, gfd_transformer* rest1,
                   gfd_transformer* intp1, gfd_transformer* intp2,
                   gfd_operator* kin0, gfd_operator* kin1, gfd_operator* kin2,
                   void* d0, void* q0, void* r1, void* d1, void* q1, void* r2,
                   void* d2, void* q2, const void* residuals, long nin,
                   double step, const double* phases, void* stream);

/* gfd_fd_precond with the multigrid chain in FP32: the work arrays q0, r1,
 * d1, q1, r2, d2, q2 are float32 (complex64 for complex objects) of the
 * same shapes; residuals and d0 stay float64 / complex128.  Every pass
 * computes and accumulates in FP32.  nn=1 transformers only.  No
 * counterpart in the reference (its preconditioner is FP64 only). */
int gfd_fd_precond32(gfd_transformer* rest0, gfd_transformer* rest1,
                     gfd_transformer* intp1, gfd_transformer* intp2,
                     gfd_operator* kin0, gfd_operator* kin1,
                     gfd_operator* kin2, void* d0, void* q0, void* r1,
                     void* d1, void* q1, void* r2, void* d2, void* q2,
                     const void* residuals, long nin, double step,
                     const double* phases, void* stream);

/* ---------------------------------------------------------------- Poisson
 * Native multigrid driver of FDPoissonSolver (gpaw/poisson.py:321-537):
 * ops[0..nlevels] are the Laplacians of levels 0..L, restrictors[l] /
 * interpolators[l] map level l <-> l+1.  Level buffers are allocated here
 * (gpaw/poisson.py:451-472). */
int gfd_poisson_new(int nlevels, gfd_operator* const* ops,
                    gfd_transformer* const* restrictors,
                    gfd_transformer* const* interpolators,
                    const int* presmooths, const int* postsmooths,
                    const double* weights, int relax_method, double dv,
                    gfd_poisson** out);
void gfd_poisson_free(gfd_poisson* p);

/* One V-cycle, FDPoissonSolver.iterate2(step) at level 0
 * (gpaw/poisson.py:500-537); *error = dv sum r^2 (synchronous). */
int gfd_poisson_iterate(gfd_poisson* p, double* phi, const double* rho,
                        double* error, void* stream);

/* The loop of FDPoissonSolver.solve_neutral (gpaw/poisson.py:474-498):
 * iterate while error > eps and niter < maxiter; returns GFD_ECONVERGE when
 * niter reaches maxiter.  remove_mean: subtract the average of phi
 * (fully periodic grids, :492-496).  Synchronous. */
int gfd_poisson_solve(gfd_poisson* p, double* phi, const double* rho,
                      double eps, int maxiter, int remove_mean, int* niter,
                      double* error, void* stream);

/* Replay each V-cycle (and each mixed-precision refinement step) as one
 * captured hipGraph on a library stream (on: default) or launch eagerly
 * (no reference counterpart).  Results are identical either way. */
int gfd_poisson_set_graph(gfd_poisson* p, int on);

/* Mixed-precision solve (no reference counterpart; the FP64 equations are
 * those of FDPoissonSolver): FP64 iterative refinement on the level-0
 * residual r = L phi - rho, each step solving L e = r with `inner` FP32
 * V-cycles over `levels` levels (pre / post Jacobi sweeps per level,
 * coarse_sweeps at the coarsest, weights w0 at level 0 and wc below; -1
 * keeps the FDPoissonSolver value; in the LDS-resident coarse solve the
 * coarsest level's coarse_sweeps are Chebyshev steps, see "coarse_cheb")
 * and phi -= e; stops when
 * |r| / |rho| <= tol_rel.  Returns GFD_ECONVERGE after max_outer steps.
 * inner_fp64 != 0 runs the same inner cycles (levels, sweeps, weights,
 * LDS coarse solve) on FP64 correction arrays instead, so that the two
 * differ in precision only.  Synchronous; *relres = final |r| / |rho|. */
int gfd_poisson_solve_mixed(gfd_poisson* p, double* phi, const double* rho,
                            double tol_rel, int levels, int pre, int post,
                            int coarse_sweeps, double w0, double wc,
                            int inner, int max_outer, int inner_fp64,
                            int* n_outer, double* relres, void* stream);

/* --------------------------------------------------------- batched grids
 * Many independent real Poisson problems of uneven shape (no counterpart
 * in the reference, which solves one grid at a time; the per-grid algebra
 * is FDPoissonSolver's, gpaw/poisson.py:366-537).  Grid g, level l of the
 * FP32 hierarchy: ops[g * L + l] (real axis-only stencils, radius <= 3),
 * rst / itp[g * (L - 1) + l] (real nn=1 transformers), weights[l] the
 * Jacobi weight of level l.  Level arrays are the grids' arrays
 * concatenated in grid order. */
typedef struct gfd_batch_poisson gfd_batch_poisson;

int gfd_batch_poisson_new(int G, int L, gfd_operator** ops,
                          gfd_transformer** rst, gfd_transformer** itp,
                          const double* weights, gfd_batch_poisson** out);
void gfd_batch_poisson_free(gfd_batch_poisson* b);

/* Run each grid's levels L-1 .. nlev[g]-1 (its whole remaining hierarchy)
 * in one LDS-resident workgroup per grid in place of the batched coarsest
 * level.  ops / rst / itp: every grid's full hierarchy, grid after grid
 * (nlev[g] operators and nlev[g] - 1 transformers each); weights[l]: the
 * Jacobi weight of level l.  nlev == NULL switches it off. */
int gfd_batch_poisson_set_coarse(gfd_batch_poisson* b, const int* nlev,
                                 gfd_operator** ops, gfd_transformer** rst,
                                 gfd_transformer** itp,
                                 const double* weights);

/* Points of all grids at one level (-1 for a bad level). */
long gfd_batch_poisson_level_points(gfd_batch_poisson* b, int level);

/* Time candidate tile shapes of both batched stencil kernels (register
 * tiles, x-march tiles) per level on this batch and keep the fastest (done
 * by the first apply / transfer / solve if never called). */
int gfd_batch_poisson_tune(gfd_batch_poisson* b, void* stream);

/* Chosen register tile of a level (-1: the FP64 fine-level residual):
 * threads along z and y, z points per thread, resident blocks (0: one
 * block per tile); 0, 0, 0 and the tile count when the level runs on
 * x-march tiles. */
int gfd_batch_poisson_tiles(gfd_batch_poisson* b, int level, int* bz,
                            int* by, int* pz, long* blocks);

/* Chosen stencil kernel and tile of a level (-1: the FP64 fine-level
 * residual), ninfo values of: kernel (0 register tiles, 1 x-march tiles),
 * x planes per tile, tiles, then for register tiles threads along z and y,
 * z points per thread, resident blocks (0: one block per tile), and for
 * x-march tiles threads per block, rows per thread, widest z part of a
 * row, planes of loads in flight; values past those are 0. */
int gfd_batch_poisson_tile_info(gfd_batch_poisson* b, int level, long* info,
                                int ninfo);

/* out = L_level in on every grid, float64 (fp32 = 0) or float32 level
 * arrays; one launch. */
int gfd_batch_poisson_apply(gfd_batch_poisson* b, int level, int fp32,
                            const void* in, void* out, void* stream);

/* Restriction level -> level + 1 (interpolate = 0), or
 * out = base - I(in) from level + 1 to level; one launch. */
int gfd_batch_poisson_transfer(gfd_batch_poisson* b, int level,
                               int interpolate, int fp32, const void* in,
                               void* out, const void* base, void* stream);

/* FP32 V-cycles (pre / post Jacobi sweeps, coarse_sweeps on the coarsest
 * level: Chebyshev steps in the LDS-resident coarse solve, see
 * "coarse_cheb") inside FP64 refinement until |L phi_g - rho_g| / |rho_g| <=
 * tol_rel for every grid; phis[g] in/out and rhos[g] are float64 device
 * arrays of grid g's fine shape.  use_graph replays one captured
 * refinement step.  relres[g] receives the final relative residuals;
 * GFD_ECONVERGE after max_outer steps. */
int gfd_batch_poisson_solve_mixed(gfd_batch_poisson* b, double* const* phis,
                                  const double* const* rhos, double tol_rel,
                                  int pre, int post, int coarse_sweeps,
                                  int max_outer, int use_graph, int* n_outer,
                                  double* relres, void* stream);

/* The FP32 kinetic preconditioner chain of gfd_fd_precond32 (reference:
 * fd_precond, c/fd_preconditioner.cpp:221-345, one grid per call) on every
 * grid of a batch at once.  b: a batch of exactly 3 levels built from
 * each grid's kinetic operators (K0 on the grid, nn=1 on its two
 * coarsenings) and nn=1 transformers; steps: host array, one
 * 0.66666666 / diag(K0) per grid (gpaw/preconditioner.py:19); res, out:
 * FP64 device arrays of every grid's fine values back to back (batch
 * order), not overlapping.  Eight launches, FP32 work arrays owned by the
 * batch; asynchronous on stream. */
int gfd_batch_precond32(gfd_batch_poisson* b, const double* steps,
                        const double* res, double* out, void* stream);

/* ------------------------------------------------------- FFT Poisson
 * FastPoissonSolver (gpaw/poisson.py:846-1028): the FD operator is
 * diagonal in the Fourier basis of periodic axes and the DST-I basis of
 * zero-boundary axes.  n: grid points (interior points on zero axes);
 * zero_bc[c]: axis c has zero boundaries (odd extension to 2 (n + 1)
 * points); inv_lambda: host array of the inverse eigenvalues on the
 * extended grid's real-to-complex half spectrum, C order, shape
 * (m0, m1, m2/2 + 1) as gfd_fft_poisson_spectrum_shape returns (0 where
 * the eigenvalue vanishes).  hipFFT plans and work arrays are made here. */
typedef struct gfd_fft_poisson gfd_fft_poisson;

long gfd_fft_poisson_spectrum_shape(const long n[3], const int zero_bc[3],
                                    long m_out[3]);
int gfd_fft_poisson_new(const long n[3], const int zero_bc[3],
                        const double* inv_lambda, gfd_fft_poisson** out);
void gfd_fft_poisson_free(gfd_fft_poisson* f);

/* phi = L^-1 rho on the grid (float64 device arrays of shape n). */
int gfd_fft_poisson_solve(gfd_fft_poisson* f, const double* rho,
                          double* phi, void* stream);

/* ---------------------------------------------------------------- bands
 * FD eigensolver step (complex128 bands, row-major nbands x npts with row
 * stride ld points). */

/* hpsi = kin(psi) + vpot psi for nbands grids: FDHamiltonian
 * .apply_local_potential (gpaw/new/fd/hamiltonian.py:23-32); vpot is one
 * real grid shared by all bands.  Real or complex operator. */
int gfd_h_apply(gfd_operator* kin, const double* vpot, const void* psi,
                void* hpsi, long nbands, const double* phases, void* stream);

/* Residuals r_n = (kin + vpot - eps[n]) psi_n of the grid part of
 * calculate_residuals (gpaw/new/pwfd/eigensolver.py:197-227, the
 * `axpy(-e, p, r)` after apply_local_potential) in one stencil pass; eps is
 * a device array of nbands doubles. */
int gfd_h_residual(gfd_operator* kin, const double* vpot, const double* eps,
                   const void* psi, void* r, long nbands,
                   const double* phases, void* stream);

/* out[n] = dv Re sum_x conj(a_n(x)) b_n(x) for nbands rows of npts points
 * (stride ld), float64 (is_complex 0) or complex128; out is a device array
 * of nbands doubles.  The band energies <psi_n|H|psi_n> of the eigensolver
 * step (gpaw/new/pwfd/eigensolver.py). */
int gfd_band_dots(int is_complex, const void* a, const void* b, int nbands,
                  long npts, long ld, double dv, double* out, void* stream);

/* RMM step length per band, lam[n] = -Re<dR_n|R_n> / <dR_n|dR_n> (0 where
 * dR_n vanishes): the a_n / b_n of block_step
 * (gpaw/new/pwfd/rmmdiis.py:150-155) in one pass over dR and R. */
int gfd_rmm_lambda(int is_complex, const void* dR, const void* R, int nbands,
                   long npts, long ld, double* lam, void* stream);

/* Per-band updates with c_n = scale * alpha[n] (alpha: device array):
 * mode 0: y_n += c_n x_n (rmmdiis.py:252-257, the final preconditioned
 * step); mode 1: y_n += c_n x_n and w_n <- v_n + c_n w_n in one pass
 * (rmmdiis.py:157-162: psi + lambda PR and R + lambda dR).  w must not
 * alias y. */
int gfd_band_update(int is_complex, int mode, const double* alpha,
                    double scale, const void* x, void* y, const void* v,
                    void* w, int nbands, long npts, long ld, void* stream);

/* S[i][j] = dv sum_x psi_i(x) conj(psi_j(x)), row-major nbands x nbands:
 * the matrix_elements(psit, psit, cc=True) of gpaw/core/arrays.py:183-262
 * used by orthonormalize (gpaw/new/pwfd/wave_functions.py:223-226).  With
 * lower_only only the 128 x 128 tiles on or below the diagonal are
 * written.  FP64 MFMA. */
int gfd_band_overlap(const void* psi, int nbands, long npts, long ld,
                     double dv, void* S, int lower_only, void* stream);

/* H[i][j] = dv sum_x conj(psi_i(x)) hpsi_j(x): matrix_elements(psit,
 * Htpsit) (gpaw/core/arrays.py:183-262).  FP64 MFMA. */
int gfd_band_matrix(const void* psi, const void* hpsi, int nbands, long npts,
                    long ld, double dv, void* H, void* stream);

/* Matrix.invcholesky (gpaw/core/matrix.py:382-417) and the rotation of
 * orthonormalize (gpaw/new/pwfd/wave_functions.py:228-238): L = lower
 * Cholesky factor of S (lower part used), S <- C = L^-1, out_i =
 * sum_j C_ij psi_j.  out must not alias psi.  Synchronous (checks that S
 * is positive definite). */
int gfd_band_cholesky_orth(void* S, int nbands, const void* psi, long npts,
                           long ld, void* out, void* stream);

/* Subspace diagonalisation and canonical rotation, grid part
 * (gpaw/new/pwfd/wave_functions.py:289-368: subspace_eigenvalues, then
 * canonical_transformation).  H: row-major nbands x nbands Hermitian
 * matrix H[i][j] = <psi_i|H|psi_j> (gfd_band_matrix; its lower part is
 * read, as Matrix.eigh reads the lower half, gpaw/core/matrix.py:419-460).
 * On return eps (device, nbands doubles) holds the eigenvalues in
 * ascending order, H[m][k] = v_m[k] with H v_m = eps_m v_m -- the content
 * of the reference's H_nn after its complex_conjugate -- and
 * out_m = sum_k v_m[k] psi_k (dense FP64 MFMA rotation).  rocSOLVER zheevd;
 * out must not alias psi.  Synchronous (checks convergence). */
int gfd_band_subspace_diag(void* H, int nbands, const void* psi, long npts,
                           long ld, void* out, double* eps, void* stream);

/* ---------------------------------------------------------------- misc */

/* *result = sum_i a_i b_i over n doubles (b == NULL: sum_i a_i); the
 * numpy dot / sum of gpaw/poisson.py:493,531.  Synchronous. */
int gfd_dot(const double* a, const double* b, long n, double* result,
            void* stream);

/* a_i += v for n doubles (the `phi -= phi_ave` of gpaw/poisson.py:496). */
int gfd_add_scalar(double* a, long n, double v, void* stream);

/* Device-to-device copy of nbytes with 16-byte accesses (no reference
 * counterpart): the measured streaming bandwidth that stencil throughput is
 * compared against.  16-byte aligned pointers, nbytes a multiple of 16. */
int gfd_stream_copy(const void* src, void* dst, long nbytes, void* stream);

/* ------------------------------------------------------ weighted operators
 * Replaces _gpaw.WOperator(nweights, weights, coefs, offsets, size, range,
 * neighbors, real, comm, cfd) -- c/woperators.cpp:422-477, bound from
 * gpaw/wfd_operators.py:63-79 (set_weights).  Stencil w has ncoefs[w]
 * points, its first at the centre; coefs and offsets (padded strides, as
 * gfd_operator_new) are concatenated over w; weights is a host array of
 * nweights device pointers to n0 n1 n2 doubles each, borrowed (not copied)
 * for the life of the handle, as the reference borrows them. */
typedef struct gfd_woperator gfd_woperator;
int gfd_woperator_new(int nweights, const int* ncoefs, const double* coefs,
                      const long* offsets, const long size[3], int range,
                      const long neighbors[6], int real,
                      const double* const* weights, gfd_woperator** out);
void gfd_woperator_free(gfd_woperator* op);

/* WOperator.apply(in, out[, phases]) -- c/woperators.cpp:212-283 with
 * c/bmgs/wfd.cpp: out = sum_w weights_w * (stencil_w applied to in). */
int gfd_woperator_apply(gfd_woperator* op, const void* in, void* out,
                        long nin, const double* phases, void* stream);

/* WOperator.relax(method, f, src, nrelax[, w]) -- c/woperators.cpp:55-94
 * with c/bmgs/wrelax.cpp.  Weighted Jacobi (method 2); Gauss-Seidel
 * (method 1) returns GFD_ENOTSUP, as gfd_operator_relax.  Real only. */
int gfd_woperator_relax(gfd_woperator* op, int method, double* f,
                        const double* src, int nrelax, double w,
                        void* stream);

/* WOperator.get_diagonal_element() -- c/woperators.cpp:341-378: max over
 * the grid of |sum_w weight_w c_{w,0}| (synchronous, *out on the host). */
int gfd_woperator_get_diagonal_element(gfd_woperator* op, double* out,
                                       void* stream);

/* Tuning knobs (no reference counterpart): key "force_path" selects the
 * stencil kernel family (0 automatic, 1 general, 2 tiled x-march, 3
 * full-row x-march; a family that cannot take a grid falls back to the
 * general kernel); "force_general" (0/1) is force_path 1/0; "xm_tile" and
 * "xr_tile" with value tz | nseg << 12 | py << 24 | xchunk << 32 override
 * the tile of the tiled / full-row x-march kernel (value 0 clears); for
 * "xr_tile", dma << 48 (1-4 planes in flight) selects the LDS-DMA variant
 * for plain passes on real data and order << 52 its block order (0
 * XCD-contiguous tiles, 1 linear, 2 bands interleaved over the XCDs);
 * "autotune" (0/1, default 1) times candidate full-row tiles on the first
 * launch of each shape (off: heuristic tiles) and, per transfer shape, the
 * x-marching kernel against the other transfer kernels (off: x-marching);
 * "xdma" (0/1, default 1)
 * lets it time the LDS-DMA variant too; "xr_nt" (-1 automatic, else
 * bit 0 non-temporal stores, bit 1 non-temporal loads);
 * "coarse_lds" (0/1, default 1) runs the coarse levels of the Poisson
 * V-cycles that fit in one workgroup's LDS as a single launch;
 * "coarse_max_points" (default and 0: 4096) caps the points of its
 * first level; "coarse_cheb" (0/1, default 1) makes the coarsest level of
 * that solve in the mixed-precision schedules (single grid and batched)
 * run Chebyshev iteration on D^-1 A (eigenvalue bounds from the radius-1
 * stencil's Fourier / sine modes) instead of Jacobi sweeps; the reference
 * schedule keeps Jacobi;
 * "xfer_passes" (0/1) runs K = 2 transfers as three 1-D passes instead of
 * the fused kernel; "xfer_march" (0/1, default 1) lets transfers use the
 * x-marching fused kernel (off: one point per thread for K = 2, three 1-D
 * passes for K > 2); "xfer_march_scale" (>= 1, default 1) sets its tiles
 * per resident block to about 4 x value; "wx_march" (0/1, default 1) runs
 * weighted operators whose stencils are all axis-only (at most 4 of them)
 * through the weighted x-march kernel (off: one point per thread);
 * "wx_tile" (py | xchunk << 8, 0 automatic) overrides its rows per thread
 * (1, 2, 4) and x planes per block; "batch_kernel" (0 timed choice, 1
 * register tiles, 2 x-march tiles) limits the candidates of the batched
 * stencil tuner (takes effect at the next tune); "batch_tile" (nt | py <<
 * 12 | zmax << 16 | xc << 28 | pf << 40, 0 timed choice) makes every level
 * tuned afterwards use that x-march tile: nt threads per block (64 .. 512,
 * multiple of 64), py rows per thread (1, 2, 4), z parts of at most zmax
 * columns, xc x planes per tile, pf planes of loads in flight (1, 2);
 * "batch_xfer" (0 timed choice, 1 one output point per thread, 2 x-march
 * tiles, or a tile code nt | py << 12 | zmax << 16 | xc << 28 with nt >= 64
 * forcing that x-march tile) sets the batched transfers of batches tuned
 * afterwards; "precond_fine" (0/1/2, default 1) runs the fine level of
 * the preconditioner chains (gfd_fd_precond32 and gfd_fd_precond, real
 * or complex) as one fused pass where the shape fits (0: interpolation and
 * update as two passes;
 * 2: fused, with non-temporal stores of the result);
 * "precond_fine_tile" (py | nys << 8 | xc << 16, 0 automatic) forces its
 * tile: py rows per thread (1, 2), nys thread rows, xc x planes per block
 * (a forced tile that does not fit the grid makes the call fail).
 * Returns GFD_EINVAL for unknown keys or values. */
int gfd_set_tuning(const char* key, long value);

/* Eigenvalue bounds [lo, hi] that the LDS-resident coarse solve's
 * Chebyshev iteration uses on a level with the radius-1 axis stencil
 * c0 + cp[c] (e_{+c} + e_{-c}) on an n[0] x n[1] x n[2] grid (periodic[c]
 * or zero boundary): the extreme nonzero eigenvalues of D^-1 A, 0.9 x the
 * lowest and 1.05 x the highest.  No counterpart in the reference (the
 * reference smooths the coarsest level with Jacobi); exported so that the
 * bounds can be checked against dense eigenvalues.  GFD_ENOTSUP when no
 * bounds apply. */
int gfd_coarse_cheb_bounds(const int* n, const int* periodic,
                           const double* cp, double c0, double* lo,
                           double* hi);

/* The full-row tiles the tuner times for a stencil radius, data type
 * (0 f64, 1 c128, 2 f32, 3 c64) and grid shape, besides its heuristic one:
 * up to max_codes of them are written to codes in the "xr_tile" encoding
 * (tz | nseg << 12 | py << 24 | xchunk << 32 | dma << 48 | order << 52 |
 * wide << 56); with_dma adds the LDS-DMA tiles.  Returns how many there
 * are (-1 on bad arguments).  No reference counterpart: it lets tests run
 * every tile the tuner may pick. */
int gfd_row_tile_candidates(int radius, int dtype, const long n[3],
                            int with_dma, long* codes, int max_codes);

#ifdef __cplusplus
}
#endif
#endif /* GPAW_FD_H */
