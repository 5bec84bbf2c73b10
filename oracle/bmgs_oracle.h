/* CPU oracle for the GPAW finite-difference grid path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that the HIP kernels are compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it; the product library never links or calls it.
 *
 * Conventions shared by every entry point
 *   - grids are C-order (n0, n1, n2) with n2 contiguous; complex data is
 *     interleaved (re, im) doubles; `nin` grids are stored back to back;
 *   - nb[6] = neighbor_cd flattened: >= 0 means the axis side wraps onto this
 *     (single) domain, -1 means a zero boundary;
 *   - phases[12] = phase_cd (3 x 2 complex) flattened, or NULL for real data;
 *   - return value 0 on success, nonzero on bad arguments.
 */
#ifndef GPAW_FD_ORACLE_H
#define GPAW_FD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* FDOperator.apply: out = sum_p coefs[p] * in(x + off3[p]) with halos. */
int oracle_fd_apply(int ncoefs, const double* coefs, const long* off3,
                    const long n[3], int range, const long nb[6], int real,
                    int cfd, const double* phases, const double* in,
                    double* out, long nin, int nthreads);

/* FDOperator.relax: nrelax sweeps of Jacobi (method 2) or Gauss-Seidel
 * (method 1) for operator(f) = src; f is updated in place.  Real only. */
int oracle_fd_relax(int method, int ncoefs, const double* coefs,
                    const long* off3, const long n[3], int range,
                    const long nb[6], int cfd, double* f, const double* src,
                    int nrelax, double w);

/* WeightedFDOperator.apply (c/bmgs/wfd.cpp): out = sum_w weights[w](x) *
 * sum_p coefs_w[p] in(x + off_w[p]); stencil w has ncoefs[w] points (the
 * first at the centre), coefs/off3 concatenated over w, weights nweights
 * grids of n0 n1 n2 doubles. */
int oracle_wfd_apply(int nweights, const int* ncoefs, const double* coefs,
                     const long* off3, const double* weights, const long n[3],
                     int range, const long nb[6], int real, int cfd,
                     const double* phases, const double* in, double* out,
                     long nin);

/* WeightedFDOperator.relax (c/bmgs/wrelax.cpp): Jacobi (method 2) or
 * Gauss-Seidel (method 1) sweeps, real data, f updated in place. */
int oracle_wfd_relax(int method, int nweights, const int* ncoefs,
                     const double* coefs, const long* off3,
                     const double* weights, const long n[3], int range,
                     const long nb[6], int cfd, double* f, const double* src,
                     int nrelax, double w);

/* Transformer.apply: restriction (interpolate == 0) or interpolation with a
 * k-tap (k = 2 nn) stencil; pad/npad/skip are the (3, 2) arrays the Python
 * front end computes. */
int oracle_transform(int interpolate, int k, const long size_in[3],
                     const long size_out[3], const long pad[6],
                     const long npad[6], const long skip[6], const long nb[6],
                     int real, const double* phases, const double* in,
                     double* out, long nin, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
