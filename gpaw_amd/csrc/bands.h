// Band contractions (overlap / subspace matrices, Cholesky orthonormalisation).
#pragma once

#include "gfd_common.h"

namespace gfd {

// Grow-only device work space of the band calls.  It lives as long as the
// process (no destructor: the HIP runtime may be gone at static teardown).
// The C ABI keeps one per stream, so only calls queued on one stream share
// it, and those run in order.
struct BandWork {
    void* partial = nullptr;
    size_t partial_bytes = 0;
    void* blas = nullptr;  // rocblas_handle
    void* info = nullptr;
    int ensure_partial(size_t bytes, hipStream_t s);
};

// out[i][j] = alpha (conj ? conj : id)(sum_k A[i][k] conj(B[j][k])) for
// i < M, j < N (only 128 x 128 tiles with tile_j <= tile_i if lower_only).
int gram(BandWork& w, const double2* A, long lda, const double2* B, long ldb,
         int M, int N, long K, double alpha, int conj, int lower_only,
         double2* out, long ldo, hipStream_t s);

// out[n] = dv Re sum_x conj(A_n(x)) B_n(x), dtype 0 f64 / 1 c128, rows of
// npts points at stride ld; out a device array of nbands doubles.
int band_dots(int dtype, const void* A, const void* B, int nbands, long npts,
              long ld, double dv, double* out, hipStream_t s);

// lam[n] = -Re<dR_n|R_n> / <dR_n|dR_n> (0 where dR_n = 0)
int rmm_lambda(int dtype, const void* dR, const void* R, int nbands,
               long npts, long ld, double* lam, hipStream_t s);

// c_n = scale alpha[n]; mode 0: y_n += c_n x_n; mode 1 also
// w_n <- v_n + c_n w_n in the same pass.
int band_update(int dtype, int mode, const double* alpha, double scale,
                const void* x, void* y, const void* v, void* w, int nbands,
                long npts, long ld, hipStream_t s);

// Lower Cholesky L of S (row-major, lower part used), C = L^-1 in S, and
// out_i = sum_j C_ij psi_j for the n rows of psi (X points each).
int cholesky_orth(BandWork& w, double2* S, int n, const double2* psi,
                  long ldpsi, long X, double2* out, long ldout, hipStream_t s);

// Subspace diagonalisation: H (row-major n x n Hermitian, lower part
// read) -> eigenvalues eps (device, ascending) and H[m][k] = v_m[k] with
// H v_m = eps_m v_m; out_m = sum_k v_m[k] psi_k.  Synchronous.
int subspace_diag(BandWork& w, double2* H, int n, const double2* psi,
                  long ldpsi, long X, double2* out, long ldout, double* eps,
                  hipStream_t s);

}  // namespace gfd
