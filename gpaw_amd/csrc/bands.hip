// Band contractions of the FD eigensolver step on gfx950 (FP64 MFMA).
//
//  * k_gram: G = A B^H for row-major complex A (M x K) and B (N x K) with a
//    long inner dimension (K = grid points): 128 x 128 output tiles, split
//    over K into chunks, 512 threads (8 waves, 2 x 4 MFMA tiles of 16 x 16
//    per wave); A and B stream through LDS in 16-point stages as separate
//    re / im planes, and every complex 16 x 16 x 4 step is four
//    v_mfma_f64_16x16x4f64 (re.re + im.im, im.re - re.im).  Chunk partials
//    are reduced in a fixed order (bit-reproducible).
//  * the Cholesky factor and its inverse come from rocSOLVER (zpotrf,
//    ztrtri) on the column-major view of the row-major matrix, and the
//    rotation Psi <- C Psi is k_rot (FP64 MFMA, the lower triangle only);
//    the subspace eigenvectors come from rocSOLVER zheevd and rotate Psi
//    through the dense form of k_rot.
#include "bands.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>

namespace gfd {

constexpr int GT = 128;  // output tile
constexpr int GK = 16;   // inner points per LDS stage
// LDS row padding (doubles): rows of the k-major stage arrays are 1 KB
// apart, so the four k rows read by one MFMA operand load fell on the same
// banks (measured: the 1024-band overlap 6 % faster with 2 doubles of pad)
constexpr int GP = GT + 2;

typedef double __attribute__((ext_vector_type(4))) d4;

// consecutive logical blocks (the band tiles of one point tile share its
// psi columns) on one XCD; hardware deals blocks to the 8 XCDs in turn
__device__ inline int xcd_remap_rot(int h, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int x = h & 7, k = h >> 3;
    return x * q + min(x, r) + k;
}

struct GramArgs {
    const double2* A;
    const double2* B;
    long lda, ldb;
    int M, N;
    long K;
    long kchunk;
    int nchunks;
    int ntiles;
    int tcols;           // tile columns (full matrix)
    int lower;           // tiles (i, j <= i) in row order
    double2* partial;    // [chunk][tile][GT][GT]
};

// Output tile t -> (tile row, tile column).
__device__ inline int2 tile_of(int t, int lower, int tcols)
{
    if (!lower) return make_int2(t / tcols, t % tcols);
    int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) i++;
    while (i * (i + 1) / 2 > t) i--;
    return make_int2(i, t - i * (i + 1) / 2);
}

__global__ void __launch_bounds__(512) k_gram(GramArgs a)
{
    __shared__ double sAr[GK][GP], sAi[GK][GP], sBr[GK][GP], sBi[GK][GP];
    const int tile = blockIdx.x % a.ntiles;
    const int chunk = blockIdx.x / a.ntiles;
    const int2 tt = tile_of(tile, a.lower, a.tcols);
    const int i0 = tt.x * GT, j0 = tt.y * GT;
    const long kb = (long)chunk * a.kchunk;
    const long ke = min(kb + a.kchunk, a.K);
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wr = w >> 1, wc = w & 1;  // 4 x 2 waves: 32 rows x 64 cols
    const int li = lane & 15, lk = lane >> 4;

    d4 accr[2][4], acci[2][4];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            accr[r][c] = d4{0.0, 0.0, 0.0, 0.0};
            acci[r][c] = d4{0.0, 0.0, 0.0, 0.0};
        }

    // stage loader: 128 rows x 16 points of A and of B, 4 rows per thread
    // (16 lanes read 16 consecutive points of a row: 256-byte segments,
    // conflict-free LDS writes); the next stage's values are fetched into
    // registers while the current stage is multiplied out of LDS
    const int lk16 = t & 15, lr = t >> 4;
    double2 va[4], vb[4];
    auto fetch = [&](long ks) {
        const long k = ks + lk16;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = lr + 32 * q;
            va[q] = make_double2(0.0, 0.0);
            vb[q] = make_double2(0.0, 0.0);
            if (k < ke) {
                if (i0 + row < a.M) va[q] = a.A[(long)(i0 + row) * a.lda + k];
                if (j0 + row < a.N) vb[q] = a.B[(long)(j0 + row) * a.ldb + k];
            }
        }
    };
    fetch(kb);
    for (long ks = kb; ks < ke; ks += GK) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            sAr[lk16][lr + 32 * q] = va[q].x;
            sAi[lk16][lr + 32 * q] = va[q].y;
            sBr[lk16][lr + 32 * q] = vb[q].x;
            sBi[lk16][lr + 32 * q] = vb[q].y;
        }
        __syncthreads();
        if (ks + GK < ke) fetch(ks + GK);
#pragma unroll
        for (int kk = 0; kk < GK / 4; kk++) {
            const int kl = kk * 4 + lk;
            double ar[2], ai[2], br[4], bi[4], nbi[4];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                ar[r] = sAr[kl][wr * 32 + r * 16 + li];
                ai[r] = sAi[kl][wr * 32 + r * 16 + li];
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
                br[c] = sBr[kl][wc * 64 + c * 16 + li];
                bi[c] = sBi[kl][wc * 64 + c * 16 + li];
                nbi[c] = -bi[c];
            }
#pragma unroll
            for (int r = 0; r < 2; r++)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    accr[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        ar[r], br[c], accr[r][c], 0, 0, 0);
                    accr[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        ai[r], bi[c], accr[r][c], 0, 0, 0);
                    acci[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        ai[r], br[c], acci[r][c], 0, 0, 0);
                    acci[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        ar[r], nbi[c], acci[r][c], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    // D register q of lane l holds row l / 16 + 4 q, column l % 16
    double2* out = a.partial + ((long)chunk * a.ntiles + tile) * GT * GT;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = wr * 32 + r * 16 + lk + 4 * q;
                const int col = wc * 64 + c * 16 + li;
                out[row * GT + col] = make_double2(accr[r][c][q],
                                                   acci[r][c][q]);
            }
}

// Rotation out[i][x] = sum_j C[i][j] psi[j][x] (row-major n x n C, psi n
// rows of X points): with `lower` C is lower triangular (entries above the
// diagonal are ignored, K = bands up to the tile's last row), else dense.
// 128 bands x 128 points per block, the same MFMA tiling as k_gram (A = C,
// B = psi without conjugation).
struct RotArgs {
    const double2* C;
    int n;
    const double2* psi;
    long ldpsi;
    long X;
    double2* out;
    long ldout;
    int nti;   // band tiles
    int lower; // C lower triangular (Cholesky), else dense (eigenvectors)
};

typedef double RotStage[GK][GP];

// One 128-band x 128-point output tile (band tile bt, point tile xt) of the
// rotation.  LOWER: K runs to the tile's last row, and an MFMA whose 16
// rows x 4 k of C lie wholly above the diagonal (all zero) is skipped.
template <bool LOWER>
__device__ inline void rot_tile(const RotArgs& a, int bt, long xt,
                                RotStage& sCr, RotStage& sCi, RotStage& sPr,
                                RotStage& sPi)
{
    const int i0 = bt * GT;
    const long x0 = xt * GT;
    const int kend = LOWER ? min(a.n, i0 + GT) : a.n;
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wr = w >> 1, wc = w & 1;
    const int li = lane & 15, lk = lane >> 4;

    d4 accr[2][4], acci[2][4];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            accr[r][c] = d4{0.0, 0.0, 0.0, 0.0};
            acci[r][c] = d4{0.0, 0.0, 0.0, 0.0};
        }
    // C tile: 128 rows x 16 columns, one column of 4 rows per thread; psi
    // tile: 16 rows x 128 points, one point of 4 rows per thread
    // (consecutive lanes on consecutive columns / points: conflict-free LDS
    // writes)
    const int ck = t & 15, cr = t >> 4;
    const int pp = t & 127, pr = t >> 7;
    double2 vc[4], vp[4];
    auto fetch = [&](int ks) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int gi = i0 + cr + 32 * q;
            const int j = ks + ck;
            vc[q] = (gi < a.n && j < kend && (!LOWER || j <= gi))
                        ? a.C[(long)gi * a.n + j]
                        : make_double2(0.0, 0.0);
            const int jp = ks + pr + 4 * q;
            const long x = x0 + pp;
            vp[q] = (jp < kend && x < a.X) ? a.psi[(long)jp * a.ldpsi + x]
                                           : make_double2(0.0, 0.0);
        }
    };
    fetch(0);
    for (int ks = 0; ks < kend; ks += GK) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            sCr[ck][cr + 32 * q] = vc[q].x;
            sCi[ck][cr + 32 * q] = vc[q].y;
            sPr[pr + 4 * q][pp] = vp[q].x;
            sPi[pr + 4 * q][pp] = vp[q].y;
        }
        __syncthreads();
        if (ks + GK < kend) fetch(ks + GK);
#pragma unroll
        for (int kk = 0; kk < GK / 4; kk++) {
            const int kl = kk * 4 + lk;
            double ar[2], ai[2], nai[2], br[4], bi[4];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                ar[r] = sCr[kl][wr * 32 + r * 16 + li];
                ai[r] = sCi[kl][wr * 32 + r * 16 + li];
                nai[r] = -ai[r];
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
                br[c] = sPr[kl][wc * 64 + c * 16 + li];
                bi[c] = sPi[kl][wc * 64 + c * 16 + li];
            }
#pragma unroll
            for (int r = 0; r < 2; r++) {
                // rows i0 + wr 32 + r 16 .. +15 against k = ks + 4 kk .. +3
                if (LOWER && ks + 4 * kk > i0 + wr * 32 + r * 16 + 15)
                    continue;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    accr[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        ar[r], br[c], accr[r][c], 0, 0, 0);
                    accr[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        nai[r], bi[c], accr[r][c], 0, 0, 0);
                    acci[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        ar[r], bi[c], acci[r][c], 0, 0, 0);
                    acci[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        ai[r], br[c], acci[r][c], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = i0 + wr * 32 + r * 16 + lk + 4 * q;
                const long col = x0 + wc * 64 + c * 16 + li;
                if (row < a.n && col < a.X)
                    a.out[(long)row * a.ldout + col] =
                        make_double2(accr[r][c][q], acci[r][c][q]);
            }
}

// Dense: one block per (band tile, point tile).  LOWER: a block computes
// band tiles p and nti - 1 - p of its point tile one after the other, so
// that every block runs the same number of K stages (the tiles of a point
// tile stay in step and share its psi rows in L2).
template <bool LOWER>
__global__ void __launch_bounds__(512) k_rot(RotArgs a)
{
    __shared__ double sCr[GK][GP], sCi[GK][GP], sPr[GK][GP], sPi[GK][GP];
    const int L = xcd_remap_rot(blockIdx.x, gridDim.x);
    if (!LOWER) {
        rot_tile<false>(a, L % a.nti, L / a.nti, sCr, sCi, sPr, sPi);
        return;
    }
    const int npair = (a.nti + 1) / 2;
    const int p = L % npair;
    const long xt = L / npair;
    rot_tile<true>(a, p, xt, sCr, sCi, sPr, sPi);
    if (a.nti - 1 - p != p) {
        __syncthreads();  // the stage arrays are reused
        rot_tile<true>(a, a.nti - 1 - p, xt, sCr, sCi, sPr, sPi);
    }
}

// out[i][j] = alpha (conj?) sum_chunks partial, for every computed tile
__global__ void __launch_bounds__(256)
k_gram_reduce(const double2* __restrict__ partial, int nchunks, int ntiles,
              int lower, int tcols, int M, int N, double alpha, int conj,
              double2* __restrict__ out, long ldo)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const long per = (long)GT * GT;
    if (e >= per * ntiles) return;
    const int tile = (int)(e / per);
    const int rc = (int)(e % per);
    const int2 tt = tile_of(tile, lower, tcols);
    const int row = tt.x * GT + rc / GT;
    const int col = tt.y * GT + rc % GT;
    if (row >= M || col >= N) return;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < nchunks; c++) {
        const double2 v = partial[((long)c * ntiles + tile) * per + rc];
        sr += v.x;
        si += v.y;
    }
    out[(long)row * ldo + col] =
        make_double2(alpha * sr, conj ? -alpha * si : alpha * si);
}

// out[n] = dv Re sum_x conj(a_n(x)) b_n(x): one block per band, a fixed
// order of partial sums (bit-reproducible)
template <typename T>
__global__ void __launch_bounds__(1024)
k_band_dot(const T* __restrict__ A, const T* __restrict__ B, long npts,
           long ld, double dv, double* __restrict__ out)
{
    const T* a = A + (long)blockIdx.x * ld;
    const T* b = B + (long)blockIdx.x * ld;
    double s = 0.0;
    for (long x = threadIdx.x; x < npts; x += 1024) {
        if constexpr (sizeof(T) == 16) {
            const double2 u = *reinterpret_cast<const double2*>(a + x);
            const double2 v = *reinterpret_cast<const double2*>(b + x);
            s = fma(u.x, v.x, s);
            s = fma(u.y, v.y, s);
        } else {
            s = fma((double)a[x], (double)b[x], s);
        }
    }
    __shared__ double red[1024];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = dv * red[0];
}

int band_dots(int dtype, const void* A, const void* B, int nbands, long npts,
              long ld, double dv, double* out, hipStream_t s)
{
    if (nbands <= 0) return OK;
    if (dtype == 1)
        hipLaunchKernelGGL(k_band_dot<double2>, dim3(nbands), dim3(1024), 0,
                           s, (const double2*)A, (const double2*)B, npts, ld,
                           dv, out);
    else
        hipLaunchKernelGGL(k_band_dot<double>, dim3(nbands), dim3(1024), 0,
                           s, (const double*)A, (const double*)B, npts, ld,
                           dv, out);
    GFD_LAUNCH_CHECK();
    return OK;
}

// Two real parts of band inner products at once: s0 = Re sum conj(a) b,
// s1 = Re sum conj(a) a.
template <typename T>
__device__ inline void dot2(const T* a, const T* b, long x, double& s0,
                            double& s1)
{
    if constexpr (sizeof(T) == 16) {
        const double2 u = *reinterpret_cast<const double2*>(a + x);
        const double2 v = *reinterpret_cast<const double2*>(b + x);
        s0 = fma(u.x, v.x, fma(u.y, v.y, s0));
        s1 = fma(u.x, u.x, fma(u.y, u.y, s1));
    } else {
        const double u = (double)a[x], v = (double)b[x];
        s0 = fma(u, v, s0);
        s1 = fma(u, u, s1);
    }
}

// RMM step length per band: lambda_n = -Re<dR_n|R_n> / <dR_n|dR_n>
// (the dv factors cancel); one block per band, fixed summation order.
template <typename T>
__global__ void __launch_bounds__(1024)
k_rmm_lambda(const T* __restrict__ dR, const T* __restrict__ R, long npts,
             long ld, double* __restrict__ lam)
{
    const T* a = dR + (long)blockIdx.x * ld;
    const T* b = R + (long)blockIdx.x * ld;
    double s0 = 0.0, s1 = 0.0;
    for (long x = threadIdx.x; x < npts; x += 1024) dot2(a, b, x, s0, s1);
    __shared__ double r0[1024], r1[1024];
    r0[threadIdx.x] = s0;
    r1[threadIdx.x] = s1;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r0[threadIdx.x] += r0[threadIdx.x + o];
            r1[threadIdx.x] += r1[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        lam[blockIdx.x] = r1[0] > 0.0 ? -r0[0] / r1[0] : 0.0;
}

int rmm_lambda(int dtype, const void* dR, const void* R, int nbands,
               long npts, long ld, double* lam, hipStream_t s)
{
    if (nbands <= 0) return OK;
    if (dtype == 1)
        hipLaunchKernelGGL(k_rmm_lambda<double2>, dim3(nbands), dim3(1024),
                           0, s, (const double2*)dR, (const double2*)R, npts,
                           ld, lam);
    else
        hipLaunchKernelGGL(k_rmm_lambda<double>, dim3(nbands), dim3(1024),
                           0, s, (const double*)dR, (const double*)R, npts,
                           ld, lam);
    GFD_LAUNCH_CHECK();
    return OK;
}

// Per-band updates with a device coefficient c_n = scale * alpha[n]:
//   mode 0:  y_n += c_n x_n
//   mode 1:  y_n += c_n x_n  and  w_n = v_n + c_n w_n   (both RMM updates:
//            psi += lambda PR, dR <- R + lambda dR, one pass)
struct BandUpd {
    const double* alpha;
    double scale;
    long npts, ld;
    int nbands, mode;
    const void* x;
    void* y;
    const void* v;
    void* w;
};

template <typename T>
__global__ void __launch_bounds__(256) k_band_update(BandUpd u)
{
    const long x = (long)blockIdx.x * 256 + threadIdx.x;
    if (x >= u.npts) return;
    for (int n = blockIdx.y; n < u.nbands; n += gridDim.y) {
        const double c = u.scale * u.alpha[n];
        const long i = (long)n * u.ld + x;
        T* y = (T*)u.y;
        const T* xs = (const T*)u.x;
        T yv = y[i];
        const T xv = xs[i];
        if constexpr (sizeof(T) == 16) {
            yv.x = fma(c, xv.x, yv.x);
            yv.y = fma(c, xv.y, yv.y);
        } else {
            yv = fma(c, xv, yv);
        }
        y[i] = yv;
        if (u.mode == 1) {
            T* w = (T*)u.w;
            const T* vs = (const T*)u.v;
            T wv = w[i];
            const T vv = vs[i];
            if constexpr (sizeof(T) == 16) {
                wv.x = fma(c, wv.x, vv.x);
                wv.y = fma(c, wv.y, vv.y);
            } else {
                wv = fma(c, wv, vv);
            }
            w[i] = wv;
        }
    }
}

int band_update(int dtype, int mode, const double* alpha, double scale,
                const void* x, void* y, const void* v, void* w, int nbands,
                long npts, long ld, hipStream_t s)
{
    if (nbands <= 0 || npts <= 0) return OK;
    BandUpd u{alpha, scale, npts, ld, nbands, mode, x, y, v, w};
    const dim3 grid((unsigned)((npts + 255) / 256),
                    (unsigned)(nbands < 65535 ? nbands : 65535));
    if (dtype == 1)
        hipLaunchKernelGGL(k_band_update<double2>, grid, dim3(256), 0, s, u);
    else
        hipLaunchKernelGGL(k_band_update<double>, grid, dim3(256), 0, s, u);
    GFD_LAUNCH_CHECK();
    return OK;
}

int gram(BandWork& w, const double2* A, long lda, const double2* B, long ldb,
         int M, int N, long K, double alpha, int conj, int lower_only,
         double2* out, long ldo, hipStream_t s)
{
    const int TM = (M + GT - 1) / GT, TN = (N + GT - 1) / GT;
    if (lower_only && TM != TN) {
        set_error("gram: lower_only needs a square matrix");
        return EINVAL_;
    }
    const int nt = lower_only ? TM * (TM + 1) / 2 : TM * TN;
    if (nt == 0) return OK;
    if (K <= 0) {
        set_error("gram: no points");
        return EINVAL_;
    }
    // chunks: about 9 blocks per CU (256 CUs), at least 4 stages each
    long chunks = std::max(1L, (long)(2304 + nt - 1) / nt);
    chunks = std::min(chunks, std::max(1L, K / (4 * GK)));
    long kchunk = (K + chunks - 1) / chunks;
    kchunk = (kchunk + GK - 1) / GK * GK;
    chunks = (K + kchunk - 1) / kchunk;
    const size_t pbytes = sizeof(double2) * (size_t)GT * GT * nt * chunks;
    GFD_TRY(w.ensure_partial(pbytes, s));
    GramArgs g;
    g.A = A;
    g.B = B;
    g.lda = lda;
    g.ldb = ldb;
    g.M = M;
    g.N = N;
    g.K = K;
    g.kchunk = kchunk;
    g.nchunks = (int)chunks;
    g.ntiles = nt;
    g.tcols = TN;
    g.lower = lower_only;
    g.partial = (double2*)w.partial;
    hipLaunchKernelGGL(k_gram, dim3((unsigned)(nt * chunks)), dim3(512), 0, s,
                       g);
    GFD_LAUNCH_CHECK();
    const long total = (long)GT * GT * nt;
    hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)((total + 255) / 256)),
                       dim3(256), 0, s, (const double2*)w.partial,
                       (int)chunks, nt, lower_only, TN, M, N, alpha, conj,
                       out, ldo);
    GFD_LAUNCH_CHECK();
    return OK;
}

int BandWork::ensure_partial(size_t bytes, hipStream_t s)
{
    if (bytes <= partial_bytes) return OK;
    if (partial) {
        GFD_CHECK_HIP(hipStreamSynchronize(s));
        GFD_CHECK_HIP(hipFree(partial));
        partial = nullptr;
        partial_bytes = 0;
    }
    if (hipMalloc(&partial, bytes) != hipSuccess) {
        (void)hipGetLastError();
        partial = nullptr;
        set_error("hipMalloc of %zu bytes failed", bytes);
        return ENOMEM_;
    }
    partial_bytes = bytes;
    return OK;
}

static int blas_ok(rocblas_status st, const char* what)
{
    if (st != rocblas_status_success) {
        set_error("%s failed: %s", what, rocblas_status_to_string(st));
        return EHIP;
    }
    return OK;
}

// the work space's rocBLAS handle (created on first use) on stream s
static int blas_handle(BandWork& w, hipStream_t s, rocblas_handle* out)
{
    if (!w.blas) {
        rocblas_handle h;
        GFD_TRY(blas_ok(rocblas_create_handle(&h), "rocblas_create_handle"));
        w.blas = h;
        GFD_CHECK_HIP(hipMalloc(&w.info, sizeof(int)));
    }
    *out = (rocblas_handle)w.blas;
    return blas_ok(rocblas_set_stream(*out, s), "rocblas_set_stream");
}

static int rotate(const double2* C, int n, int lower, const double2* psi,
                  long ldpsi, long X, double2* out, long ldout, hipStream_t s)
{
    RotArgs r;
    r.C = C;
    r.n = n;
    r.psi = psi;
    r.ldpsi = ldpsi;
    r.X = X;
    r.out = out;
    r.ldout = ldout;
    r.nti = (n + GT - 1) / GT;
    r.lower = lower;
    const long per = lower ? (r.nti + 1) / 2 : r.nti;
    const long nb = per * ((X + GT - 1) / GT);
    if (nb > 0x7fffffffL) {
        set_error("band rotation: too many blocks");
        return ENOTSUP_;
    }
    if (lower)
        hipLaunchKernelGGL(k_rot<true>, dim3((unsigned)nb), dim3(512), 0, s,
                           r);
    else
        hipLaunchKernelGGL(k_rot<false>, dim3((unsigned)nb), dim3(512), 0, s,
                           r);
    GFD_LAUNCH_CHECK();
    return OK;
}

__global__ void __launch_bounds__(256)
k_conj(double2* __restrict__ a, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long)gridDim.x * 256)
        a[i].y = -a[i].y;
}

int subspace_diag(BandWork& w, double2* H, int n, const double2* psi,
                  long ldpsi, long X, double2* out, long ldout, double* eps,
                  hipStream_t s)
{
    rocblas_handle h;
    GFD_TRY(blas_handle(w, s, &h));
    // off-diagonal work of the tridiagonal reduction
    GFD_TRY(w.ensure_partial(sizeof(double) * (size_t)n, s));
    // the column-major view of row-major H is H^T = conj(H); its upper
    // triangle is the lower triangle of H, the part the reference reads
    // (scipy eigh lower=True)
    GFD_TRY(blas_ok(rocsolver_zheevd(h, rocblas_evect_original,
                                     rocblas_fill_upper, n,
                                     (rocblas_double_complex*)H, n, eps,
                                     (double*)w.partial,
                                     (rocblas_int*)w.info),
                    "rocsolver_zheevd"));
    int info = 0;
    GFD_CHECK_HIP(hipMemcpyAsync(&info, w.info, sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    GFD_CHECK_HIP(hipStreamSynchronize(s));
    if (info != 0) {
        set_error("subspace eigensolver did not converge (zheevd info %d)",
                  info);
        return ECONVERGE;
    }
    // column m of the view holds an eigenvector of conj(H), i.e. conj(v_m)
    // with H v_m = eps_m v_m: row m of H is conj(v_m); conjugate to v_m
    const long nn = (long)n * n;
    const unsigned g = (unsigned)std::min<long>((nn + 255) / 256, 4096);
    hipLaunchKernelGGL(k_conj, dim3(g), dim3(256), 0, s, H, nn);
    GFD_LAUNCH_CHECK();
    return rotate(H, n, 0, psi, ldpsi, X, out, ldout, s);
}

int cholesky_orth(BandWork& w, double2* S, int n, const double2* psi,
                  long ldpsi, long X, double2* out, long ldout, hipStream_t s)
{
    rocblas_handle h;
    GFD_TRY(blas_handle(w, s, &h));
    // column-major view of the row-major lower triangle = upper triangle
    GFD_TRY(blas_ok(rocsolver_zpotrf(h, rocblas_fill_upper, n,
                                     (rocblas_double_complex*)S, n,
                                     (rocblas_int*)w.info),
                    "rocsolver_zpotrf"));
    int info = 0;
    GFD_CHECK_HIP(hipMemcpyAsync(&info, w.info, sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    GFD_CHECK_HIP(hipStreamSynchronize(s));
    if (info != 0) {
        set_error("overlap matrix is not positive definite (zpotrf info %d)",
                  info);
        return EINVAL_;
    }
    GFD_TRY(blas_ok(rocsolver_ztrtri(h, rocblas_fill_upper,
                                     rocblas_diagonal_non_unit, n,
                                     (rocblas_double_complex*)S, n,
                                     (rocblas_int*)w.info),
                    "rocsolver_ztrtri"));
    // psi_i <- sum_{j <= i} C_ij psi_j (C = L^-1 in the row-major lower
    // triangle of S)
    return rotate(S, n, 1, psi, ldpsi, X, out, ldout, s);
}

}  // namespace gfd
