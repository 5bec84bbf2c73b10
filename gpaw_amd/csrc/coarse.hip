// LDS-resident coarse V-cycle (see coarse.h).
//
// The coarse levels of a multigrid hierarchy hold few points (24^3 FP32 =
// 54 KB) but cost a dozen launches each when run like the fine levels;
// here one workgroup of 1024 threads keeps every coarse level's arrays in
// LDS and runs the down-sweep (Jacobi pre-smoothing, residual, restriction),
// the coarsest-level sweeps and the up-sweep (interpolation + correction,
// post-smoothing) with barriers between passes.  The arithmetic per point
// is that of the global kernels: Jacobi (1 - w) e + w (r - sum_nb) / c0,
// residual (c0 e + sum_nb) - r, and the K = 2 transfer weights in the
// same association order (z innermost, then y, then x).
#include "../../include/gpaw_fd.h"
#include "coarse.h"

#include <algorithm>
#include <cmath>

#include "stencil_dev.h"

namespace gfd {

// index along an axis with periodic wrap, -1 outside a zero boundary
__device__ inline int cw(int i, int n, int periodic)
{
    if (i >= 0 && i < n) return i;
    if (!periodic) return -1;
    i %= n;
    if (i < 0) i += n;
    return i;
}

template <typename T>
__device__ inline T cat(const T* a, int idx)
{
    return idx >= 0 ? a[idx] : zero_of<T>();
}

// Threads own columns (i1, i2) of a level and march along x; when a plane
// has fewer columns than threads, the x range is split between `parts`
// thread groups.
struct CWork {
    int q0, qstep;   // first column, column stride
    int x0, x1;      // x range
};

__device__ inline CWork cwork(int plane, int n0)
{
    CWork w;
    const int nt = blockDim.x, t = threadIdx.x;
    const int parts = plane >= nt ? 1 : min(nt / plane, n0);
    if (t >= parts * (plane >= nt ? nt : plane)) {
        w.q0 = plane;  // idle
        w.qstep = 1;
        w.x0 = w.x1 = 0;
        return w;
    }
    if (plane >= nt) {
        w.q0 = t;
        w.qstep = nt;
        w.x0 = 0;
        w.x1 = n0;
    } else {
        const int part = t / plane;
        w.q0 = t % plane;
        w.qstep = plane;  // one column
        w.x0 = part * n0 / parts;
        w.x1 = (part + 1) * n0 / parts;
    }
    return w;
}

// plane offsets of the z and y neighbours of column q (-1: zero boundary)
struct CCol {
    int zo[2][CMAXR];
    int yo[2][CMAXR];
};

__device__ inline void ccol(const CLevel& L, int q, CCol& c)
{
    const int n2 = L.n[2];
    const int i2 = q % n2, i1 = q / n2;
#pragma unroll
    for (int k = 1; k <= CMAXR; k++) {
#pragma unroll
        for (int sg = 0; sg < 2; sg++) {
            const int d = sg ? -k : k;
            const int j = cw(i2 + d, n2, L.periodic[2]);
            c.zo[sg][k - 1] = j < 0 ? -1 : i1 * n2 + j;
            const int y = cw(i1 + d, L.n[1], L.periodic[1]);
            c.yo[sg][k - 1] = y < 0 ? -1 : y * n2 + i2;
        }
    }
}

// off-centre stencil terms at (i0, column q): z, then y, then x
template <typename T>
__device__ inline T cneigh(const T* e, const CLevel& L, const CCol& c,
                           int q, int i0, int plane)
{
    typedef typename ScalarOf<T>::type S;
    const T* row = e + i0 * plane;
    T acc = zero_of<T>();
    if (L.has[2]) {
#pragma unroll
        for (int k = 1; k <= CMAXR; k++) {
            if (k > L.R) break;
            acc = fmac(acc, (S)L.cp[2][k - 1], cat(row, c.zo[0][k - 1]));
            acc = fmac(acc, (S)L.cm[2][k - 1], cat(row, c.zo[1][k - 1]));
        }
    }
    if (L.has[1]) {
#pragma unroll
        for (int k = 1; k <= CMAXR; k++) {
            if (k > L.R) break;
            acc = fmac(acc, (S)L.cp[1][k - 1], cat(row, c.yo[0][k - 1]));
            acc = fmac(acc, (S)L.cm[1][k - 1], cat(row, c.yo[1][k - 1]));
        }
    }
    if (L.has[0]) {
#pragma unroll
        for (int k = 1; k <= CMAXR; k++) {
            if (k > L.R) break;
            const int xp = cw(i0 + k, L.n[0], L.periodic[0]);
            const int xm = cw(i0 - k, L.n[0], L.periodic[0]);
            acc = fmac(acc, (S)L.cp[0][k - 1],
                       cat(e, xp < 0 ? -1 : xp * plane + q));
            acc = fmac(acc, (S)L.cm[0][k - 1],
                       cat(e, xm < 0 ? -1 : xm * plane + q));
        }
    }
    return acc;
}

// `sweeps` Jacobi sweeps on level L; bit `lbit` of cur says which of A / B
// holds the current iterate.
template <typename T>
__device__ void crelax(T* sm, const CLevel& L, const T* rl, int sweeps,
                       unsigned& cur, int lbit)
{
    typedef typename ScalarOf<T>::type S;
    const int plane = L.n[1] * L.n[2];
    const CWork w = cwork(plane, L.n[0]);
    for (int s = 0; s < sweeps; s++) {
        const bool inb = (cur >> lbit) & 1u;
        const T* src = sm + (inb ? L.b : L.a);
        T* dst = sm + (inb ? L.a : L.b);
        for (int q = w.q0; q < plane; q += w.qstep) {
            CCol c;
            ccol(L, q, c);
            for (int i0 = w.x0; i0 < w.x1; i0++) {
                const int p = i0 * plane + q;
                const T acc = cneigh(src, L, c, q, i0, plane);
                dst[p] = ((S)1.0 - (S)L.w) * src[p] +
                         (S)L.w * (rl[p] - acc) / (S)L.c0;
            }
            if (w.qstep == plane) break;
        }
        __syncthreads();
        cur ^= 1u << lbit;
    }
}

// `steps` steps of Chebyshev iteration for A e = r with the Jacobi
// preconditioner D = c0 (Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed., Alg. 12.1): with the eigenvalues of D^-1 A in
// [lo, hi], theta = (hi + lo) / 2, delta = (hi - lo) / 2, sigma =
// theta / delta, z = D^-1 (r - A e):
//   d_0 = z_0 / theta,  rho_0 = 1 / sigma
//   rho_k = 1 / (2 sigma - rho_{k-1}),
//   d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k / delta) z_k,  e += d_k.
// After m steps the error is p_m(D^-1 A) e_0 with p_m the scaled
// Chebyshev polynomial, |p_m| <= 2 ((sqrt k - 1) / (sqrt k + 1))^m on
// [lo, hi] (k = hi / lo), where m Jacobi sweeps give (1 - w lo)^m.
template <typename T>
__device__ void ccheb(T* sm, const CLevel& L, const T* rl, int steps,
                      unsigned& cur, int lbit)
{
    typedef typename ScalarOf<T>::type S;
    const int plane = L.n[1] * L.n[2];
    const CWork w = cwork(plane, L.n[0]);
    T* D = sm + L.dbuf;
    const double theta = 0.5 * (L.lam_hi + L.lam_lo);
    const double delta = 0.5 * (L.lam_hi - L.lam_lo);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    for (int s = 0; s < steps; s++) {
        double ca = 0.0, cb = 1.0 / theta;
        if (s > 0) {
            const double rn = 1.0 / (2.0 * sigma - rho);
            ca = rn * rho;
            cb = 2.0 * rn / delta;
            rho = rn;
        }
        const S a = (S)ca, bq = (S)cb;
        const bool inb = (cur >> lbit) & 1u;
        const T* src = sm + (inb ? L.b : L.a);
        T* dst = sm + (inb ? L.a : L.b);
        for (int q = w.q0; q < plane; q += w.qstep) {
            CCol c;
            ccol(L, q, c);
            for (int i0 = w.x0; i0 < w.x1; i0++) {
                const int p = i0 * plane + q;
                const T acc = cneigh(src, L, c, q, i0, plane);
                const T z = (rl[p] - acc) / (S)L.c0 - src[p];
                const T dd = s > 0 ? fmac(bq * z, a, D[p]) : bq * z;
                D[p] = dd;
                dst[p] = src[p] + dd;
            }
            if (w.qstep == plane) break;
        }
        __syncthreads();
        cur ^= 1u << lbit;
    }
}

template <typename T>
__global__ void __launch_bounds__(1024)
k_coarse(const CGrid* __restrict__ grids, const T* __restrict__ rhs,
         T* __restrict__ out)
{
    typedef typename ScalarOf<T>::type S;
    extern __shared__ __align__(16) unsigned char csm_raw[];
    T* sm = reinterpret_cast<T*>(csm_raw);
    // the descriptor in LDS: its fields are re-read after every barrier,
    // which from global memory costs a round trip each time
    __shared__ CGrid G;
    {
        const int* src = reinterpret_cast<const int*>(grids + blockIdx.x);
        int* dst = reinterpret_cast<int*>(&G);
        for (int i = threadIdx.x; i < (int)(sizeof(CGrid) / sizeof(int));
             i += blockDim.x)
            dst[i] = src[i];
        __syncthreads();
    }
    const int nl = G.nlev;
    const T* r0 = rhs + G.rhs_off;
    unsigned cur = 0;   // bit l: level l's iterate is in B

    // first level of the block: e = 0, and its rhs staged in LDS when the
    // plan left room (every pass then stays in LDS)
    {
        const CLevel& L = G.lev[0];
        const int n = L.n[0] * L.n[1] * L.n[2];
        for (int p = threadIdx.x; p < n; p += blockDim.x) {
            sm[L.a + p] = zero_of<T>();
            if (L.r >= 0) sm[L.r + p] = r0[p];
        }
        __syncthreads();
    }
    // down
    for (int l = 0; l + 1 < nl; l++) {
        const CLevel& L = G.lev[l];
        const CLevel& N = G.lev[l + 1];
        const T* rl = L.r >= 0 ? sm + L.r : r0;
        crelax(sm, L, rl, L.pre, cur, l);
        const bool inb = (cur >> l) & 1u;
        const T* E = sm + (inb ? L.b : L.a);
        T* W = sm + (inb ? L.a : L.b);
        const int plane = L.n[1] * L.n[2];
        {
            const CWork w = cwork(plane, L.n[0]);
            for (int q = w.q0; q < plane; q += w.qstep) {
                CCol c;
                ccol(L, q, c);
                for (int i0 = w.x0; i0 < w.x1; i0++) {
                    const int p = i0 * plane + q;
                    const T acc = cneigh(E, L, c, q, i0, plane);
                    W[p] = fmac(acc, (S)L.c0, E[p]) - rl[p];
                }
                if (w.qstep == plane) break;
            }
        }
        __syncthreads();
        // restriction of W into N's rhs; N's iterate starts at zero
        const int nplane = N.n[1] * N.n[2];
        const CWork w = cwork(nplane, N.n[0]);
        for (int q = w.q0; q < nplane; q += w.qstep) {
            const int o2 = q % N.n[2], o1 = q / N.n[2];
            int off[3][3];  // [dy + 1][dz + 1]: plane offset in L, -1: zero
#pragma unroll
            for (int dy = -1; dy <= 1; dy++) {
                const int y = cw(2 * o1 + L.roff[1] + dy, L.n[1],
                                 L.periodic[1]);
#pragma unroll
                for (int dz = -1; dz <= 1; dz++) {
                    const int z = cw(2 * o2 + L.roff[2] + dz, L.n[2],
                                     L.periodic[2]);
                    off[dy + 1][dz + 1] =
                        (y < 0 || z < 0) ? -1 : y * L.n[2] + z;
                }
            }
            for (int o0 = w.x0; o0 < w.x1; o0++) {
                T px[3];
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    const int x = cw(2 * o0 + L.roff[0] + dx, L.n[0],
                                     L.periodic[0]);
                    T py[3];
#pragma unroll
                    for (int dy = 0; dy < 3; dy++) {
                        T a0 = zero_of<T>(), am = a0, ap = a0;
                        if (x >= 0) {
                            const T* row = W + x * plane;
                            am = cat(row, off[dy][0]);
                            a0 = cat(row, off[dy][1]);
                            ap = cat(row, off[dy][2]);
                        }
                        py[dy] = (S)0.5 * fmac(a0, (S)0.5, ap + am);
                    }
                    px[dx + 1] = (S)0.5 * fmac(py[1], (S)0.5, py[2] + py[0]);
                }
                const int qq = o0 * nplane + q;
                sm[N.r + qq] = (S)0.5 * fmac(px[1], (S)0.5, px[2] + px[0]);
                sm[N.a + qq] = zero_of<T>();
            }
            if (w.qstep == nplane) break;
        }
        __syncthreads();
    }
    // coarsest
    {
        const CLevel& L = G.lev[nl - 1];
        const T* rl = L.r >= 0 ? sm + L.r : r0;
        if (L.cheb)
            ccheb(sm, L, rl, L.post, cur, nl - 1);
        else
            crelax(sm, L, rl, L.post, cur, nl - 1);
    }
    // up
    for (int l = nl - 2; l >= 0; l--) {
        const CLevel& L = G.lev[l];
        const CLevel& N = G.lev[l + 1];
        const T* EN = sm + (((cur >> (l + 1)) & 1u) ? N.b : N.a);
        T* E = sm + (((cur >> l) & 1u) ? L.b : L.a);
        const int plane = L.n[1] * L.n[2];
        const int nplane = N.n[1] * N.n[2];
        const CWork w = cwork(plane, L.n[0]);
        for (int q = w.q0; q < plane; q += w.qstep) {
            const int o2 = q % L.n[2], o1 = q / L.n[2];
            const int q1 = o1 + L.iskip0[1], q2 = o2 + L.iskip0[2];
            const int c1 = (q1 >> 1) + L.ioff[1], c2 = (q2 >> 1) + L.ioff[2];
            const int m1 = (q1 & 1) + 1, m2 = (q2 & 1) + 1;
            int off[2][2];  // [dy][dz]: plane offset in N, -1: zero
#pragma unroll
            for (int dy = 0; dy < 2; dy++) {
                const int y = cw(c1 + dy, N.n[1], N.periodic[1]);
#pragma unroll
                for (int dz = 0; dz < 2; dz++) {
                    const int z = cw(c2 + dz, N.n[2], N.periodic[2]);
                    off[dy][dz] = (y < 0 || z < 0) ? -1 : y * N.n[2] + z;
                }
            }
            for (int o0 = w.x0; o0 < w.x1; o0++) {
                const int q0 = o0 + L.iskip0[0];
                const int c0 = (q0 >> 1) + L.ioff[0];
                const int m0 = (q0 & 1) + 1;
                T vx[2];
#pragma unroll
                for (int dx = 0; dx < 2; dx++) {
                    if (dx >= m0) break;
                    const int x = cw(c0 + dx, N.n[0], N.periodic[0]);
                    T vy[2];
#pragma unroll
                    for (int dy = 0; dy < 2; dy++) {
                        if (dy >= m1) break;
                        T vz = zero_of<T>(), vz1 = vz;
                        if (x >= 0) {
                            const T* row = EN + x * nplane;
                            vz = cat(row, off[dy][0]);
                            if (m2 == 2) vz1 = cat(row, off[dy][1]);
                        }
                        if (m2 == 2) vz = (S)0.5 * (vz + vz1);
                        vy[dy] = vz;
                    }
                    vx[dx] = (m1 == 2) ? (S)0.5 * (vy[0] + vy[1]) : vy[0];
                }
                const T v = (m0 == 2) ? (S)0.5 * (vx[0] + vx[1]) : vx[0];
                const int p = o0 * plane + q;
                E[p] = E[p] - v;
            }
            if (w.qstep == plane) break;
        }
        __syncthreads();
        const T* rl = L.r >= 0 ? sm + L.r : r0;
        crelax(sm, L, rl, L.post, cur, l);
    }
    // result
    {
        const CLevel& L = G.lev[0];
        const T* E = sm + ((cur & 1u) ? L.b : L.a);
        const int n = L.n[0] * L.n[1] * L.n[2];
        T* o = out + G.out_off;
        for (int p = threadIdx.x; p < n; p += blockDim.x) o[p] = E[p];
    }
}

static bool g_cheb = true;

void coarse_set_cheb(bool on) { g_cheb = on; }
bool coarse_cheb_enabled() { return g_cheb; }

bool cheb_bounds(const CLevel& L, double* lo, double* hi)
{
    if (L.R != 1 || L.c0 == 0.0) return false;
    // lambda(theta) = 1 + sum_c g_c(theta_c), g_c = (cp e^{i theta} +
    // cm e^{-i theta}) / c0 = 2 cp cos(theta) / c0 for cp = cm
    double gmin[3], gnext[3], gmax[3];
    for (int c = 0; c < 3; c++) {
        const double cp = L.has[c] ? L.cp[c][0] : 0.0;
        const double cm = L.has[c] ? L.cm[c][0] : 0.0;
        if (cp != cm) return false;
        const int n = L.n[c];
        gmin[c] = 1e300;
        gnext[c] = 1e300;
        gmax[c] = -1e300;
        for (int j = 0; j < n; j++) {
            const double th = L.periodic[c] ? 2.0 * M_PI * j / n
                                            : M_PI * (j + 1) / (n + 1);
            const double g = 2.0 * cp * std::cos(th) / L.c0;
            gmax[c] = std::max(gmax[c], g);
            if (g < gmin[c]) {
                gnext[c] = gmin[c];
                gmin[c] = g;
            } else if (g < gnext[c]) {
                gnext[c] = g;
            }
        }
    }
    double smin = 1.0, smax = 1.0;
    for (int c = 0; c < 3; c++) {
        smin += gmin[c];
        smax += gmax[c];
    }
    double l0 = smin;
    if (l0 <= 1e-9 * smax) {
        // the constant mode of a periodic operator (eigenvalue 0): the
        // smallest nonzero one moves one axis to its next mode
        l0 = 1e300;
        for (int c = 0; c < 3; c++)
            if (gnext[c] < 1e299) l0 = std::min(l0, smin + gnext[c] - gmin[c]);
    }
    if (!(l0 > 0.0) || !(smax > l0) || l0 >= 1e299) return false;
    // rounding and non-exact modes: a lower bound too low only slows the
    // iteration, an upper bound too low lets the top modes grow
    *lo = 0.9 * l0;
    *hi = 1.05 * smax;
    return true;
}

int launch_coarse(const CGrid* d_grids, int G, size_t lds_bytes, int dtype,
                  const void* rhs, void* out, hipStream_t s)
{
    if (G <= 0) return OK;
    if (lds_bytes > COARSE_LDS_MAX) {
        set_error("coarse solve: %zu bytes of LDS requested", lds_bytes);
        return EINVAL_;
    }
    static size_t attr[3] = {64 * 1024, 64 * 1024, 64 * 1024};
    const void* fn = dtype == 2 ? (const void*)k_coarse<float>
                                : (const void*)k_coarse<double>;
    size_t& a = attr[dtype == 2 ? 2 : 0];
    if (lds_bytes > a) {
        const hipError_t e = hipFuncSetAttribute(
            fn, hipFuncAttributeMaxDynamicSharedMemorySize,
            (int)COARSE_LDS_MAX);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("coarse solve: LDS attribute failed (%s)",
                      hipGetErrorString(e));
            return ENOTSUP_;
        }
        a = COARSE_LDS_MAX;
    }
    if (dtype == 2)
        hipLaunchKernelGGL(k_coarse<float>, dim3(G), dim3(1024), lds_bytes,
                           s, d_grids, (const float*)rhs, (float*)out);
    else
        hipLaunchKernelGGL(k_coarse<double>, dim3(G), dim3(1024), lds_bytes,
                           s, d_grids, (const double*)rhs, (double*)out);
    GFD_LAUNCH_CHECK();
    return OK;
}

}  // namespace gfd

extern "C" int gfd_coarse_cheb_bounds(const int* n, const int* periodic,
                                      const double* cp, double c0,
                                      double* lo, double* hi)
{
    if (!n || !periodic || !cp || !lo || !hi) {
        gfd::set_error("gfd_coarse_cheb_bounds: null argument");
        return gfd::EINVAL_;
    }
    gfd::CLevel L{};
    L.R = 1;
    L.c0 = c0;
    for (int c = 0; c < 3; c++) {
        if (n[c] < 1) {
            gfd::set_error("gfd_coarse_cheb_bounds: bad shape");
            return gfd::EINVAL_;
        }
        L.n[c] = n[c];
        L.periodic[c] = periodic[c];
        L.has[c] = cp[c] != 0.0;
        L.cp[c][0] = L.cm[c][0] = cp[c];
    }
    if (!gfd::cheb_bounds(L, lo, hi)) {
        gfd::set_error("gfd_coarse_cheb_bounds: no bounds for this stencil");
        return gfd::ENOTSUP_;
    }
    return gfd::OK;
}
