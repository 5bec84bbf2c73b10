// Fused fine level of the FP32 kinetic preconditioner for complex data
// (c/fd_preconditioner.cpp:311-345 with the nn=1 interpolation of
// c/bmgs/interpolate.cpp:70-99, K = 2):
//
//     d0' = I1(-d1)                       (never stored)
//     d0  = -(d0' - step (K0 d0' - res))  (FP64 out, FP32 arithmetic)
//
// The two-pass chain writes d0' (8 bytes per fine point) and reads it back
// with its stencil halo; here every fine plane of d0' is built in
// registers from the coarse d1 (1/8 of the points, read through L2).
//
// Block: TY fine rows x the whole fine z row (lane tx owns the complex z
// pair 2 tx, 2 tx + 1; PY rows per thread) marching along an x chunk.
// * Interpolation, per point: the yz-interpolated values Y_c of the coarse
//   planes c and c + 1 (z pass, then y pass, the reference's order, with
//   the Bloch factor of each wrapped coarse element applied first as the
//   marching transfer kernel does), advanced one coarse plane every two
//   fine planes (real data issues the raw loads of plane c + 2 at the
//   advance to c, two fine planes before they are used; complex data loads
//   at the advance, keeping registers for occupancy).  Fine plane p: Y_c
//   (even) or (Y_c + Y_{c+1}) / 2 (odd), negated.
// * Stencil, as the full-row kernel: x terms from a register queue of the
//   own points of planes x - R .. x + R, y and z terms from the current
//   plane in LDS (double buffered, one barrier per plane), whose 2R halo
//   rows are interpolated the same way; periodic z columns are mirrored
//   into the row padding with their Bloch factor.  Accumulation order z,
//   y, x and the UPDATE epilogue of the row kernel, so the result has the
//   bits of the two passes.
#include "precond_fine.h"

#include <algorithm>
#include <type_traits>

#include "batch.h"
#include "stencil_dev.h"

namespace gfd {

constexpr int PF_RP = 4;      // LDS columns each side of a row (even)
constexpr int PF_MAXT = 512;  // threads per block (256 or 512 variants)

struct PfArgs {
    const void* d1;        // T (c64 / f32)
    const void* res;       // FP64 counterpart of T
    void* out;
    int n[3];             // fine grid
    int m[3];             // coarse grid
    int off[3], skip0[3];  // interpolation plan
    int fper[3], cper[3];  // periodic axes, fine / coarse
    // Bloch factors: coarse element of sides (x, y, z) at cf[9 x + y + 3 z]
    // (the product x, then y, then z in FP64, as the marching transfer
    // kernel forms it; 1 below, 2 above), fine axis c at ff[c][0] above,
    // [1] below; rounded to FP32 for FP32 data (cf, ff), FP64 (cfd, ffd)
    float2 cf[27];
    float2 ff[3][2];
    double2 cfd[27];
    double2 ffd[3][2];
    int phased;            // some factor differs from 1
    double cp[3][3], cm[3][3];
    int has[3];
    double c0, step;
    int tiles_y, chunks_x, xchunk, nblocks;
    long fgrid, cgrid;     // points per band
    int nt_out;            // non-temporal result stores
};

// Data type T (float2: complex, float: real), its FP64 counterpart W and
// the vector P of a lane's two z points
// and the type F of its Bloch factors (FP32 data takes them rounded to
// FP32, as the transfer kernels do)
struct __align__(16) PfD2x2 {
    double2 a, b;
};
template <typename T> struct PfT;
template <> struct PfT<float2> {
    typedef float4 P;
    typedef double2 W;
    typedef float2 F;
    static constexpr bool cplx = true;
};
template <> struct PfT<float> {
    typedef float2 P;
    typedef double W;
    typedef float2 F;
    static constexpr bool cplx = false;
};
template <> struct PfT<double2> {
    typedef PfD2x2 P;
    typedef double2 W;
    typedef double2 F;
    static constexpr bool cplx = true;
};
template <> struct PfT<double> {
    typedef double2 P;
    typedef double W;
    typedef double2 F;
    static constexpr bool cplx = false;
};
__device__ inline float4 pf_pack(float2 a, float2 b)
{
    return make_float4(a.x, a.y, b.x, b.y);
}
__device__ inline float2 pf_pack(float a, float b) { return make_float2(a, b); }
__device__ inline void pf_unpack(float4 u, float2& a, float2& b)
{
    a = make_float2(u.x, u.y);
    b = make_float2(u.z, u.w);
}
__device__ inline void pf_unpack(float2 u, float& a, float& b)
{
    a = u.x;
    b = u.y;
}
__device__ inline PfD2x2 pf_pack(double2 a, double2 b) { return {a, b}; }
__device__ inline double2 pf_pack(double a, double b)
{
    return make_double2(a, b);
}
__device__ inline void pf_unpack(PfD2x2 u, double2& a, double2& b)
{
    a = u.a;
    b = u.b;
}
__device__ inline void pf_unpack(double2 u, double& a, double& b)
{
    a = u.x;
    b = u.y;
}
// the FP64 residual as the compute type, the result as FP64
__device__ inline float pf_in(double v) { return (float)v; }
__device__ inline float2 pf_in(double2 v) { return to_narrow(v); }
__device__ inline double pf_out(float v) { return (double)v; }
__device__ inline double2 pf_out(float2 v) { return to_wide(v); }
__device__ inline double pf_out(double v) { return v; }
__device__ inline double2 pf_out(double2 v) { return v; }

// Bloch factor times a value (real data: the factor is 1)
__device__ inline float2 pf_cmul(float2 f, float2 v) { return cmul(f, v); }
__device__ inline float pf_cmul(float2, float v) { return v; }
__device__ inline double2 pf_cmul(double2 f, double2 v) { return cmul(f, v); }
__device__ inline double pf_cmul(double2, double v) { return v; }

// Bloch factors in the precision of the data
template <typename F> __device__ inline F pf_fac(const PfArgs& a, int c, int d);
template <>
__device__ inline float2 pf_fac<float2>(const PfArgs& a, int c, int d)
{
    return a.ff[c][d];
}
template <>
__device__ inline double2 pf_fac<double2>(const PfArgs& a, int c, int d)
{
    return a.ffd[c][d];
}
template <typename F> __device__ inline F pf_cf(const PfArgs& a, int t);
template <> __device__ inline float2 pf_cf<float2>(const PfArgs& a, int t)
{
    return a.cf[t];
}
template <> __device__ inline double2 pf_cf<double2>(const PfArgs& a, int t)
{
    return a.cfd[t];
}

// index inside [0, n) after one wrap; false outside a zero boundary
__device__ inline bool pf_wrap(int& i, int n, int periodic)
{
    if (i >= 0 && i < n) return true;
    if (!periodic) return false;
    i += i < 0 ? n : -n;
    return true;
}

template <typename T, int R, int PY, int HM, int MAXT, bool PRE>
__global__ void __launch_bounds__(MAXT) k_pfine(PfArgs a)
{
    typedef typename PfT<T>::P P;
    typedef typename PfT<T>::W W;
    typedef typename PfT<T>::F F;
    typedef typename ScalarOf<T>::type S;
    constexpr bool CPLX = PfT<T>::cplx;
    extern __shared__ __align__(16) unsigned char pf_smem[];
    T* lds = reinterpret_cast<T*>(pf_smem);
    constexpr int Q = 2 * R + 1;

    const int L = xcd_remap(blockIdx.x, a.nblocks);
    const int yt = L % a.tiles_y;
    const int xt = (L / a.tiles_y) % a.chunks_x;
    const long b = L / (a.tiles_y * a.chunks_x);
    const int tx = threadIdx.x, sg = threadIdx.y;
    const int NZ = blockDim.x, NT = NZ * blockDim.y;
    const int tid = sg * NZ + tx;
    const int TY = blockDim.y * PY;  // divides n1 (host)
    const int n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    const int m0 = a.m[0], m1 = a.m[1], m2 = a.m[2];
    const int LW = n2 + 2 * PF_RP;
    const int BUF = (TY + 2 * R) * LW;
    const int y0 = yt * TY;
    const int xs = xt * a.xchunk, xe = min(xs + a.xchunk, n0);
    const T* __restrict__ gc = (const T*)a.d1 + b * a.cgrid;
    const long cplane = (long)m1 * m2;
    const unsigned cpb = (unsigned)(cplane * (long)sizeof(T));
    const bool zodd = (a.skip0[2] & 1) != 0;
    // the coarse Bloch factors (complex data; 2 BUF is even, so the table
    // is 8-byte aligned)
    F* ctab = reinterpret_cast<F*>(lds + 2 * BUF);
    if (CPLX)
        for (int t = tid; t < 27; t += NT) ctab[t] = pf_cf<F>(a, t);
    __syncthreads();

    // coarse elements of a point (fine row yf inside the grid, z pair t):
    // byte offsets in a coarse plane of rows cy, cy + 1 x columns cz,
    // cz + 1 (BXM_OOB reads 0: outside a zero boundary, or a second row
    // the point does not use) and their Bloch sides (y + 3 z, 4 bits each)
    auto setup = [&](int yf, int t, unsigned (&o)[4], unsigned& codes,
                     bool& two_y, bool both) {
        const int qy = yf + a.skip0[1];
        const int cy = (qy >> 1) + a.off[1];
        two_y = (qy & 1) != 0;
        const int cz = ((2 * t + a.skip0[2]) >> 1) + a.off[2];
        codes = 0;
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                int row = cy + r, col = cz + k;
                const int ys = row < 0 ? 1 : (row >= m1 ? 2 : 0);
                const int zs = col < 0 ? 1 : (col >= m2 ? 2 : 0);
                const bool ok = (r == 0 || two_y || both) &&
                                pf_wrap(row, m1, a.cper[1]) &&
                                pf_wrap(col, m2, a.cper[2]);
                o[2 * r + k] =
                    ok ? (unsigned)(row * m2 + col) * (unsigned)sizeof(T)
                       : BXM_OOB;
                codes |= (unsigned)(ys + 3 * zs) << (4 * (2 * r + k));
            }
    };
    // coarse plane c: buffer resource (empty outside a zero boundary) and
    // its x side
    auto cres = [&](int c, int& xside) {
        xside = c < 0 ? 1 : (c >= m0 ? 2 : 0);
        int cw = c;
        const bool ok = pf_wrap(cw, m0, a.cper[0]);
        return brsrc(gc + (ok ? (long)cw * cplane : 0L), ok ? cpb : 0u);
    };
    auto raw = [&](const __amdgpu_buffer_rsrc_t r, const unsigned (&o)[4],
                   T (&v)[4]) {
#pragma unroll
        for (int e = 0; e < 4; e++) v[e] = bload<T>(r, o[e]);
    };
    // z pass of both coarse rows of a point from its raw values (Bloch
    // factors first): z[r][k] at fine z 2t + k
    auto zrows = [&](T (&v)[4], unsigned codes, int xside,
                     T (&z)[2][2]) {
        if (CPLX && a.phased) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const unsigned t = 9u * (unsigned)xside +
                                   ((codes >> (4 * e)) & 15u);
                if (t) v[e] = pf_cmul(ctab[t], v[e]);
            }
        }
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const T h = (S)0.5 * (v[2 * r] + v[2 * r + 1]);
            z[r][0] = zodd ? h : v[2 * r];
            z[r][1] = zodd ? v[2 * r + 1] : h;
        }
    };
    // Y_c of a point (y pass: one coarse row, or the mean of two)
    auto yz = [&](T (&v)[4], unsigned codes, int xside, bool two_y,
                  T (&Y)[2]) {
        T z[2][2];
        zrows(v, codes, xside, z);
#pragma unroll
        for (int k = 0; k < 2; k++)
            Y[k] = two_y ? (S)0.5 * (z[0][k] + z[1][k]) : z[0][k];
    };

    // ---- own points: rows y0 + sg PY + j, pair tx.  Two rows per thread
    // (an even fine row and the next) always read the same two coarse rows,
    // so they share one group of four loads; oty then says whether the
    // first row is the odd one (it takes the mean, the second the upper
    // coarse row) or the even one (the lower row; the second the mean).
    constexpr int NG = PY == 2 ? 1 : PY;
    unsigned oo[NG][4], ocode[NG];
    bool oty[NG];
#pragma unroll
    for (int g = 0; g < NG; g++)
        setup(y0 + sg * PY + g, tx, oo[g], ocode[g], oty[g], PY == 2);
    // Y of the own rows from the raw values of their groups
    auto own_y = [&](T (&v)[NG][4], int xside, T (&Y)[PY][2]) {
        if constexpr (PY == 2) {
            T z[2][2];
            zrows(v[0], ocode[0], xside, z);
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const T mid = (S)0.5 * (z[0][k] + z[1][k]);
                Y[0][k] = oty[0] ? mid : z[0][k];
                Y[1][k] = oty[0] ? z[1][k] : mid;
            }
        } else {
#pragma unroll
            for (int g = 0; g < NG; g++) yz(v[g], ocode[g], xside, oty[g], Y[g]);
        }
    };
    // ---- halo slots: 2R rows x NZ pairs over the block
    const int H = 2 * R * NZ;
    unsigned ho[HM][4], hcode[HM];
    bool hty[HM];
    int hl[HM], hside[HM];  // LDS index (-1: none); fine y side (3: zero)
#pragma unroll
    for (int m = 0; m < HM; m++) {
        const int h = tid + m * NT;
        hl[m] = -1;
        hside[m] = 3;
        ho[m][0] = ho[m][1] = ho[m][2] = ho[m][3] = BXM_OOB;
        hcode[m] = 0;
        hty[m] = false;
        if (h < H) {
            const int hr = h / NZ, hc = h - hr * NZ;
            const int rr = hr < R ? hr - R : TY + hr - R;
            hl[m] = (rr + R) * LW + PF_RP + 2 * hc;
            int yw = y0 + rr;
            const int side = yw < 0 ? 1 : (yw >= n1 ? 2 : 0);
            if (pf_wrap(yw, n1, a.fper[1])) {
                hside[m] = side;
                setup(yw, hc, ho[m], hcode[m], hty[m], false);
            }
        }
    }

    // incremental coarse state of each point set: Y of planes c, c + 1
    // and the raw values of c + 2 in flight
    T oA[PY][2], oB[PY][2], oN[NG][4];
    T hA[HM][2], hB[HM][2], hN[HM][4];
    int oc = -0x40000000, hc0 = -0x40000000, oNx = 0, hNx = 0;

    // d0' of fine plane p (inside the grid) at the own points
    auto own_plane = [&](int p, T (&dst)[PY][2]) {
        const int qx = p + a.skip0[0];
        const int c = (qx >> 1) + a.off[0];
        if (c == oc + 1) {
#pragma unroll
            for (int j = 0; j < PY; j++) {
                oA[j][0] = oB[j][0];
                oA[j][1] = oB[j][1];
            }
            if constexpr (PRE) {
                own_y(oN, oNx, oB);
            } else {
                int xb;
                const __amdgpu_buffer_rsrc_t rb = cres(c + 1, xb);
                T vb[NG][4];
#pragma unroll
                for (int g = 0; g < NG; g++) raw(rb, oo[g], vb[g]);
                own_y(vb, xb, oB);
            }
        } else if (c != oc) {
            int xa, xb;
            const __amdgpu_buffer_rsrc_t ra = cres(c, xa);
            const __amdgpu_buffer_rsrc_t rb = cres(c + 1, xb);
            T va[NG][4], vb[NG][4];
#pragma unroll
            for (int g = 0; g < NG; g++) {
                raw(ra, oo[g], va[g]);
                raw(rb, oo[g], vb[g]);
            }
            own_y(va, xa, oA);
            own_y(vb, xb, oB);
        }
        if (c != oc) {
            if constexpr (PRE) {
                const __amdgpu_buffer_rsrc_t rn = cres(c + 2, oNx);
#pragma unroll
                for (int g = 0; g < NG; g++) raw(rn, oo[g], oN[g]);
            }
            oc = c;
        }
        const bool odd = (qx & 1) != 0;
#pragma unroll
        for (int j = 0; j < PY; j++)
#pragma unroll
            for (int k = 0; k < 2; k++)
                dst[j][k] = -(odd ? (S)0.5 * (oA[j][k] + oB[j][k]) : oA[j][k]);
    };
    // any fine plane: wrapped with its Bloch factor, zero outside a zero
    // boundary
    auto own_any = [&](int p, T (&dst)[PY][2]) {
        int pw = p;
        const int side = p < 0 ? 1 : (p >= n0 ? 2 : 0);
        if (!pf_wrap(pw, n0, a.fper[0])) {
#pragma unroll
            for (int j = 0; j < PY; j++)
                dst[j][0] = dst[j][1] = zero_of<T>();
            return;
        }
        own_plane(pw, dst);
        if (side) {
            const F f = pf_fac<F>(a, 0, side == 1 ? 1 : 0);
#pragma unroll
            for (int j = 0; j < PY; j++)
#pragma unroll
                for (int k = 0; k < 2; k++) dst[j][k] = pf_cmul(f, dst[j][k]);
        }
    };
    // d0' of fine plane p (inside the grid) at the halo slots, with the
    // fine y factor
    auto halo_plane = [&](int p, T (&dst)[HM][2]) {
        const int qx = p + a.skip0[0];
        const int c = (qx >> 1) + a.off[0];
        if (c == hc0 + 1) {
#pragma unroll
            for (int m = 0; m < HM; m++) {
                hA[m][0] = hB[m][0];
                hA[m][1] = hB[m][1];
                if constexpr (PRE) {
                    yz(hN[m], hcode[m], hNx, hty[m], hB[m]);
                } else {
                    int xb;
                    const __amdgpu_buffer_rsrc_t rb = cres(c + 1, xb);
                    T vb[4];
                    raw(rb, ho[m], vb);
                    yz(vb, hcode[m], xb, hty[m], hB[m]);
                }
            }
        } else if (c != hc0) {
            int xa, xb;
            const __amdgpu_buffer_rsrc_t ra = cres(c, xa);
            const __amdgpu_buffer_rsrc_t rb = cres(c + 1, xb);
#pragma unroll
            for (int m = 0; m < HM; m++) {
                T va[4], vb[4];
                raw(ra, ho[m], va);
                raw(rb, ho[m], vb);
                yz(va, hcode[m], xa, hty[m], hA[m]);
                yz(vb, hcode[m], xb, hty[m], hB[m]);
            }
        }
        if (c != hc0) {
            if constexpr (PRE) {
                const __amdgpu_buffer_rsrc_t rn = cres(c + 2, hNx);
#pragma unroll
                for (int m = 0; m < HM; m++) raw(rn, ho[m], hN[m]);
            }
            hc0 = c;
        }
        const bool odd = (qx & 1) != 0;
#pragma unroll
        for (int m = 0; m < HM; m++) {
#pragma unroll
            for (int k = 0; k < 2; k++) {
                T v = -(odd ? (S)0.5 * (hA[m][k] + hB[m][k]) : hA[m][k]);
                if (hside[m] == 3) v = zero_of<T>();
                if (hside[m] == 1 || hside[m] == 2)
                    v = pf_cmul(pf_fac<F>(a, 1, hside[m] == 1 ? 1 : 0), v);
                dst[m][k] = v;
            }
        }
    };

    S cpx[R], cmx[R], cpy[R], cmy[R], cpz[R], cmz[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        cpx[k] = (S)a.cp[0][k];
        cmx[k] = (S)a.cm[0][k];
        cpy[k] = (S)a.cp[1][k];
        cmy[k] = (S)a.cm[1][k];
        cpz[k] = (S)a.cp[2][k];
        cmz[k] = (S)a.cm[2][k];
    }
    const S c0 = (S)a.c0, ms = -(S)a.step;

    // z padding of the own rows: Bloch mirrors (periodic, every plane) or
    // zeros (written once, both buffers; ordered by the first barrier)
    const int col = PF_RP + 2 * tx;
    const bool zmir = a.fper[2] && a.has[2];
    const bool zlo = 2 * tx >= n2 - PF_RP;  // also at column - n2
    const bool zhi = 2 * tx < PF_RP;        // also at column + n2
    if (!a.fper[2] && (zlo || zhi)) {
#pragma unroll
        for (int j = 0; j < PY; j++)
#pragma unroll
            for (int bb = 0; bb < 2; bb++) {
                T* row = lds + bb * BUF + (R + sg * PY + j) * LW;
                const int cc = zhi ? col + n2 : col - n2;
                row[cc] = zero_of<T>();
                row[cc + 1] = zero_of<T>();
            }
    }

    T q[Q][PY][2];
#pragma unroll
    for (int s = 0; s < Q; s++) own_any(xs - R + s, q[s]);

    const long fplane = (long)n1 * n2;
    const long obase = b * a.fgrid + (long)(y0 + sg * PY) * n2 + 2 * tx;
    auto ld_res = [&](int x, W (&rs)[PY][2]) {
        const W* rp = (const W*)a.res + obase + (long)x * fplane;
#pragma unroll
        for (int j = 0; j < PY; j++) {
            rs[j][0] = rp[(long)j * n2];
            rs[j][1] = rp[(long)j * n2 + 1];
        }
    };
    for (int x = xs; x < xe; x++) {
        T* buf = lds + ((x - xs) & 1) * BUF;
        // the residual of this plane: in flight over the halo, the LDS
        // stores and the barrier
        W rc[PY][2];
        ld_res(x, rc);
        T hv[HM][2];
        halo_plane(x, hv);
#pragma unroll
        for (int j = 0; j < PY; j++) {
            T* row = buf + (R + sg * PY + j) * LW;
            *reinterpret_cast<P*>(row + col) =
                pf_pack(q[R][j][0], q[R][j][1]);
            if (zmir && zhi) {
                const T u0 = pf_cmul(pf_fac<F>(a, 2, 0), q[R][j][0]);
                const T u1 = pf_cmul(pf_fac<F>(a, 2, 0), q[R][j][1]);
                *reinterpret_cast<P*>(row + col + n2) =
                    pf_pack(u0, u1);
            }
            if (zmir && zlo) {
                const T u0 = pf_cmul(pf_fac<F>(a, 2, 1), q[R][j][0]);
                const T u1 = pf_cmul(pf_fac<F>(a, 2, 1), q[R][j][1]);
                *reinterpret_cast<P*>(row + col - n2) =
                    pf_pack(u0, u1);
            }
        }
#pragma unroll
        for (int m = 0; m < HM; m++)
            if (hl[m] >= 0)
                *reinterpret_cast<P*>(buf + hl[m]) =
                    pf_pack(hv[m][0], hv[m][1]);
        __syncthreads();

        W* op = (W*)a.out + obase + (long)x * fplane;
#pragma unroll
        for (int j = 0; j < PY; j++) {
            const int lrow = (R + sg * PY + j) * LW + col;
            T acc[2] = {zero_of<T>(), zero_of<T>()};
            if (a.has[2]) {
                // z - 4 .. z + 5 around the pair
                T w[10];
#pragma unroll
                for (int c = 0; c < 5; c++) {
                    if (c == 2) {
                        w[4] = q[R][j][0];
                        w[5] = q[R][j][1];
                    } else {
                        const P u = *reinterpret_cast<const P*>(
                            buf + lrow + 2 * (c - 2));
                        pf_unpack(u, w[2 * c], w[2 * c + 1]);
                    }
                }
#pragma unroll
                for (int k = 0; k < 2; k++)
#pragma unroll
                    for (int kk = 1; kk <= R; kk++) {
                        acc[k] = fmac(acc[k], cpz[kk - 1], w[4 + k + kk]);
                        acc[k] = fmac(acc[k], cmz[kk - 1], w[4 + k - kk]);
                    }
            }
            if (a.has[1]) {
#pragma unroll
                for (int kk = 1; kk <= R; kk++) {
                    T up[2], dn[2];
                    if (j + kk < PY) {
                        up[0] = q[R][j + kk][0];
                        up[1] = q[R][j + kk][1];
                    } else {
                        const P u = *reinterpret_cast<const P*>(
                            buf + lrow + kk * LW);
                        pf_unpack(u, up[0], up[1]);
                    }
                    if (j - kk >= 0) {
                        dn[0] = q[R][j - kk][0];
                        dn[1] = q[R][j - kk][1];
                    } else {
                        const P u = *reinterpret_cast<const P*>(
                            buf + lrow - kk * LW);
                        pf_unpack(u, dn[0], dn[1]);
                    }
#pragma unroll
                    for (int k = 0; k < 2; k++) {
                        acc[k] = fmac(acc[k], cpy[kk - 1], up[k]);
                        acc[k] = fmac(acc[k], cmy[kk - 1], dn[k]);
                    }
                }
            }
            if (a.has[0]) {
#pragma unroll
                for (int kk = 1; kk <= R; kk++)
#pragma unroll
                    for (int k = 0; k < 2; k++) {
                        acc[k] = fmac(acc[k], cpx[kk - 1], q[R + kk][j][k]);
                        acc[k] = fmac(acc[k], cmx[kk - 1], q[R - kk][j][k]);
                    }
            }
            W o[2];
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const T v = q[R][j][k];
                T r = fmac(acc[k], c0, v);
                if constexpr (std::is_same<T, W>::value)
                    r = r - rc[j][k];
                else
                    r = r - pf_in(rc[j][k]);
                const T d = fmac(v, ms, r);
                o[k] = pf_out(-d);
            }
            if (a.nt_out) {
                // streamed result: non-temporal stores
                typedef double nt2 __attribute__((ext_vector_type(2)));
                if constexpr (CPLX) {
                    __builtin_nontemporal_store(
                        nt2{o[0].x, o[0].y},
                        reinterpret_cast<nt2*>(op + (long)j * n2));
                    __builtin_nontemporal_store(
                        nt2{o[1].x, o[1].y},
                        reinterpret_cast<nt2*>(op + (long)j * n2 + 1));
                } else {
                    __builtin_nontemporal_store(
                        nt2{o[0], o[1]},
                        reinterpret_cast<nt2*>(op + (long)j * n2));
                }
            } else {
                op[(long)j * n2] = o[0];
                op[(long)j * n2 + 1] = o[1];
            }
        }
#pragma unroll
        for (int s = 0; s + 1 < Q; s++)
#pragma unroll
            for (int j = 0; j < PY; j++) {
                q[s][j][0] = q[s + 1][j][0];
                q[s][j][1] = q[s + 1][j][1];
            }
        if (x + 1 < xe) own_any(x + R + 1, q[Q - 1]);
    }
}

// ------------------------------------------------------------------ host

static bool g_pf_on = true;
static int g_pf_py = 0, g_pf_nys = 0, g_pf_xc = 0, g_pf_nt = 0;

void precond_fine_enable(bool on) { g_pf_on = on; }

void precond_fine_nt(bool on) { g_pf_nt = on ? 1 : 0; }

void precond_fine_force(int py, int nys, int xc)
{
    g_pf_py = py;
    g_pf_nys = nys;
    g_pf_xc = xc;
}

// Raw coarse loads two fine planes ahead for real data; complex data
// loads at the advance (16 fewer VGPRs; measured 2 % faster there, and the
// prefetch 2 % faster for real data)
template <typename T, int R, int PY, int HM>
static void pf_launch(const PfArgs& a, int NZ, int NYS, size_t lds,
                      hipStream_t s)
{
    constexpr bool PRE = !PfT<T>::cplx;
    if (NZ * NYS > 256)
        hipLaunchKernelGGL((k_pfine<T, R, PY, HM, 512, PRE>),
                           dim3((unsigned)a.nblocks), dim3(NZ, NYS), lds, s,
                           a);
    else
        hipLaunchKernelGGL((k_pfine<T, R, PY, HM, 256, PRE>),
                           dim3((unsigned)a.nblocks), dim3(NZ, NYS), lds, s,
                           a);
}

template <typename T>
static int pf_dispatch(const PfArgs& a, int R, int PY, int HM, int NZ,
                       int NYS, size_t lds, hipStream_t s)
{
#define GFD_PF(RR, PP, HH)                                                     \
    if (R == RR && PY == PP && HM == HH) {                                     \
        pf_launch<T, RR, PP, HH>(a, NZ, NYS, lds, s);                          \
        GFD_LAUNCH_CHECK();                                                    \
        return OK;                                                             \
    }
    GFD_PF(1, 1, 1) GFD_PF(1, 1, 2) GFD_PF(1, 2, 1) GFD_PF(1, 2, 2)
    GFD_PF(2, 1, 1) GFD_PF(2, 1, 2) GFD_PF(2, 2, 1) GFD_PF(2, 2, 2)
    GFD_PF(3, 1, 1) GFD_PF(3, 1, 2) GFD_PF(3, 2, 1) GFD_PF(3, 2, 2)
#undef GFD_PF
    return ENOTSUP_;
}

int launch_precond_fine(const AxisStencil& ax, double c0, const int n[3],
                        const int fine_periodic[3], const XferPlan& ip,
                        const int coarse_periodic[3], const double* phases,
                        const void* d1, const void* res, void* out, long nin,
                        double step, int real, int fp64, hipStream_t s)
{
    if (!g_pf_on || !ip.interpolate || ip.K != 2 || ax.R < 1 || ax.R > 3)
        return ENOTSUP_;
    for (int c = 0; c < 3; c++)
        if (ip.nout[c] != n[c]) return ENOTSUP_;
    const int R = ax.R;
    const int n0 = n[0], n1 = n[1], n2 = n[2];
    if (n2 % 2 || n2 / 2 < 4 || n2 / 2 > PF_MAXT || n0 < R || n1 < R)
        return ENOTSUP_;
    const int NZ = n2 / 2;
    // LDS of a tile TY rows high: two plane buffers of T, then the 27
    // complex factors (aligned: the buffers hold an even number of
    // elements)
    const size_t es = (real ? 4 : 8) * (fp64 ? 2 : 1);
    auto lds_of = [&](int ty) {
        return 2 * (size_t)(ty + 2 * R) * (n2 + 2 * PF_RP) * es + 28 * 16;
    };
    // tile: TY = NYS PY dividing n1, the tallest of at most 512 threads
    // whose LDS fits 64 KB; two rows per thread first for complex FP32
    // data, one row first otherwise (measured on 128 bands of 128^3:
    // complex FP32 3.66 against 3.75 ms, real FP32 1.80 against 1.86 ms,
    // real FP64 2.17 against 2.21 ms, complex FP64 5.34 against 6.26 ms)
    int PY = 0, NYS = 0;
    const int pys[2] = {(!fp64 && !real) ? 2 : 1, (!fp64 && !real) ? 1 : 2};
    for (int py : pys) {
        if (g_pf_py && py != g_pf_py) continue;
        for (int nys = PF_MAXT / NZ; nys >= 1; nys--) {
            if (g_pf_nys && nys != g_pf_nys) continue;
            if (n1 % (nys * py) == 0 && nys * py >= R &&
                lds_of(nys * py) <= 64 * 1024) {
                PY = py;
                NYS = nys;
                break;
            }
        }
        if (PY) break;
    }
    const bool forced = g_pf_py || g_pf_nys;
    const int TY = NYS * PY, NT = NZ * NYS;
    const int H = 2 * R * NZ;
    const int HM = PY ? (H + NT - 1) / NT : 0;
    if (!PY || HM > 2) {
        if (!forced) return ENOTSUP_;
        set_error("fused preconditioner: forced tile (%d rows, %d thread "
                  "rows) does not fit the %d x %d x %d grid",
                  g_pf_py, g_pf_nys, n0, n1, n2);
        return EINVAL_;
    }
    const size_t lds = lds_of(TY);
    const long cplane = (long)ip.nin[1] * ip.nin[2];
    if (cplane * (long)es >= 0x7fffffffL) return ENOTSUP_;

    PfArgs a;
    a.d1 = d1;
    a.res = res;
    a.out = out;
    a.phased = 0;
    double2 ph[3][2];
    for (int c = 0; c < 3; c++) {
        a.n[c] = n[c];
        a.m[c] = ip.nin[c];
        a.off[c] = ip.ax[c].off;
        a.skip0[c] = ip.ax[c].skip0;
        a.fper[c] = fine_periodic[c];
        a.cper[c] = coarse_periodic[c];
        for (int d = 0; d < 2; d++) {
            ph[c][d] = phases ? make_double2(phases[4 * c + 2 * d],
                                             phases[4 * c + 2 * d + 1])
                              : make_double2(1.0, 0.0);
            a.ff[c][d] = make_float2((float)ph[c][d].x, (float)ph[c][d].y);
            a.ffd[c][d] = ph[c][d];
            if ((fine_periodic[c] || coarse_periodic[c]) &&
                (ph[c][d].x != 1.0 || ph[c][d].y != 0.0))
                a.phased = 1;
        }
        for (int k = 0; k < 3; k++) {
            a.cp[c][k] = k < R ? ax.cp[c][k] : 0.0;
            a.cm[c][k] = k < R ? ax.cm[c][k] : 0.0;
        }
        a.has[c] = ax.has[c];
    }
    for (int xs = 0; xs < 3; xs++)
        for (int zs = 0; zs < 3; zs++)
            for (int ys = 0; ys < 3; ys++) {
                double2 f = xs ? ph[0][xs == 1 ? 1 : 0] : make_double2(1.0, 0.0);
                if (ys) f = cmul(f, ph[1][ys == 1 ? 1 : 0]);
                if (zs) f = cmul(f, ph[2][zs == 1 ? 1 : 0]);
                a.cf[9 * xs + ys + 3 * zs] = make_float2((float)f.x, (float)f.y);
                a.cfd[9 * xs + ys + 3 * zs] = f;
            }
    a.c0 = c0;
    a.step = step;
    a.nt_out = g_pf_nt;
    a.fgrid = (long)n0 * n1 * n2;
    a.cgrid = (long)ip.nin[0] * cplane;
    a.tiles_y = n1 / TY;
    const long tiles = nin * a.tiles_y;
    // x chunks only to reach about 1024 blocks (four per CU): each chunk
    // start rebuilds 2R + 1 planes and the coarse state from scratch
    int chunks = (int)std::max(1L, (1024 + tiles - 1) / tiles);
    chunks = std::min(chunks, std::max(1, n0 / 8));
    a.xchunk = g_pf_xc > 0 ? std::min(g_pf_xc, n0) : (n0 + chunks - 1) / chunks;
    a.chunks_x = (n0 + a.xchunk - 1) / a.xchunk;
    const long nb = tiles * a.chunks_x;
    if (nb > 0x7fffffffL) return ENOTSUP_;
    a.nblocks = (int)nb;

    if (fp64)
        return real ? pf_dispatch<double>(a, R, PY, HM, NZ, NYS, lds, s)
                    : pf_dispatch<double2>(a, R, PY, HM, NZ, NYS, lds, s);
    return real ? pf_dispatch<float>(a, R, PY, HM, NZ, NYS, lds, s)
                : pf_dispatch<float2>(a, R, PY, HM, NZ, NYS, lds, s);
}

}  // namespace gfd
