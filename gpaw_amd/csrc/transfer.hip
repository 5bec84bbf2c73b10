// Multigrid transfer kernels for gfx950: K-tap restriction and
// interpolation (K = 2 nn = 2, 4, 6, 8), real and complex, periodic / zero /
// Bloch-phase boundaries.
//
// A 3-D transfer is the product of three 1-D transfers (z, then y, then x,
// the order of c/bmgs/restrict.cpp:153-155 and interpolate.cpp:86-97).
// Each pass is one kernel with one thread per output point; the boundary of
// the pass's axis is taken by index wrap, the other two axes are in range.
//
// 1-D semantics in interior-index form (derived from the reference's padded
// buffers, c/bmgs/restrict.cpp:18-63 and interpolate.cpp:10-46):
//   restrict:    out[i] = 1/2 (a[c] + sum_j w_j (a[c+j] + a[c-j])),
//                c = 2 i + off, j = 1, 3, .., K-1
//   interpolate: with p = (i + skip0) / 2, c = p + off:
//                even (i + skip0): out[i] = a[c]
//                odd:              out[i] = sum_j w_j (a[c-j] + a[c+1+j]),
//                j = 0 .. K/2-1
// The weights are exact dyadic rationals.
#include "transfer.h"

#include <algorithm>
#include <map>
#include <mutex>

#include <type_traits>

namespace gfd {

template <int K> struct W;
template <> struct W<2> {
    static constexpr double w[1] = {0.5};
};
template <> struct W<4> {
    static constexpr double w[2] = {0.5625, -0.0625};
};
template <> struct W<6> {
    static constexpr double w[3] = {0.58593750, -0.09765625, 0.01171875};
};
template <> struct W<8> {
    static constexpr double w[4] = {0.59814453125, -0.11962890625,
                                    0.02392578125, -0.00244140625};
};

template <typename T> struct XIsCplx { static constexpr bool value = false; };
template <> struct XIsCplx<double2> { static constexpr bool value = true; };
template <> struct XIsCplx<float2> { static constexpr bool value = true; };

// input value (possibly the wider FP64 type) as the compute type T
template <typename T, typename TI>
__device__ inline T narrow(TI v)
{
    if constexpr (std::is_same<T, TI>::value) {
        return v;
    } else if constexpr (XIsCplx<T>::value) {
        return T{(float)v.x, (float)v.y};
    } else {
        return (T)v;
    }
}

template <typename T>
struct XArgs {
    const T* in;
    T* out;
    T* out2;
    const T* base;
    int d[3];      // output dims of this pass
    int ilen;      // input length along the axis
    long istride;  // input stride along the axis
    long in_grid;  // input points per grid
    long out_grid; // output points per grid
    long total;
    int off, skip0;
    int periodic;
    double2 ph_hi, ph_lo;
    int negate;
    double alpha;
};

template <typename T>
__device__ inline T xload(const XArgs<T>& a, const T* __restrict__ g,
                          long pos, int j)
{
    if (j >= 0 && j < a.ilen) return g[pos + (long)j * a.istride];
    if (!a.periodic) return zero_of<T>();
    double2 f = make_double2(1.0, 0.0);
    bool first = true;
    while (j < 0) {
        j += a.ilen;
        f = first ? a.ph_lo : cmul(a.ph_lo, f);
        first = false;
    }
    while (j >= a.ilen) {
        j -= a.ilen;
        f = first ? a.ph_hi : cmul(a.ph_hi, f);
        first = false;
    }
    T v = g[pos + (long)j * a.istride];
    if constexpr (XIsCplx<T>::value) v = cmul(f, v);
    return v;
}

template <typename T, int AX, int K, bool INTERP, int EPI>
__global__ void __launch_bounds__(256) k_xfer1d(XArgs<T> a)
{
    typedef typename ScalarOf<T>::type S;
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= a.total) return;
    long b;
    int o0, o1, o2;
    flat_split(gid, a.total, a.out_grid, a.d[1], a.d[2], b, o0, o1, o2);
    // input position with the pass axis at 0
    long pos;
    int i;
    if (AX == 2) {
        pos = ((long)o0 * a.d[1] + o1) * a.ilen;
        i = o2;
    } else if (AX == 1) {
        pos = (long)o0 * a.ilen * a.d[2] + o2;
        i = o1;
    } else {
        pos = (long)o1 * a.d[2] + o2;
        i = o0;
    }
    const T* __restrict__ g = a.in + b * a.in_grid;
    T v;
    if constexpr (!INTERP) {
        const int c = 2 * i + a.off;
        T acc = xload(a, g, pos, c);
#pragma unroll
        for (int m = 0; m < K / 2; m++) {
            const int j = 2 * m + 1;
            T pair = xload(a, g, pos, c + j) + xload(a, g, pos, c - j);
            acc = fmac(acc, (S)W<K>::w[m], pair);
        }
        v = (S)0.5 * acc;
    } else {
        const int q = i + a.skip0;
        const int c = (q >> 1) + a.off;
        if ((q & 1) == 0) {
            v = xload(a, g, pos, c);
        } else {
            T pair0 = xload(a, g, pos, c) + xload(a, g, pos, c + 1);
            T acc = (S)W<K>::w[0] * pair0;
#pragma unroll
            for (int m = 1; m < K / 2; m++) {
                T pair = xload(a, g, pos, c - m) + xload(a, g, pos, c + 1 + m);
                acc = fmac(acc, (S)W<K>::w[m], pair);
            }
            v = acc;
        }
    }
    if constexpr (EPI == XEPI_PLAIN) {
        a.out[gid] = v;
    } else if constexpr (EPI == XEPI_NEG) {
        a.out[gid] = -v;
    } else if constexpr (EPI == XEPI_SCALE2) {
        if (a.negate) v = -v;
        a.out[gid] = v;
        a.out2[gid] = (S)a.alpha * v;
    } else {
        a.out[gid] = a.base[gid] - v;
    }
}

// ------------------------------------------------------------ fused K = 2
// One thread per output point; the 3 x 3 x 3 fine neighbourhood
// (restriction) or the <= 8 coarse neighbours (interpolation) come through
// the caches.  Sums are associated z innermost, then y, then x, as the
// reference's pass order.

template <typename T, typename TI = T>
struct F3Args {
    const TI* in;
    T* out;
    T* out2;
    const T* base;
    int ni[3];      // input grid
    int no[3];      // output grid
    int off[3];     // interior index of the first centre per axis
    int skip0[3];
    int periodic[3];
    double2 ph[3][2];
    long in_grid, out_grid, total;
    int negate;
    double alpha;
};

// value of the input at (i0, i1, i2) with per-axis wrap / zero; the
// phase is applied at the input's precision, then the value is narrowed
template <typename T, typename TI>
__device__ inline T f3load(const F3Args<T, TI>& a, const TI* __restrict__ g,
                           int i0, int i1, int i2)
{
    int ii[3] = {i0, i1, i2};
    double2 f = make_double2(1.0, 0.0);
    bool phased = false;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (ii[c] >= 0 && ii[c] < a.ni[c]) continue;
        if (!a.periodic[c]) return zero_of<T>();
        while (ii[c] < 0) {
            ii[c] += a.ni[c];
            f = phased ? cmul(a.ph[c][1], f) : a.ph[c][1];
            phased = true;
        }
        while (ii[c] >= a.ni[c]) {
            ii[c] -= a.ni[c];
            f = phased ? cmul(a.ph[c][0], f) : a.ph[c][0];
            phased = true;
        }
    }
    TI v = g[((long)ii[0] * a.ni[1] + ii[1]) * a.ni[2] + ii[2]];
    if constexpr (XIsCplx<TI>::value)
        if (phased) v = apply_phase(v, f);
    return narrow<T, TI>(v);
}

template <typename T, int EPI, typename TI>
__device__ inline void f3store(const F3Args<T, TI>& a, long gid, T v)
{
    typedef typename ScalarOf<T>::type S;
    if constexpr (EPI == XEPI_PLAIN) {
        a.out[gid] = v;
    } else if constexpr (EPI == XEPI_NEG) {
        a.out[gid] = -v;
    } else if constexpr (EPI == XEPI_SCALE2) {
        if (a.negate) v = -v;
        a.out[gid] = v;
        a.out2[gid] = (S)a.alpha * v;
    } else {
        a.out[gid] = a.base[gid] - v;
    }
}

template <typename T, int EPI, typename TI>
__global__ void __launch_bounds__(256) k_restrict2(F3Args<T, TI> a)
{
    typedef typename ScalarOf<T>::type S;
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= a.total) return;
    long b;
    int o0, o1, o2;
    flat_split(gid, a.total, a.out_grid, a.no[1], a.no[2], b, o0, o1, o2);
    const int c0 = 2 * o0 + a.off[0], c1 = 2 * o1 + a.off[1],
              c2 = 2 * o2 + a.off[2];
    const TI* __restrict__ g = a.in + b * a.in_grid;
    T px[3];
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
        T py[3];
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            // 1/2 (a0 + 1/2 (a1 + a-1)) along z
            const T zp = f3load(a, g, c0 + dx, c1 + dy, c2 + 1) +
                         f3load(a, g, c0 + dx, c1 + dy, c2 - 1);
            py[dy + 1] = (S)0.5 * fmac(f3load(a, g, c0 + dx, c1 + dy, c2),
                                       (S)0.5, zp);
        }
        px[dx + 1] = (S)0.5 * fmac(py[1], (S)0.5, py[2] + py[0]);
    }
    const T v = (S)0.5 * fmac(px[1], (S)0.5, px[2] + px[0]);
    f3store<T, EPI, TI>(a, gid, v);
}

template <typename T, int EPI, typename TI>
__global__ void __launch_bounds__(256) k_interp2(F3Args<T, TI> a)
{
    typedef typename ScalarOf<T>::type S;
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= a.total) return;
    long b;
    int o0, o1, o2;
    flat_split(gid, a.total, a.out_grid, a.no[1], a.no[2], b, o0, o1, o2);
    const int q0 = o0 + a.skip0[0], q1 = o1 + a.skip0[1], q2 = o2 + a.skip0[2];
    const int c0 = (q0 >> 1) + a.off[0], c1 = (q1 >> 1) + a.off[1],
              c2 = (q2 >> 1) + a.off[2];
    const int n0 = (q0 & 1) + 1, n1 = (q1 & 1) + 1, n2 = (q2 & 1) + 1;
    const TI* __restrict__ g = a.in + b * a.in_grid;
    T vx[2];
#pragma unroll
    for (int dx = 0; dx < 2; dx++) {
        if (dx >= n0) break;
        T vy[2];
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            if (dy >= n1) break;
            T vz = f3load(a, g, c0 + dx, c1 + dy, c2);
            if (n2 == 2)
                vz = (S)0.5 * (vz + f3load(a, g, c0 + dx, c1 + dy, c2 + 1));
            vy[dy] = vz;
        }
        vx[dx] = (n1 == 2) ? (S)0.5 * (vy[0] + vy[1]) : vy[0];
    }
    const T v = (n0 == 2) ? (S)0.5 * (vx[0] + vx[1]) : vx[0];
    f3store<T, EPI, TI>(a, gid, v);
}

// value of interpolated output point (o0, o1, o2) of grid g
template <typename T, typename TI>
__device__ inline T interp2_point(const F3Args<T, TI>& a,
                                  const TI* __restrict__ g, int o0, int o1,
                                  int o2)
{
    typedef typename ScalarOf<T>::type S;
    const int q0 = o0 + a.skip0[0], q1 = o1 + a.skip0[1], q2 = o2 + a.skip0[2];
    const int c0 = (q0 >> 1) + a.off[0], c1 = (q1 >> 1) + a.off[1],
              c2 = (q2 >> 1) + a.off[2];
    const int n0 = (q0 & 1) + 1, n1 = (q1 & 1) + 1, n2 = (q2 & 1) + 1;
    T vx[2];
#pragma unroll
    for (int dx = 0; dx < 2; dx++) {
        if (dx >= n0) break;
        T vy[2];
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            if (dy >= n1) break;
            T vz = f3load(a, g, c0 + dx, c1 + dy, c2);
            if (n2 == 2)
                vz = (S)0.5 * (vz + f3load(a, g, c0 + dx, c1 + dy, c2 + 1));
            vy[dy] = vz;
        }
        vx[dx] = (n1 == 2) ? (S)0.5 * (vy[0] + vy[1]) : vy[0];
    }
    return (n0 == 2) ? (S)0.5 * (vx[0] + vx[1]) : vx[0];
}

// Two neighbouring outputs along z per thread (even output rows): the
// index arithmetic is shared and the coarse values both need are loaded
// once
template <typename T, int EPI, typename TI>
__global__ void __launch_bounds__(256) k_interp2p(F3Args<T, TI> a)
{
    const long gp = (long)blockIdx.x * 256 + threadIdx.x;
    if (2 * gp >= a.total) return;
    const long gid = 2 * gp;
    long b;
    int o0, o1, o2;
    flat_split(gid, a.total, a.out_grid, a.no[1], a.no[2], b, o0, o1, o2);
    const TI* __restrict__ g = a.in + b * a.in_grid;
    const T v0 = interp2_point(a, g, o0, o1, o2);
    const T v1 = interp2_point(a, g, o0, o1, o2 + 1);
    f3store<T, EPI, TI>(a, gid, v0);
    f3store<T, EPI, TI>(a, gid + 1, v1);
}

template <typename T, int EPI, typename TI>
static int launch_fused2_t(const XferPlan& p, const BC& bc, const void* in,
                           void* out, long nin, const XEpiArgs& ea,
                           hipStream_t s)
{
    F3Args<T, TI> a;
    a.in = (const TI*)in;
    a.out = (T*)out;
    a.out2 = (T*)ea.out2;
    a.base = (const T*)ea.base;
    for (int c = 0; c < 3; c++) {
        a.ni[c] = p.nin[c];
        a.no[c] = p.nout[c];
        a.off[c] = p.ax[c].off;
        a.skip0[c] = p.ax[c].skip0;
        a.periodic[c] = bc.periodic[c];
        a.ph[c][0] = bc.phase[c][0];
        a.ph[c][1] = bc.phase[c][1];
    }
    a.in_grid = (long)p.nin[0] * p.nin[1] * p.nin[2];
    a.out_grid = (long)p.nout[0] * p.nout[1] * p.nout[2];
    a.total = a.out_grid * nin;
    a.negate = ea.negate;
    a.alpha = ea.alpha;
    const long nb = (a.total + 255) / 256;
    if (nb <= 0) return OK;
    if (nb > 0x7fffffffL) {
        set_error("transfer: too many blocks");
        return ENOTSUP_;
    }
    if (p.interpolate && p.nout[2] % 2 == 0)
        hipLaunchKernelGGL((k_interp2p<T, EPI, TI>),
                           dim3((unsigned)((a.total / 2 + 255) / 256)),
                           dim3(256), 0, s, a);
    else if (p.interpolate)
        hipLaunchKernelGGL((k_interp2<T, EPI, TI>), dim3((unsigned)nb),
                           dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_restrict2<T, EPI, TI>), dim3((unsigned)nb),
                           dim3(256), 0, s, a);
    GFD_LAUNCH_CHECK();
    return OK;
}

template <typename T, typename TI = T>
static int launch_fused2(const XferPlan& p, const BC& bc, const void* in,
                         void* out, long nin, const XEpiArgs& ea,
                         hipStream_t s)
{
    switch (ea.epi) {
    case XEPI_PLAIN: return launch_fused2_t<T, XEPI_PLAIN, TI>(p, bc, in, out, nin, ea, s);
    case XEPI_NEG: return launch_fused2_t<T, XEPI_NEG, TI>(p, bc, in, out, nin, ea, s);
    case XEPI_SCALE2: return launch_fused2_t<T, XEPI_SCALE2, TI>(p, bc, in, out, nin, ea, s);
    case XEPI_SUBFROM: return launch_fused2_t<T, XEPI_SUBFROM, TI>(p, bc, in, out, nin, ea, s);
    }
    set_error("transfer: bad epilogue %d", ea.epi);
    return EINVAL_;
}

long xfer_scratch_elems(const XferPlan& p, long nin)
{
    // pass z output + pass y output
    const long t1 = (long)p.nin[0] * p.nin[1] * p.nout[2];
    const long t2 = (long)p.nin[0] * p.nout[1] * p.nout[2];
    return (t1 + t2) * nin;
}

template <typename T, int AX, int K, bool INTERP, int EPI>
static int launch_pass(XArgs<T> a, hipStream_t s)
{
    const long nb = (a.total + 255) / 256;
    if (nb <= 0) return OK;
    if (nb > 0x7fffffffL) {
        set_error("transfer: too many blocks");
        return ENOTSUP_;
    }
    hipLaunchKernelGGL((k_xfer1d<T, AX, K, INTERP, EPI>), dim3((unsigned)nb),
                       dim3(256), 0, s, a);
    GFD_LAUNCH_CHECK();
    return OK;
}

template <typename T, int AX, int K, bool INTERP>
static int launch_pass_epi(XArgs<T> a, int epi, hipStream_t s)
{
    switch (epi) {
    case XEPI_PLAIN: return launch_pass<T, AX, K, INTERP, XEPI_PLAIN>(a, s);
    case XEPI_NEG: return launch_pass<T, AX, K, INTERP, XEPI_NEG>(a, s);
    case XEPI_SCALE2: return launch_pass<T, AX, K, INTERP, XEPI_SCALE2>(a, s);
    case XEPI_SUBFROM: return launch_pass<T, AX, K, INTERP, XEPI_SUBFROM>(a, s);
    }
    set_error("transfer: bad epilogue %d", epi);
    return EINVAL_;
}

template <typename T, int K, bool INTERP>
static int run_transfer(const XferPlan& p, const BC& bc, const void* in,
                        void* out, long nin, void* scratch,
                        const XEpiArgs& ea, hipStream_t s)
{
    T* t1 = (T*)scratch;
    T* t2 = t1 + (long)p.nin[0] * p.nin[1] * p.nout[2] * nin;
    const T* src = (const T*)in;
    int cur[3] = {p.nin[0], p.nin[1], p.nin[2]};
    // z, y, x
    for (int step = 0; step < 3; step++) {
        const int ax = 2 - step;
        XArgs<T> a;
        a.in = src;
        a.ilen = cur[ax];
        int d[3] = {cur[0], cur[1], cur[2]};
        d[ax] = p.nout[ax];
        a.d[0] = d[0];
        a.d[1] = d[1];
        a.d[2] = d[2];
        a.istride = ax == 2 ? 1 : (ax == 1 ? (long)cur[2] : (long)cur[1] * cur[2]);
        a.in_grid = (long)cur[0] * cur[1] * cur[2];
        a.out_grid = (long)d[0] * d[1] * d[2];
        a.total = a.out_grid * nin;
        a.off = p.ax[ax].off;
        a.skip0 = p.ax[ax].skip0;
        a.periodic = bc.periodic[ax];
        a.ph_hi = bc.phase[ax][0];
        a.ph_lo = bc.phase[ax][1];
        a.negate = ea.negate;
        a.alpha = ea.alpha;
        a.out2 = (T*)ea.out2;
        a.base = (const T*)ea.base;
        const bool last = (step == 2);
        a.out = last ? (T*)out : (step == 0 ? t1 : t2);
        const int epi = last ? ea.epi : XEPI_PLAIN;
        int st;
        if (ax == 2) st = launch_pass_epi<T, 2, K, INTERP>(a, epi, s);
        else if (ax == 1) st = launch_pass_epi<T, 1, K, INTERP>(a, epi, s);
        else st = launch_pass_epi<T, 0, K, INTERP>(a, epi, s);
        if (st != OK) return st;
        src = a.out;
        cur[ax] = p.nout[ax];
    }
    return OK;
}

template <typename T>
static int run_transfer_k(const XferPlan& p, const BC& bc, const void* in,
                          void* out, long nin, void* scratch,
                          const XEpiArgs& ea, hipStream_t s)
{
    const bool I = p.interpolate != 0;
    switch (p.K) {
    case 2: return I ? run_transfer<T, 2, true>(p, bc, in, out, nin, scratch, ea, s)
                     : run_transfer<T, 2, false>(p, bc, in, out, nin, scratch, ea, s);
    case 4: return I ? run_transfer<T, 4, true>(p, bc, in, out, nin, scratch, ea, s)
                     : run_transfer<T, 4, false>(p, bc, in, out, nin, scratch, ea, s);
    case 6: return I ? run_transfer<T, 6, true>(p, bc, in, out, nin, scratch, ea, s)
                     : run_transfer<T, 6, false>(p, bc, in, out, nin, scratch, ea, s);
    case 8: return I ? run_transfer<T, 8, true>(p, bc, in, out, nin, scratch, ea, s)
                     : run_transfer<T, 8, false>(p, bc, in, out, nin, scratch, ea, s);
    }
    set_error("transfer: unsupported K=%d", p.K);
    return EINVAL_;
}

static bool g_xfer_force_passes = false;
static bool g_xfer_march = true;

bool xfer_passes_forced() { return g_xfer_force_passes; }

void xfer_force_passes(bool on) { g_xfer_force_passes = on; }
void xfer_set_march(bool on) { g_xfer_march = on; }

// ------------------------------------------------- kernel choice (timed)
//
// Per transfer shape (plan, data type, batch, input width) the faster of
// the x-marching kernel and the one-point / separable-pass kernels, timed
// once on the call's own input into a scratch output.  Without an entry the
// marching kernel runs.

struct XKey {
    int v[12];
    bool operator<(const XKey& o) const
    {
        return std::lexicographical_compare(v, v + 12, o.v, o.v + 12);
    }
};

static std::mutex g_xc_mu;
static std::map<XKey, int> g_xchoice;  // 0 marching, 1 other kernels
static bool g_xtune = true;

void xfer_set_autotune(bool on)
{
    std::lock_guard<std::mutex> lk(g_xc_mu);
    g_xtune = on;
    if (!on) g_xchoice.clear();
}

static XKey xkey(const XferPlan& p, int dtype, long nin, int in_wide)
{
    XKey k;
    k.v[0] = p.K;
    k.v[1] = p.interpolate;
    for (int c = 0; c < 3; c++) {
        k.v[2 + c] = p.nin[c];
        k.v[5 + c] = p.nout[c];
    }
    k.v[8] = dtype;
    k.v[9] = (int)std::min<long>(nin, 0x7fffffffL);
    k.v[10] = in_wide;
    k.v[11] = 0;
    return k;
}

static int launch_other(const XferPlan& p, int dtype, const BC& bc,
                        const void* in, void* out, long nin, void* scratch,
                        const XEpiArgs& ea, hipStream_t stream);

bool xfer_needs_tuning(const XferPlan& p, int dtype, long nin, int in_wide)
{
    std::lock_guard<std::mutex> lk(g_xc_mu);
    return g_xtune && g_xfer_march && !g_xfer_force_passes &&
           g_xchoice.find(xkey(p, dtype, nin, in_wide)) == g_xchoice.end();
}

int xfer_tune(const XferPlan& p, int dtype, const BC& bc, const void* in,
              void* tune_out, long nin, void* scratch, int in_wide,
              hipStream_t s)
{
    XEpiArgs ea;
    ea.in_wide = in_wide;
    hipEvent_t e0, e1;
    GFD_CHECK_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("transfer tuner: hipEventCreate failed");
        return EHIP;
    }
    float best = 1e30f;
    int choice = 0;
    for (int c = 0; c < 2; c++) {
        auto run = [&]() {
            return c == 0 ? launch_transfer_march(p, dtype, bc, in, tune_out,
                                                  nin, ea, s)
                          : launch_other(p, dtype, bc, in, tune_out, nin,
                                         scratch, ea, s);
        };
        if (run() != OK) {  // this kernel cannot take the shape
            set_error("");
            continue;
        }
        (void)hipEventRecord(e0, s);
        int st = OK;
        for (int r = 0; r < 3 && st == OK; r++) st = run();
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess || st != OK) {
            (void)hipGetLastError();
            set_error("");
            continue;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
            best = ms;
            choice = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(g_xc_mu);
    g_xchoice[xkey(p, dtype, nin, in_wide)] = choice;
    return OK;
}

int launch_transfer(const XferPlan& p, int dtype, const BC& bc,
                    const void* in, void* out, long nin, void* scratch,
                    const XEpiArgs& ea, hipStream_t stream)
{
    if (ea.in_wide && (p.K != 2 || (dtype != 2 && dtype != 3))) {
        set_error("transfer: FP64 input only for FP32 data and K = 2");
        return EINVAL_;
    }
    bool march = g_xfer_march && !g_xfer_force_passes;
    if (march) {
        std::lock_guard<std::mutex> lk(g_xc_mu);
        auto it = g_xchoice.find(xkey(p, dtype, nin, ea.in_wide));
        if (it != g_xchoice.end() && it->second == 1) march = false;
    }
    if (march) {
        // the x-marching fused kernel (xfer_march.hip), any K
        const int st =
            launch_transfer_march(p, dtype, bc, in, out, nin, ea, stream);
        if (st != ENOTSUP_) return st;
        set_error("");
    }
    return launch_other(p, dtype, bc, in, out, nin, scratch, ea, stream);
}

// the one-point-per-thread K = 2 kernels, or the three separable passes
static int launch_other(const XferPlan& p, int dtype, const BC& bc,
                        const void* in, void* out, long nin, void* scratch,
                        const XEpiArgs& ea, hipStream_t stream)
{
    if (dtype == 3 && p.K != 2) {
        set_error("transfer: complex64 data only for K = 2");
        return ENOTSUP_;
    }
    // K = 2 on the separable passes only when forced and given their work
    // space (caller-owned work spaces are sized for the fused kernel)
    if (p.K == 2 &&
        (!g_xfer_force_passes || dtype == 3 || ea.in_wide || !scratch)) {
        if (dtype == 0)
            return launch_fused2<double>(p, bc, in, out, nin, ea, stream);
        if (dtype == 1)
            return launch_fused2<double2>(p, bc, in, out, nin, ea, stream);
        if (dtype == 2)
            return ea.in_wide
                       ? launch_fused2<float, double>(p, bc, in, out, nin, ea,
                                                      stream)
                       : launch_fused2<float>(p, bc, in, out, nin, ea, stream);
        if (dtype == 3)
            return ea.in_wide
                       ? launch_fused2<float2, double2>(p, bc, in, out, nin,
                                                        ea, stream)
                       : launch_fused2<float2>(p, bc, in, out, nin, ea,
                                               stream);
    }
    if (dtype == 0)
        return run_transfer_k<double>(p, bc, in, out, nin, scratch, ea, stream);
    if (dtype == 1)
        return run_transfer_k<double2>(p, bc, in, out, nin, scratch, ea, stream);
    if (dtype == 2)
        return run_transfer_k<float>(p, bc, in, out, nin, scratch, ea, stream);
    set_error("transfer: bad dtype %d", dtype);
    return EINVAL_;
}

}  // namespace gfd
