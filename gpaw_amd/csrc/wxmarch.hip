// Weighted operators whose stencils are all axis-only (the dielectric
// operator eps lap + grad eps . grad of the solvation Poisson solver):
// out(x) = sum_w W_w(x) (c0_w in(x) + sum_{axis, k} c_{w,axis,+-k}
// in(x +- k e_axis)), and the weighted Jacobi sweep
// nxt = (1 - w) cur + w (src - sum_w W_w acc_w) / sum_w W_w c0_w
// (c/bmgs/wfd.cpp, c/bmgs/wrelax.cpp).
//
// The data movement is that of one axis stencil on the tiled x-march: a
// block owns a (TY x TZ) tile of one grid and marches along x; each thread
// keeps a register queue of 2R + 1 planes for its PY rows at one z column
// (the x terms), the current plane sits in LDS (double buffered) for the y
// and z terms, halo rows and columns are read once per plane.  Every
// neighbour value is read once and applied to the NW stencils' partial
// sums, so the weighted operator costs one pass over the data (plus the
// weights, one read per point each) instead of one generic load per
// stencil point.
#include "wstencil.h"

#include "stencil_dev.h"

namespace gfd {

constexpr int WX_HMAX = 6;    // halo elements per thread per plane
constexpr int WX_MAXT = 512;  // threads per block

template <typename T>
struct WxArgs {
    const T* in;
    T* out;
    const double* src;  // relax: right-hand side (real data)
    int n0, n1, n2;
    long ng;
    int tiles_y, tiles_z, chunks_x, xchunk;
    int nblocks;
    int has_lds;
    int hasx, hasy, hasz;
    double w;           // relax weight
    WAxis st;
    BC bc;
};

template <typename T, int R, int PY, int NW, bool RELAX>
__global__ void __launch_bounds__(WX_MAXT) k_wxmarch(WxArgs<T> a)
{
    typedef typename ScalarOf<T>::type S;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);

    const int L = xcd_remap(blockIdx.x, a.nblocks);
    int rem = L;
    const int tz = rem % a.tiles_z;
    rem /= a.tiles_z;
    const int ty = rem % a.tiles_y;
    rem /= a.tiles_y;
    const int cx = rem % a.chunks_x;
    const long b = rem / a.chunks_x;

    const int tx = threadIdx.x, sg = threadIdx.y;
    const int TZ = blockDim.x, NSEG = blockDim.y;
    const int TY = NSEG * PY;
    const int NT = TZ * NSEG;
    const int LW = TZ + 2 * R, LH = TY + 2 * R;
    const int n0 = a.n0, n1 = a.n1, n2 = a.n2;
    const long plane = (long)n1 * n2;
    const int z0 = tz * TZ, y0 = ty * TY;
    const int xs = cx * a.xchunk;
    const int xe = min(xs + a.xchunk, n0);
    const T* __restrict__ gin = a.in + b * a.ng;
    const BC& bc = a.bc;

    // own column and rows (wrapped for ragged tiles)
    const int zc = z0 + tx;
    int zw = zc;
    double2 fz = make_double2(1.0, 0.0);
    bool phz = false;
    const bool okz = wrap_full(zw, n2, bc.periodic[2], bc.phase[2][0],
                               bc.phase[2][1], fz, phz);
    int rowoff[PY];
    bool valid[PY], active[PY], phyz[PY];
    double2 fyz[PY];
#pragma unroll
    for (int j = 0; j < PY; j++) {
        const int yr = y0 + sg * PY + j;
        int yw = yr;
        double2 f = fz;
        bool ph = phz;
        const bool oky = wrap_full(yw, n1, bc.periodic[1], bc.phase[1][0],
                                   bc.phase[1][1], f, ph);
        valid[j] = oky && okz;
        rowoff[j] = yw * n2 + zw;
        fyz[j] = f;
        phyz[j] = ph;
        active[j] = (yr < n1) && (zc < n2);
    }

    // halo elements of this thread (rows above/below, columns left/right)
    const int H = a.has_lds ? (2 * R * TZ + 2 * R * TY) : 0;
    const int t = sg * TZ + tx;
    int h_lds[WX_HMAX], h_off[WX_HMAX];
    bool h_ok[WX_HMAX], h_ph[WX_HMAX];
    double2 h_f[WX_HMAX];
#pragma unroll
    for (int m = 0; m < WX_HMAX; m++) {
        const int h = t + m * NT;
        h_lds[m] = -1;
        h_ok[m] = false;
        h_ph[m] = false;
        h_off[m] = 0;
        h_f[m] = make_double2(1.0, 0.0);
        if (h < H) {
            int row, col;
            if (h < 2 * R * TZ) {
                const int hr = h / TZ, hc = h % TZ;
                row = hr < R ? hr - R : TY + hr - R;
                col = hc;
            } else {
                const int h2 = h - 2 * R * TZ;
                const int hr = h2 / (2 * R), hc = h2 % (2 * R);
                row = hr;
                col = hc < R ? hc - R : TZ + hc - R;
            }
            h_lds[m] = (row + R) * LW + (col + R);
            int yw = y0 + row, zw2 = z0 + col;
            double2 f = make_double2(1.0, 0.0);
            bool ph = false;
            bool ok = wrap_full(zw2, n2, bc.periodic[2], bc.phase[2][0],
                                bc.phase[2][1], f, ph);
            ok = ok && wrap_full(yw, n1, bc.periodic[1], bc.phase[1][0],
                                 bc.phase[1][1], f, ph);
            h_ok[m] = ok;
            h_off[m] = yw * n2 + zw2;
            h_f[m] = f;
            h_ph[m] = ph;
        }
    }

    auto plane_base = [&](int x, bool& okx, double2& fx, bool& phx) -> long {
        int xw = x;
        fx = make_double2(1.0, 0.0);
        phx = false;
        okx = wrap_full(xw, n0, bc.periodic[0], bc.phase[0][0],
                        bc.phase[0][1], fx, phx);
        return (long)xw * plane;
    };
    auto load_own = [&](int x, T* dst) {
        bool okx, phx;
        double2 fx;
        const long base = plane_base(x, okx, fx, phx);
#pragma unroll
        for (int j = 0; j < PY; j++) {
            T v = zero_of<T>();
            if (okx && valid[j]) {
                v = gin[base + rowoff[j]];
                if constexpr (IsCplx<T>::value) {
                    if (phx || phyz[j]) {
                        double2 f = phx ? (phyz[j] ? cmul(fyz[j], fx) : fx)
                                        : fyz[j];
                        v = scale_by<T>(v, f, true);
                    }
                }
            }
            dst[j] = v;
        }
    };
    auto load_halo = [&](int x, T* dst) {
        bool okx, phx;
        double2 fx;
        const long base = plane_base(x, okx, fx, phx);
#pragma unroll
        for (int m = 0; m < WX_HMAX; m++) {
            T v = zero_of<T>();
            if (okx && h_ok[m]) {
                v = gin[base + h_off[m]];
                if constexpr (IsCplx<T>::value) {
                    if (phx || h_ph[m]) {
                        double2 f = phx ? (h_ph[m] ? cmul(h_f[m], fx) : fx)
                                        : h_f[m];
                        v = scale_by<T>(v, f, true);
                    }
                }
            }
            dst[m] = v;
        }
    };

    T q[2 * R + 1][PY];
    T qn[PY];
    T hv[WX_HMAX];
#pragma unroll
    for (int k = 0; k < 2 * R + 1; k++) load_own(xs - R + k, q[k]);
    if (a.has_lds) load_halo(xs, hv);

    const int lcol = tx + R;
    for (int x = xs; x < xe; ++x) {
        T* buf = lds + (x & 1) * (LH * LW);
        if (a.has_lds) {
#pragma unroll
            for (int j = 0; j < PY; j++)
                buf[(sg * PY + j + R) * LW + lcol] = q[R][j];
#pragma unroll
            for (int m = 0; m < WX_HMAX; m++)
                if (h_lds[m] >= 0) buf[h_lds[m]] = hv[m];
        }
        if (x + 1 < xe) {
            load_own(x + R + 1, qn);
            if (a.has_lds) load_halo(x + 1, hv);
        }
        if (a.has_lds) __syncthreads();

#pragma unroll
        for (int j = 0; j < PY; j++) {
            const int lrow = sg * PY + j + R;
            T acc[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) acc[w] = zero_of<T>();
            // each neighbour value once, into every stencil's partial sum
            if (a.hasz) {
#pragma unroll
                for (int k = 1; k <= R; k++) {
                    const T up = buf[lrow * LW + lcol + k];
                    const T dn = buf[lrow * LW + lcol - k];
#pragma unroll
                    for (int w = 0; w < NW; w++)
                        if (a.st.hasw[w][2]) {
                            acc[w] = fmac(acc[w], (S)a.st.cp[w][2][k - 1], up);
                            acc[w] = fmac(acc[w], (S)a.st.cm[w][2][k - 1], dn);
                        }
                }
            }
            if (a.hasy) {
#pragma unroll
                for (int k = 1; k <= R; k++) {
                    const T up = (j + k < PY) ? q[R][(j + k) % PY]
                                              : buf[(lrow + k) * LW + lcol];
                    const T dn = (j - k >= 0) ? q[R][(j - k + PY) % PY]
                                              : buf[(lrow - k) * LW + lcol];
#pragma unroll
                    for (int w = 0; w < NW; w++)
                        if (a.st.hasw[w][1]) {
                            acc[w] = fmac(acc[w], (S)a.st.cp[w][1][k - 1], up);
                            acc[w] = fmac(acc[w], (S)a.st.cm[w][1][k - 1], dn);
                        }
                }
            }
            if (a.hasx) {
#pragma unroll
                for (int k = 1; k <= R; k++) {
#pragma unroll
                    for (int w = 0; w < NW; w++)
                        if (a.st.hasw[w][0]) {
                            acc[w] = fmac(acc[w], (S)a.st.cp[w][0][k - 1],
                                          q[R + k][j]);
                            acc[w] = fmac(acc[w], (S)a.st.cm[w][0][k - 1],
                                          q[R - k][j]);
                        }
                }
            }
            if (active[j]) {
                const long g = (long)x * plane + (long)(y0 + sg * PY + j) * n2 +
                               zc;  // point in the grid (weights)
                const T v = q[R][j];
                if constexpr (RELAX) {
                    // real data only (rejected on the host otherwise)
                    double xs_ = 0.0, coef = 0.0;
#pragma unroll
                    for (int w = 0; w < NW; w++) {
                        const double wt = a.st.wts[w][g];
                        xs_ = fma(wt, (double)acc[w], xs_);
                        coef = fma(wt, a.st.c0[w], coef);
                    }
                    a.out[b * a.ng + g] =
                        (T)((1.0 - a.w) * (double)v +
                            a.w * (a.src[g] - xs_) / coef);
                } else {
                    T r = zero_of<T>();
#pragma unroll
                    for (int w = 0; w < NW; w++) {
                        const T tw = fmac(acc[w], (S)a.st.c0[w], v);
                        r = fmac(r, (S)a.st.wts[w][g], tw);
                    }
                    a.out[b * a.ng + g] = r;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * R; k++)
#pragma unroll
            for (int j = 0; j < PY; j++) q[k][j] = q[k + 1][j];
#pragma unroll
        for (int j = 0; j < PY; j++) q[2 * R][j] = qn[j];
    }
}

// ------------------------------------------------------------ host side

static bool wx_tile_ok(const XmTile& t, int R, int es, const int n[3])
{
    const int NT = t.tz * t.nseg;
    const int TY = t.nseg * t.py;
    return NT <= WX_MAXT && NT >= 64 &&
           2 * R * (t.tz + TY) <= WX_HMAX * NT && t.tz <= n[2] &&
           R <= n[0] && R <= n[1] && R <= n[2] &&
           2L * (TY + 2 * R) * (t.tz + 2 * R) * es <= 64 * 1024;
}

// Tile: whole z rows up to 128 points, then the most rows that keep the
// halo within WX_HMAX per thread and the LDS within 64 KB; x chunks for
// at least ~2048 blocks.
static int g_wx_py = 0, g_wx_xchunk = 0;

void wx_set_tile(int py, int xchunk)
{
    g_wx_py = py;
    g_wx_xchunk = xchunk;
}

static bool wx_tile(int R, int es, const int n[3], long nin, XmTile& t)
{
    const int ntz = (n[2] + 127) / 128;
    t.tz = (n[2] + ntz - 1) / ntz;
    const int py0 = g_wx_py > 0 ? g_wx_py : 2;
    for (int py = py0; py >= 1; py = py == 4 ? 2 : py - 1) {
        t.py = py;
        const int maxseg = std::max(1, WX_MAXT / t.tz);
        const int want = (n[1] + t.py - 1) / t.py;
        t.nseg = std::max(1, std::min(maxseg, want));
        while (t.nseg > 1 && !wx_tile_ok(t, R, es, n)) t.nseg--;
        if (wx_tile_ok(t, R, es, n)) break;
    }
    if (!wx_tile_ok(t, R, es, n)) return false;
    const long tiles = (long)ceil_div(n[1], t.nseg * t.py) *
                       ceil_div(n[2], t.tz) * nin;
    int chunks = 1;
    if (tiles < 2048)
        chunks = std::max(1, (int)std::min<long>(ceil_div(2048, tiles),
                                                 std::max(1, n[0] / (4 * R))));
    t.xchunk = g_wx_xchunk > 0 ? std::min(g_wx_xchunk, n[0])
                               : ceil_div(n[0], chunks);
    return true;
}

template <typename T, int R, int PY, int NW, bool RELAX>
static int wx_launch_t(const WAxis& st, const XmTile& tl, const int n[3],
                       const BC& bc, const void* in, void* out,
                       const double* src, long nin, double w, hipStream_t s)
{
    WxArgs<T> a;
    a.in = (const T*)in;
    a.out = (T*)out;
    a.src = src;
    a.n0 = n[0];
    a.n1 = n[1];
    a.n2 = n[2];
    a.ng = (long)n[0] * n[1] * n[2];
    const int TY = tl.nseg * PY;
    a.tiles_y = ceil_div(n[1], TY);
    a.tiles_z = ceil_div(n[2], tl.tz);
    a.xchunk = tl.xchunk;
    a.chunks_x = ceil_div(n[0], tl.xchunk);
    const long nb = (long)a.tiles_y * a.tiles_z * a.chunks_x * nin;
    if (nb <= 0) return OK;
    if (nb > 0x7fffffffL) {
        set_error("weighted x-march: too many blocks (%ld)", nb);
        return ENOTSUP_;
    }
    a.nblocks = (int)nb;
    a.hasx = st.has[0];
    a.hasy = st.has[1];
    a.hasz = st.has[2];
    a.has_lds = st.has[1] || st.has[2];
    a.w = w;
    a.st = st;
    a.bc = bc;
    const size_t shm =
        a.has_lds ? 2 * (size_t)(TY + 2 * R) * (tl.tz + 2 * R) * sizeof(T) : 0;
    hipLaunchKernelGGL((k_wxmarch<T, R, PY, NW, RELAX>), dim3((unsigned)nb),
                       dim3(tl.tz, tl.nseg), shm, s, a);
    GFD_LAUNCH_CHECK();
    return OK;
}

template <typename T, int NW, bool RELAX>
static int wx_launch_n(const WAxis& st, const XmTile& tl, const int n[3],
                       const BC& bc, const void* in, void* out,
                       const double* src, long nin, double w, hipStream_t s)
{
#define GFD_WX(RR, PP)                                                        \
    if (st.R == RR && tl.py == PP)                                            \
        return wx_launch_t<T, RR, PP, NW, RELAX>(st, tl, n, bc, in, out, src, \
                                                 nin, w, s);
    GFD_WX(1, 1) GFD_WX(1, 2) GFD_WX(1, 4) GFD_WX(2, 1) GFD_WX(2, 2)
    GFD_WX(2, 4) GFD_WX(3, 1) GFD_WX(3, 2) GFD_WX(3, 4) GFD_WX(4, 1)
    GFD_WX(4, 2) GFD_WX(4, 4)
#undef GFD_WX
    set_error("weighted x-march: radius %d / %d rows not built", st.R, tl.py);
    return ENOTSUP_;
}

template <typename T, bool RELAX>
static int wx_launch_w(const WAxis& st, const XmTile& tl, const int n[3],
                       const BC& bc, const void* in, void* out,
                       const double* src, long nin, double w, hipStream_t s)
{
    switch (st.nw) {
    case 1: return wx_launch_n<T, 1, RELAX>(st, tl, n, bc, in, out, src, nin, w, s);
    case 2: return wx_launch_n<T, 2, RELAX>(st, tl, n, bc, in, out, src, nin, w, s);
    case 3: return wx_launch_n<T, 3, RELAX>(st, tl, n, bc, in, out, src, nin, w, s);
    case 4: return wx_launch_n<T, 4, RELAX>(st, tl, n, bc, in, out, src, nin, w, s);
    }
    set_error("weighted x-march: %d stencils", st.nw);
    return ENOTSUP_;
}

static bool g_wx_off = false;

void wx_set_enabled(bool on) { g_wx_off = !on; }

int launch_wxmarch(const WAxis& st, int real, const int n[3], const BC& bc,
                   const void* in, void* out, const double* src, long nin,
                   double w, bool relax, hipStream_t s)
{
    if (g_wx_off || st.nw < 1 || st.nw > WX_NW || st.R < 1 ||
        st.R > MAXR_FAST || (relax && !real)) {
        set_error("weighted x-march: not for this operator");
        return ENOTSUP_;
    }
    XmTile tl;
    if (!wx_tile(st.R, real ? 8 : 16, n, nin, tl)) {
        set_error("weighted x-march: no tile fits this grid");
        return ENOTSUP_;
    }
    if (real)
        return relax ? wx_launch_w<double, true>(st, tl, n, bc, in, out, src,
                                                 nin, w, s)
                     : wx_launch_w<double, false>(st, tl, n, bc, in, out,
                                                  src, nin, w, s);
    return wx_launch_w<double2, false>(st, tl, n, bc, in, out, src, nin, w,
                                       s);
}

}  // namespace gfd
