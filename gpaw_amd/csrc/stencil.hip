// Stencil kernels for gfx950.
//
// Two kernel families compute `out = sum_p c_p in(x + o_p)` over batches of
// grids with the halo taken by index wrap (no padded copies):
//
//  * k_xmarch: axis-only stencils (the reference's cfd == 1 case,
//    gpaw/fd_operators.py:48-52).  A block owns a (TY x TZ) tile of one grid
//    and marches along x (the slowest axis).  Each thread keeps a register
//    queue of 2R+1 planes for its PY rows at one z column, so the x terms
//    come from registers; the current plane is staged in LDS (double
//    buffered, one barrier per plane) for the y and z terms, with halo rows
//    and columns read once per plane.  Rows of the thread's own segment
//    are taken from registers.  Blocks are remapped so that the tiles of one
//    grid land on the same XCD (shared halo lines in that XCD's L2).
//
//  * k_general: any offsets (non-orthorhombic GUC Laplacians, Mehrstellen,
//    gradients in skew cells, tiny grids); one output point per thread with
//    direct, cache-served neighbour loads.
//
// Both finish with one of the epilogues of stencil.h, which fuse the Jacobi
// update, the residual (+ its squared norm) and the preconditioner's
// residual-update into the same pass.
#include "stencil.h"
#include "stencil_dev.h"

#include <algorithm>
#include <cstdarg>
#include <mutex>

namespace gfd {

// ------------------------------------------------------------ x-march

template <typename T>
struct XmArgs {
    const T* in;
    T* out;
    const T* aux;
    int n0, n1, n2;
    long ng;
    int tiles_y, tiles_z, chunks_x, xchunk;
    int nblocks;
    int has_lds;
    int hasx, hasy, hasz;
    double c0, w, s;
    int negate;
    double cp[3][MAXR_FAST];
    double cm[3][MAXR_FAST];
    BC bc;
    double* partial;
};

constexpr int XM_HMAX = 6;    // halo elements per thread per plane
constexpr int XM_MAXT = 512;  // threads per block

template <typename T, int R, int PY, int EPI>
__global__ void __launch_bounds__(XM_MAXT) k_xmarch(XmArgs<T> a)
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
    int h_lds[XM_HMAX], h_off[XM_HMAX];
    bool h_ok[XM_HMAX], h_ph[XM_HMAX];
    double2 h_f[XM_HMAX];
#pragma unroll
    for (int m = 0; m < XM_HMAX; m++) {
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

    // plane loader
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
        for (int m = 0; m < XM_HMAX; m++) {
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

    T q[2 * R + 1][PY];
    T qn[PY];
    T hv[XM_HMAX];
#pragma unroll
    for (int k = 0; k < 2 * R + 1; k++) load_own(xs - R + k, q[k]);
    if (a.has_lds) load_halo(xs, hv);

    double nrm = 0.0;
    const int lcol = tx + R;
    for (int x = xs; x < xe; ++x) {
        T* buf = lds + (x & 1) * (LH * LW);
        if (a.has_lds) {
#pragma unroll
            for (int j = 0; j < PY; j++)
                buf[(sg * PY + j + R) * LW + lcol] = q[R][j];
#pragma unroll
            for (int m = 0; m < XM_HMAX; m++)
                if (h_lds[m] >= 0) buf[h_lds[m]] = hv[m];
        }
        if (x + 1 < xe) {
            load_own(x + R + 1, qn);
            if (a.has_lds) load_halo(x + 1, hv);
        }
        if (a.has_lds) __syncthreads();

        const long obase = b * a.ng + (long)x * plane;
#pragma unroll
        for (int j = 0; j < PY; j++) {
            const int lrow = sg * PY + j + R;
            T acc = zero_of<T>();
            if (a.hasz) {
#pragma unroll
                for (int k = 1; k <= R; k++) {
                    acc = fmac(acc, cpz[k - 1], buf[lrow * LW + lcol + k]);
                    acc = fmac(acc, cmz[k - 1], buf[lrow * LW + lcol - k]);
                }
            }
            if (a.hasy) {
#pragma unroll
                for (int k = 1; k <= R; k++) {
                    const T up = (j + k < PY) ? q[R][(j + k) % PY]
                                              : buf[(lrow + k) * LW + lcol];
                    const T dn = (j - k >= 0) ? q[R][(j - k + PY) % PY]
                                              : buf[(lrow - k) * LW + lcol];
                    acc = fmac(acc, cpy[k - 1], up);
                    acc = fmac(acc, cmy[k - 1], dn);
                }
            }
            if (a.hasx) {
#pragma unroll
                for (int k = 1; k <= R; k++) {
                    acc = fmac(acc, cpx[k - 1], q[R + k][j]);
                    acc = fmac(acc, cmx[k - 1], q[R - k][j]);
                }
            }
            if (active[j]) {
                const long idx = obase + (long)(y0 + sg * PY + j) * n2 + zc;
                double nl = 0.0;
                T r = epilogue<T, EPI>(acc, q[R][j], a.aux, idx, a.c0, a.w,
                                       a.s, a.negate, nl);
                a.out[idx] = r;
                nrm += nl;
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * R; k++)
#pragma unroll
            for (int j = 0; j < PY; j++) q[k][j] = q[k + 1][j];
#pragma unroll
        for (int j = 0; j < PY; j++) q[2 * R][j] = qn[j];
    }
    if (EPI == EPI_RESID && a.partial) block_sum_store(nrm, a.partial, L);
}

// ------------------------------------------------------------ general

template <typename T>
struct GenArgs {
    const T* in;
    T* out;
    const T* aux;
    int n0, n1, n2;
    long ng, total;
    int P, centre;
    const double* coef;
    const int* off;
    double c0, w, s;
    int negate;
    BC bc;
    double* partial;
    const double* pot;
    const double* shift;
};

constexpr int GEN_BLOCK = 256;

template <typename T, int EPI>
__global__ void __launch_bounds__(GEN_BLOCK) k_general(GenArgs<T> a)
{
    typedef typename ScalarOf<T>::type S;
    const long gid = (long)blockIdx.x * GEN_BLOCK + threadIdx.x;
    double nrm = 0.0;
    if (gid < a.total) {
        long b;
        int i0, i1, i2;
        flat_split(gid, a.total, a.ng, a.n1, a.n2, b, i0, i1, i2);
        const T* __restrict__ g = a.in + b * a.ng;
        T acc = zero_of<T>();
        for (int c = 0; c < a.P; c++) {
            if (c == a.centre) continue;
            int j0 = i0 + a.off[3 * c], j1 = i1 + a.off[3 * c + 1],
                j2 = i2 + a.off[3 * c + 2];
            double2 f = make_double2(1.0, 0.0);
            bool ph = false;
            bool ok = wrap_full(j0, a.n0, a.bc.periodic[0], a.bc.phase[0][0],
                                a.bc.phase[0][1], f, ph);
            ok = ok && wrap_full(j1, a.n1, a.bc.periodic[1],
                                 a.bc.phase[1][0], a.bc.phase[1][1], f, ph);
            ok = ok && wrap_full(j2, a.n2, a.bc.periodic[2],
                                 a.bc.phase[2][0], a.bc.phase[2][1], f, ph);
            if (!ok) continue;
            T v = g[((long)j0 * a.n1 + j1) * a.n2 + j2];
            v = scale_by<T>(v, f, ph);
            acc = fmac(acc, (S)a.coef[c], v);
        }
        const T v = g[((long)i0 * a.n1 + i1) * a.n2 + i2];
        if constexpr (EPI == EPI_POT) {
            typedef typename ScalarOf<T>::type S;
            T r = fmac(fmac(acc, (S)a.c0, v), (S)a.pot[gid - b * a.ng], v);
            if (a.shift) r = fmac(r, -(S)a.shift[b], v);
            a.out[gid] = r;
        } else if constexpr (EPI == EPI_RESN) {
            ((float*)(void*)a.out)[gid] = (float)epilogue<T, EPI_RESID>(
                acc, v, a.aux, gid, a.c0, a.w, a.s, a.negate, nrm);
        } else if constexpr (EPI == EPI_UPDW) {
            typedef typename WideOf<T>::type WT;
            const T av = to_narrow(((const WT*)(const void*)a.aux)[gid]);
            const T r = epi_apply<T, EPI_UPDATE>(acc, v, av, a.c0, a.w, a.s,
                                                 a.negate, nrm);
            ((WT*)(void*)a.out)[gid] = to_wide(r);
        } else {
            a.out[gid] = epilogue<T, EPI>(acc, v, a.aux, gid, a.c0, a.w, a.s,
                                          a.negate, nrm);
        }
    }
    if ((EPI == EPI_RESID || EPI == EPI_RESN) && a.partial)
        block_sum_store(nrm, a.partial, blockIdx.x);
}

// ------------------------------------------------------------ reduction

__global__ void __launch_bounds__(1024)
k_sum_partials(const double* __restrict__ p, long np, double* out)
{
    double s = 0.0;
    for (long i = threadIdx.x; i < np; i += 1024) s += p[i];
    __shared__ double red[16];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < 16; i++) t += red[i];
        *out = t;
    }
}

int launch_sum_partials(const double* partial, long np, double* result,
                        hipStream_t stream)
{
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, stream,
                       partial, np, result);
    GFD_LAUNCH_CHECK();
    return OK;
}

// ------------------------------------------------------------ host side

static std::mutex g_tile_mu;
static bool g_tile_override = false;
static XmTile g_tile;

void xm_set_override(const XmTile* t)
{
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile_override = (t != nullptr);
    if (t) g_tile = *t;
}

// rows per thread: complex R >= 3 with 4 rows exceeds 256 VGPRs
static int py_for(int R, int dtype) { return (dtype == 1 && R >= 2) ? 2 : 4; }

static bool xm_tile_ok(const XmTile& t, int R, int dtype, const int n[3])
{
    const int NT = t.tz * t.nseg;
    const int TY = t.nseg * t.py;
    const int es = dtype == 1 ? 16 : (dtype == 2 ? 4 : 8);
    return NT <= XM_MAXT && 2 * R * (t.tz + TY) <= XM_HMAX * NT &&
           t.tz <= n[2] && R <= n[0] && R <= n[1] && R <= n[2] &&
           2L * (TY + 2 * R) * (t.tz + 2 * R) * es <= 150 * 1024;
}

XmTile xm_default_tile(int R, int dtype, const int n[3], long nin)
{
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        if (g_tile_override) return g_tile;
    }
    XmTile t;
    const int ntz = (n[2] + 127) / 128;
    t.tz = (n[2] + ntz - 1) / ntz;
    const int es = dtype == 1 ? 16 : (dtype == 2 ? 4 : 8);
    // rows per thread: the largest that fits the halo budget and registers
    for (int py = py_for(R, dtype); py >= 1; py >>= 1) {
        t.py = py;
        const int maxseg = std::max(1, XM_MAXT / t.tz);
        const int want = (n[1] + t.py - 1) / t.py;
        t.nseg = std::max(1, std::min(maxseg, want));
        while (t.nseg > 1 &&
               2L * (t.nseg * t.py + 2 * R) * (t.tz + 2 * R) * es > 96 * 1024)
            t.nseg--;
        t.xchunk = n[0];
        if (xm_tile_ok(t, R, dtype, n)) break;
    }
    const long tiles = (long)ceil_div(n[1], t.nseg * t.py) *
                       ceil_div(n[2], t.tz) * nin;
    int chunks = 1;
    if (tiles < 2048) {
        chunks = (int)std::min<long>(ceil_div(2048, tiles),
                                     std::max(1, n[0] / (4 * R)));
        chunks = std::max(1, chunks);
    }
    t.xchunk = ceil_div(n[0], chunks);
    return t;
}

bool axis_stencil_fits(const AxisStencil& st, int dtype, const int n[3],
                       long nin)
{
    return xm_tile_ok(xm_default_tile(st.R, dtype, n, nin), st.R, dtype, n);
}

template <typename T, int R, int PY, int EPI>
static int launch_xm_t(const AxisStencil& st, const XmTile& tl,
                       const int n[3], const BC& bc, const void* in,
                       void* out, long nin, const EpiArgs& ea,
                       hipStream_t stream, int* nblocks_out)
{
    XmArgs<T> a;
    a.in = (const T*)in;
    a.out = (T*)out;
    a.aux = (const T*)ea.aux;
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
        set_error("too many blocks (%ld)", nb);
        return ENOTSUP_;
    }
    a.nblocks = (int)nb;
    a.hasx = st.has[0];
    a.hasy = st.has[1];
    a.hasz = st.has[2];
    a.has_lds = st.has[1] || st.has[2];
    a.c0 = ea.c0;
    a.w = ea.w;
    a.s = ea.s;
    a.negate = ea.negate;
    for (int ax = 0; ax < 3; ax++)
        for (int k = 0; k < MAXR_FAST; k++) {
            a.cp[ax][k] = st.cp[ax][k];
            a.cm[ax][k] = st.cm[ax][k];
        }
    a.bc = bc;
    a.partial = ea.partial;
    const int NT = tl.tz * tl.nseg;
    const int H = 2 * R * tl.tz + 2 * R * TY;
    if (NT > XM_MAXT || (a.has_lds && H > XM_HMAX * NT) || tl.tz > n[2] + 0 ||
        R > n[0] || R > n[1] || R > n[2]) {
        set_error("x-march tile (tz=%d nseg=%d py=%d) does not fit", tl.tz,
                  tl.nseg, PY);
        return ENOTSUP_;
    }
    const size_t shm =
        a.has_lds ? 2 * (size_t)(TY + 2 * R) * (tl.tz + 2 * R) * sizeof(T) : 0;
    // more than the default 64 KB of dynamic LDS must be opted into
    static size_t attr_bytes = 64 * 1024;
    if (shm > attr_bytes) {
        hipError_t e = hipFuncSetAttribute(
            (const void*)k_xmarch<T, R, PY, EPI>,
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("x-march: %zu bytes of LDS not available (%s)", shm,
                      hipGetErrorString(e));
            return ENOTSUP_;
        }
        attr_bytes = shm;
    }
    (void)hipGetLastError();
    hipLaunchKernelGGL((k_xmarch<T, R, PY, EPI>), dim3((unsigned)nb),
                       dim3(tl.tz, tl.nseg), shm, stream, a);
    GFD_LAUNCH_CHECK();
    if (nblocks_out) *nblocks_out = (int)nb;
    return OK;
}

template <typename T, int R, int EPI>
static int launch_xm_py(const AxisStencil& st, const XmTile& tl,
                        const int n[3], const BC& bc, const void* in,
                        void* out, long nin, const EpiArgs& ea,
                        hipStream_t stream, int* nbo)
{
    switch (tl.py) {
    case 1: return launch_xm_t<T, R, 1, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 2: return launch_xm_t<T, R, 2, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 4: return launch_xm_t<T, R, 4, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    }
    set_error("unsupported rows per thread %d", tl.py);
    return ENOTSUP_;
}

template <typename T, int EPI>
static int launch_xm_r(const AxisStencil& st, const XmTile& tl,
                       const int n[3], const BC& bc, const void* in,
                       void* out, long nin, const EpiArgs& ea,
                       hipStream_t stream, int* nbo)
{
    switch (st.R) {
    case 1: return launch_xm_py<T, 1, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 2: return launch_xm_py<T, 2, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 3: return launch_xm_py<T, 3, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 4: return launch_xm_py<T, 4, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    }
    set_error("unsupported stencil radius %d", st.R);
    return ENOTSUP_;
}

long axis_stencil_max_blocks(const AxisStencil& st, int dtype, const int n[3],
                             long nin)
{
    XmTile tl = xm_default_tile(st.R, dtype, n, nin);
    return (long)ceil_div(n[1], tl.nseg * tl.py) * ceil_div(n[2], tl.tz) *
           ceil_div(n[0], tl.xchunk) * nin;
}

long general_stencil_max_blocks(const int n[3], long nin)
{
    return ((long)n[0] * n[1] * n[2] * nin + GEN_BLOCK - 1) / GEN_BLOCK;
}

int launch_axis_stencil(const AxisStencil& st, int dtype, const int n[3],
                        const BC& bc, const void* in, void* out, long nin,
                        int epi, const EpiArgs& ea, hipStream_t stream,
                        int* nblocks_out)
{
    const XmTile tl = xm_default_tile(st.R, dtype, n, nin);
    if (dtype == 0) {
        switch (epi) {
        case EPI_PLAIN: return launch_xm_r<double, EPI_PLAIN>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RELAX: return launch_xm_r<double, EPI_RELAX>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RESID: return launch_xm_r<double, EPI_RESID>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_UPDATE: return launch_xm_r<double, EPI_UPDATE>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        }
    } else if (dtype == 1) {
        switch (epi) {
        case EPI_PLAIN: return launch_xm_r<double2, EPI_PLAIN>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RESID: return launch_xm_r<double2, EPI_RESID>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_UPDATE: return launch_xm_r<double2, EPI_UPDATE>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        }
    } else if (dtype == 2) {
        switch (epi) {
        case EPI_PLAIN: return launch_xm_r<float, EPI_PLAIN>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RELAX: return launch_xm_r<float, EPI_RELAX>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RESID: return launch_xm_r<float, EPI_RESID>(st, tl, n, bc, in, out, nin, ea, stream, nblocks_out);
        }
    }
    set_error("x-march: unsupported dtype %d / epilogue %d", dtype, epi);
    return ENOTSUP_;
}

template <typename T, int EPI>
static int launch_gen_t(const GeneralStencil& st, const int n[3],
                        const BC& bc, const void* in, void* out, long nin,
                        const EpiArgs& ea, hipStream_t stream, int* nbo)
{
    GenArgs<T> a;
    a.in = (const T*)in;
    a.out = (T*)out;
    a.aux = (const T*)ea.aux;
    a.n0 = n[0];
    a.n1 = n[1];
    a.n2 = n[2];
    a.ng = (long)n[0] * n[1] * n[2];
    a.total = a.ng * nin;
    a.P = st.P;
    a.centre = st.centre;
    a.coef = st.coef;
    a.off = st.off;
    a.c0 = ea.c0;
    a.w = ea.w;
    a.s = ea.s;
    a.negate = ea.negate;
    a.bc = bc;
    a.partial = ea.partial;
    a.pot = ea.pot;
    a.shift = ea.shift;
    const long nb = (a.total + GEN_BLOCK - 1) / GEN_BLOCK;
    if (nb <= 0) return OK;
    if (nb > 0x7fffffffL) {
        set_error("too many blocks (%ld)", nb);
        return ENOTSUP_;
    }
    hipLaunchKernelGGL((k_general<T, EPI>), dim3((unsigned)nb),
                       dim3(GEN_BLOCK), 0, stream, a);
    GFD_LAUNCH_CHECK();
    if (nbo) *nbo = (int)nb;
    return OK;
}

int launch_general_stencil(const GeneralStencil& st, int dtype,
                           const int n[3], const BC& bc, const void* in,
                           void* out, long nin, int epi, const EpiArgs& ea,
                           hipStream_t stream, int* nblocks_out)
{
    if (dtype == 0) {
        switch (epi) {
        case EPI_PLAIN: return launch_gen_t<double, EPI_PLAIN>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RELAX: return launch_gen_t<double, EPI_RELAX>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RESID: return launch_gen_t<double, EPI_RESID>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_UPDATE: return launch_gen_t<double, EPI_UPDATE>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_POT: return launch_gen_t<double, EPI_POT>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RESN: return launch_gen_t<double, EPI_RESN>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        }
    } else if (dtype == 1) {
        switch (epi) {
        case EPI_POT: return launch_gen_t<double2, EPI_POT>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_PLAIN: return launch_gen_t<double2, EPI_PLAIN>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RESID: return launch_gen_t<double2, EPI_RESID>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_UPDATE: return launch_gen_t<double2, EPI_UPDATE>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        }
    } else if (dtype == 2) {
        switch (epi) {
        case EPI_PLAIN: return launch_gen_t<float, EPI_PLAIN>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RELAX: return launch_gen_t<float, EPI_RELAX>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RESID: return launch_gen_t<float, EPI_RESID>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_UPDATE: return launch_gen_t<float, EPI_UPDATE>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_UPDW: return launch_gen_t<float, EPI_UPDW>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        }
    } else if (dtype == 3) {
        switch (epi) {
        case EPI_PLAIN: return launch_gen_t<float2, EPI_PLAIN>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_RESID: return launch_gen_t<float2, EPI_RESID>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_UPDATE: return launch_gen_t<float2, EPI_UPDATE>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        case EPI_UPDW: return launch_gen_t<float2, EPI_UPDW>(st, n, bc, in, out, nin, ea, stream, nblocks_out);
        }
    }
    set_error("general stencil: unsupported dtype %d / epilogue %d", dtype,
              epi);
    return ENOTSUP_;
}

}  // namespace gfd
