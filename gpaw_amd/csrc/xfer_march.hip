// Fused 3-D multigrid transfers that march along x (every K = 2, 4, 6, 8):
// the restriction and interpolation of c/bmgs/restrict.cpp:137-156 and
// c/bmgs/interpolate.cpp:70-99 in one kernel, with the reference's order
// of the separable passes (z, then y, then x).
//
// A block owns a tile of TY x TZ coarse (y, z) positions and a chunk of x
// planes.  Plane by plane along x it stages one plane of the input tile,
// halo included, in LDS (wrap and Bloch phases applied by the loader; a
// zero boundary reads 0) and reduces it over y and z for its tile; the x
// pass combines the last few of those plane results, which each thread
// keeps in registers.  Every input value is read from HBM once per chunk
// (plus the tile halo), every output written once.
//
// Restriction (per axis, c = 2 o + off the centre of coarse point o):
//   R a(o) = 1/2 (a[c] + sum_j w_j (a[c - j] + a[c + j])), j = 1, 3, .., K-1
// Interpolation (q = fine index + skip0, c = q / 2 + off):
//   even q: a[c];  odd q: sum_m w_m (a[c - m] + a[c + 1 + m]), m < K / 2
// with the exact dyadic weights of transfer.hip.
#include <algorithm>
#include <type_traits>

#include "transfer.h"

namespace gfd {

template <int K> struct MW;
template <> struct MW<2> {
    static constexpr double w[1] = {0.5};
};
template <> struct MW<4> {
    static constexpr double w[2] = {0.5625, -0.0625};
};
template <> struct MW<6> {
    static constexpr double w[3] = {0.58593750, -0.09765625, 0.01171875};
};
template <> struct MW<8> {
    static constexpr double w[4] = {0.59814453125, -0.11962890625,
                                    0.02392578125, -0.00244140625};
};

template <typename T> struct MIsCplx { static constexpr bool value = false; };
template <> struct MIsCplx<double2> { static constexpr bool value = true; };
template <> struct MIsCplx<float2> { static constexpr bool value = true; };

template <typename T, typename TI>
__device__ inline T mnarrow(TI v)
{
    if constexpr (std::is_same<T, TI>::value) {
        return v;
    } else if constexpr (MIsCplx<T>::value) {
        return T{(float)v.x, (float)v.y};
    } else {
        return (T)v;
    }
}

template <typename T, typename TI>
struct MArgs {
    const TI* in;
    T* out;
    T* out2;
    const T* base;
    int ni[3], no[3], off[3], skip0[3], periodic[3];
    double2 ph[3][2];
    long in_grid, out_grid;
    int tiles_y, tiles_z;  // tiles of the (y, z) plane
    int xchunk, nchunks;   // x planes per block (coarse planes or fine
                           // plane pairs), chunks per grid
    int fy, fz;            // staged input tile (rows, columns)
    int negate;
    double alpha;
};

// Index along one axis mapped into the grid with the accumulated Bloch
// factor; false for a zero boundary.  Any number of periods.
__device__ inline bool mwrap(int& i, int n, int periodic, double2 ph_hi,
                             double2 ph_lo, double2& f, bool& phased)
{
    if (i >= 0 && i < n) return true;
    if (!periodic) return false;
    while (i < 0) {
        i += n;
        f = phased ? cmul(ph_lo, f) : ph_lo;
        phased = true;
    }
    while (i >= n) {
        i -= n;
        f = phased ? cmul(ph_hi, f) : ph_hi;
        phased = true;
    }
    return true;
}

constexpr int MAXE = 12;  // staged elements per thread and plane

// Load this thread's share of input plane x, rows y0 .. y0 + fy - 1,
// columns z0 .. z0 + fz - 1 of grid g (row-major fy x fz, element
// e = tid + k nthreads) into registers: the next plane is fetched while
// the current one is reduced.
template <typename T, typename TI>
__device__ inline void mfetch(const MArgs<T, TI>& a, const TI* __restrict__ g,
                              int x, int y0, int z0, T (&v)[MAXE])
{
    double2 fx = make_double2(1.0, 0.0);
    bool px = false;
    int xi = x;
    const bool xok = mwrap(xi, a.ni[0], a.periodic[0], a.ph[0][0],
                           a.ph[0][1], fx, px);
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    int r = tid / a.fz, c = tid - r * a.fz;
    const int dr = nt / a.fz, dc = nt - dr * a.fz;
#pragma unroll
    for (int k = 0; k < MAXE; k++) {
        T u = zero_of<T>();
        if (xok && r < a.fy) {
            double2 f = fx;
            bool p = px;
            int yi = y0 + r, zi = z0 + c;
            if (mwrap(yi, a.ni[1], a.periodic[1], a.ph[1][0], a.ph[1][1], f,
                      p) &&
                mwrap(zi, a.ni[2], a.periodic[2], a.ph[2][0], a.ph[2][1], f,
                      p)) {
                TI w = g[((long)xi * a.ni[1] + yi) * a.ni[2] + zi];
                if constexpr (MIsCplx<TI>::value)
                    if (p) w = apply_phase(w, f);
                u = mnarrow<T, TI>(w);
            }
        }
        v[k] = u;
        r += dr;
        c += dc;
        if (c >= a.fz) {
            c -= a.fz;
            r++;
        }
    }
}

template <typename T, typename TI>
__device__ inline void mput(const MArgs<T, TI>& a, const T (&v)[MAXE],
                            T* lds)
{
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int n = a.fy * a.fz;
#pragma unroll
    for (int k = 0; k < MAXE; k++) {
        const int e = tid + k * nt;
        if (e < n) lds[e] = v[k];
    }
}

template <typename T, int EPI, typename TI>
__device__ inline void mstore(const MArgs<T, TI>& a, long idx, T v)
{
    typedef typename ScalarOf<T>::type S;
    if constexpr (EPI == XEPI_PLAIN) {
        a.out[idx] = v;
    } else if constexpr (EPI == XEPI_NEG) {
        a.out[idx] = -v;
    } else if constexpr (EPI == XEPI_SCALE2) {
        if (a.negate) v = -v;
        a.out[idx] = v;
        a.out2[idx] = (S)a.alpha * v;
    } else {
        a.out[idx] = a.base[idx] - v;
    }
}

// block -> (grid, x chunk, y tile, z tile)
template <typename T, typename TI>
__device__ inline void mblock(const MArgs<T, TI>& a, int& b, int& chunk,
                              int& ty, int& tz)
{
    long k = blockIdx.x;
    tz = (int)(k % a.tiles_z);
    k /= a.tiles_z;
    ty = (int)(k % a.tiles_y);
    k /= a.tiles_y;
    chunk = (int)(k % a.nchunks);
    b = (int)(k / a.nchunks);
}

// 1-D restriction of the K + 1 taps around the centre at p[0] (stride s)
template <typename T, int K>
__device__ inline T mrestrict1(const T* p, int s)
{
    typedef typename ScalarOf<T>::type S;
    T acc = p[0];
#pragma unroll
    for (int m = 0; m < K / 2; m++) {
        const int j = 2 * m + 1;
        acc = fmac(acc, (S)MW<K>::w[m], p[j * s] + p[-j * s]);
    }
    return (S)0.5 * acc;
}

template <typename T, int K, int EPI, typename TI>
__global__ void __launch_bounds__(256) k_restrict_march(MArgs<T, TI> a)
{
    typedef typename ScalarOf<T>::type S;
    extern __shared__ __align__(16) unsigned char m_smem[];
    T* lds = (T*)m_smem;
    int b, chunk, tyi, tzi;
    mblock(a, b, chunk, tyi, tzi);
    const int TY = blockDim.y, TZ = blockDim.x;
    const int oy = tyi * TY + threadIdx.y, oz = tzi * TZ + threadIdx.x;
    // first input row / column of the staged tile
    const int y0 = 2 * tyi * TY + a.off[1] - (K - 1);
    const int z0 = 2 * tzi * TZ + a.off[2] - (K - 1);
    const int xa = chunk * a.xchunk;
    const int xb = min(a.no[0], xa + a.xchunk);
    if (xa >= xb) return;
    const TI* __restrict__ g = a.in + (long)b * a.in_grid;
    const bool mine = oy < a.no[1] && oz < a.no[2];
    // centre of this thread's taps in the tile
    const T* cen = lds + (2 * threadIdx.y + K - 1) * a.fz +
                   (2 * threadIdx.x + K - 1);
    // the yz-restricted values of the last 2K - 1 input planes
    T q[2 * K - 1];
    const int fx0 = 2 * xa + a.off[0] - (K - 1);
    const int fx1 = 2 * (xb - 1) + a.off[0] + (K - 1);
    T nxt[MAXE];
    mfetch(a, g, fx0, y0, z0, nxt);
    for (int fx = fx0; fx <= fx1; fx++) {
        __syncthreads();  // previous plane fully read
        mput(a, nxt, lds);
        __syncthreads();
        if (fx < fx1) mfetch(a, g, fx + 1, y0, z0, nxt);
        // z pass on the K + 1 rows of the y taps, then the y pass
        T zr[K + 1];
        zr[0] = mrestrict1<T, K>(cen, 1);
#pragma unroll
        for (int m = 0; m < K / 2; m++) {
            const int j = 2 * m + 1;
            zr[1 + 2 * m] = mrestrict1<T, K>(cen + j * a.fz, 1);
            zr[2 + 2 * m] = mrestrict1<T, K>(cen - j * a.fz, 1);
        }
        T acc = zr[0];
#pragma unroll
        for (int m = 0; m < K / 2; m++)
            acc = fmac(acc, (S)MW<K>::w[m], zr[1 + 2 * m] + zr[2 + 2 * m]);
#pragma unroll
        for (int i = 0; i < 2 * K - 2; i++) q[i] = q[i + 1];
        q[2 * K - 2] = (S)0.5 * acc;
        const int done = fx - fx0 - (2 * K - 2);
        if (done >= 0 && (done & 1) == 0 && mine) {
            // centre plane q[K - 1]
            T v = q[K - 1];
#pragma unroll
            for (int m = 0; m < K / 2; m++) {
                const int j = 2 * m + 1;
                v = fmac(v, (S)MW<K>::w[m], q[K - 1 + j] + q[K - 1 - j]);
            }
            v = (S)0.5 * v;
            const int ox = xa + done / 2;
            const long idx = (long)b * a.out_grid +
                             ((long)ox * a.no[1] + oy) * a.no[2] + oz;
            mstore<T, EPI, TI>(a, idx, v);
        }
    }
}

// 1-D interpolation (even, odd) from the K taps p[(1 - K/2) s] ..
// p[(K/2) s] around the centre p[0]
template <typename T, int K>
__device__ inline void minterp1(const T* p, int s, T& even, T& odd)
{
    typedef typename ScalarOf<T>::type S;
    even = p[0];
    T acc = (S)MW<K>::w[0] * (p[0] + p[s]);
#pragma unroll
    for (int m = 1; m < K / 2; m++)
        acc = fmac(acc, (S)MW<K>::w[m], p[-m * s] + p[(1 + m) * s]);
    odd = acc;
}

template <typename T, int K, int EPI, typename TI>
__global__ void __launch_bounds__(256) k_interp_march(MArgs<T, TI> a)
{
    typedef typename ScalarOf<T>::type S;
    extern __shared__ __align__(16) unsigned char m_smem[];
    T* lds = (T*)m_smem;
    int b, chunk, tyi, tzi;
    mblock(a, b, chunk, tyi, tzi);
    const int TY = blockDim.y, TZ = blockDim.x;
    // this thread's coarse cell: fine q = 2 p, 2 p + 1 on each axis
    const int p1 = tyi * TY + threadIdx.y, p2 = tzi * TZ + threadIdx.x;
    const int y0 = tyi * TY + a.off[1] - K / 2 + 1;
    const int z0 = tzi * TZ + a.off[2] - K / 2 + 1;
    // fine plane pairs of this chunk
    const int pa = chunk * a.xchunk;
    const int pe = (a.no[0] + a.skip0[0] + 1) / 2;  // pairs in the grid
    const int pb = min(pe, pa + a.xchunk);
    if (pa >= pb) return;
    const TI* __restrict__ g = a.in + (long)b * a.in_grid;
    const T* cen = lds + (threadIdx.y + K / 2 - 1) * a.fz +
                   (threadIdx.x + K / 2 - 1);
    int fi[2][3];  // fine index of (q even, q odd) per axis, -1: none
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const int q1 = 2 * p1 + e - a.skip0[1];
        const int q2 = 2 * p2 + e - a.skip0[2];
        fi[e][1] = (q1 >= 0 && q1 < a.no[1]) ? q1 : -1;
        fi[e][2] = (q2 >= 0 && q2 < a.no[2]) ? q2 : -1;
    }
    // yz-interpolated values [ye][ze] of the last K coarse planes
    T q[K][4];
    const int cx0 = pa + a.off[0] - K / 2 + 1;
    const int cx1 = pb - 1 + a.off[0] + K / 2;
    T nxt[MAXE];
    mfetch(a, g, cx0, y0, z0, nxt);
    for (int cx = cx0; cx <= cx1; cx++) {
        __syncthreads();
        mput(a, nxt, lds);
        __syncthreads();
        if (cx < cx1) mfetch(a, g, cx + 1, y0, z0, nxt);
        // z pass on the K rows of the y taps
        T ze[K], zo[K];
#pragma unroll
        for (int r = 0; r < K; r++)
            minterp1<T, K>(cen + (r - K / 2 + 1) * a.fz, 1, ze[r], zo[r]);
        // y pass on the z results (tap r - K/2 + 1 at index r)
        T ye_e, ye_o, yo_e, yo_o;
        minterp1<T, K>(ze + K / 2 - 1, 1, ye_e, yo_e);
        minterp1<T, K>(zo + K / 2 - 1, 1, ye_o, yo_o);
#pragma unroll
        for (int i = 0; i < K - 1; i++)
#pragma unroll
            for (int k = 0; k < 4; k++) q[i][k] = q[i + 1][k];
        q[K - 1][0] = ye_e;
        q[K - 1][1] = ye_o;
        q[K - 1][2] = yo_e;
        q[K - 1][3] = yo_o;
        const int p0 = cx - K / 2 - a.off[0];  // pair whose taps are complete
        if (p0 < pa) continue;
#pragma unroll
        for (int e0 = 0; e0 < 2; e0++) {
            const int i0 = 2 * p0 + e0 - a.skip0[0];
            if (i0 < 0 || i0 >= a.no[0]) continue;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int e1 = k >> 1, e2 = k & 1;
                if (fi[e1][1] < 0 || fi[e2][2] < 0) continue;
                T v;
                if (e0 == 0) {
                    v = q[K / 2 - 1][k];
                } else {
                    v = (S)MW<K>::w[0] * (q[K / 2 - 1][k] + q[K / 2][k]);
#pragma unroll
                    for (int m = 1; m < K / 2; m++)
                        v = fmac(v, (S)MW<K>::w[m],
                                 q[K / 2 - 1 - m][k] + q[K / 2 + m][k]);
                }
                const long idx = (long)b * a.out_grid +
                                 ((long)i0 * a.no[1] + fi[e1][1]) * a.no[2] +
                                 fi[e2][2];
                mstore<T, EPI, TI>(a, idx, v);
            }
        }
    }
}

// ------------------------------------------------------------------ host

// Tile and chunk of a transfer: TZ threads along z (coarse points or
// cells; a whole row where it has at most 64, else equal tiles), TY along
// y, up to 256 threads, the staged tile within MAXE elements per thread;
// x chunks so that the batch gives every CU several blocks.
template <typename T, typename TI>
static int mplan(MArgs<T, TI>& a, int K, bool interp, long nin, int& TY,
                 int& TZ)
{
    const int cz = interp ? (a.no[2] + a.skip0[2] + 1) / 2 : a.no[2];
    const int cy = interp ? (a.no[1] + a.skip0[1] + 1) / 2 : a.no[1];
    const int tz_tiles = (cz + 63) / 64;
    TZ = (cz + tz_tiles - 1) / tz_tiles;
    TY = std::max(1, std::min(cy, 256 / TZ));
    for (;; TY--) {
        a.fy = interp ? TY + K - 1 : 2 * TY + 2 * K - 3;
        a.fz = interp ? TZ + K - 1 : 2 * TZ + 2 * K - 3;
        if ((long)a.fy * a.fz <= (long)MAXE * TY * TZ || TY == 1) break;
    }
    if ((long)a.fy * a.fz > (long)MAXE * TY * TZ) {
        set_error("transfer: tile does not fit the marching kernel");
        return ENOTSUP_;
    }
    a.tiles_z = tz_tiles;
    a.tiles_y = (cy + TY - 1) / TY;
    const int cx = interp ? (a.no[0] + a.skip0[0] + 1) / 2 : a.no[0];
    // at least ~8 blocks per CU (256 CUs) over the batch, chunks of at
    // least 4 x planes
    const long tiles = (long)a.tiles_y * a.tiles_z * nin;
    long chunks = std::max<long>(1, (2048 + tiles - 1) / tiles);
    chunks = std::min<long>(chunks, std::max(1, cx / 4));
    a.xchunk = (int)((cx + chunks - 1) / chunks);
    a.nchunks = (cx + a.xchunk - 1) / a.xchunk;
    return OK;
}

template <typename T, int K, int EPI, typename TI>
static int launch_march_t(const XferPlan& p, const BC& bc, const void* in,
                          void* out, long nin, const XEpiArgs& ea,
                          hipStream_t s)
{
    MArgs<T, TI> a;
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
    a.negate = ea.negate;
    a.alpha = ea.alpha;
    int TY, TZ;
    GFD_TRY(mplan(a, K, p.interpolate != 0, nin, TY, TZ));
    const size_t lds = sizeof(T) * (size_t)a.fy * a.fz;
    const long nb = (long)a.nchunks * a.tiles_y * a.tiles_z * nin;
    if (nb <= 0) return OK;
    if (nb > 0x7fffffffL || lds > 64 * 1024) {
        set_error("transfer: grid too large for the marching kernel");
        return ENOTSUP_;
    }
    if (p.interpolate)
        hipLaunchKernelGGL((k_interp_march<T, K, EPI, TI>),
                           dim3((unsigned)nb), dim3(TZ, TY), lds, s, a);
    else
        hipLaunchKernelGGL((k_restrict_march<T, K, EPI, TI>),
                           dim3((unsigned)nb), dim3(TZ, TY), lds, s, a);
    GFD_LAUNCH_CHECK();
    return OK;
}

template <typename T, int K, typename TI>
static int launch_march_k(const XferPlan& p, const BC& bc, const void* in,
                          void* out, long nin, const XEpiArgs& ea,
                          hipStream_t s)
{
    switch (ea.epi) {
    case XEPI_PLAIN: return launch_march_t<T, K, XEPI_PLAIN, TI>(p, bc, in, out, nin, ea, s);
    case XEPI_NEG: return launch_march_t<T, K, XEPI_NEG, TI>(p, bc, in, out, nin, ea, s);
    case XEPI_SCALE2: return launch_march_t<T, K, XEPI_SCALE2, TI>(p, bc, in, out, nin, ea, s);
    case XEPI_SUBFROM: return launch_march_t<T, K, XEPI_SUBFROM, TI>(p, bc, in, out, nin, ea, s);
    }
    set_error("transfer: bad epilogue %d", ea.epi);
    return EINVAL_;
}

template <typename T, typename TI = T>
static int launch_march_dt(const XferPlan& p, const BC& bc, const void* in,
                           void* out, long nin, const XEpiArgs& ea,
                           hipStream_t s)
{
    switch (p.K) {
    case 2: return launch_march_k<T, 2, TI>(p, bc, in, out, nin, ea, s);
    case 4: return launch_march_k<T, 4, TI>(p, bc, in, out, nin, ea, s);
    case 6: return launch_march_k<T, 6, TI>(p, bc, in, out, nin, ea, s);
    case 8: return launch_march_k<T, 8, TI>(p, bc, in, out, nin, ea, s);
    }
    set_error("transfer: unsupported K=%d", p.K);
    return EINVAL_;
}

int launch_transfer_march(const XferPlan& p, int dtype, const BC& bc,
                          const void* in, void* out, long nin,
                          const XEpiArgs& ea, hipStream_t s)
{
    if (ea.in_wide) {
        if (dtype == 2)
            return launch_march_dt<float, double>(p, bc, in, out, nin, ea, s);
        if (dtype == 3)
            return launch_march_dt<float2, double2>(p, bc, in, out, nin, ea,
                                                    s);
        set_error("transfer: FP64 input only for FP32 data");
        return EINVAL_;
    }
    switch (dtype) {
    case 0: return launch_march_dt<double>(p, bc, in, out, nin, ea, s);
    case 1: return launch_march_dt<double2>(p, bc, in, out, nin, ea, s);
    case 2: return launch_march_dt<float>(p, bc, in, out, nin, ea, s);
    case 3: return launch_march_dt<float2>(p, bc, in, out, nin, ea, s);
    }
    set_error("transfer: bad dtype %d", dtype);
    return EINVAL_;
}

}  // namespace gfd
