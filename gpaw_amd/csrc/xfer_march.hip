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
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

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
    int xchunk, nchunks;   // x planes per tile (coarse planes or fine
                           // plane pairs), chunks per grid
    long ntiles;           // chunks x y tiles x z tiles x grids
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

constexpr int MAXE = 12;  // largest staged elements per thread and plane

// Index along one axis mapped into the grid (any number of periods); false
// outside a zero boundary.
__device__ inline bool mwrap_i(int& i, int n, int periodic)
{
    if (i >= 0 && i < n) return true;
    if (!periodic) return false;
    while (i < 0) i += n;
    while (i >= n) i -= n;
    return true;
}

// Issue this thread's loads of input plane x, rows y0 .. y0 + fy - 1,
// columns z0 .. z0 + fz - 1 of grid g (row-major fy x fz tile, element
// e = tid + k nthreads), raw: zero outside a zero boundary, no Bloch factor
// or narrowing yet -- those are applied when the values go to LDS (mput),
// so the loads stay in flight while earlier planes are reduced.
template <typename T, typename TI, int E>
__device__ inline void mfetch(const MArgs<T, TI>& a, const TI* __restrict__ g,
                              int x, int y0, int z0, TI (&v)[E])
{
    int xi = x;
    const bool xok = mwrap_i(xi, a.ni[0], a.periodic[0]);
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    int r = tid / a.fz, c = tid - r * a.fz;
    const int dr = nt / a.fz, dc = nt - dr * a.fz;
    const TI* __restrict__ pl = g + (long)xi * a.ni[1] * a.ni[2];
#pragma unroll
    for (int k = 0; k < E; k++) {
        TI u;
        if constexpr (MIsCplx<TI>::value)
            u = TI{0, 0};
        else
            u = (TI)0;
        if (xok && r < a.fy) {
            int yi = y0 + r, zi = z0 + c;
            if (mwrap_i(yi, a.ni[1], a.periodic[1]) &&
                mwrap_i(zi, a.ni[2], a.periodic[2]))
                u = pl[(long)yi * a.ni[2] + zi];
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

// Bloch wrap code of each staged element of this thread (complex data): 0
// inside the grid along y and z, else 1 + (y side) + 3 (z side), side 0
// none, 1 below (phase[.][1]), 2 above (phase[.][0]).  The same for every
// plane of a tile (one wrap at most: periodic axes have n >= K).
template <typename T, typename TI, int E>
__device__ inline void mcodes(const MArgs<T, TI>& a, int y0, int z0,
                              unsigned char (&code)[E])
{
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    int r = tid / a.fz, c = tid - r * a.fz;
    const int dr = nt / a.fz, dc = nt - dr * a.fz;
#pragma unroll
    for (int k = 0; k < E; k++) {
        const int yi = y0 + r, zi = z0 + c;
        const int ys = yi < 0 ? 1 : (yi >= a.ni[1] ? 2 : 0);
        const int zs = zi < 0 ? 1 : (zi >= a.ni[2] ? 2 : 0);
        code[k] = (unsigned char)(ys + 3 * zs);
        r += dr;
        c += dc;
        if (c >= a.fz) {
            c -= a.fz;
            r++;
        }
    }
}

// Store the loaded plane x into the LDS tile: the Bloch factor of wrapped
// elements (complex data), narrowing to T.
template <typename T, typename TI, int E>
__device__ inline void mput(const MArgs<T, TI>& a, int x,
                            const unsigned char (&code)[E], const TI (&v)[E],
                            T* lds)
{
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int n = a.fy * a.fz;
    if constexpr (MIsCplx<TI>::value) {
        // x factor of this plane
        const int xs = x < 0 ? 1 : (x >= a.ni[0] ? 2 : 0);
        const double2 fx = xs == 1 ? a.ph[0][1] : a.ph[0][0];
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int e = tid + k * nt;
            if (e < n) {
                // narrowed first: the factor is applied in T's precision
                T u = mnarrow<T, TI>(v[k]);
                const int cd = code[k];
                if (cd | xs) {
                    typedef typename ScalarOf<T>::type S;
                    const int ys = cd % 3, zs = cd / 3;
                    double2 f = xs ? fx : make_double2(1.0, 0.0);
                    if (ys) f = cmul(f, ys == 1 ? a.ph[1][1] : a.ph[1][0]);
                    if (zs) f = cmul(f, zs == 1 ? a.ph[2][1] : a.ph[2][0]);
                    u = cmul(T{(S)f.x, (S)f.y}, u);
                }
                lds[e] = u;
            }
        }
    } else {
        (void)x;
        (void)code;
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int e = tid + k * nt;
            if (e < n) lds[e] = mnarrow<T, TI>(v[k]);
        }
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

// tile -> (grid, x chunk, y tile, z tile); consecutive tiles are
// neighbours in z, then y, of one chunk (they share halo rows in L2)
template <typename T, typename TI>
__device__ inline void mblock(const MArgs<T, TI>& a, long k, int& b,
                              int& chunk, int& ty, int& tz)
{
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

// Persistent blocks: each walks the tiles with a grid stride.  Per tile it
// marches the input planes of its x chunk through LDS; plane x + 1 loads
// while plane x is reduced.
template <typename T, int K, int EPI, typename TI, int E>
__global__ void __launch_bounds__(256) k_restrict_march(MArgs<T, TI> a)
{
    typedef typename ScalarOf<T>::type S;
    extern __shared__ __align__(16) unsigned char m_smem[];
    T* lds = (T*)m_smem;
    const int TY = blockDim.y, TZ = blockDim.x;
    for (long tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        int b, chunk, tyi, tzi;
        mblock(a, tile, b, chunk, tyi, tzi);
        const int oy = tyi * TY + threadIdx.y, oz = tzi * TZ + threadIdx.x;
        // first input row / column of the staged tile
        const int y0 = 2 * tyi * TY + a.off[1] - (K - 1);
        const int z0 = 2 * tzi * TZ + a.off[2] - (K - 1);
        const int xa = chunk * a.xchunk;
        const int xb = min(a.no[0], xa + a.xchunk);
        if (xa >= xb) continue;  // uniform over the block
        const TI* __restrict__ g = a.in + (long)b * a.in_grid;
        const bool mine = oy < a.no[1] && oz < a.no[2];
        // centre of this thread's taps in the tile
        const T* cen = lds + (2 * threadIdx.y + K - 1) * a.fz +
                       (2 * threadIdx.x + K - 1);
        // the yz-restricted values of the last 2K - 1 input planes
        T q[2 * K - 1];
        const int fx0 = 2 * xa + a.off[0] - (K - 1);
        const int fx1 = 2 * (xb - 1) + a.off[0] + (K - 1);
        unsigned char code[E];
        if constexpr (MIsCplx<TI>::value) mcodes<T, TI, E>(a, y0, z0, code);
        TI buf[E];
        mfetch<T, TI, E>(a, g, fx0, y0, z0, buf);
        for (int fx = fx0; fx <= fx1; fx++) {
            __syncthreads();  // the previous plane is fully read
            mput<T, TI, E>(a, fx, code, buf, lds);
            __syncthreads();
            if (fx < fx1) mfetch<T, TI, E>(a, g, fx + 1, y0, z0, buf);
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
                acc = fmac(acc, (S)MW<K>::w[m],
                           zr[1 + 2 * m] + zr[2 + 2 * m]);
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
        __syncthreads();  // LDS is reused by the next tile
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

template <typename T, int K, int EPI, typename TI, int E>
__global__ void __launch_bounds__(256) k_interp_march(MArgs<T, TI> a)
{
    typedef typename ScalarOf<T>::type S;
    extern __shared__ __align__(16) unsigned char m_smem[];
    T* lds = (T*)m_smem;
    const int TY = blockDim.y, TZ = blockDim.x;
    for (long tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        int b, chunk, tyi, tzi;
        mblock(a, tile, b, chunk, tyi, tzi);
        // this thread's coarse cell: fine q = 2 p, 2 p + 1 on each axis
        const int p1 = tyi * TY + threadIdx.y, p2 = tzi * TZ + threadIdx.x;
        const int y0 = tyi * TY + a.off[1] - K / 2 + 1;
        const int z0 = tzi * TZ + a.off[2] - K / 2 + 1;
        // fine plane pairs of this chunk
        const int pa = chunk * a.xchunk;
        const int pe = (a.no[0] + a.skip0[0] + 1) / 2;  // pairs in the grid
        const int pb = min(pe, pa + a.xchunk);
        if (pa >= pb) continue;  // uniform over the block
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
        unsigned char code[E];
        if constexpr (MIsCplx<TI>::value) mcodes<T, TI, E>(a, y0, z0, code);
        TI buf[E];
        mfetch<T, TI, E>(a, g, cx0, y0, z0, buf);
        for (int cx = cx0; cx <= cx1; cx++) {
            __syncthreads();
            mput<T, TI, E>(a, cx, code, buf, lds);
            __syncthreads();
            if (cx < cx1) mfetch<T, TI, E>(a, g, cx + 1, y0, z0, buf);
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
            const int p0 = cx - K / 2 - a.off[0];  // pair with all its taps
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
                                     ((long)i0 * a.no[1] + fi[e1][1]) *
                                         a.no[2] +
                                     fi[e2][2];
                    mstore<T, EPI, TI>(a, idx, v);
                }
            }
        }
        __syncthreads();  // LDS is reused by the next tile
    }
}

// ------------------------------------------------------------------ host

// Staged elements per thread and plane of each kernel (the full-row tiles
// of 64 z threads need 4.6 / 6.8 / 9.1 / 11.4 per thread for a restriction
// with K = 2 / 4 / 6 / 8 and 1.3 / 1.8 / 2.4 / 3.1 for an interpolation;
// narrower rows get fewer rows per tile).
constexpr int m_elems(int K, bool interp)
{
    return interp ? (K == 2 ? 2 : K == 4 ? 2 : K == 6 ? 3 : 4)
                  : (K == 2 ? 5 : K == 4 ? 8 : K == 6 ? 10 : MAXE);
}

// Tile and chunk of a transfer: TZ threads along z (coarse points or
// cells; a whole row where it has at most 64, else equal tiles), TY along
// y, up to 256 threads, the staged tile within MAXE elements per thread.
template <typename T, typename TI>
static int mplan(MArgs<T, TI>& a, int K, bool interp, int E, int& TY,
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
        if ((long)a.fy * a.fz <= (long)E * TY * TZ || TY == 1) break;
    }
    if ((long)a.fy * a.fz > (long)E * TY * TZ) {
        set_error("transfer: tile does not fit the marching kernel");
        return ENOTSUP_;
    }
    a.tiles_z = tz_tiles;
    a.tiles_y = (cy + TY - 1) / TY;
    return OK;
}

// x chunks for `capacity` resident blocks: about four tiles per block
// (scale multiplies that), chunks of at least min_xc x planes.
template <typename T, typename TI>
static void mchunks(MArgs<T, TI>& a, bool interp, long nin, long capacity,
                    int min_xc, int scale)
{
    const int cx = interp ? (a.no[0] + a.skip0[0] + 1) / 2 : a.no[0];
    const long yz = (long)a.tiles_y * a.tiles_z * nin;
    long chunks = std::max<long>(1, (4L * scale * capacity + yz - 1) / yz);
    chunks = std::min<long>(chunks, std::max(1, cx / min_xc));
    a.xchunk = (int)((cx + chunks - 1) / chunks);
    a.nchunks = (cx + a.xchunk - 1) / a.xchunk;
    a.ntiles = (long)a.nchunks * yz;
}

// resident blocks of one kernel instantiation on the current device
static std::mutex g_occ_mu;
static std::map<std::pair<const void*, long>, long> g_occ;

static long resident_blocks(const void* fn, int threads, size_t lds)
{
    const std::pair<const void*, long> key(fn, (long)threads << 32 |
                                                   (long)lds);
    {
        std::lock_guard<std::mutex> lk(g_occ_mu);
        auto it = g_occ.find(key);
        if (it != g_occ.end()) return it->second;
    }
    int per_cu = 0, dev = 0, ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads,
                                                     lds) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    hipDeviceProp_t pr;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipGetDeviceProperties(&pr, dev) == hipSuccess)
        ncu = pr.multiProcessorCount;
    (void)hipGetLastError();
    const long cap = (long)per_cu * ncu;
    std::lock_guard<std::mutex> lk(g_occ_mu);
    g_occ[key] = cap;
    return cap;
}

static int g_march_scale = 1;  // tiles per resident block, in units of 4

void xfer_set_march_scale(int s) { g_march_scale = std::max(1, s); }

template <typename T, int K, int EPI, typename TI, int E, bool INTERP>
static int launch_march_e(MArgs<T, TI>& a, long nin, int TY, int TZ,
                          size_t lds, hipStream_t s)
{
    const void* fn;
    if constexpr (INTERP)
        fn = (const void*)k_interp_march<T, K, EPI, TI, E>;
    else
        fn = (const void*)k_restrict_march<T, K, EPI, TI, E>;
    const long cap = resident_blocks(fn, TZ * TY, lds);
    mchunks(a, INTERP, nin, cap, K == 2 ? 8 : 16, g_march_scale);
    if (a.ntiles < cap / 2) {
        // a small problem: shorter chunks (more halo planes, served by the
        // caches) to reach more CUs
        mchunks(a, INTERP, nin, cap, 2, g_march_scale);
    }
    if (a.ntiles <= 0) return OK;
    const long nb = std::min(a.ntiles, cap);
    if constexpr (INTERP)
        hipLaunchKernelGGL((k_interp_march<T, K, EPI, TI, E>),
                           dim3((unsigned)nb), dim3(TZ, TY), lds, s, a);
    else
        hipLaunchKernelGGL((k_restrict_march<T, K, EPI, TI, E>),
                           dim3((unsigned)nb), dim3(TZ, TY), lds, s, a);
    GFD_LAUNCH_CHECK();
    return OK;
}

template <typename T, int K, int EPI, typename TI, bool INTERP>
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
    // one wrap at most per staged element (the Bloch codes): periodic axes
    // of at least K input points
    for (int c = 0; c < 3; c++)
        if (a.periodic[c] && a.ni[c] < K) {
            set_error("transfer: grid too small for the marching kernel");
            return ENOTSUP_;
        }
    int TY, TZ;
    constexpr int E = m_elems(K, INTERP);
    GFD_TRY(mplan(a, K, INTERP, E, TY, TZ));
    const size_t lds = sizeof(T) * (size_t)a.fy * a.fz;
    if (lds > 64 * 1024) {
        set_error("transfer: tile too large for the marching kernel");
        return ENOTSUP_;
    }
    return launch_march_e<T, K, EPI, TI, E, INTERP>(a, nin, TY, TZ, lds, s);
}

// The epilogues the callers use: restriction PLAIN (apply, V-cycle) and
// SCALE2 (preconditioner); interpolation PLAIN, SUBFROM (V-cycle,
// preconditioner) and NEG (preconditioner).  Others fall back (ENOTSUP_).
template <typename T, int K, typename TI>
static int launch_march_k(const XferPlan& p, const BC& bc, const void* in,
                          void* out, long nin, const XEpiArgs& ea,
                          hipStream_t s)
{
    if (!p.interpolate) {
        if (ea.epi == XEPI_PLAIN)
            return launch_march_t<T, K, XEPI_PLAIN, TI, false>(p, bc, in, out, nin,
                                                        ea, s);
        if (ea.epi == XEPI_SCALE2)
            return launch_march_t<T, K, XEPI_SCALE2, TI, false>(p, bc, in, out, nin,
                                                         ea, s);
    } else if constexpr (std::is_same<T, TI>::value) {
        if (ea.epi == XEPI_PLAIN)
            return launch_march_t<T, K, XEPI_PLAIN, TI, true>(p, bc, in, out, nin,
                                                        ea, s);
        if (ea.epi == XEPI_SUBFROM)
            return launch_march_t<T, K, XEPI_SUBFROM, TI, true>(p, bc, in, out,
                                                          nin, ea, s);
        if (ea.epi == XEPI_NEG)
            return launch_march_t<T, K, XEPI_NEG, TI, true>(p, bc, in, out, nin, ea,
                                                      s);
    }
    set_error("transfer: epilogue %d not built for the marching kernel",
              ea.epi);
    return ENOTSUP_;
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
