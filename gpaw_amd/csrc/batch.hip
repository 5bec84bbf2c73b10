// Batched multigrid over many independent real grids of uneven shape.
//
// Every grid of a hierarchy level lives in one concatenated level array
// (grid g at element offset off_g).  Each stencil stage of a level is ONE
// persistent launch over the tiles of all grids: a block walks the tile
// list with a grid stride, finds the tile's grid by binary search over the
// per-grid first-tile table and applies that grid's axis stencil and
// boundary (periodic by index wrap, zero by skipping).  Transfers are one
// persistent launch per level pair with one thread per output point.  The
// tile shape (threads along z and y, z points per thread) and the number
// of resident blocks are picked per level by timing candidates on the
// actual batch (gfd_batch_poisson_tune).
//
// The mixed-precision solve runs FP32 V-cycles (Jacobi smoothing, K = 2
// transfers) on all grids together inside FP64 residual refinement; one
// refinement step (FP64 residual + per-grid norms, FP32 V-cycle, FP64
// update) is captured once as a hipGraph and replayed.
#include "../../include/gpaw_fd.h"

#include <algorithm>
#include <cmath>
#include <vector>

#include "batch.h"
#include "coarse.h"
#include "gfd_common.h"
#include "handles.h"
#include "stencil_dev.h"
#include "vector.h"

namespace gfd {

// T: data read, TO: data written (float output of an FP64 residual).
// EPI_PLAIN, EPI_RELAX (aux = right-hand side), EPI_RESID (aux = rhs,
// partial[tile] = sum r^2, when partial is set), EPI_UPDATE
// (out = +-(v - step (L v - aux)), the preconditioner's smoothing step) or
// EPI_UPDW (EPI_UPDATE on FP32 data with FP64 aux and out: aux narrowed
// to FP32 first, the FP32 result widened).  A tile is BZ x PZ z points by BY rows of one
// grid and marches over an x chunk: the centre values of planes x-R .. x+R
// stay in a register queue, the z and y neighbours come through L1 / L2.
template <typename T, typename TO, int R, int PZ, int EPI>
__global__ void __launch_bounds__(256) k_bstencil(BSArgs a)
{
    typedef typename ScalarOf<T>::type S;
    constexpr int W = PZ + 2 * R;
    const int BZ = blockDim.x, BY = blockDim.y;
    for (long t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        // by value: stores to out could alias a reference into the table,
        // forcing a reload of every field after each store
        const BSt d = a.g[bfind_tile(a.g, a.G, t)];
        long lt = t - d.tile0;
        const int zt = (int)(lt % d.tz);
        lt /= d.tz;
        const int yt = (int)(lt % d.ty);
        const int xt = (int)(lt / d.ty);
        const int i1 = yt * BY + threadIdx.y;
        const int z0 = (zt * BZ + threadIdx.x) * PZ;
        const int n0 = d.n[0], n1 = d.n[1], n2 = d.n[2];
        const int xs = xt * d.xc, xe = min(xs + d.xc, n0);
        double nrm = 0.0;
        if (i1 < n1 && z0 < n2) {
            const T* in = (const T*)a.in + d.off;
            const long plane = (long)n1 * n2;
            const long own = (long)i1 * n2;
            // z window indices (wrapped; -1: zero or beyond the row)
            int zi[W];
#pragma unroll
            for (int k = 0; k < W; k++) {
                int j = z0 - R + k;
                if (k >= R && k < R + PZ) {
                    zi[k] = j < n2 ? j : -1;
                } else {
                    zi[k] = bwrap(j, n2, d.periodic[2]) ? j : -1;
                }
            }
            // y neighbour rows (plane offsets; -1: zero boundary)
            long yr[2][R];
#pragma unroll
            for (int k = 1; k <= R; k++) {
                int yp = i1 + k, ym = i1 - k;
                yr[0][k - 1] = bwrap(yp, n1, d.periodic[1]) ? (long)yp * n2 : -1;
                yr[1][k - 1] = bwrap(ym, n1, d.periodic[1]) ? (long)ym * n2 : -1;
            }
            // queue of centre chunks of planes x - R .. x + R
            T q[2 * R + 1][PZ];
            auto load_c = [&](int x, T* dst) {
                int xx = x;
                const bool ok = bwrap(xx, n0, d.periodic[0]);
                const T* row = in + (long)xx * plane + own;
#pragma unroll
                for (int p = 0; p < PZ; p++)
                    dst[p] = (ok && zi[R + p] >= 0) ? row[zi[R + p]]
                                                    : zero_of<T>();
            };
            // z-window ends and y rows of plane x (loaded one plane ahead so
            // their latency overlaps the previous plane's arithmetic)
            T zx[2 * R], yv[2 * R][PZ];
            auto load_side = [&](int x, T* zo, T (*yo)[PZ]) {
                const T* pl = in + (long)x * plane;
#pragma unroll
                for (int k = 0; k < R; k++) {
                    zo[k] = (d.has[2] && zi[k] >= 0) ? pl[own + zi[k]]
                                                     : zero_of<T>();
                    zo[R + k] = (d.has[2] && zi[R + PZ + k] >= 0)
                                    ? pl[own + zi[R + PZ + k]]
                                    : zero_of<T>();
                }
#pragma unroll
                for (int k = 0; k < R; k++)
#pragma unroll
                    for (int sg = 0; sg < 2; sg++) {
                        const long ro = yr[sg][k];
#pragma unroll
                        for (int p = 0; p < PZ; p++)
                            yo[2 * k + sg][p] =
                                (d.has[1] && ro >= 0 && zi[R + p] >= 0)
                                    ? pl[ro + zi[R + p]]
                                    : zero_of<T>();
                    }
            };
#pragma unroll
            for (int k = 0; k < 2 * R + 1; k++) load_c(xs - R + k, q[k]);
            load_side(xs, zx, yv);
            TO* out = (TO*)a.out;
            for (int x = xs; x < xe; x++) {
                T nzx[2 * R], nyv[2 * R][PZ];
                if (x + 1 < xe) load_side(x + 1, nzx, nyv);
                T acc[PZ];
#pragma unroll
                for (int p = 0; p < PZ; p++) acc[p] = zero_of<T>();
                if (d.has[2]) {
                    T wz[W];
#pragma unroll
                    for (int k = 0; k < W; k++) {
                        if (k < R)
                            wz[k] = zx[k];
                        else if (k < R + PZ)
                            wz[k] = q[R][k - R];
                        else
                            wz[k] = zx[R + k - R - PZ];
                    }
#pragma unroll
                    for (int p = 0; p < PZ; p++)
#pragma unroll
                        for (int k = 1; k <= R; k++) {
                            acc[p] = fmac(acc[p], (S)d.cp[2][k - 1],
                                          wz[R + p + k]);
                            acc[p] = fmac(acc[p], (S)d.cm[2][k - 1],
                                          wz[R + p - k]);
                        }
                }
                if (d.has[1]) {
#pragma unroll
                    for (int k = 1; k <= R; k++)
#pragma unroll
                        for (int p = 0; p < PZ; p++) {
                            acc[p] = fmac(acc[p], (S)d.cp[1][k - 1],
                                          yv[2 * (k - 1)][p]);
                            acc[p] = fmac(acc[p], (S)d.cm[1][k - 1],
                                          yv[2 * (k - 1) + 1][p]);
                        }
                }
                if (d.has[0]) {
#pragma unroll
                    for (int k = 1; k <= R; k++)
#pragma unroll
                        for (int p = 0; p < PZ; p++) {
                            acc[p] = fmac(acc[p], (S)d.cp[0][k - 1],
                                          q[R + k][p]);
                            acc[p] = fmac(acc[p], (S)d.cm[0][k - 1],
                                          q[R - k][p]);
                        }
                }
                const long obase = d.off + (long)x * plane + own + z0;
#pragma unroll
                for (int p = 0; p < PZ; p++) {
                    if (z0 + p >= n2) break;
                    const T v = q[R][p];
                    T r;
                    if constexpr (EPI == EPI_PLAIN) {
                        r = fmac(acc[p], (S)d.c0, v);
                    } else if constexpr (EPI == EPI_RELAX) {
                        const T f = ((const T*)a.aux)[obase + p];
                        r = ((S)1.0 - (S)d.w) * v +
                            (S)d.w * (f - acc[p]) / (S)d.c0;
                    } else if constexpr (EPI == EPI_UPDATE ||
                                         EPI == EPI_UPDW) {
                        T f;
                        if constexpr (EPI == EPI_UPDW)
                            f = (T)((const double*)a.aux)[obase + p];
                        else
                            f = ((const T*)a.aux)[obase + p];
                        const T q = fmac(acc[p], (S)d.c0, v) - f;
                        r = fmac(v, -(S)d.step, q);
                        if (a.negate) r = -r;
                    } else {
                        const T f = ((const T*)a.aux)[obase + p];
                        r = fmac(acc[p], (S)d.c0, v) - f;
                        nrm += (double)r * (double)r;
                    }
                    out[obase + p] = (TO)r;
                }
#pragma unroll
                for (int k = 0; k < 2 * R; k++)
#pragma unroll
                    for (int p = 0; p < PZ; p++) q[k][p] = q[k + 1][p];
                if (x + 1 < xe) {
                    load_c(x + R + 1, q[2 * R]);
#pragma unroll
                    for (int k = 0; k < 2 * R; k++) {
                        zx[k] = nzx[k];
#pragma unroll
                        for (int p = 0; p < PZ; p++) yv[k][p] = nyv[k][p];
                    }
                }
            }
        }
        if (EPI == EPI_RESID && a.partial) block_sum_store(nrm, a.partial, t);
    }
}

// e = (w_g / c0_g) f on every grid g: a Jacobi sweep from e = 0 (the
// x-march kernel's Jacobi update multiplies by the same FP32 w / c0)
__global__ void __launch_bounds__(256)
k_bscale0(const BSt* __restrict__ g, int G, long total,
          const float* __restrict__ f, float* __restrict__ e)
{
    for (long b0 = (long)blockIdx.x * 1024; b0 < total;
         b0 += (long)gridDim.x * 1024) {
        int lo = 0, hi = G - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (g[mid].off <= b0)
                lo = mid;
            else
                hi = mid - 1;
        }
        int gi = lo;
        float wc = (float)(g[gi].w / g[gi].c0);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const long i = b0 + k * 256 + threadIdx.x;
            if (i >= total) break;
            while (gi + 1 < G && g[gi + 1].off <= i) {
                gi++;
                wc = (float)(g[gi].w / g[gi].c0);
            }
            e[i] = wc * f[i];
        }
    }
}

// per-grid sums of the tile partials: one block per grid
__global__ void __launch_bounds__(256)
k_bnorms(const double* __restrict__ partial, const BSt* __restrict__ g, int G,
         long ntiles, double* __restrict__ out)
{
    const int gi = blockIdx.x;
    const long t0 = g[gi].tile0;
    const long t1 = gi + 1 < G ? g[gi + 1].tile0 : ntiles;
    double s = 0.0;
    for (long t = t0 + threadIdx.x; t < t1; t += 256) s += partial[t];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[gi] = red[0];
}

__device__ inline int bfind_point(const BXf* g, int G, long p)
{
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g[mid].p0 <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// input value (type TI, narrowed to T) at a wrapped index; 0 outside a
// zero boundary
template <typename T, typename TI>
__device__ inline T bxload(const BXf& d, const TI* __restrict__ g, int i0,
                           int i1, int i2)
{
    if (!bwrap(i0, d.ni[0], d.periodic[0]) ||
        !bwrap(i1, d.ni[1], d.periodic[1]) ||
        !bwrap(i2, d.ni[2], d.periodic[2]))
        return zero_of<T>();
    return (T)g[((long)i0 * d.ni[1] + i1) * d.ni[2] + i2];
}

// Restriction (INTERP false) or interpolation (INTERP true) with K = 2
// weights, the association order of the single-grid fused kernels (z
// innermost, then y, then x); input of type TI (FP64 residuals into an
// FP32 chain) narrowed on load; result T stored by the transfer epilogue
// XE (transfer.h).
template <typename T, typename TI, bool INTERP, int XE>
__global__ void __launch_bounds__(256) k_bxfer2(BXArgs a)
{
    typedef typename ScalarOf<T>::type S;
    for (long b0 = (long)blockIdx.x * 256; b0 < a.total;
         b0 += (long)gridDim.x * 256) {
        const long gid = b0 + threadIdx.x;
        if (gid >= a.total) break;
        // the block's first point by binary search, then forward to the
        // thread's own grid (a block spans few grid boundaries)
        int gi = bfind_point(a.g, a.G, b0);
        while (gi + 1 < a.G && a.g[gi + 1].p0 <= gid) gi++;
        const BXf d = a.g[gi];
        long r = gid - d.p0;
        const int o2 = (int)(r % d.no[2]);
        r /= d.no[2];
        const int o1 = (int)(r % d.no[1]);
        const int o0 = (int)(r / d.no[1]);
        const TI* __restrict__ g = (const TI*)a.in + d.in_off;
        const long oidx =
            d.out_off + ((long)o0 * d.no[1] + o1) * d.no[2] + o2;
        T v;
        if constexpr (!INTERP) {
            const int c0 = 2 * o0 + d.off[0], c1 = 2 * o1 + d.off[1],
                      c2 = 2 * o2 + d.off[2];
            T px[3];
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                T py[3];
#pragma unroll
                for (int dy = -1; dy <= 1; dy++) {
                    const T zp = bxload<T, TI>(d, g, c0 + dx, c1 + dy, c2 + 1) +
                                 bxload<T, TI>(d, g, c0 + dx, c1 + dy, c2 - 1);
                    py[dy + 1] = (S)0.5 * fmac(bxload<T, TI>(d, g, c0 + dx, c1 + dy,
                                                      c2),
                                               (S)0.5, zp);
                }
                px[dx + 1] = (S)0.5 * fmac(py[1], (S)0.5, py[2] + py[0]);
            }
            v = (S)0.5 * fmac(px[1], (S)0.5, px[2] + px[0]);
        } else {
            const int q0 = o0 + d.skip0[0], q1 = o1 + d.skip0[1],
                      q2 = o2 + d.skip0[2];
            const int c0 = (q0 >> 1) + d.off[0], c1 = (q1 >> 1) + d.off[1],
                      c2 = (q2 >> 1) + d.off[2];
            const int m0 = (q0 & 1) + 1, m1 = (q1 & 1) + 1, m2 = (q2 & 1) + 1;
            T vx[2];
#pragma unroll
            for (int dx = 0; dx < 2; dx++) {
                if (dx >= m0) break;
                T vy[2];
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    if (dy >= m1) break;
                    T vz = bxload<T, TI>(d, g, c0 + dx, c1 + dy, c2);
                    if (m2 == 2)
                        vz = (S)0.5 *
                             (vz + bxload<T, TI>(d, g, c0 + dx, c1 + dy, c2 + 1));
                    vy[dy] = vz;
                }
                vx[dx] = (m1 == 2) ? (S)0.5 * (vy[0] + vy[1]) : vy[0];
            }
            v = (m0 == 2) ? (S)0.5 * (vx[0] + vx[1]) : vx[0];
        }
        if constexpr (XE == XEPI_PLAIN) {
            ((T*)a.out)[oidx] = v;
        } else if constexpr (XE == XEPI_NEG) {
            ((T*)a.out)[oidx] = -v;
        } else if constexpr (XE == XEPI_SCALE2) {
            if (a.negate) v = -v;
            ((T*)a.out)[oidx] = v;
            ((T*)a.out2)[oidx] = (S)d.alpha * v;
        } else {
            ((T*)a.out)[oidx] = ((const T*)a.base)[oidx] - v;
        }
    }
}

// ------------------------------------------------------------ host side

// Tile shape of one level: z threads, y threads, z points per thread,
// resident blocks (0: one block per tile).
struct BTile {
    int kind = 0;                   // 0: register tiles, 1: x-march tiles
    int bz = 64, by = 4, pz = 1;
    int xc = 16;                    // x planes per tile (0: whole grid)
    long blocks = 0;
    // x-march: threads per block, rows per thread, widest z part
    int nt = 256, py = 4, zmax = 128, pf = 1;   // pf: planes ahead
    long lds_elems = 0;             // x-march: LDS elements per block
};

struct BLevel {
    int R = 1;
    long total = 0;                 // points of all grids
    std::vector<BSt> st;            // per grid (tile0 for `tile`)
    BTile tile;
    long ntiles = 0;
    DevBuf d_st;
};

// x-march transfer tiles of one direction (nt 0: one point per thread)
struct BXTile {
    int nt = 0, py = 1, zmax = 64, xc = 8;
    long ntiles = 0;
};

struct BPair {                      // transfer from level l to l + 1
    std::vector<BXf> rst, itp;
    long nr = 0, ni = 0;            // output points
    DevBuf d_rst, d_itp;
    BXTile trst, titp;
};

}  // namespace gfd

using namespace gfd;

struct gfd_batch_poisson {
    int G = 0, L = 0;
    std::vector<BLevel> lev;
    std::vector<BPair> pair;
    BLevel l0_64;                   // level 0 for the FP64 residual
    bool tuned = false;
    // FP64 level-0 arrays, FP32 arrays per level
    DevBuf phi, rho;
    std::vector<DevBuf> e32, t32, r32, s32;
    DevBuf partial, norms;
    std::vector<double> rn2;
    hipStream_t gs = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t graph = nullptr;
    int gkey[4] = {-1, -1, -1, -1};
    int num_cu = 256;
    // per-grid levels L-1 .. nlev-1 in LDS-resident workgroups (one per
    // grid) in place of the coarsest batched level
    bool coarse = false;
    std::vector<int> nlev;
    std::vector<std::vector<gfd_operator*>> cops;
    std::vector<std::vector<gfd_transformer*>> crst, citp;
    std::vector<double> cw;
    std::vector<CGrid> cgrids;
    size_t clds = 0;
    int ckey[4] = {-1, -1, -1, -1};
    DevBuf d_cgrids;
};

namespace gfd {

// Tile counts of every grid for tile shape t; false if an x-march tile
// does not fit (halo per thread, LDS for FP64 data: any level may be
// applied in FP64).
static bool fill_tiles(BLevel& lv, const BTile& t)
{
    lv.tile = t;
    long t0 = 0;
    long lds = 0;
    bool ok = true;
    const int R = lv.R;
    for (BSt& s : lv.st) {
        s.xc = t.xc > 0 ? std::min(t.xc, s.n[0]) : s.n[0];
        if (t.kind == 1) {
            // the row in equal parts of at most zmax (and nt) columns
            const int zw = std::min(t.zmax, t.nt);
            s.tz = (s.n[2] + zw - 1) / zw;
            s.mz = (s.n[2] + s.tz - 1) / s.tz;
            s.mseg = std::max(1, std::min(t.nt / s.mz,
                                          (s.n[1] + t.py - 1) / t.py));
            const int TY = s.mseg * t.py;
            s.ty = (s.n[1] + TY - 1) / TY;
            // halo loads: rows, and columns unless the tile is whole rows
            const int hc = s.tz > 1 ? TY : 0;
            ok = ok && 2 * R * (s.mz + hc) <= BXM_HMAX * t.nt;
            // two plane buffers, each with a spare element
            lds = std::max(lds, 2L * ((TY + 2 * R) * (s.mz + 2 * R) + 1));
        } else {
            s.tz = (s.n[2] + t.bz * t.pz - 1) / (t.bz * t.pz);
            s.ty = (s.n[1] + t.by - 1) / t.by;
            s.mz = s.mseg = 0;
        }
        s.tile0 = t0;
        t0 += (long)((s.n[0] + s.xc - 1) / s.xc) * s.ty * s.tz;
    }
    lv.ntiles = t0;
    lv.tile.lds_elems = lds;
    // x-march buffer resources span one plane: 32-bit byte offsets of
    // FP64 data below the out-of-range marker 2^31
    long pmax = 0;
    for (const BSt& s : lv.st) pmax = std::max(pmax, (long)s.n[1] * s.n[2]);
    if (t.kind == 1 && pmax * 8 >= 0x7fffffffL) ok = false;
    return ok && lds * 8 <= 64 * 1024 && t0 <= 0x7fffffffL;
}

static int upload_level(BLevel& lv, hipStream_t s)
{
    GFD_TRY(lv.d_st.ensure(sizeof(BSt) * lv.st.size(), s));
    GFD_CHECK_HIP(hipMemcpyAsync(lv.d_st.p, lv.st.data(),
                                 sizeof(BSt) * lv.st.size(),
                                 hipMemcpyHostToDevice, s));
    return hipStreamSynchronize(s) == hipSuccess ? OK : EHIP;
}

template <typename T, typename TO, int R, int PZ, int EPI>
static int bst_launch_t(const BLevel& lv, const void* in, void* out,
                        const void* aux, double* partial, hipStream_t s,
                        int negate)
{
    BSArgs a;
    a.g = (const BSt*)lv.d_st.p;
    a.G = (int)lv.st.size();
    a.ntiles = lv.ntiles;
    a.in = in;
    a.out = out;
    a.aux = aux;
    a.partial = partial;
    a.negate = negate;
    long nb = lv.tile.blocks > 0 ? std::min(lv.tile.blocks, lv.ntiles)
                                 : lv.ntiles;
    if (nb <= 0) return OK;
    if (nb > 0x7fffffffL) nb = 0x7fffffffL;
    hipLaunchKernelGGL((k_bstencil<T, TO, R, PZ, EPI>), dim3((unsigned)nb),
                       dim3(lv.tile.bz, lv.tile.by), 0, s, a);
    GFD_LAUNCH_CHECK();
    return OK;
}

template <typename T, typename TO, int EPI>
static int bst_launch_r(const BLevel& lv, const void* in, void* out,
                        const void* aux, double* partial, hipStream_t s,
                        int negate = 0)
{
#define GFD_BST(RR, PP)                                                       \
    if (lv.R == RR && lv.tile.pz == PP)                                       \
        return bst_launch_t<T, TO, RR, PP, EPI>(lv, in, out, aux, partial, s, \
                                                negate);
    GFD_BST(1, 1) GFD_BST(1, 2) GFD_BST(1, 4)
    GFD_BST(2, 1) GFD_BST(2, 2) GFD_BST(2, 4)
    GFD_BST(3, 1) GFD_BST(3, 2) GFD_BST(3, 4)
#undef GFD_BST
    set_error("batched stencil: radius %d / %d z points not built", lv.R,
              lv.tile.pz);
    return ENOTSUP_;
}

// dtype 0: f64 -> f64, 2: f32 -> f32 (EPI_UPDW: f32 -> f64), 4: f64 ->
// f32 (RESID only)
static int bst(const BLevel& lv, int dtype, int epi, const void* in,
               void* out, const void* aux, double* partial, hipStream_t s,
               int negate = 0)
{
    if (lv.tile.kind == 1) {
        BSArgs a;
        a.g = (const BSt*)lv.d_st.p;
        a.G = (int)lv.st.size();
        a.ntiles = lv.ntiles;
        a.in = in;
        a.out = out;
        a.aux = aux;
        a.partial = partial;
        a.negate = negate;
        // element read: FP32 for dtype 2 (EPI_UPDW included), else FP64
        const size_t es = dtype == 2 ? 4 : 8;
        return launch_bxmarch(a, lv.R, lv.tile.py, lv.tile.pf, lv.tile.nt,
                              lv.tile.lds_elems * es, dtype, epi, s);
    }
    if (dtype == 0 && epi == EPI_PLAIN)
        return bst_launch_r<double, double, EPI_PLAIN>(lv, in, out, aux,
                                                       partial, s);
    if (dtype == 2 && epi == EPI_PLAIN)
        return bst_launch_r<float, float, EPI_PLAIN>(lv, in, out, aux,
                                                     partial, s);
    if (dtype == 2 && epi == EPI_RELAX)
        return bst_launch_r<float, float, EPI_RELAX>(lv, in, out, aux,
                                                     partial, s);
    if (dtype == 2 && epi == EPI_RESID)
        return bst_launch_r<float, float, EPI_RESID>(lv, in, out, aux,
                                                     partial, s);
    if (dtype == 4 && epi == EPI_RESID)
        return bst_launch_r<double, float, EPI_RESID>(lv, in, out, aux,
                                                      partial, s);
    if (dtype == 2 && epi == EPI_UPDATE)
        return bst_launch_r<float, float, EPI_UPDATE>(lv, in, out, aux,
                                                      partial, s, negate);
    if (dtype == 2 && epi == EPI_UPDW)
        return bst_launch_r<float, double, EPI_UPDW>(lv, in, out, aux,
                                                     partial, s, negate);
    set_error("batched stencil: dtype %d / epilogue %d not built", dtype,
              epi);
    return ENOTSUP_;
}

// x-march tiles of every grid's output for tile shape t; false if a grid's
// input plane is too large for 32-bit offsets
static bool fill_xtiles(std::vector<BXf>& v, BXTile& t)
{
    long t0 = 0;
    bool ok = true;
    for (BXf& x : v) {
        if (t.nt <= 0) {
            x.tile0 = 0;
            x.tz = x.ty = x.xc = x.mz = x.mseg = 1;
            continue;
        }
        const int zw = std::min(t.zmax, t.nt);
        x.tz = (x.no[2] + zw - 1) / zw;
        x.mz = (x.no[2] + x.tz - 1) / x.tz;
        x.mseg = std::max(1, std::min(t.nt / x.mz, (x.no[1] + t.py - 1) / t.py));
        x.ty = (x.no[1] + x.mseg * t.py - 1) / (x.mseg * t.py);
        x.xc = std::min(t.xc, x.no[0]);
        x.tile0 = t0;
        t0 += (long)((x.no[0] + x.xc - 1) / x.xc) * x.ty * x.tz;
        const long pl = std::max((long)x.ni[1] * x.ni[2], (long)x.no[1] * x.no[2]);
        if (pl * 8 >= 0x7fffffffL) ok = false;
    }
    t.ntiles = t0;
    return ok && t0 <= 0x7fffffffL;
}

template <typename T, typename TI, bool I, int XE>
static void bxfer_launch(const BXArgs& a, long nb, hipStream_t s)
{
    hipLaunchKernelGGL((k_bxfer2<T, TI, I, XE>), dim3((unsigned)nb),
                       dim3(256), 0, s, a);
}

// dtype 0: FP64, 2: FP32, 4: FP32 from FP64 input (restriction).
// Restriction epilogue xe (PLAIN, SCALE2 with the grids' alpha); the
// interpolation stores base - I(in) (SUBFROM) or -I(in) (NEG).
static int bxfer(const BPair& pr, bool interp, int dtype, const void* in,
                 void* out, const void* base, int num_cu, hipStream_t s,
                 int xe = -1, void* out2 = nullptr, int negate = 0)
{
    BXArgs a;
    a.g = (const BXf*)(interp ? pr.d_itp.p : pr.d_rst.p);
    a.G = (int)(interp ? pr.itp.size() : pr.rst.size());
    a.total = interp ? pr.ni : pr.nr;
    a.in = in;
    a.out = out;
    a.base = base;
    a.out2 = out2;
    a.negate = negate;
    if (xe < 0) xe = interp ? XEPI_SUBFROM : XEPI_PLAIN;
    if (a.total <= 0) return OK;
    const BXTile& xt = interp ? pr.titp : pr.trst;
    if (xt.nt > 0) {
        a.ntiles = xt.ntiles;
        return launch_bxfer_march(a, interp, xt.py, xt.nt, dtype, xe, s);
    }
    const long nb = std::min((a.total + 255) / 256, (long)num_cu * 16);
    if (dtype == 0 && !interp && xe == XEPI_PLAIN)
        bxfer_launch<double, double, false, XEPI_PLAIN>(a, nb, s);
    else if (dtype == 0 && interp && xe == XEPI_SUBFROM)
        bxfer_launch<double, double, true, XEPI_SUBFROM>(a, nb, s);
    else if (dtype == 2 && !interp && xe == XEPI_PLAIN)
        bxfer_launch<float, float, false, XEPI_PLAIN>(a, nb, s);
    else if (dtype == 2 && !interp && xe == XEPI_SCALE2)
        bxfer_launch<float, float, false, XEPI_SCALE2>(a, nb, s);
    else if (dtype == 4 && !interp && xe == XEPI_SCALE2)
        bxfer_launch<float, double, false, XEPI_SCALE2>(a, nb, s);
    else if (dtype == 2 && interp && xe == XEPI_SUBFROM)
        bxfer_launch<float, float, true, XEPI_SUBFROM>(a, nb, s);
    else if (dtype == 2 && interp && xe == XEPI_NEG)
        bxfer_launch<float, float, true, XEPI_NEG>(a, nb, s);
    else {
        set_error("batched transfer: dtype %d / epilogue %d not built", dtype,
                  xe);
        return ENOTSUP_;
    }
    GFD_LAUNCH_CHECK();
    return OK;
}

static int bnorms(const BLevel& lv, const double* partial, double* out,
                  hipStream_t s)
{
    hipLaunchKernelGGL(k_bnorms, dim3((unsigned)lv.st.size()), dim3(256), 0,
                       s, partial, (const BSt*)lv.d_st.p, (int)lv.st.size(),
                       lv.ntiles, out);
    GFD_LAUNCH_CHECK();
    return OK;
}

static int g_bkind = 0;  // 0: timed choice, 1: register tiles, 2: x-march
static bool g_bforced = false;
static BTile g_bforce;   // x-march tile every level must use when forced

void batch_set_kernel(int kind) { g_bkind = kind; }

void batch_force_tile(int nt, int py, int zmax, int xc, int pf)
{
    g_bforced = nt > 0;
    g_bforce = BTile();
    g_bforce.kind = 1;
    g_bforce.nt = nt;
    g_bforce.py = py;
    g_bforce.zmax = zmax;
    g_bforce.xc = xc;
    g_bforce.pf = pf;
}

// Candidate tile shapes of both kernels, timed on the level's dominant
// pass: the FP32 Jacobi sweep (epi EPI_RELAX, dtype 2) on the FP32 levels,
// the FP64 residual with FP32 output (EPI_RESID, dtype 4) on the FP64 one.
static int tune_level(gfd_batch_poisson* b, BLevel& lv, int dtype, int epi,
                      const void* in, const void* aux, void* out,
                      hipStream_t s)
{
    std::vector<BTile> cand;
    if (g_bforced) {
        if (!fill_tiles(lv, g_bforce)) {
            set_error("batched stencil: forced x-march tile does not fit");
            return ENOTSUP_;
        }
        return upload_level(lv, s);
    }
    if (g_bkind != 2) {
        static const int shapes[][2] = {{64, 4}, {32, 8}, {16, 16}, {64, 2},
                                        {32, 4}, {16, 8}};
        static const int pzs[] = {1, 2, 4};
        static const int xcs[] = {16, 32};
        const long res[] = {0, (long)b->num_cu * 8};
        for (const auto& sh : shapes)
            for (int pz : pzs)
                for (int ci = 0; ci < 4; ci++) {
                    BTile t;
                    t.bz = sh[0];
                    t.by = sh[1];
                    t.pz = pz;
                    t.xc = xcs[ci & 1];
                    t.blocks = res[ci >> 1];
                    cand.push_back(t);
                }
    }
    if (g_bkind != 1) {
        for (int nt : {256, 512})
            for (int py : {1, 2, 4})
                for (int zmax : {64, 128})
                    for (int xc : {8, 16, 32})
                        for (int pf : {1, 2}) {
                            BTile t;
                            t.kind = 1;
                            t.nt = nt;
                            t.py = py;
                            t.zmax = zmax;
                            t.xc = xc;
                            t.pf = pf;
                            cand.push_back(t);
                        }
    }
    hipEvent_t e0, e1;
    GFD_CHECK_HIP(hipEventCreate(&e0));
    GFD_CHECK_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    BTile bt;
    bool found = false;
    int st = OK;
    for (const BTile& t : cand) {
        if (!fill_tiles(lv, t)) continue;
        if ((st = upload_level(lv, s)) != OK) break;
        if ((st = bst(lv, dtype, epi, in, out, aux, nullptr, s)) != OK) break;
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < 3 && st == OK; r++)
            st = bst(lv, dtype, epi, in, out, aux, nullptr, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        if (st != OK) break;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
            best = ms;
            bt = lv.tile;
            found = true;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != OK) return st;
    if (!found) {
        set_error("batched stencil: no tile shape fits the level");
        return ENOTSUP_;
    }
    fill_tiles(lv, bt);
    return upload_level(lv, s);
}

}  // namespace gfd

extern "C" int gfd_batch_poisson_new(int G, int L, gfd_operator** ops,
                                     gfd_transformer** rst,
                                     gfd_transformer** itp,
                                     const double* weights,
                                     gfd_batch_poisson** out)
{
    if (!out || G < 1 || L < 1 || !ops || !weights ||
        (L > 1 && (!rst || !itp))) {
        set_error("gfd_batch_poisson_new: bad arguments");
        return EINVAL_;
    }
    *out = nullptr;
    auto* b = new gfd_batch_poisson;
    b->G = G;
    b->L = L;
    b->lev.resize(L);
    b->pair.resize(L > 1 ? L - 1 : 0);
    for (int l = 0; l < L; l++) {
        BLevel& lv = b->lev[l];
        long off = 0;
        for (int g = 0; g < G; g++) {
            gfd_operator* op = ops[g * L + l];
            if (!op || !op->real || !op->axis_ok || op->ax.R > BMAXR) {
                set_error("batched Poisson: grid %d level %d needs a real "
                          "axis-only stencil of radius <= %d",
                          g, l, BMAXR);
                delete b;
                return EINVAL_;
            }
            if (g == 0) lv.R = op->ax.R;
            if (op->ax.R != lv.R) {
                set_error("batched Poisson: level %d mixes stencil radii", l);
                delete b;
                return EINVAL_;
            }
            BSt s{};
            s.off = off;
            for (int c = 0; c < 3; c++) {
                s.n[c] = op->n[c];
                s.periodic[c] = op->periodic[c];
                s.has[c] = op->ax.has[c];
                for (int k = 0; k < BMAXR; k++) {
                    s.cp[c][k] = k < op->ax.R ? op->ax.cp[c][k] : 0.0;
                    s.cm[c][k] = k < op->ax.R ? op->ax.cm[c][k] : 0.0;
                }
                if (op->n[c] < op->ax.R) {
                    set_error("batched Poisson: grid %d level %d smaller "
                              "than the stencil",
                              g, l);
                    delete b;
                    return EINVAL_;
                }
            }
            s.c0 = op->ax.c0;
            s.w = weights[l];
            lv.st.push_back(s);
            off += op->ng();
        }
        lv.total = off;
        fill_tiles(lv, BTile());
    }
    b->l0_64 = BLevel();
    b->l0_64.R = b->lev[0].R;
    b->l0_64.st = b->lev[0].st;
    b->l0_64.total = b->lev[0].total;
    fill_tiles(b->l0_64, BTile());
    for (int l = 0; l + 1 < L; l++) {
        BPair& pr = b->pair[l];
        long ioff = 0, ooff = 0, p0r = 0, p0i = 0;
        for (int g = 0; g < G; g++) {
            gfd_transformer* r = rst[g * (L - 1) + l];
            gfd_transformer* i = itp[g * (L - 1) + l];
            if (!r || !i || !r->real || !i->real || r->plan.K != 2 ||
                i->plan.K != 2 || r->plan.interpolate ||
                !i->plan.interpolate) {
                set_error("batched Poisson: grid %d level %d needs real nn=1 "
                          "restriction / interpolation",
                          g, l);
                delete b;
                return EINVAL_;
            }
            const BSt& f = b->lev[l].st[g];
            const BSt& c = b->lev[l + 1].st[g];
            for (int d = 0; d < 3; d++)
                if (r->plan.nin[d] != f.n[d] || r->plan.nout[d] != c.n[d] ||
                    i->plan.nin[d] != c.n[d] || i->plan.nout[d] != f.n[d]) {
                    set_error("batched Poisson: grid %d level %d transfer "
                              "shapes do not match the operators",
                              g, l);
                    delete b;
                    return EINVAL_;
                }
            BXf x{};
            x.in_off = f.off;
            x.out_off = c.off;
            x.p0 = p0r;
            for (int d = 0; d < 3; d++) {
                x.ni[d] = f.n[d];
                x.no[d] = c.n[d];
                x.off[d] = r->plan.ax[d].off;
                x.skip0[d] = r->plan.ax[d].skip0;
                x.periodic[d] = r->periodic[d];
            }
            pr.rst.push_back(x);
            p0r += c.n[0] * (long)c.n[1] * c.n[2];
            BXf y{};
            y.in_off = c.off;
            y.out_off = f.off;
            y.p0 = p0i;
            for (int d = 0; d < 3; d++) {
                y.ni[d] = c.n[d];
                y.no[d] = f.n[d];
                y.off[d] = i->plan.ax[d].off;
                y.skip0[d] = i->plan.ax[d].skip0;
                y.periodic[d] = i->periodic[d];
            }
            pr.itp.push_back(y);
            p0i += f.n[0] * (long)f.n[1] * f.n[2];
            (void)ioff;
            (void)ooff;
        }
        pr.nr = p0r;
        pr.ni = p0i;
    }
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, dev) == hipSuccess)
            b->num_cu = pr.multiProcessorCount;
    }
    (void)hipGetLastError();
    b->e32.resize(L);
    b->t32.resize(L);
    b->r32.resize(L);
    b->s32.resize(L);
    *out = b;
    return OK;
}

extern "C" void gfd_batch_poisson_free(gfd_batch_poisson* b)
{
    if (!b) return;
    if (b->gs) (void)hipStreamSynchronize(b->gs);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (auto& lv : b->lev) lv.d_st.release();
    b->l0_64.d_st.release();
    for (auto& pr : b->pair) {
        pr.d_rst.release();
        pr.d_itp.release();
    }
    for (int l = 0; l < b->L; l++) {
        b->e32[l].release();
        b->t32[l].release();
        b->r32[l].release();
        b->s32[l].release();
    }
    b->phi.release();
    b->rho.release();
    b->d_cgrids.release();
    b->partial.release();
    b->norms.release();
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->gs) (void)hipStreamDestroy(b->gs);
    delete b;
}

extern "C" int gfd_batch_poisson_set_coarse(gfd_batch_poisson* b,
                                            const int* nlev,
                                            gfd_operator** ops,
                                            gfd_transformer** rst,
                                            gfd_transformer** itp,
                                            const double* weights)
{
    if (!b || (nlev && (!ops || !weights))) {
        set_error("gfd_batch_poisson_set_coarse: bad arguments");
        return EINVAL_;
    }
    b->coarse = false;
    b->ckey[0] = b->ckey[1] = b->ckey[2] = -1;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    if (!nlev) return OK;
    const int lc = b->L - 1;
    b->nlev.assign(nlev, nlev + b->G);
    b->cops.assign(b->G, {});
    b->crst.assign(b->G, {});
    b->citp.assign(b->G, {});
    int maxlev = 0;
    long o = 0, t = 0;
    for (int g = 0; g < b->G; g++) {
        const int n = nlev[g];
        if (n < b->L || n - lc > CMAXLEV || (n > 1 && (!rst || !itp))) {
            set_error("batched Poisson: grid %d has %d levels (need %d .. "
                      "%d)",
                      g, n, b->L, lc + CMAXLEV);
            return EINVAL_;
        }
        maxlev = std::max(maxlev, n);
        for (int l = 0; l < n; l++) {
            gfd_operator* op = ops[o + l];
            if (!op || !op->real || !op->axis_ok || op->ax.R > CMAXR) {
                set_error("batched Poisson: grid %d level %d needs a real "
                          "axis-only stencil",
                          g, l);
                return EINVAL_;
            }
            b->cops[g].push_back(op);
        }
        for (int l = 0; l + 1 < n; l++) {
            gfd_transformer* r = rst[t + l];
            gfd_transformer* i = itp[t + l];
            if (!r || !i || r->plan.K != 2 || i->plan.K != 2) {
                set_error("batched Poisson: grid %d level %d needs nn=1 "
                          "transfers",
                          g, l);
                return EINVAL_;
            }
            b->crst[g].push_back(r);
            b->citp[g].push_back(i);
        }
        o += n;
        t += n - 1;
    }
    b->cw.assign(weights, weights + maxlev);
    b->coarse = true;
    return OK;
}

extern "C" long gfd_batch_poisson_level_points(gfd_batch_poisson* b,
                                               int level)
{
    if (!b || level < 0 || level >= b->L) return -1;
    return b->lev[level].total;
}

namespace gfd {

static int b_alloc(gfd_batch_poisson* b, hipStream_t s)
{
    GFD_TRY(b->phi.ensure(8 * b->lev[0].total, s));
    GFD_TRY(b->rho.ensure(8 * b->lev[0].total, s));
    for (int l = 0; l < b->L; l++) {
        const size_t n = 4 * (size_t)b->lev[l].total + 16;
        GFD_TRY(b->e32[l].ensure(n, s));
        GFD_TRY(b->t32[l].ensure(n, s));
        GFD_TRY(b->r32[l].ensure(n, s));
        GFD_TRY(b->s32[l].ensure(n, s));
    }
    for (auto& pr : b->pair) {
        GFD_TRY(pr.d_rst.ensure(sizeof(BXf) * pr.rst.size(), s));
        GFD_TRY(pr.d_itp.ensure(sizeof(BXf) * pr.itp.size(), s));
        GFD_CHECK_HIP(hipMemcpyAsync(pr.d_rst.p, pr.rst.data(),
                                     sizeof(BXf) * pr.rst.size(),
                                     hipMemcpyHostToDevice, s));
        GFD_CHECK_HIP(hipMemcpyAsync(pr.d_itp.p, pr.itp.data(),
                                     sizeof(BXf) * pr.itp.size(),
                                     hipMemcpyHostToDevice, s));
    }
    GFD_TRY(b->norms.ensure(sizeof(double) * b->G, s));
    GFD_CHECK_HIP(hipStreamSynchronize(s));
    return OK;
}

static int g_bxkind = 0;   // 0: timed choice, 1: one point per thread,
                          // 2: x-march (timed), 3: the forced x-march tile
static BXTile g_bxforce;

void batch_set_xfer(long code)
{
    g_bxkind = code >= 64 ? 3 : (int)code;
    if (code >= 64) {
        g_bxforce = BXTile();
        g_bxforce.nt = (int)(code & 0xfff);
        g_bxforce.py = (int)((code >> 12) & 0xf);
        g_bxforce.zmax = (int)((code >> 16) & 0xfff);
        g_bxforce.xc = (int)((code >> 28) & 0xfff);
    }
}

static int upload_xtable(const std::vector<BXf>& v, DevBuf& d, hipStream_t s)
{
    GFD_TRY(d.ensure(sizeof(BXf) * v.size(), s));
    GFD_CHECK_HIP(hipMemcpyAsync(d.p, v.data(), sizeof(BXf) * v.size(),
                                 hipMemcpyHostToDevice, s));
    return hipStreamSynchronize(s) == hipSuccess ? OK : EHIP;
}

// Transfer kernel and tile per direction of pair l, timed on the V-cycle's
// FP32 passes (restriction PLAIN, interpolation SUBFROM).
static int tune_pair(gfd_batch_poisson* b, int l, hipStream_t s)
{
    BPair& pr = b->pair[l];
    std::vector<BXTile> cand;
    if (g_bxkind == 0 || g_bxkind == 1) cand.push_back(BXTile());
    if (g_bxkind == 3) cand.push_back(g_bxforce);
    if (g_bxkind == 0 || g_bxkind == 2)
        for (int nt : {256, 512})
            for (int py : {1, 2, 4})
                for (int zmax : {64, 128})
                    for (int xc : {4, 8, 16}) {
                        BXTile t;
                        t.nt = nt;
                        t.py = py;
                        t.zmax = zmax;
                        t.xc = xc;
                        cand.push_back(t);
                    }
    hipEvent_t e0, e1;
    GFD_CHECK_HIP(hipEventCreate(&e0));
    GFD_CHECK_HIP(hipEventCreate(&e1));
    int st = OK;
    for (int interp = 0; interp < 2 && st == OK; interp++) {
        std::vector<BXf>& v = interp ? pr.itp : pr.rst;
        DevBuf& dv = interp ? pr.d_itp : pr.d_rst;
        BXTile& cur = interp ? pr.titp : pr.trst;
        const void* in = interp ? b->e32[l + 1].p : b->s32[l].p;
        void* out = interp ? b->t32[l].p : b->r32[l + 1].p;
        const void* base = interp ? b->e32[l].p : nullptr;
        float best = 1e30f;
        BXTile bt;
        bool found = false;
        for (BXTile t : cand) {
            if (!fill_xtiles(v, t)) continue;
            if ((st = upload_xtable(v, dv, s)) != OK) break;
            cur = t;
            if ((st = bxfer(pr, interp, 2, in, out, base, b->num_cu, s)) != OK)
                break;
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < 3 && st == OK; r++)
                st = bxfer(pr, interp, 2, in, out, base, b->num_cu, s);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            if (st != OK) break;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) {
                best = ms;
                bt = t;
                found = true;
            }
        }
        if (st != OK) break;
        if (!found) {
            set_error("batched transfer: no tile fits level %d", l);
            st = ENOTSUP_;
            break;
        }
        fill_xtiles(v, bt);
        cur = bt;
        st = upload_xtable(v, dv, s);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

static int b_tune(gfd_batch_poisson* b, hipStream_t s)
{
    GFD_TRY(b_alloc(b, s));
    for (int l = 0; l < b->L; l++) {
        BLevel& lv = b->lev[l];
        GFD_TRY(vec_zero_bytes(b->e32[l].p, 4 * lv.total, s));
        GFD_TRY(vec_zero_bytes(b->r32[l].p, 4 * lv.total, s));
        GFD_TRY(tune_level(b, lv, 2, EPI_RELAX, b->e32[l].p, b->r32[l].p,
                           b->t32[l].p, s));
    }
    GFD_TRY(vec_zero((double*)b->phi.p, b->lev[0].total, s));
    GFD_TRY(vec_zero((double*)b->rho.p, b->lev[0].total, s));
    GFD_TRY(tune_level(b, b->l0_64, 4, EPI_RESID, b->phi.p, b->rho.p,
                       b->r32[0].p, s));
    for (int l = 0; l + 1 < b->L; l++) GFD_TRY(tune_pair(b, l, s));
    long mt = 0;
    for (auto& lv : b->lev) mt = std::max(mt, lv.ntiles);
    mt = std::max(mt, b->l0_64.ntiles);
    GFD_TRY(b->partial.ensure(sizeof(double) * (mt + 1), s));
    b->tuned = true;
    return OK;
}

// Jacobi sweeps at level l (FP32), result in *cur
static int b_relax(gfd_batch_poisson* b, int l, void** cur, void** alt,
                   const void* rhs, int n, hipStream_t s)
{
    for (int i = 0; i < n; i++) {
        GFD_TRY(bst(b->lev[l], 2, EPI_RELAX, *cur, *alt, rhs, nullptr, s));
        std::swap(*cur, *alt);
    }
    return OK;
}

// n Jacobi sweeps from e = 0 (*cur is not read): the first is the scaled
// right-hand side; n = 0 leaves e = 0
static int b_relax0(gfd_batch_poisson* b, int l, void** cur, void** alt,
                    const void* rhs, int n, hipStream_t s)
{
    const BLevel& lv = b->lev[l];
    if (n <= 0) return vec_zero_bytes(*cur, 4 * lv.total, s);
    const long nb = std::min((lv.total + 1023) / 1024, (long)b->num_cu * 16);
    hipLaunchKernelGGL(k_bscale0, dim3((unsigned)std::max(1L, nb)), dim3(256),
                       0, s, (const BSt*)lv.d_st.p, (int)lv.st.size(),
                       lv.total, (const float*)rhs, (float*)*cur);
    GFD_LAUNCH_CHECK();
    return b_relax(b, l, cur, alt, rhs, n - 1, s);
}

// FP32 V-cycle on e_l for right-hand side r_l, from e_l = 0 (*cur is not
// read).  phi (level 0 only, may be null): when the level's stencil runs
// on x-march tiles, the last post-smoothing sweep subtracts its result
// from the FP64 phi in the same pass and *sub is set (e_0 is then not
// stored).
static int b_vcycle(gfd_batch_poisson* b, int l, void** cur, void** alt,
                    const void* rhs, int pre, int post, int coarse,
                    hipStream_t s, double* phi = nullptr,
                    bool* sub = nullptr)
{
    if (l + 1 < b->L) {
        GFD_TRY(b_relax0(b, l, cur, alt, rhs, pre, s));
        GFD_TRY(bst(b->lev[l], 2, EPI_RESID, *cur, b->s32[l].p, rhs, nullptr,
                    s));
        GFD_TRY(bxfer(b->pair[l], false, 2, b->s32[l].p, b->r32[l + 1].p,
                      nullptr, b->num_cu, s));
        void* c1 = b->e32[l + 1].p;
        void* a1 = b->t32[l + 1].p;
        GFD_TRY(b_vcycle(b, l + 1, &c1, &a1, b->r32[l + 1].p, pre, post,
                         coarse, s));
        GFD_TRY(bxfer(b->pair[l], true, 2, c1, *alt, *cur, b->num_cu, s));
        std::swap(*cur, *alt);
        if (phi && post >= 1 && b->lev[l].tile.kind == 1) {
            GFD_TRY(b_relax(b, l, cur, alt, rhs, post - 1, s));
            GFD_TRY(bst(b->lev[l], 2, BEPI_RELAXSUB, *cur, phi, rhs, nullptr,
                        s));
            *sub = true;
            return OK;
        }
        return b_relax(b, l, cur, alt, rhs, post, s);
    }
    if (b->coarse)  // starts from e = 0 in LDS
        return launch_coarse((const CGrid*)b->d_cgrids.p, b->G, b->clds, 2,
                             rhs, *cur, s);
    return b_relax0(b, l, cur, alt, rhs, coarse, s);
}

// CGrid of every grid for levels L-1 .. nlev_g - 1 (pre / post sweeps,
// `coarse` sweeps on each grid's coarsest level); uploaded when changed.
static int b_plan_coarse(gfd_batch_poisson* b, int pre, int post, int coarse,
                         hipStream_t s, bool* changed)
{
    *changed = false;
    if (!b->coarse) return OK;
    const int key[4] = {pre, post, coarse, coarse_cheb_enabled() ? 1 : 0};
    if (std::equal(key, key + 4, b->ckey)) return OK;
    const int lc = b->L - 1;
    b->cgrids.assign(b->G, CGrid{});
    size_t maxb = 0;
    for (int g = 0; g < b->G; g++) {
        CGrid& cg = b->cgrids[g];
        cg.nlev = b->nlev[g] - lc;
        int off = 0;
        for (int i = 0; i < cg.nlev; i++) {
            const int l = lc + i;
            const gfd_operator* op = b->cops[g][l];
            CLevel& L = cg.lev[i];
            for (int d = 0; d < 3; d++) {
                L.n[d] = op->n[d];
                L.periodic[d] = op->periodic[d];
                L.has[d] = op->ax.has[d];
                for (int k = 0; k < CMAXR; k++) {
                    L.cp[d][k] = k < op->ax.R ? op->ax.cp[d][k] : 0.0;
                    L.cm[d][k] = k < op->ax.R ? op->ax.cm[d][k] : 0.0;
                }
            }
            L.R = op->ax.R;
            L.c0 = op->ax.c0;
            L.w = b->cw[l];
            L.pre = pre;
            L.post = i + 1 == cg.nlev ? coarse : post;
            const int n = (int)op->ng();
            L.a = off;
            off += n;
            L.b = off;
            off += n;
            L.r = off;  // first level: staged, dropped below if no room
            off += n;
            if (i + 1 < cg.nlev)
                for (int d = 0; d < 3; d++) {
                    L.roff[d] = b->crst[g][l]->plan.ax[d].off;
                    L.ioff[d] = b->citp[g][l]->plan.ax[d].off;
                    L.iskip0[d] = b->citp[g][l]->plan.ax[d].skip0;
                }
        }
        if ((size_t)off * 4 > COARSE_LDS_MAX) {
            // no room to stage the first level's rhs: read it from global
            // memory in every pass (it sits last in the layout)
            const int n0 = cg.lev[0].n[0] * cg.lev[0].n[1] * cg.lev[0].n[2];
            const int r0 = cg.lev[0].r;
            for (int i = 1; i < cg.nlev; i++) {
                CLevel& L = cg.lev[i];
                if (L.a > r0) L.a -= n0;
                if (L.b > r0) L.b -= n0;
                if (L.r > r0) L.r -= n0;
            }
            cg.lev[0].r = -1;
            off -= n0;
        }
        // Chebyshev iteration on the coarsest level when its update array
        // fits (the batch's inner cycles are the mixed schedule's)
        {
            CLevel& C = cg.lev[cg.nlev - 1];
            const int nc = C.n[0] * C.n[1] * C.n[2];
            double lo, hi;
            C.cheb = 0;
            C.dbuf = -1;
            if (coarse_cheb_enabled() && (size_t)(off + nc) * 4 <= COARSE_LDS_MAX &&
                cheb_bounds(C, &lo, &hi)) {
                C.cheb = 1;
                C.dbuf = off;
                C.lam_lo = lo;
                C.lam_hi = hi;
                off += nc;
            }
        }
        cg.lds_elems = off;
        cg.rhs_off = b->lev[lc].st[g].off;
        cg.out_off = b->lev[lc].st[g].off;
        maxb = std::max(maxb, (size_t)off * 4);
    }
    if (maxb > COARSE_LDS_MAX) {
        set_error("batched Poisson: coarse levels need %zu bytes of LDS",
                  maxb);
        return ENOTSUP_;
    }
    GFD_TRY(b->d_cgrids.ensure(sizeof(CGrid) * b->G, s));
    GFD_CHECK_HIP(hipMemcpyAsync(b->d_cgrids.p, b->cgrids.data(),
                                 sizeof(CGrid) * b->G, hipMemcpyHostToDevice,
                                 s));
    GFD_CHECK_HIP(hipStreamSynchronize(s));
    b->clds = maxb;
    std::copy(key, key + 4, b->ckey);
    *changed = true;
    return OK;
}

// One refinement step: e = V(r32), phi -= e, r = L phi - rho (FP64, per-grid
// norms), r32 = (float) r.  Entry: r32[0] holds the current residual.
static int b_step(gfd_batch_poisson* b, int pre, int post, int coarse,
                  hipStream_t s)
{
    const long n0 = b->lev[0].total;
    void* cur = b->e32[0].p;
    void* alt = b->t32[0].p;
    bool sub = false;
    GFD_TRY(b_vcycle(b, 0, &cur, &alt, b->r32[0].p, pre, post, coarse, s,
                     (double*)b->phi.p, &sub));
    if (!sub) GFD_TRY(vec_sub_f((double*)b->phi.p, (const float*)cur, n0, s));
    GFD_TRY(bst(b->l0_64, 4, EPI_RESID, b->phi.p, b->r32[0].p, b->rho.p,
                (double*)b->partial.p, s));
    return bnorms(b->l0_64, (const double*)b->partial.p,
                  (double*)b->norms.p, s);
}

static int b_read_norms(gfd_batch_poisson* b, std::vector<double>& h,
                        hipStream_t s)
{
    h.resize(b->G);
    GFD_CHECK_HIP(hipMemcpyAsync(h.data(), b->norms.p, sizeof(double) * b->G,
                                 hipMemcpyDeviceToHost, s));
    GFD_CHECK_HIP(hipStreamSynchronize(s));
    return OK;
}

}  // namespace gfd

extern "C" int gfd_batch_poisson_tune(gfd_batch_poisson* b, void* stream)
{
    if (!b) {
        set_error("gfd_batch_poisson_tune: null handle");
        return EINVAL_;
    }
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return b_tune(b, (hipStream_t)stream);
}

extern "C" int gfd_batch_poisson_tiles(gfd_batch_poisson* b, int level,
                                       int* bz, int* by, int* pz,
                                       long* blocks)
{
    if (!b || level < -1 || level >= b->L || !bz || !by || !pz || !blocks) {
        set_error("gfd_batch_poisson_tiles: bad arguments");
        return EINVAL_;
    }
    const BTile& t = level < 0 ? b->l0_64.tile : b->lev[level].tile;
    const bool regs = t.kind == 0;
    *bz = regs ? t.bz : 0;
    *by = regs ? t.by : 0;
    *pz = regs ? t.pz : 0;
    *blocks = regs ? t.blocks : (level < 0 ? b->l0_64 : b->lev[level]).ntiles;
    return OK;
}

extern "C" int gfd_batch_poisson_tile_info(gfd_batch_poisson* b, int level,
                                           long* info, int ninfo)
{
    if (!b || level < -1 || level >= b->L || !info || ninfo < 1) {
        set_error("gfd_batch_poisson_tile_info: bad arguments");
        return EINVAL_;
    }
    const BLevel& lv = level < 0 ? b->l0_64 : b->lev[level];
    const BTile& t = lv.tile;
    const long v[] = {t.kind, t.xc,     lv.ntiles, t.bz,   t.by, t.pz,
                      t.blocks, t.nt, t.py,      t.zmax, t.pf};
    const int n = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < ninfo; i++) info[i] = i < n ? v[i] : 0;
    return OK;
}

extern "C" int gfd_batch_poisson_apply(gfd_batch_poisson* b, int level,
                                       int fp32, const void* in, void* out,
                                       void* stream)
{
    if (!b || level < 0 || level >= b->L || !in || !out || in == out) {
        set_error("gfd_batch_poisson_apply: bad arguments");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    if (!b->tuned) GFD_TRY(b_tune(b, s));
    return bst(b->lev[level], fp32 ? 2 : 0, EPI_PLAIN, in, out, nullptr,
               nullptr, s);
}

extern "C" int gfd_batch_poisson_transfer(gfd_batch_poisson* b, int level,
                                          int interpolate, int fp32,
                                          const void* in, void* out,
                                          const void* base, void* stream)
{
    if (!b || level < 0 || level + 1 >= b->L || !in || !out ||
        (interpolate && !base)) {
        set_error("gfd_batch_poisson_transfer: bad arguments");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    if (!b->tuned) GFD_TRY(b_tune(b, s));
    return bxfer(b->pair[level], interpolate != 0, fp32 ? 2 : 0, in, out,
                 base, b->num_cu, s);
}

extern "C" int gfd_batch_poisson_solve_mixed(
    gfd_batch_poisson* b, double* const* phis, const double* const* rhos,
    double tol_rel, int pre, int post, int coarse_sweeps, int max_outer,
    int use_graph, int* n_outer, double* relres, void* stream)
{
    if (!b || !phis || !rhos || !n_outer || !relres || max_outer < 1 ||
        pre < 0 || post < 0 || coarse_sweeps < 0) {
        set_error("gfd_batch_poisson_solve_mixed: bad arguments");
        return EINVAL_;
    }
    for (int g = 0; g < b->G; g++)
        if (!phis[g] || !rhos[g]) {
            set_error("gfd_batch_poisson_solve_mixed: null grid %d", g);
            return EINVAL_;
        }
    hipStream_t cs = (hipStream_t)stream;
    if (!b->gs) {
        GFD_CHECK_HIP(hipStreamCreateWithFlags(&b->gs, hipStreamNonBlocking));
        GFD_CHECK_HIP(hipEventCreateWithFlags(&b->ev_in,
                                              hipEventDisableTiming));
        GFD_CHECK_HIP(hipEventCreateWithFlags(&b->ev_out,
                                              hipEventDisableTiming));
    }
    hipStream_t s = b->gs;
    GFD_CHECK_HIP(hipEventRecord(b->ev_in, cs));
    GFD_CHECK_HIP(hipStreamWaitEvent(s, b->ev_in, 0));
    if (!b->tuned) GFD_TRY(b_tune(b, s));
    // gather phi, rho; |rho_g|^2
    const BLevel& l0 = b->lev[0];
    std::vector<double> rn2(b->G);
    GFD_TRY(b->norms.ensure(sizeof(double) * b->G, s));
    for (int g = 0; g < b->G; g++) {
        const long ng = (long)l0.st[g].n[0] * l0.st[g].n[1] * l0.st[g].n[2];
        double* pd = (double*)b->phi.p + l0.st[g].off;
        double* rd = (double*)b->rho.p + l0.st[g].off;
        GFD_CHECK_HIP(hipMemcpyAsync(pd, phis[g], 8 * ng,
                                     hipMemcpyDeviceToDevice, s));
        GFD_CHECK_HIP(hipMemcpyAsync(rd, rhos[g], 8 * ng,
                                     hipMemcpyDeviceToDevice, s));
    }
    // r = L phi - rho with norms; rn2 from the same pass on phi = 0 is not
    // needed: |rho|^2 by a dot per grid
    {
        DevBuf tmp;
        const long ng0 = l0.total;
        GFD_TRY(tmp.ensure(sizeof(double) * (vec_dot_partials(ng0) + 8), s));
        for (int g = 0; g < b->G; g++) {
            const long ng =
                (long)l0.st[g].n[0] * l0.st[g].n[1] * l0.st[g].n[2];
            const double* rd = (const double*)b->rho.p + l0.st[g].off;
            GFD_TRY(vec_dot(rd, rd, ng, (double*)tmp.p + 1, (double*)tmp.p,
                            s));
            GFD_CHECK_HIP(hipMemcpyAsync(&rn2[g], tmp.p, sizeof(double),
                                         hipMemcpyDeviceToHost, s));
            GFD_CHECK_HIP(hipStreamSynchronize(s));
        }
        tmp.release();
    }
    GFD_TRY(bst(b->l0_64, 4, EPI_RESID, b->phi.p, b->r32[0].p, b->rho.p,
                (double*)b->partial.p, s));
    GFD_TRY(bnorms(b->l0_64, (const double*)b->partial.p,
                   (double*)b->norms.p, s));
    std::vector<double> err2;
    GFD_TRY(b_read_norms(b, err2, s));
    auto done = [&]() {
        for (int g = 0; g < b->G; g++)
            if (err2[g] > tol_rel * tol_rel * rn2[g]) return false;
        return true;
    };
    bool cchanged = false;
    GFD_TRY(b_plan_coarse(b, pre, post, coarse_sweeps, s, &cchanged));
    const int key[4] = {pre, post, coarse_sweeps, b->coarse ? 1 : 0};
    if (b->graph && (cchanged || !std::equal(key, key + 4, b->gkey))) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    int outer = 0;
    while (!done() && outer < max_outer) {
        if (use_graph) {
            if (!b->graph) {
                hipGraph_t gr = nullptr;
                GFD_CHECK_HIP(hipStreamBeginCapture(
                    s, hipStreamCaptureModeThreadLocal));
                const int st = b_step(b, pre, post, coarse_sweeps, s);
                const hipError_t e = hipStreamEndCapture(s, &gr);
                if (st != OK) {
                    if (gr) (void)hipGraphDestroy(gr);
                    return st;
                }
                if (e != hipSuccess) {
                    set_error("graph capture failed: %s",
                              hipGetErrorString(e));
                    return EHIP;
                }
                const hipError_t e2 =
                    hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e2 != hipSuccess) {
                    b->graph = nullptr;
                    set_error("graph instantiation failed: %s",
                              hipGetErrorString(e2));
                    return EHIP;
                }
                std::copy(key, key + 4, b->gkey);
            }
            GFD_CHECK_HIP(hipGraphLaunch(b->graph, s));
        } else {
            GFD_TRY(b_step(b, pre, post, coarse_sweeps, s));
        }
        GFD_TRY(b_read_norms(b, err2, s));
        outer++;
    }
    for (int g = 0; g < b->G; g++) {
        const long ng = (long)l0.st[g].n[0] * l0.st[g].n[1] * l0.st[g].n[2];
        GFD_CHECK_HIP(hipMemcpyAsync(phis[g],
                                     (double*)b->phi.p + l0.st[g].off, 8 * ng,
                                     hipMemcpyDeviceToDevice, s));
        relres[g] = rn2[g] > 0 ? std::sqrt(err2[g] / rn2[g])
                               : std::sqrt(err2[g]);
    }
    GFD_CHECK_HIP(hipEventRecord(b->ev_out, s));
    GFD_CHECK_HIP(hipStreamWaitEvent(cs, b->ev_out, 0));
    *n_outer = outer;
    if (!done()) {
        set_error("batched mixed-precision Poisson: not converged to %g in "
                  "%d refinement steps",
                  tol_rel, max_outer);
        return ECONVERGE;
    }
    return OK;
}

// ------------------------------------------------- batched preconditioner

// The FP32 kinetic preconditioner chain of gfd_fd_precond32
// (c/fd_preconditioner.cpp:221-345) on every grid of a three-level batch
// at once: eight persistent launches, each over all grids.  Work arrays
// (level 0: q0; levels 1, 2: r, d, q) are the batch's FP32 level arrays.
// steps[g] is grid g's 0.66666666 / diag(K0); levels 1 and 2 use 4 and 16
// times it, as the reference does.
extern "C" int gfd_batch_precond32(gfd_batch_poisson* b, const double* steps,
                                   const double* res, double* out,
                                   void* stream)
{
    if (!b || !steps || !res || !out || (const void*)res == (void*)out) {
        set_error("gfd_batch_precond32: bad arguments");
        return EINVAL_;
    }
    if (b->L != 3) {
        set_error("gfd_batch_precond32: the batch needs exactly 3 levels "
                  "(has %d)", b->L);
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    bool changed = false;
    for (int g = 0; g < b->G; g++)
        if (!std::isfinite(steps[g]) || steps[g] <= 0) {
            set_error("gfd_batch_precond32: step %d must be positive", g);
            return EINVAL_;
        }
    for (int g = 0; g < b->G; g++) {
        const double st = steps[g];
        if (b->lev[0].st[g].step != st) changed = true;
        b->lev[0].st[g].step = st;
        b->lev[1].st[g].step = 4 * st;
        b->lev[2].st[g].step = 16 * st;
        b->pair[0].rst[g].alpha = 4 * st;
        b->pair[1].rst[g].alpha = 16 * st;
    }
    if (!b->tuned) {
        GFD_TRY(b_tune(b, s));  // uploads the level and transfer tables
    } else if (changed) {
        GFD_TRY(b_alloc(b, s));
        for (auto& lv : b->lev) GFD_TRY(upload_level(lv, s));
    }
    void* q0 = b->e32[0].p;
    void* r1 = b->r32[1].p;
    void* d1 = b->e32[1].p;
    void* q1 = b->t32[1].p;
    void* r2 = b->r32[2].p;
    void* d2 = b->e32[2].p;
    void* q2 = b->t32[2].p;
    const int cu = b->num_cu;
    // r1 = R0(-res) (narrowed), d1 = 4 step r1
    GFD_TRY(bxfer(b->pair[0], false, 4, res, r1, nullptr, cu, s, XEPI_SCALE2,
                  d1, 1));
    // q1 = K1 d1 - r1
    GFD_TRY(bst(b->lev[1], 2, EPI_RESID, d1, q1, r1, nullptr, s));
    // r2 = R1(q1), d2 = 16 step r2
    GFD_TRY(bxfer(b->pair[1], false, 2, q1, r2, nullptr, cu, s, XEPI_SCALE2,
                  d2, 0));
    // q2 = d2 - 16 step (K2 d2 - r2)
    GFD_TRY(bst(b->lev[2], 2, EPI_UPDATE, d2, q2, r2, nullptr, s));
    // d1 -= I2(q2)
    GFD_TRY(bxfer(b->pair[1], true, 2, q2, d1, d1, cu, s, XEPI_SUBFROM));
    // q1 = d1 - 4 step (K1 d1 - r1)
    GFD_TRY(bst(b->lev[1], 2, EPI_UPDATE, d1, q1, r1, nullptr, s));
    // q0 = -I1(q1)
    GFD_TRY(bxfer(b->pair[0], true, 2, q1, q0, nullptr, cu, s, XEPI_NEG));
    // out = -(q0 - step (K0 q0 - res)), FP64 residual and result
    return bst(b->lev[0], 2, EPI_UPDW, q0, out, res, nullptr, s, 1);
}
