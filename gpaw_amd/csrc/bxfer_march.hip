// Batched K = 2 restriction and interpolation on x-march tiles (the
// transfers of the batched multigrid of batch.hip over many uneven real
// grids).
//
// One block per tile of one grid's OUTPUT: TZ = mz output columns by
// mseg * PY output rows over a 1-D block (thread t: column t % mz, row
// segment t / mz; threads past the last full segment repeat its loads and
// store nothing), marching along an x chunk of output planes.  The tile's
// grid comes from a binary search over the table's first-tile column.
//
//  * interpolation: output plane o0 averages input planes c0 (and c0 + 1
//    for odd q0 = o0 + skip0); the (y, z) average of an input plane
//    (V2, up to 4 input points per output point, cache-served) is kept for
//    the two current input planes, so each input plane is averaged once
//    per output column and row while the march passes it;
//  * restriction: output plane o0 is 1/2 (P(c) + 1/2 (P(c+1) + P(c-1)))
//    with c = 2 o0 + off and P the (y, z) weighted sum of a fine plane (9
//    points); P(c + 1) carries over as the next output's P(c - 1), so each
//    fine plane is summed once.
//
// Weights and association order are those of the one-point-per-thread
// kernel of batch.hip (z innermost, then y, then x), so both give the same
// bits.  Global traffic goes through per-plane buffer resources (batch.h):
// zero boundaries and idle lanes are out-of-range offsets.
#include "batch.h"
#include "stencil_dev.h"
#include "transfer.h"

namespace gfd {

__device__ inline int bfind_xtile(const BXf* g, int G, long t)
{
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g[mid].tile0 <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// wrapped index along an axis, -1 beyond a zero boundary
__device__ inline int bxw(int i, int n, int periodic)
{
    return bwrap(i, n, periodic) ? i : -1;
}

// byte offset of input point (row base yb, column z) or BXM_OOB
template <typename TI>
__device__ inline unsigned bxo(int yb, int z)
{
    return (yb < 0 || z < 0) ? BXM_OOB : (unsigned)(yb + z) * sizeof(TI);
}

struct BxTile {
    int tx, sg, TY, z0, y0, xs, xe;
    bool lane_on;
};

__device__ inline BxTile bx_tile(const BXf& d, int L, int PY)
{
    BxTile t;
    long lt = L - d.tile0;
    const int zt = (int)(lt % d.tz);
    lt /= d.tz;
    const int yt = (int)(lt % d.ty);
    const int xt = (int)(lt / d.ty);
    const int tid = threadIdx.x;
    t.tx = tid % d.mz;
    const int sg0 = tid / d.mz;
    t.lane_on = sg0 < d.mseg;
    t.sg = t.lane_on ? sg0 : d.mseg - 1;
    t.TY = d.mseg * PY;
    t.z0 = zt * d.mz;
    t.y0 = yt * t.TY;
    t.xs = xt * d.xc;
    t.xe = min(t.xs + d.xc, d.no[0]);
    return t;
}

template <typename T, typename TI, int PY, int XE>
__global__ void __launch_bounds__(512) k_bxinterp(BXArgs a)
{
    typedef typename ScalarOf<T>::type S;
    const int L = xcd_remap(blockIdx.x, (int)a.ntiles);
    const BXf d = a.g[bfind_xtile(a.g, a.G, L)];
    const BxTile t = bx_tile(d, L, PY);
    const int ni1 = d.ni[1], ni2 = d.ni[2];
    const int no1 = d.no[1], no2 = d.no[2];
    const long iplane = (long)ni1 * ni2, oplane = (long)no1 * no2;

    // input columns of the thread's output column (clamped for lanes past
    // the row: their outputs are dropped)
    const int o2 = min(t.z0 + t.tx, no2 - 1);
    const int q2 = o2 + d.skip0[2];
    const int c2 = (q2 >> 1) + d.off[2];
    const bool two_z = (q2 & 1) != 0;
    const int zi0 = bxw(c2, ni2, d.periodic[2]);
    const int zi1 = two_z ? bxw(c2 + 1, ni2, d.periodic[2]) : -1;
    // per output row: offsets of the (up to) 4 input points of a plane
    unsigned o00[PY], o01[PY], o10[PY], o11[PY], oo[PY];
    bool two_y[PY];
#pragma unroll
    for (int j = 0; j < PY; j++) {
        const int o1r = t.y0 + t.sg * PY + j;
        const int o1 = min(o1r, no1 - 1);
        const int q1 = o1 + d.skip0[1];
        const int c1 = (q1 >> 1) + d.off[1];
        two_y[j] = (q1 & 1) != 0;
        const int y0i = bxw(c1, ni1, d.periodic[1]);
        const int y1i = two_y[j] ? bxw(c1 + 1, ni1, d.periodic[1]) : -1;
        const int yb0 = y0i < 0 ? -1 : y0i * ni2;
        const int yb1 = y1i < 0 ? -1 : y1i * ni2;
        o00[j] = bxo<TI>(yb0, zi0);
        o01[j] = bxo<TI>(yb0, zi1);
        o10[j] = bxo<TI>(yb1, zi0);
        o11[j] = bxo<TI>(yb1, zi1);
        oo[j] = (t.lane_on && o1r < no1 && t.z0 + t.tx < no2)
                    ? (unsigned)(o1r * no2 + t.z0 + t.tx)
                    : BXM_OOB;
    }
    const TI* gin = (const TI*)a.in + d.in_off;
    const unsigned ipb = (unsigned)(iplane * (long)sizeof(TI));
    // (y, z) average of input plane c
    auto v2 = [&](int c, T* v) {
        const int cw = bxw(c, d.ni[0], d.periodic[0]);
        const __amdgpu_buffer_rsrc_t r =
            brsrc(gin + (cw < 0 ? 0 : (long)cw * iplane), cw < 0 ? 0u : ipb);
#pragma unroll
        for (int j = 0; j < PY; j++) {
            const T a00 = (T)bload<TI>(r, o00[j]);
            const T a01 = (T)bload<TI>(r, o01[j]);
            const T a10 = (T)bload<TI>(r, o10[j]);
            const T a11 = (T)bload<TI>(r, o11[j]);
            const T vz0 = two_z ? (S)0.5 * (a00 + a01) : a00;
            const T vz1 = two_z ? (S)0.5 * (a10 + a11) : a10;
            v[j] = two_y[j] ? (S)0.5 * (vz0 + vz1) : vz0;
        }
    };
    const T* gbase = (const T*)a.base + d.out_off;
    T* gout = (T*)a.out + d.out_off;
    const unsigned opb = (unsigned)(oplane * (long)sizeof(T));
    auto ld_base = [&](int o0, T* v) {
        const __amdgpu_buffer_rsrc_t r = brsrc(gbase + (long)o0 * oplane, opb);
#pragma unroll
        for (int j = 0; j < PY; j++)
            v[j] = bload<T>(r, oo[j] == BXM_OOB ? BXM_OOB
                                                : oo[j] * (unsigned)sizeof(T));
    };

    int ca = ((t.xs + d.skip0[0]) >> 1) + d.off[0];
    T VA[PY], VB[PY], bb[PY], bn[PY];
    v2(ca, VA);
    v2(ca + 1, VB);
    if constexpr (XE == XEPI_SUBFROM) ld_base(t.xs, bb);
    for (int o0 = t.xs; o0 < t.xe; o0++) {
        const int q = o0 + d.skip0[0];
        const int c = (q >> 1) + d.off[0];
        if (c != ca) {  // c == ca + 1
#pragma unroll
            for (int j = 0; j < PY; j++) VA[j] = VB[j];
            ca = c;
            v2(ca + 1, VB);
        }
        if constexpr (XE == XEPI_SUBFROM)
            ld_base(min(o0 + 1, t.xe - 1), bn);
        const bool two_x = (q & 1) != 0;
        const __amdgpu_buffer_rsrc_t ro = brsrc(gout + (long)o0 * oplane, opb);
#pragma unroll
        for (int j = 0; j < PY; j++) {
            const T v = two_x ? (S)0.5 * (VA[j] + VB[j]) : VA[j];
            T r;
            if constexpr (XE == XEPI_SUBFROM)
                r = bb[j] - v;
            else if constexpr (XE == XEPI_NEG)
                r = -v;
            else
                r = v;
            bstore<T>(r, ro,
                      oo[j] == BXM_OOB ? BXM_OOB : oo[j] * (unsigned)sizeof(T));
        }
        if constexpr (XE == XEPI_SUBFROM) {
#pragma unroll
            for (int j = 0; j < PY; j++) bb[j] = bn[j];
        }
    }
}

template <typename T, typename TI, int PY, int XE>
__global__ void __launch_bounds__(512) k_bxrestrict(BXArgs a)
{
    typedef typename ScalarOf<T>::type S;
    const int L = xcd_remap(blockIdx.x, (int)a.ntiles);
    const BXf d = a.g[bfind_xtile(a.g, a.G, L)];
    const BxTile t = bx_tile(d, L, PY);
    const int ni1 = d.ni[1], ni2 = d.ni[2];
    const int no1 = d.no[1], no2 = d.no[2];
    const long iplane = (long)ni1 * ni2, oplane = (long)no1 * no2;

    // input columns 2 o2 + off - 1 .. + 1 and rows 2 o1 + off - 1 .. + 1
    const int o2 = min(t.z0 + t.tx, no2 - 1);
    const int c2 = 2 * o2 + d.off[2];
    int zi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) zi[k] = bxw(c2 - 1 + k, ni2, d.periodic[2]);
    int yb[PY][3];
    unsigned oo[PY];
#pragma unroll
    for (int j = 0; j < PY; j++) {
        const int o1r = t.y0 + t.sg * PY + j;
        const int o1 = min(o1r, no1 - 1);
        const int c1 = 2 * o1 + d.off[1];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int y = bxw(c1 - 1 + k, ni1, d.periodic[1]);
            yb[j][k] = y < 0 ? -1 : y * ni2;
        }
        oo[j] = (t.lane_on && o1r < no1 && t.z0 + t.tx < no2)
                    ? (unsigned)(o1r * no2 + t.z0 + t.tx)
                    : BXM_OOB;
    }
    const TI* gin = (const TI*)a.in + d.in_off;
    const unsigned ipb = (unsigned)(iplane * (long)sizeof(TI));
    // (y, z) weighted sum of input plane c
    auto p2 = [&](int c, T* v) {
        const int cw = bxw(c, d.ni[0], d.periodic[0]);
        const __amdgpu_buffer_rsrc_t r =
            brsrc(gin + (cw < 0 ? 0 : (long)cw * iplane), cw < 0 ? 0u : ipb);
#pragma unroll
        for (int j = 0; j < PY; j++) {
            T py[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const T am = (T)bload<TI>(r, bxo<TI>(yb[j][k], zi[0]));
                const T a0 = (T)bload<TI>(r, bxo<TI>(yb[j][k], zi[1]));
                const T ap = (T)bload<TI>(r, bxo<TI>(yb[j][k], zi[2]));
                py[k] = (S)0.5 * fmac(a0, (S)0.5, ap + am);
            }
            v[j] = (S)0.5 * fmac(py[1], (S)0.5, py[2] + py[0]);
        }
    };
    T* gout = (T*)a.out + d.out_off;
    T* gout2 = XE == XEPI_SCALE2 ? (T*)a.out2 + d.out_off : nullptr;
    const unsigned opb = (unsigned)(oplane * (long)sizeof(T));
    const T alpha = (T)d.alpha;

    T Pm[PY], P0[PY], Pp[PY];
    p2(2 * t.xs + d.off[0] - 1, Pm);
    for (int o0 = t.xs; o0 < t.xe; o0++) {
        const int c = 2 * o0 + d.off[0];
        p2(c, P0);
        p2(c + 1, Pp);
        const __amdgpu_buffer_rsrc_t ro = brsrc(gout + (long)o0 * oplane, opb);
#pragma unroll
        for (int j = 0; j < PY; j++) {
            T v = (S)0.5 * fmac(P0[j], (S)0.5, Pp[j] + Pm[j]);
            const unsigned ob =
                oo[j] == BXM_OOB ? BXM_OOB : oo[j] * (unsigned)sizeof(T);
            if constexpr (XE == XEPI_SCALE2) {
                if (a.negate) v = -v;
                bstore<T>(v, ro, ob);
                const __amdgpu_buffer_rsrc_t r2 =
                    brsrc(gout2 + (long)o0 * oplane, opb);
                bstore<T>(alpha * v, r2, ob);
            } else {
                bstore<T>(v, ro, ob);
            }
            Pm[j] = Pp[j];
        }
    }
}

template <typename T, typename TI, int XE>
static int bxf_launch(const BXArgs& a, bool interp, int py, int nt,
                      hipStream_t s)
{
#define GFD_BXF(PP)                                                            \
    if (py == PP) {                                                            \
        if (interp)                                                            \
            hipLaunchKernelGGL((k_bxinterp<T, TI, PP, XE>),                    \
                               dim3((unsigned)a.ntiles), dim3(nt), 0, s, a);   \
        else                                                                   \
            hipLaunchKernelGGL((k_bxrestrict<T, TI, PP, XE>),                  \
                               dim3((unsigned)a.ntiles), dim3(nt), 0, s, a);   \
        GFD_LAUNCH_CHECK();                                                    \
        return OK;                                                             \
    }
    GFD_BXF(1) GFD_BXF(2) GFD_BXF(4)
#undef GFD_BXF
    set_error("batched x-march transfer: %d rows per thread not built", py);
    return ENOTSUP_;
}

int launch_bxfer_march(const BXArgs& a, bool interp, int py, int nt,
                       int dtype, int xe, hipStream_t s)
{
    if (a.ntiles <= 0) return OK;
    if (a.ntiles > 0x7fffffffL || nt < 64 || nt > 512 || nt % 64) {
        set_error("batched x-march transfer: %ld tiles / %d threads",
                  a.ntiles, nt);
        return ENOTSUP_;
    }
    if (interp) {
        if (dtype == 0 && xe == XEPI_SUBFROM)
            return bxf_launch<double, double, XEPI_SUBFROM>(a, true, py, nt, s);
        if (dtype == 2 && xe == XEPI_SUBFROM)
            return bxf_launch<float, float, XEPI_SUBFROM>(a, true, py, nt, s);
        if (dtype == 2 && xe == XEPI_NEG)
            return bxf_launch<float, float, XEPI_NEG>(a, true, py, nt, s);
    } else {
        if (dtype == 0 && xe == XEPI_PLAIN)
            return bxf_launch<double, double, XEPI_PLAIN>(a, false, py, nt, s);
        if (dtype == 2 && xe == XEPI_PLAIN)
            return bxf_launch<float, float, XEPI_PLAIN>(a, false, py, nt, s);
        if (dtype == 2 && xe == XEPI_SCALE2)
            return bxf_launch<float, float, XEPI_SCALE2>(a, false, py, nt, s);
        if (dtype == 4 && xe == XEPI_SCALE2)
            return bxf_launch<float, double, XEPI_SCALE2>(a, false, py, nt, s);
    }
    set_error("batched x-march transfer: dtype %d / epilogue %d not built",
              dtype, xe);
    return ENOTSUP_;
}

}  // namespace gfd
