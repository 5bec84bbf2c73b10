// Batched axis stencils on the tiled x-march (the stencil pass of the
// batched multigrid of batch.hip over many uneven real grids).
//
// One block per tile; the tile's grid comes from a binary search over the
// grid table's first-tile column.  Each grid has its own tile geometry in
// the table: TZ = mz columns (the row split into equal parts of at most the
// level's z width, so ragged rows waste little) by mseg * PY rows, laid
// over a 1-D block of `nt` threads (thread t: column t % mz, row segment
// t / mz; threads past mseg segments repeat the last segment's loads and
// store nothing).  The block
// marches along an x chunk: each thread keeps a register queue of 2R + 1
// planes for its PY rows (the x terms), the current plane sits in LDS,
// double buffered with one barrier per plane, for the y and z terms, and
// the tile's 2R halo rows and columns are read once per plane.  Logical
// tiles are remapped so that neighbouring tiles of one grid share an XCD.
//
// Epilogues (per-grid c0, Jacobi weight and step from the table):
// EPI_PLAIN, EPI_RELAX, EPI_RESID (+ per-tile sum of squares), EPI_UPDATE,
// EPI_UPDW (FP32 data, FP64 aux and out) and BEPI_RELAXSUB (a Jacobi sweep
// subtracted from an FP64 array: the last sweep of a refinement's
// correction applied to phi in the same pass), with the accumulation order
// of the register-tile kernel (z, then y, then x terms), so both kernels
// give the same bits except for the Jacobi update, which multiplies by
// w / c0 (rounded once per grid) where the other divides by c0.
//
// All global traffic goes through buffer instructions (the loop is
// instruction-issue bound, not memory bound): no 64-bit vector addresses,
// and masking by out-of-range offsets instead of branches and selects.
#include <type_traits>

#include "batch.h"
#include "stencil.h"
#include "stencil_dev.h"

namespace gfd {

template <typename T, typename TO, int R, int PY, int PF, int EPI>
__global__ void __launch_bounds__(BXM_MAXT) k_bxmarch(BSArgs a)
{
    // aux: right-hand side (RELAX, RESID, UPDATE) in T, FP64 for UPDW
    typedef typename std::conditional<EPI == EPI_UPDW, double, T>::type TA;
    constexpr bool AUX = EPI != EPI_PLAIN;
    // RELAXSUB: the FP64 output array is read at the output points too
    constexpr bool SUB = EPI == BEPI_RELAXSUB;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);

    const int L = xcd_remap(blockIdx.x, (int)a.ntiles);
    // by value: stores to out could alias a reference into the table
    const BSt d = a.g[bfind_tile(a.g, a.G, L)];
    long lt = L - d.tile0;
    const int zt = (int)(lt % d.tz);
    lt /= d.tz;
    const int yt = (int)(lt % d.ty);
    const int xt = (int)(lt / d.ty);

    const int NT = blockDim.x, tid = threadIdx.x;
    const int TZ = d.mz, NSEG = d.mseg, TY = NSEG * PY;
    const int tx = tid % TZ;
    // threads past the last full segment repeat the last segment's work
    // (same loads and LDS values) and store nothing
    const int sg0 = tid / TZ;
    const bool lane_on = sg0 < NSEG;
    const int sg = lane_on ? sg0 : NSEG - 1;
    const int LW = TZ + 2 * R;
    const int n0 = d.n[0], n1 = d.n[1], n2 = d.n[2];
    const long plane = (long)n1 * n2;
    const int z0 = zt * TZ, y0 = yt * TY;
    const int xs = xt * d.xc, xe = min(xs + d.xc, n0);
    const T* __restrict__ gin = (const T*)a.in + d.off;
    const TA* __restrict__ gaux = (const TA*)a.aux + d.off;
    TO* __restrict__ gout = (TO*)a.out + d.off;
    const bool has_lds = d.has[1] || d.has[2];
    const unsigned pbytes = (unsigned)(plane * (long)sizeof(T));

    // own column and rows: wrapped (periodic) or zero beyond the grid, so
    // the LDS plane of a ragged tile holds what the neighbours must read.
    // off_own: byte offset in the plane (BXM_OOB: zero); off_out: output
    // offset in elements (BXM_OOB: no output)
    const int zc = z0 + tx;
    int zw = zc;
    const bool okz = bwrap(zw, n2, d.periodic[2]);
    unsigned off_own[PY], off_out[PY];
#pragma unroll
    for (int j = 0; j < PY; j++) {
        const int yr = y0 + sg * PY + j;
        int yw = yr;
        const bool oky = bwrap(yw, n1, d.periodic[1]);
        off_own[j] = (oky && okz) ? (unsigned)(yw * n2 + zw) * sizeof(T)
                                  : BXM_OOB;
        off_out[j] = (lane_on && yr < n1 && zc < n2)
                         ? (unsigned)(yr * n2 + zc) : BXM_OOB;
    }

    // halo elements of this thread (rows above / below, columns left /
    // right of the tile): LDS index (unused ones write a spare element
    // past the plane) and byte offset (BXM_OOB: zero or unused).  A tile
    // of whole rows (one z part) loads no halo columns: the edge threads
    // copy their value to the wrapped column in LDS (periodic z) or the
    // columns stay zero.
    const bool fullrow = d.tz == 1;
    const bool mirror = fullrow && d.periodic[2] && d.has[2];
    const int H = has_lds ? 2 * R * TZ + (fullrow ? 0 : 2 * R * TY) : 0;
    const int LSZ = (TY + 2 * R) * LW + 1;   // + the spare element
    int h_lds[BXM_HMAX];
    unsigned h_off[BXM_HMAX];
#pragma unroll
    for (int m = 0; m < BXM_HMAX; m++) {
        const int h = tid + m * NT;
        h_lds[m] = LSZ - 1;
        h_off[m] = BXM_OOB;
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
            const bool okh = bwrap(zw2, n2, d.periodic[2]);
            const bool okv = bwrap(yw, n1, d.periodic[1]);
            if (okh && okv) h_off[m] = (unsigned)(yw * n2 + zw2) * sizeof(T);
        }
    }

    // resource of input plane x (wrapped on a periodic axis, empty for a
    // zero plane)
    auto in_plane = [&](int x) {
        int xw = x % n0;
        if (xw < 0) xw += n0;
        const bool ok = (x >= 0 && x < n0) || d.periodic[0] != 0;
        return brsrc(gin + (long)xw * plane, ok ? pbytes : 0u);
    };
    auto ld_own = [&](int x, T* dst) {
        const __amdgpu_buffer_rsrc_t r = in_plane(x);
#pragma unroll
        for (int j = 0; j < PY; j++) dst[j] = bload<T>(r, off_own[j]);
    };
    auto ld_halo = [&](int x, T* dst) {
        const __amdgpu_buffer_rsrc_t r = in_plane(x);
#pragma unroll
        for (int m = 0; m < BXM_HMAX; m++) dst[m] = bload<T>(r, h_off[m]);
    };
    // right-hand side of the output rows of plane x (inside the chunk)
    // the FP64 array RELAXSUB updates, at the output points of plane x
    auto ld_sub = [&](int x, double* dst) {
        const __amdgpu_buffer_rsrc_t r = brsrc(
            (const double*)a.out + d.off + (long)x * plane,
            (unsigned)(plane * (long)sizeof(double)));
#pragma unroll
        for (int j = 0; j < PY; j++)
            dst[j] = bload<double>(r, off_out[j] == BXM_OOB
                                          ? BXM_OOB
                                          : off_out[j] * 8u);
    };
    auto ld_aux = [&](int x, TA* dst) {
        const __amdgpu_buffer_rsrc_t r = brsrc(
            gaux + (long)x * plane, (unsigned)(plane * (long)sizeof(TA)));
#pragma unroll
        for (int j = 0; j < PY; j++)
            dst[j] = bload<TA>(r, off_out[j] == BXM_OOB
                                      ? BXM_OOB
                                      : off_out[j] * (unsigned)sizeof(TA));
    };

    T cpx[R], cmx[R], cpy[R], cmy[R], cpz[R], cmz[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        cpx[k] = (T)d.cp[0][k];
        cmx[k] = (T)d.cm[0][k];
        cpy[k] = (T)d.cp[1][k];
        cmy[k] = (T)d.cm[1][k];
        cpz[k] = (T)d.cp[2][k];
        cmz[k] = (T)d.cm[2][k];
    }
    const T c0 = (T)d.c0, step = (T)d.step;
    // Jacobi: (1 - w) v + (w / c0) (f - acc)
    const T w1 = (T)(1.0 - d.w), wc = (T)(d.w / d.c0);

    // q: own values of planes x - R .. x + R.  In flight, slot u of plane
    // x = x0 + u: own values of plane x + R + 1 (qp), halo values of plane
    // x (hp), right-hand side of plane x (ap); each slot is reloaded for
    // plane x + PF as soon as it is consumed (clamped to the chunk at its
    // end).  The x loop is unrolled by PF so the buffers are renamed
    // statically.
    T q[2 * R + 1][PY];
    T qp[PF][PY];
    T hp[PF][BXM_HMAX];
    TA ap[AUX ? PF : 1][PY];
    double sp[SUB ? PF : 1][PY];
#pragma unroll
    for (int k = 0; k < 2 * R + 1; k++) ld_own(xs - R + k, q[k]);
#pragma unroll
    for (int u = 0; u < PF; u++) {
        const int x = min(xs + u, xe - 1);
        if (has_lds) ld_halo(x, hp[u]);
        ld_own(min(x + R + 1, xe + R - 1), qp[u]);
        if constexpr (AUX) ld_aux(x, ap[u]);
        if constexpr (SUB) ld_sub(x, sp[u]);
    }

    double nrm = 0.0;
    const int lcol = tx + R;
    if (fullrow && !mirror && tx < R) {
        // zero halo columns of the own rows, both buffers (never written
        // again; the first barrier of the x loop orders these stores)
#pragma unroll
        for (int j = 0; j < PY; j++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                T* row = lds + b * LSZ + (sg * PY + j + R) * LW;
                row[tx] = zero_of<T>();
                row[TZ + R + tx] = zero_of<T>();
            }
    }

    auto body = [&](const int x, T* qpu, T* hpu, TA* apu, double* spu) {
        T* buf = lds + (x & 1) * LSZ;
        if (has_lds) {
#pragma unroll
            for (int j = 0; j < PY; j++)
                buf[(sg * PY + j + R) * LW + lcol] = q[R][j];
            // column tx is also column tx + n2 (left edge) and / or
            // tx - n2 (right edge) of the periodic row
            if (mirror && tx < R) {
#pragma unroll
                for (int j = 0; j < PY; j++)
                    buf[(sg * PY + j + R) * LW + lcol + TZ] = q[R][j];
            }
            if (mirror && tx >= TZ - R) {
#pragma unroll
                for (int j = 0; j < PY; j++)
                    buf[(sg * PY + j + R) * LW + lcol - TZ] = q[R][j];
            }
#pragma unroll
            for (int m = 0; m < BXM_HMAX; m++) buf[h_lds[m]] = hpu[m];
            ld_halo(min(x + PF, xe - 1), hpu);
            __syncthreads();
        }
        TA f[PY];
        if constexpr (AUX) {
#pragma unroll
            for (int j = 0; j < PY; j++) f[j] = apu[j];
            ld_aux(min(x + PF, xe - 1), apu);
        }
        double ph[PY];
        if constexpr (SUB) {
#pragma unroll
            for (int j = 0; j < PY; j++) ph[j] = spu[j];
            ld_sub(min(x + PF, xe - 1), spu);
        }
        const __amdgpu_buffer_rsrc_t ro = brsrc(
            gout + (long)x * plane, (unsigned)(plane * (long)sizeof(TO)));
#pragma unroll
        for (int j = 0; j < PY; j++) {
            const int lrow = sg * PY + j + R;
            T acc = zero_of<T>();
            if (d.has[2]) {
#pragma unroll
                for (int k = 1; k <= R; k++) {
                    acc = fmac(acc, cpz[k - 1], buf[lrow * LW + lcol + k]);
                    acc = fmac(acc, cmz[k - 1], buf[lrow * LW + lcol - k]);
                }
            }
            if (d.has[1]) {
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
            if (d.has[0]) {
#pragma unroll
                for (int k = 1; k <= R; k++) {
                    acc = fmac(acc, cpx[k - 1], q[R + k][j]);
                    acc = fmac(acc, cmx[k - 1], q[R - k][j]);
                }
            }
            const T v = q[R][j];
            T r;
            if constexpr (EPI == EPI_PLAIN) {
                r = fmac(acc, c0, v);
            } else if constexpr (EPI == EPI_RELAX || SUB) {
                r = fmac(w1 * v, wc, f[j] - acc);
            } else if constexpr (EPI == EPI_UPDATE || EPI == EPI_UPDW) {
                const T qv = fmac(acc, c0, v) - (T)f[j];
                r = fmac(v, -step, qv);
                if (a.negate) r = -r;
            } else {
                r = fmac(acc, c0, v) - f[j];
                if (off_out[j] != BXM_OOB) nrm += (double)r * (double)r;
            }
            const unsigned ob = off_out[j] == BXM_OOB
                                    ? BXM_OOB
                                    : off_out[j] * (unsigned)sizeof(TO);
            if constexpr (SUB)
                bstore<double>(ph[j] - (double)r, ro, ob);
            else
                bstore<TO>((TO)r, ro, ob);
        }
#pragma unroll
        for (int k = 0; k < 2 * R; k++)
#pragma unroll
            for (int j = 0; j < PY; j++) q[k][j] = q[k + 1][j];
#pragma unroll
        for (int j = 0; j < PY; j++) q[2 * R][j] = qpu[j];
        ld_own(min(x + R + 1 + PF, xe + R - 1), qpu);
    };

    // whole groups of PF planes, then the rest (no branch inside a group)
    const int xm = xs + (xe - xs) / PF * PF;
    for (int x0 = xs; x0 < xm; x0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++)
            body(x0 + u, qp[u], hp[u], ap[AUX ? u : 0], sp[SUB ? u : 0]);
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (xm + u < xe)
            body(xm + u, qp[u], hp[u], ap[AUX ? u : 0], sp[SUB ? u : 0]);
    if (EPI == EPI_RESID && a.partial) block_sum_store(nrm, a.partial, L);
}

template <typename T, typename TO, int R, int PY, int PF, int EPI>
static int bxm_launch_t(const BSArgs& a, int nt, size_t lds, hipStream_t s)
{
    static size_t attr_bytes = 64 * 1024;
    if (lds > attr_bytes) {
        hipError_t e = hipFuncSetAttribute(
            (const void*)k_bxmarch<T, TO, R, PY, PF, EPI>,
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("batched x-march: %zu bytes of LDS not available (%s)",
                      lds, hipGetErrorString(e));
            return ENOTSUP_;
        }
        attr_bytes = lds;
    }
    hipLaunchKernelGGL((k_bxmarch<T, TO, R, PY, PF, EPI>),
                       dim3((unsigned)a.ntiles), dim3(nt), lds, s, a);
    GFD_LAUNCH_CHECK();
    return OK;
}

template <typename T, typename TO, int R, int EPI>
static int bxm_launch_p(const BSArgs& a, int py, int pf, int nt, size_t lds,
                        hipStream_t s)
{
#define GFD_BXM(PP, FF)                                                        \
    if (py == PP && pf == FF)                                                  \
        return bxm_launch_t<T, TO, R, PP, FF, EPI>(a, nt, lds, s);
    GFD_BXM(1, 1) GFD_BXM(1, 2)
    GFD_BXM(2, 1) GFD_BXM(2, 2)
    GFD_BXM(4, 1) GFD_BXM(4, 2)
#undef GFD_BXM
    set_error("batched x-march: %d rows / %d planes ahead not built", py, pf);
    return ENOTSUP_;
}

template <typename T, typename TO, int EPI>
static int bxm_launch_r(const BSArgs& a, int R, int py, int pf, int nt,
                        size_t lds, hipStream_t s)
{
    switch (R) {
    case 1: return bxm_launch_p<T, TO, 1, EPI>(a, py, pf, nt, lds, s);
    case 2: return bxm_launch_p<T, TO, 2, EPI>(a, py, pf, nt, lds, s);
    case 3: return bxm_launch_p<T, TO, 3, EPI>(a, py, pf, nt, lds, s);
    }
    set_error("batched x-march: radius %d not built", R);
    return ENOTSUP_;
}

int launch_bxmarch(const BSArgs& a, int R, int py, int pf, int nt,
                   size_t lds, int dtype, int epi, hipStream_t s)
{
    if (a.ntiles <= 0) return OK;
    if (a.ntiles > 0x7fffffffL || nt < 64 || nt > BXM_MAXT || nt % 64) {
        set_error("batched x-march: %ld tiles / %d threads not supported",
                  a.ntiles, nt);
        return ENOTSUP_;
    }
    if (dtype == 0 && epi == EPI_PLAIN)
        return bxm_launch_r<double, double, EPI_PLAIN>(a, R, py, pf, nt, lds, s);
    if (dtype == 2 && epi == EPI_PLAIN)
        return bxm_launch_r<float, float, EPI_PLAIN>(a, R, py, pf, nt, lds, s);
    if (dtype == 2 && epi == EPI_RELAX)
        return bxm_launch_r<float, float, EPI_RELAX>(a, R, py, pf, nt, lds, s);
    if (dtype == 2 && epi == BEPI_RELAXSUB)
        return bxm_launch_r<float, double, BEPI_RELAXSUB>(a, R, py, pf, nt,
                                                          lds, s);
    if (dtype == 2 && epi == EPI_RESID)
        return bxm_launch_r<float, float, EPI_RESID>(a, R, py, pf, nt, lds, s);
    if (dtype == 4 && epi == EPI_RESID)
        return bxm_launch_r<double, float, EPI_RESID>(a, R, py, pf, nt, lds, s);
    if (dtype == 2 && epi == EPI_UPDATE)
        return bxm_launch_r<float, float, EPI_UPDATE>(a, R, py, pf, nt, lds, s);
    if (dtype == 2 && epi == EPI_UPDW)
        return bxm_launch_r<float, double, EPI_UPDW>(a, R, py, pf, nt, lds, s);
    set_error("batched x-march: dtype %d / epilogue %d not built", dtype,
              epi);
    return ENOTSUP_;
}

}  // namespace gfd
