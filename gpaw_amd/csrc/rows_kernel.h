// Full-row x-marching stencil kernel (k_xrow) for axis-only stencils.
//
// The tile of a block spans whole z rows (n2 points) and TY rows in y; the
// block marches along x.  Each thread owns PY rows x PZ consecutive z points
// (PZ * sizeof(T) == 16 bytes), so every global access, LDS write and LDS
// read of the z window is one 16-byte vector operation.
//
//  * x terms: a register queue of 2R+2 plane slots per owned point; the
//    plane loop is unrolled by the queue length so slots are renamed at
//    compile time (no register moves) and the load of plane x+R+1 issued
//    in step x is not waited for before step x+1's arithmetic.
//  * y terms: rows of the thread's own segment from registers, the others
//    from the LDS copy of the current plane (double buffered, one barrier
//    per plane); the 2R halo rows of the tile are loaded from global memory
//    (L2-resident: the neighbouring tile of the same grid streams them as
//    its own rows on the same XCD).
//  * z terms: from LDS through a window of aligned chunks whose (wrapped)
//    columns and boundary factors are fixed per thread, so the LDS plane
//    holds no halo columns and rows never branch on the boundary.
//  * the tile height divides n1 and every thread of the first 2R segments
//
// This header holds the kernel template and its launch helpers; it is
// included only by the rows_*.hip launchers.
#pragma once

#include "rows.h"
#include "stencil_dev.h"

namespace gfd {

template <typename T, int PZ>
struct __align__(16) Chunk {
    T v[PZ];
};

template <typename T>
struct XrArgs {
    const T* in;
    T* out;
    const T* aux;
    int n0, n1, n2;
    long ng;
    int tiles_y, chunks_x, xchunk;
    int nblocks;
    int hasx, hasy, hasz;
    double c0, w, s;
    int negate;
    double cp[3][MAXR_FAST];
    double cm[3][MAXR_FAST];
    BC bc;
    double* partial;
    const double* pot;
    const double* shift;
    int nt;              // non-temporal hints (xr_nt_flags)
};


template <typename T, int PZ>
__device__ inline Chunk<T, PZ> zero_chunk()
{
    Chunk<T, PZ> c;
#pragma unroll
    for (int i = 0; i < PZ; i++) c.v[i] = zero_of<T>();
    return c;
}

// 16-byte loads / stores with the non-temporal hint (streamed data that
// no other block reads again)
typedef double nt16 __attribute__((ext_vector_type(2)));
template <typename C>
__device__ inline C ld16nt(const C* p)
{
    static_assert(sizeof(C) == 16, "16-byte chunks only");
    const nt16 v = __builtin_nontemporal_load(reinterpret_cast<const nt16*>(p));
    C c;
    __builtin_memcpy(&c, &v, 16);
    return c;
}
template <typename C>
__device__ inline void st16nt(const C& c, C* p)
{
    static_assert(sizeof(C) == 16, "16-byte chunks only");
    nt16 v;
    __builtin_memcpy(&v, &c, 16);
    __builtin_nontemporal_store(v, reinterpret_cast<nt16*>(p));
}

template <typename T, int PZ>
__device__ inline Chunk<T, PZ> phase_chunk(Chunk<T, PZ> c, double2 f)
{
    if constexpr (IsCplx<T>::value) {
#pragma unroll
        for (int i = 0; i < PZ; i++) c.v[i] = scale_by<T>(c.v[i], f, true);
    }
    return c;
}

template <typename T, int R, int PY, int PZ, int EPI>
__global__ void __launch_bounds__(XR_MAXT) k_xrow(XrArgs<T> a)
{
    typedef typename ScalarOf<T>::type S;
    typedef Chunk<T, PZ> C;
    constexpr int RC = (R + PZ - 1) / PZ;  // z-window chunks on each side
    constexpr int NW = 2 * RC + 1;         // z-window chunks
    constexpr int RP = RC * PZ;
    // queue slots = unroll length: planes x-R .. x+R plus XR_LOOK planes
    // in flight
    constexpr int U = 2 * R + 1 + XR_LOOK;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);

    const int L = xcd_remap(blockIdx.x, a.nblocks);
    const int ty = L % a.tiles_y;
    const int cx = (L / a.tiles_y) % a.chunks_x;
    const long b = L / (a.tiles_y * a.chunks_x);

    const int tx = threadIdx.x, sg = threadIdx.y;
    const int TZT = blockDim.x;
    const int TY = blockDim.y * PY;  // divides n1 (host)
    const int n0 = a.n0, n1 = a.n1, n2 = a.n2;
    const int BUF = (TY + 2 * R) * n2;
    const long plane = (long)n1 * n2;
    const int z0 = tx * PZ;
    const int y0 = ty * TY;
    const int xs = cx * a.xchunk;
    const int xe = min(xs + a.xchunk, n0);
    const T* __restrict__ gin = a.in + b * a.ng;
    const BC& bc = a.bc;

    // z window: LDS column of chunk c (wrapped) and its boundary factor
    int zcol[NW];
    bool zok[NW], zph[NW];
    double2 zf[NW];
#pragma unroll
    for (int c = 0; c < NW; c++) {
        int z = z0 + (c - RC) * PZ;
        double2 f = make_double2(1.0, 0.0);
        bool ph = false;
        zok[c] = wrap_full(z, n2, bc.periodic[2], bc.phase[2][0],
                           bc.phase[2][1], f, ph);
        zcol[c] = z;
        zf[c] = f;
        zph[c] = ph;
    }
    // own rows (never wrapped: TY divides n1) and the LDS row of row 0
    const int own0 = (y0 + sg * PY) * n2 + z0;
    const int lrow0 = (sg * PY + R) * n2;
    // halo rows of the tile: segment sg loads rows sg, sg + NSEG, ... < 2R
    const int NSEG = blockDim.y;
    const bool anyhalo = a.hasy || a.hasz;
    bool hload[XR_HROWS], hok[XR_HROWS], hph[XR_HROWS];
    int hoff[XR_HROWS], hlds[XR_HROWS];
    double2 hf[XR_HROWS];
#pragma unroll
    for (int m = 0; m < XR_HROWS; m++) {
        const int hr = sg + m * NSEG;
        hload[m] = anyhalo && hr < 2 * R;
        hok[m] = false;
        hph[m] = false;
        hoff[m] = 0;
        hlds[m] = 0;
        hf[m] = make_double2(1.0, 0.0);
        if (hload[m]) {
            const int row = hr < R ? hr - R : TY + hr - R;
            int yw = y0 + row;
            hok[m] = wrap_full(yw, n1, bc.periodic[1], bc.phase[1][0],
                               bc.phase[1][1], hf[m], hph[m]);
            hoff[m] = yw * n2 + z0;
            hlds[m] = (row + R) * n2 + z0;
        }
    }

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

    // plane x (possibly wrapped): base offset, or -1 when zero
    auto plane_of = [&](int x, double2& fx, bool& phx) -> long {
        int xw = x;
        fx = make_double2(1.0, 0.0);
        phx = false;
        if (!wrap_full(xw, n0, bc.periodic[0], bc.phase[0][0], bc.phase[0][1],
                       fx, phx))
            return -1;
        return (long)xw * plane;
    };
    auto load_own = [&](int x, C* dst) {
        double2 fx;
        bool phx;
        const long base = plane_of(x, fx, phx);
        if (base < 0) {
#pragma unroll
            for (int j = 0; j < PY; j++) dst[j] = zero_chunk<T, PZ>();
            return;
        }
        const T* p = gin + base + own0;
#pragma unroll
        for (int j = 0; j < PY; j++) {
            if (a.nt & 2)
                if constexpr (sizeof(C) == 16)
                    dst[j] = ld16nt(reinterpret_cast<const C*>(p + j * n2));
                else
                    dst[j] = *reinterpret_cast<const C*>(p + j * n2);
            else
                dst[j] = *reinterpret_cast<const C*>(p + j * n2);
            if constexpr (IsCplx<T>::value)
                if (phx) dst[j] = phase_chunk<T, PZ>(dst[j], fx);
        }
    };
    auto load_halo = [&](int x, C* dst) {
        double2 fx;
        bool phx;
        const long base = plane_of(x, fx, phx);
#pragma unroll
        for (int m = 0; m < XR_HROWS; m++) {
            dst[m] = zero_chunk<T, PZ>();
            if (hload[m] && base >= 0 && hok[m]) {
                dst[m] = *reinterpret_cast<const C*>(gin + base + hoff[m]);
                if constexpr (IsCplx<T>::value) {
                    if (phx || hph[m]) {
                        const double2 f =
                            phx ? (hph[m] ? cmul(hf[m], fx) : fx) : hf[m];
                        dst[m] = phase_chunk<T, PZ>(dst[m], f);
                    }
                }
            }
        }
    };
    auto ldsc = [&](const T* buf, int idx) -> C {
        return *reinterpret_cast<const C*>(buf + idx);
    };

    C q[U][PY];
    C hv[XR_HROWS];
#pragma unroll
    for (int k = -R; k < R + XR_LOOK; k++) load_own(xs + k, q[(k + U) % U]);
    if (anyhalo) load_halo(xs, hv);

    double nrm = 0.0;
    for (int xg = xs; xg < xe; xg += U) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int x = xg + u;
            if (x < xe) {
                // alternate LDS planes by step parity (U may be odd)
                T* buf = lds + ((x - xs) & 1) * BUF;
                if (a.hasy || a.hasz) {
#pragma unroll
                    for (int j = 0; j < PY; j++)
                        *reinterpret_cast<C*>(buf + lrow0 + j * n2 + z0) =
                            q[u][j];
#pragma unroll
                    for (int m = 0; m < XR_HROWS; m++)
                        if (hload[m])
                            *reinterpret_cast<C*>(buf + hlds[m]) = hv[m];
                }
                // halo first: its wait at the next step leaves the own-plane
                // loads issued after it in flight
                if (x + 1 < xe && anyhalo) load_halo(x + 1, hv);
                if (x + XR_LOOK < xe)
                    load_own(x + R + XR_LOOK, q[(u + R + XR_LOOK) % U]);
                if (a.hasy || a.hasz) __syncthreads();

                const long obase = b * a.ng + (long)x * plane + own0;
#pragma unroll
                for (int j = 0; j < PY; j++) {
                    const int lrow = lrow0 + j * n2;
                    T acc[PZ];
#pragma unroll
                    for (int p = 0; p < PZ; p++) acc[p] = zero_of<T>();
                    if (a.hasz) {
                        T wz[NW * PZ];
#pragma unroll
                        for (int c = 0; c < NW; c++) {
                            C ch;
                            if (c == RC) {
                                ch = q[u][j];
                            } else {
                                ch = ldsc(buf, lrow + zcol[c]);
                                if (!zok[c]) ch = zero_chunk<T, PZ>();
                                if constexpr (IsCplx<T>::value)
                                    if (zph[c]) ch = phase_chunk<T, PZ>(ch, zf[c]);
                            }
#pragma unroll
                            for (int i = 0; i < PZ; i++)
                                wz[c * PZ + i] = ch.v[i];
                        }
#pragma unroll
                        for (int p = 0; p < PZ; p++)
#pragma unroll
                            for (int k = 1; k <= R; k++) {
                                acc[p] = fmac(acc[p], cpz[k - 1],
                                              wz[RP + p + k]);
                                acc[p] = fmac(acc[p], cmz[k - 1],
                                              wz[RP + p - k]);
                            }
                    }
                    if (a.hasy) {
#pragma unroll
                        for (int k = 1; k <= R; k++) {
                            const C up = (j + k < PY)
                                             ? q[u][(j + k) % PY]
                                             : ldsc(buf, lrow + k * n2 + z0);
                            const C dn = (j - k >= 0)
                                             ? q[u][(j - k + PY) % PY]
                                             : ldsc(buf, lrow - k * n2 + z0);
#pragma unroll
                            for (int p = 0; p < PZ; p++) {
                                acc[p] = fmac(acc[p], cpy[k - 1], up.v[p]);
                                acc[p] = fmac(acc[p], cmy[k - 1], dn.v[p]);
                            }
                        }
                    }
                    if (a.hasx) {
#pragma unroll
                        for (int k = 1; k <= R; k++)
#pragma unroll
                            for (int p = 0; p < PZ; p++) {
                                acc[p] = fmac(acc[p], cpx[k - 1],
                                              q[(u + k) % U][j].v[p]);
                                acc[p] = fmac(acc[p], cmx[k - 1],
                                              q[(u - k + U) % U][j].v[p]);
                            }
                    }
                    const long idx = obase + (long)j * n2;
                    if constexpr (EPI == EPI_UPDW) {
                        // FP32 arithmetic, FP64 aux and out
                        typedef typename WideOf<T>::type WT;
                        typedef Chunk<WT, PZ> CW;
                        const CW aw = *reinterpret_cast<const CW*>(
                            (const WT*)(const void*)a.aux + idx);
                        CW rw;
#pragma unroll
                        for (int p = 0; p < PZ; p++)
                            rw.v[p] = to_wide(epi_apply<T, EPI_UPDATE>(
                                acc[p], q[u][j].v[p], to_narrow(aw.v[p]),
                                a.c0, a.w, a.s, a.negate, nrm));
                        *reinterpret_cast<CW*>((WT*)(void*)a.out + idx) = rw;
                        continue;
                    }
                    if constexpr (EPI == EPI_RESN) {
                        const C an = *reinterpret_cast<const C*>(a.aux + idx);
                        float* o = (float*)(void*)a.out + idx;
#pragma unroll
                        for (int p = 0; p < PZ; p++)
                            o[p] = (float)epi_apply<T, EPI_RESID>(
                                acc[p], q[u][j].v[p], an.v[p], a.c0, a.w,
                                a.s, a.negate, nrm);
                        continue;
                    }
                    C av = zero_chunk<T, PZ>();
                    if constexpr (EPI != EPI_PLAIN && EPI != EPI_POT)
                        av = *reinterpret_cast<const C*>(a.aux + idx);
                    C r;
                    if constexpr (EPI == EPI_POT) {
                        // + V psi, V real and shared by all grids
                        const double* pv =
                            a.pot + (long)x * plane + own0 + j * n2;
#pragma unroll
                        for (int p = 0; p < PZ; p++)
                            r.v[p] = fmac(fmac(acc[p], (S)a.c0, q[u][j].v[p]),
                                          (S)pv[p], q[u][j].v[p]);
                        if (a.shift) {
                            // residual: - eps_b psi
                            const S e = -(S)a.shift[b];
#pragma unroll
                            for (int p = 0; p < PZ; p++)
                                r.v[p] = fmac(r.v[p], e, q[u][j].v[p]);
                        }
                    } else {
#pragma unroll
                        for (int p = 0; p < PZ; p++)
                            r.v[p] = epi_apply<T, EPI>(acc[p], q[u][j].v[p],
                                                       av.v[p], a.c0, a.w,
                                                       a.s, a.negate, nrm);
                    }
                    if (a.nt & 1)
                        if constexpr (sizeof(C) == 16)
                            st16nt(r, reinterpret_cast<C*>(a.out + idx));
                        else
                            *reinterpret_cast<C*>(a.out + idx) = r;
                    else
                        *reinterpret_cast<C*>(a.out + idx) = r;
                }
            }
        }
    }
    if ((EPI == EPI_RESID || EPI == EPI_RESN) && a.partial)
        block_sum_store(nrm, a.partial, L);
}

template <typename T> struct PzOf { static constexpr int value = 16 / sizeof(T); };

// WIDE = 2: two 16-byte points per thread along z for 16-byte data
// (complex128), so a 128-point row takes 64 threads and a block can hold
// taller tiles
template <typename T, int R, int PY, int EPI, int WIDE = 1>
static int launch_xr_t(const AxisStencil& st, const XmTile& tl,
                       const int n[3], const BC& bc, const void* in,
                       void* out, long nin, const EpiArgs& ea,
                       hipStream_t stream, int* nblocks_out)
{
    constexpr int PZ = PzOf<T>::value * WIDE;
    XrArgs<T> a;
    a.in = (const T*)in;
    a.out = (T*)out;
    a.aux = (const T*)ea.aux;
    a.n0 = n[0];
    a.n1 = n[1];
    a.n2 = n[2];
    a.ng = (long)n[0] * n[1] * n[2];
    const int TY = tl.nseg * PY;
    a.tiles_y = (n[1] + TY - 1) / TY;
    a.xchunk = tl.xchunk;
    a.chunks_x = (n[0] + tl.xchunk - 1) / tl.xchunk;
    const long nb = (long)a.tiles_y * a.chunks_x * nin;
    if (nb <= 0) return OK;
    if (nb > 0x7fffffffL) {
        set_error("too many blocks (%ld)", nb);
        return ENOTSUP_;
    }
    a.nblocks = (int)nb;
    a.hasx = st.has[0];
    a.hasy = st.has[1];
    a.hasz = st.has[2];
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
    a.pot = ea.pot;
    a.shift = ea.shift;
    a.nt = xr_nt_flags((long)n[0] * n[1] * n[2] * nin * (long)sizeof(T));
    const int TZT = n[2] / PZ;
    const size_t shm =
        (st.has[1] || st.has[2]) ? 2 * (size_t)(TY + 2 * R) * n[2] * sizeof(T)
                                 : 0;
    static size_t attr_bytes = 64 * 1024;
    if (shm > attr_bytes) {
        hipError_t e = hipFuncSetAttribute(
            (const void*)k_xrow<T, R, PY, PZ, EPI>,
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("x-row: %zu bytes of LDS not available (%s)", shm,
                      hipGetErrorString(e));
            return ENOTSUP_;
        }
        attr_bytes = shm;
    }
    (void)hipGetLastError();
    hipLaunchKernelGGL((k_xrow<T, R, PY, PZ, EPI>), dim3((unsigned)nb),
                       dim3(TZT, tl.nseg), shm, stream, a);
    GFD_LAUNCH_CHECK();
    if (nblocks_out) *nblocks_out = (int)nb;
    return OK;
}

template <typename T, int R, int EPI>
static int launch_xr_py(const AxisStencil& st, const XmTile& tl,
                        const int n[3], const BC& bc, const void* in,
                        void* out, long nin, const EpiArgs& ea,
                        hipStream_t stream, int* nbo)
{
    if constexpr (sizeof(T) == 16) {
        if (tl.wide == 2) {
            switch (tl.py) {
            case 1: return launch_xr_t<T, R, 1, EPI, 2>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
            case 2: return launch_xr_t<T, R, 2, EPI, 2>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
            }
            set_error("x-row: unsupported rows per thread %d (wide)", tl.py);
            return ENOTSUP_;
        }
    }
    switch (tl.py) {
    case 1: return launch_xr_t<T, R, 1, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 2: return launch_xr_t<T, R, 2, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 4: return launch_xr_t<T, R, 4, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    }
    set_error("x-row: unsupported rows per thread %d", tl.py);
    return ENOTSUP_;
}

template <typename T, int EPI>
static int launch_xr_r(const AxisStencil& st, const XmTile& tl,
                       const int n[3], const BC& bc, const void* in,
                       void* out, long nin, const EpiArgs& ea,
                       hipStream_t stream, int* nbo)
{
    switch (st.R) {
    case 1: return launch_xr_py<T, 1, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 2: return launch_xr_py<T, 2, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 3: return launch_xr_py<T, 3, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    case 4: return launch_xr_py<T, 4, EPI>(st, tl, n, bc, in, out, nin, ea, stream, nbo);
    }
    set_error("x-row: unsupported radius %d", st.R);
    return ENOTSUP_;
}

}  // namespace gfd
