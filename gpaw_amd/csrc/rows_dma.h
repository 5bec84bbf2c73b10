// Full-row x-marching stencil with LDS-DMA plane prefetch (k_xdma), for
// axis-only stencils on real data.
//
// Same tile as k_xrow (whole z rows, TY = nseg * PY rows in y, a chunk of
// x planes per block), but planes reach LDS by `global_load_lds_dwordx4`:
// the tile's TY + 2R rows of a plane are contiguous in the grid (up to the
// periodic wrap in y), so each lane copies 16-byte chunks straight into a
// ring of A + 1 plane slots without passing through VGPRs.  A planes
// are in flight while one is consumed, and the vector registers hold no
// prefetched data, so the wave count per CU stays high while the bytes in
// flight per CU grow with A.
//
//  * x terms by an accumulator queue: when plane p lands, its centre values
//    are added with the x coefficients to the partial sums of outputs
//    p - R .. p + R (2R + 1 register slots, renamed at compile time by
//    unrolling the plane loop); output p - R is complete and stored.
//  * y and z terms of output p are read from the slot of plane p when it
//    lands (y rows of the slot; z through the wrapped window of k_xrow).
//  * zero-boundary rows and planes: the lanes still issue their DMA (from a
//    valid dummy address) and overwrite their chunks with zeros after it
//    has landed, so every wave issues the same number of DMAs per plane
//    and the vmcnt bound of each wait is exact.
//  * synchronisation per plane: wait until this wave's DMA of plane i has
//    landed (vmcnt = DMAs issued after it, an immediate per count of
//    planes in flight: A = 1 .. 4), write zero chunks, one barrier,
//    then issue the DMA of plane i + A into the slot read at step i - 1
//    (every wave is past that step), so the ring needs A + 1 slots.
//    Vector memory operations complete in issue order, so stores issued
//    after a DMA only make the wait stricter.
#pragma once

#include <algorithm>

#include "rows_kernel.h"

namespace gfd {

constexpr int XD_KDMAX = 8;  // DMA instructions per wave per plane

template <typename T>
struct XdArgs {
    const T* in;
    T* out;
    int n0, n1, n2;
    long ng;
    int tiles_y, chunks_x, xchunk;
    int nblocks;
    int ahead;          // planes in flight (A); ring of A + 1 slots
    int slot_chunks;    // kd * threads: 16-byte chunks per slot
    double c0;
    double cp[3][MAXR_FAST];
    double cm[3][MAXR_FAST];
    int periodic[3];
    int nt;
    int order;          // block order: 0 XCD-contiguous tiles, 1 linear,
                        // 2 bands interleaved over the 8 XCDs
};

// Logical tile of hardware block h (blocks reach the 8 XCDs round robin).
// 0: consecutive tiles share an XCD (xcd_remap); 1: identity, so the
// blocks in flight cover one compact address window; 2: all tiles of band
// b on XCD b % 8, the XCDs on neighbouring bands (needs nin % 8 == 0).
__device__ inline int xd_block(int h, int nb, int order, int per_band)
{
    if (order == 1) return h;
    if (order == 2) {
        const int x = h & 7, k = h >> 3;
        return (x + 8 * (k / per_band)) * per_band + k % per_band;
    }
    return xcd_remap(h, nb);
}

typedef __attribute__((address_space(3))) unsigned char xd_lds_t;

// One 16-byte chunk per lane from global memory into LDS at lds + 16 *
// lane (lds wave-uniform).  Written as asm so that the compiler, which
// would otherwise wait for every outstanding DMA before any LDS read, leaves
// the waits to the kernel.
__device__ inline void xd_dma16(const void* g, unsigned lds)
{
    asm volatile("global_load_lds_dwordx4 %0, off" ::"v"(g), "{m0}"(lds)
                 : "memory");
}

#define XD_VMCNT(n) ((((n) & 0xF) | (((n) >> 4) << 14)) | 0x70 | 0xF00)

// H = ceil(2R / nseg): halo chunks per thread per plane, so that each
// thread issues KD = PY + H DMAs per plane (a compile-time count: the vmcnt
// bounds are immediates and the DMA loop has no guards).
template <typename T, int R, int PY, int PZ, int H>
__global__ void __launch_bounds__(XR_MAXT) k_xdma(XdArgs<T> a)
{
    typedef Chunk<T, PZ> C;
    constexpr int KD = PY + H;
    constexpr int RC = (R + PZ - 1) / PZ;
    constexpr int NW = 2 * RC + 1;
    constexpr int RP = RC * PZ;
    constexpr int QU = 2 * R + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];

    const int L = xd_block(blockIdx.x, a.nblocks, a.order,
                           a.tiles_y * a.chunks_x);
    const int ty = L % a.tiles_y;
    const int cx = (L / a.tiles_y) % a.chunks_x;
    const long b = L / (a.tiles_y * a.chunks_x);

    const int tx = threadIdx.x, sg = threadIdx.y;
    const int NT = blockDim.x * blockDim.y;
    const int tid = sg * blockDim.x + tx;
    const int TY = blockDim.y * PY;
    const int n0 = a.n0, n1 = a.n1, n2 = a.n2;
    const int CPR = n2 / PZ;
    const int TH = TY + 2 * R;
    const long plane = (long)n1 * n2;
    const int z0 = tx * PZ;
    const int y0 = ty * TY;
    const int xs = cx * a.xchunk;
    const int xe = min(xs + a.xchunk, n0);
    const int NL = xe - xs + 2 * R;   // landings: planes xs-R .. xe+R-1
    const int A = a.ahead;
    const int D = A + 1;
    const int SLOT = a.slot_chunks * 16;   // bytes
    const T* __restrict__ gin = a.in + b * a.ng;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wave_base = wave * 64;       // first chunk of the wave per k

    // DMA chunks of this thread: c = k * NT + tid, row c / CPR of the slot
    int doff[KD];
    bool dzero[KD];
#pragma unroll
    for (int k = 0; k < KD; k++) {
        doff[k] = 0;
        dzero[k] = false;
        {
            const int c = k * NT + tid;
            const int r = c / CPR, col = c - (c / CPR) * CPR;
            if (r < TH) {
                int yw = y0 - R + r;
                if (yw < 0 || yw >= n1) {
                    if (a.periodic[1])
                        yw = yw < 0 ? yw + n1 : yw - n1;
                    else
                        dzero[k] = true;
                }
                if (!dzero[k]) doff[k] = yw * n2 + col * PZ;
            }
        }
    }

    auto slot_ptr = [&](int si) -> unsigned char* {
        return smem_raw + si * SLOT;
    };
    // plane of landing i: offset, or -1 for a zero plane
    auto plane_off = [&](int i) -> long {
        int x = xs - R + i;
        if (x < 0 || x >= n0) {
            if (!a.periodic[0]) return -1;
            x = x < 0 ? x + n0 : x - n0;
        }
        return (long)x * plane;
    };
    auto issue = [&](int i, int si) {
        const long po = plane_off(i);
        const T* src = gin + (po < 0 ? 0 : po);
        unsigned char* s = slot_ptr(si);
#pragma unroll
        for (int k = 0; k < KD; k++)
                xd_dma16(src + doff[k],
                         (unsigned)(size_t)(xd_lds_t*)(s + (k * NT + wave_base) * 16));
    };
    // zero rows exist only in the edge tiles of a zero-boundary y axis
    const bool zrows = !a.periodic[1] && (y0 < R || y0 + TY + R > n1);
    auto zero_fill = [&](int i, int si) {
        const bool zp = plane_off(i) < 0;
        if (!zp && !zrows) return;
        unsigned char* s = slot_ptr(si);
#pragma unroll
        for (int k = 0; k < KD; k++)
            if (zp || dzero[k])
                *reinterpret_cast<C*>(s + (k * NT + tid) * 16) =
                    zero_chunk<T, PZ>();
    };

    // z window columns (wrapped) of the thread's chunk
    int zcol[NW];
    bool zok[NW];
#pragma unroll
    for (int c = 0; c < NW; c++) {
        int z = z0 + (c - RC) * PZ;
        zok[c] = true;
        if (z < 0 || z >= n2) {
            if (a.periodic[2])
                z = z < 0 ? z + n2 : z - n2;
            else
                zok[c] = false;
        }
        zcol[c] = z;
    }
    typedef typename ScalarOf<T>::type S;
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
    const S c0 = (S)a.c0;
    // slot row of the thread's first own row
    const int lrow0 = sg * PY + R;

    C acc[QU][PY];
#pragma unroll
    for (int q = 0; q < QU; q++)
#pragma unroll
        for (int j = 0; j < PY; j++) acc[q][j] = zero_chunk<T, PZ>();

    for (int i = 0; i < A && i < NL; i++) issue(i, i);
    int s_cur = 0;        // slot of landing i
    int s_nxt = A;        // slot of landing i + A

    for (int ig = 0; ig < NL; ig += QU) {
#pragma unroll
        for (int u = 0; u < QU; u++) {
            const int i = ig + u;
            if (i < NL) {
                // DMAs issued after landing i: planes i+1 .. min(i+A-1,
                // NL-1), KD each (A is 1 to 4)
                const int m = min(A - 1, NL - 1 - i);
                if (m <= 0)
                    __builtin_amdgcn_s_waitcnt(XD_VMCNT(0) & ~0xF00);
                else if (m == 1)
                    __builtin_amdgcn_s_waitcnt(XD_VMCNT(KD) & ~0xF00);
                else if (m == 2)
                    __builtin_amdgcn_s_waitcnt(XD_VMCNT(2 * KD) & ~0xF00);
                else
                    __builtin_amdgcn_s_waitcnt(XD_VMCNT(3 * KD) & ~0xF00);
                asm volatile("" ::: "memory");
                zero_fill(i, s_cur);
                __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
                // the waits are explicit; keep the compiler from moving
                // LDS accesses across the barrier
                asm volatile("" ::: "memory");
                __builtin_amdgcn_s_barrier();
                asm volatile("" ::: "memory");
                if (i + A < NL) issue(i + A, s_nxt);

                const T* s = reinterpret_cast<const T*>(slot_ptr(s_cur));
                s_cur = s_cur + 1 == D ? 0 : s_cur + 1;
                s_nxt = s_nxt + 1 == D ? 0 : s_nxt + 1;
                const int p = xs - R + i;
                const bool own = p >= xs && p < xe;
                // column chunks of rows lrow0 - R .. lrow0 + PY - 1 + R
                C col[PY + 2 * R];
                if (own) {
#pragma unroll
                    for (int r = 0; r < PY + 2 * R; r++)
                        col[r] = *reinterpret_cast<const C*>(
                            s + (lrow0 - R + r) * n2 + z0);
                } else {
#pragma unroll
                    for (int j = 0; j < PY; j++)
                        col[R + j] = *reinterpret_cast<const C*>(
                            s + (lrow0 + j) * n2 + z0);
                }
#pragma unroll
                for (int j = 0; j < PY; j++) {
                    const C v = col[R + j];
                    // x terms into outputs p - d (d = -R..R, d != 0); axes
                    // a stencil lacks have zero coefficients
                    {
#pragma unroll
                        for (int k = 1; k <= R; k++) {
                            // plane p is x + k for output x = p - k
                            C& lo = acc[(u - R - k + 2 * QU) % QU][j];
                            C& hi = acc[(u - R + k + 2 * QU) % QU][j];
#pragma unroll
                            for (int q = 0; q < PZ; q++) {
                                lo.v[q] = fmac(lo.v[q], cpx[k - 1], v.v[q]);
                                hi.v[q] = fmac(hi.v[q], cmx[k - 1], v.v[q]);
                            }
                        }
                    }
                    if (own) {
                        C& me = acc[(u - R + 2 * QU) % QU][j];
#pragma unroll
                        for (int q = 0; q < PZ; q++)
                            me.v[q] = fmac(me.v[q], c0, v.v[q]);
                        {
#pragma unroll
                            for (int k = 1; k <= R; k++)
#pragma unroll
                                for (int q = 0; q < PZ; q++) {
                                    me.v[q] = fmac(me.v[q], cpy[k - 1],
                                                   col[R + j + k].v[q]);
                                    me.v[q] = fmac(me.v[q], cmy[k - 1],
                                                   col[R + j - k].v[q]);
                                }
                        }
                        {
                            const int lrow = (lrow0 + j) * n2;
                            T wz[NW * PZ];
#pragma unroll
                            for (int c = 0; c < NW; c++) {
                                C ch = v;
                                if (c != RC) {
                                    ch = *reinterpret_cast<const C*>(
                                        s + lrow + zcol[c]);
                                    if (!zok[c]) ch = zero_chunk<T, PZ>();
                                }
#pragma unroll
                                for (int q = 0; q < PZ; q++)
                                    wz[c * PZ + q] = ch.v[q];
                            }
#pragma unroll
                            for (int q = 0; q < PZ; q++)
#pragma unroll
                                for (int k = 1; k <= R; k++) {
                                    me.v[q] = fmac(me.v[q], cpz[k - 1],
                                                   wz[RP + q + k]);
                                    me.v[q] = fmac(me.v[q], cmz[k - 1],
                                                   wz[RP + q - k]);
                                }
                        }
                    }
                }
                // output p - R is complete
                const int xd = p - R;
                C(&done)[PY] = acc[(u + 1) % QU];
                if (xd >= xs && xd < xe) {
                    const long obase =
                        b * a.ng + (long)xd * plane + (long)(y0 + sg * PY) * n2 + z0;
#pragma unroll
                    for (int j = 0; j < PY; j++) {
                        C* dst = reinterpret_cast<C*>(a.out + obase + (long)j * n2);
                        if (a.nt & 1)
                            st16nt(done[j], dst);
                        else
                            *dst = done[j];
                    }
                }
#pragma unroll
                for (int j = 0; j < PY; j++) done[j] = zero_chunk<T, PZ>();
            }
        }
    }
}

// Tile check of the DMA kernel: slot chunks per thread, ring size in LDS.
// Returns the LDS bytes, 0 if the tile does not fit.
inline size_t xd_lds_bytes(const XmTile& t, int R, int es, const int n[3],
                           int* kd_out = nullptr)
{
    const int PZ = 16 / es;
    if (t.dma < 1 || t.dma > 4 || n[2] % PZ != 0 || 2 * R > 2 * t.nseg) return 0;
    const int CPR = n[2] / PZ;
    const int NT = CPR * t.nseg;
    const int TH = t.nseg * t.py + 2 * R;
    if (NT > XR_MAXT || NT < 64) return 0;
    const int kd = (TH * CPR + NT - 1) / NT;
    if (kd > XD_KDMAX) return 0;
    const size_t bytes = (size_t)(t.dma + 1) * kd * NT * 16;
    if (bytes > 160 * 1024) return 0;
    if (kd_out) *kd_out = kd;
    return bytes;
}

template <typename T, int R, int PY, int H>
static int launch_xd_t(const AxisStencil& st, const XmTile& tl,
                       const int n[3], const BC& bc, const void* in,
                       void* out, long nin, hipStream_t stream,
                       int* nblocks_out)
{
    constexpr int PZ = PzOf<T>::value;
    int kd = 0;
    const size_t shm = xd_lds_bytes(tl, R, (int)sizeof(T), n, &kd);
    if (tl.order < 0 || tl.order > 2 || (tl.order == 2 && nin % 8 != 0)) {
        set_error("x-dma: block order %d not usable for %ld grids", tl.order,
                  nin);
        return ENOTSUP_;
    }
    if (shm != 0 && kd != PY + H) {
        set_error("x-dma: %d DMA chunks per thread, kernel built for %d", kd,
                  PY + H);
        return ENOTSUP_;
    }
    if (shm == 0 || n[1] % (tl.nseg * PY) != 0 || R > n[0] || R > n[1] ||
        n[2] < ((R + PZ - 1) / PZ) * PZ) {
        set_error("x-dma: tile does not fit this grid");
        return ENOTSUP_;
    }
    XdArgs<T> a;
    a.in = (const T*)in;
    a.out = (T*)out;
    a.n0 = n[0];
    a.n1 = n[1];
    a.n2 = n[2];
    a.ng = (long)n[0] * n[1] * n[2];
    const int TY = tl.nseg * PY;
    a.tiles_y = n[1] / TY;
    a.xchunk = tl.xchunk;
    a.chunks_x = (n[0] + tl.xchunk - 1) / tl.xchunk;
    const long nb = (long)a.tiles_y * a.chunks_x * nin;
    if (nb <= 0) return OK;
    if (nb > 0x7fffffffL) {
        set_error("too many blocks (%ld)", nb);
        return ENOTSUP_;
    }
    a.nblocks = (int)nb;
    a.ahead = tl.dma;
    a.order = tl.order;
    const int NT = (n[2] / PZ) * tl.nseg;
    a.slot_chunks = kd * NT;
    a.c0 = st.c0;
    for (int ax = 0; ax < 3; ax++) {
        a.periodic[ax] = bc.periodic[ax];
        for (int k = 0; k < MAXR_FAST; k++) {
            a.cp[ax][k] = st.cp[ax][k];
            a.cm[ax][k] = st.cm[ax][k];
        }
    }
    a.nt = xr_nt_flags((long)n[0] * n[1] * n[2] * nin * (long)sizeof(T));
    static size_t attr_bytes = 64 * 1024;
    if (shm > attr_bytes) {
        hipError_t e = hipFuncSetAttribute(
            (const void*)k_xdma<T, R, PY, PZ, H>,
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("x-dma: %zu bytes of LDS not available (%s)", shm,
                      hipGetErrorString(e));
            return ENOTSUP_;
        }
        attr_bytes = shm;
    }
    (void)hipGetLastError();
    hipLaunchKernelGGL((k_xdma<T, R, PY, PZ, H>), dim3((unsigned)nb),
                       dim3(n[2] / PZ, tl.nseg), shm, stream, a);
    GFD_LAUNCH_CHECK();
    if (nblocks_out) *nblocks_out = (int)nb;
    return OK;
}

template <typename T>
static int launch_xd(const AxisStencil& st, const XmTile& tl, const int n[3],
                     const BC& bc, const void* in, void* out, long nin,
                     hipStream_t stream, int* nbo)
{
    const int h = (2 * st.R + tl.nseg - 1) / std::max(1, tl.nseg);
#define XD_CASE(RR, PP, HH)                                                  \
    if (st.R == RR && tl.py == PP && h == HH)                                \
        return launch_xd_t<T, RR, PP, HH>(st, tl, n, bc, in, out, nin,       \
                                          stream, nbo);
    XD_CASE(1, 1, 1) XD_CASE(1, 2, 1) XD_CASE(1, 4, 1)
    XD_CASE(1, 1, 2) XD_CASE(1, 2, 2) XD_CASE(1, 4, 2)
    XD_CASE(2, 1, 1) XD_CASE(2, 2, 1) XD_CASE(2, 4, 1)
    XD_CASE(2, 1, 2) XD_CASE(2, 2, 2) XD_CASE(2, 4, 2)
    XD_CASE(3, 1, 1) XD_CASE(3, 2, 1) XD_CASE(3, 4, 1)
    XD_CASE(3, 1, 2) XD_CASE(3, 2, 2) XD_CASE(3, 4, 2)
    XD_CASE(4, 1, 1) XD_CASE(4, 2, 1) XD_CASE(4, 4, 1)
    XD_CASE(4, 1, 2) XD_CASE(4, 2, 2) XD_CASE(4, 4, 2)
    // 3 rows per thread: 24-row tiles with 8 full waves of 48-thread rows
    XD_CASE(1, 3, 1) XD_CASE(1, 3, 2) XD_CASE(2, 3, 1) XD_CASE(2, 3, 2)
    XD_CASE(3, 3, 1) XD_CASE(3, 3, 2) XD_CASE(4, 3, 1) XD_CASE(4, 3, 2)
#undef XD_CASE
    set_error("x-dma: unsupported radius %d / rows %d", st.R, tl.py);
    return ENOTSUP_;
}

}  // namespace gfd
