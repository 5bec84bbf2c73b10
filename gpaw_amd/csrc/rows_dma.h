// Full-row x-marching stencil with LDS-DMA plane prefetch (k_xdma), for
// axis-only stencils on real data.
//
// Same tile as k_xrow (whole z rows, TY = nseg * PY rows in y, a chunk of
// x planes per block), but planes reach LDS by `global_load_lds_dwordx4`:
// the tile's TY + 2R rows of a plane are contiguous in the grid (up to the
// periodic wrap in y), so each lane copies 16-byte chunks straight into a
// ring of D = A + 2 plane slots without passing through VGPRs.  A planes
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
//  * synchronisation per plane: issue the DMA of plane i + A into the slot
//    last read at step i - 2, wait until this wave's DMA of plane i has
//    landed (vmcnt = DMAs issued after it), write zero chunks, one
//    barrier.  Vector memory operations complete in issue order, so later
//    stores only make the wait stricter.
#pragma once

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
    int ahead;          // planes in flight (A); ring of A + 2 slots
    int kd;             // DMA chunks per thread per plane
    int slot_chunks;    // kd * threads: 16-byte chunks per slot
    int hasx, hasy, hasz;
    double c0;
    double cp[3][MAXR_FAST];
    double cm[3][MAXR_FAST];
    int periodic[3];
    int nt;
};

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

// s_waitcnt vmcnt(n) lgkmcnt(0) for a runtime n (the immediate must be a
// constant: one branch per plane)
__device__ inline void xd_wait(int n)
{
    switch (n) {
#define XD_W(k) case k: __builtin_amdgcn_s_waitcnt(XD_VMCNT(k) & ~0xF00); break;
        XD_W(0) XD_W(1) XD_W(2) XD_W(3) XD_W(4) XD_W(5) XD_W(6) XD_W(7)
        XD_W(8) XD_W(9) XD_W(10) XD_W(11) XD_W(12) XD_W(13) XD_W(14)
        XD_W(15) XD_W(16) XD_W(17) XD_W(18) XD_W(19) XD_W(20) XD_W(21)
        XD_W(22) XD_W(23) XD_W(24) XD_W(25) XD_W(26) XD_W(27) XD_W(28)
        XD_W(29) XD_W(30) XD_W(31) XD_W(32)
#undef XD_W
    default: __builtin_amdgcn_s_waitcnt(0); break;
    }
}

template <typename T, int R, int PY, int PZ>
__global__ void __launch_bounds__(XR_MAXT) k_xdma(XdArgs<T> a)
{
    typedef Chunk<T, PZ> C;
    constexpr int RC = (R + PZ - 1) / PZ;
    constexpr int NW = 2 * RC + 1;
    constexpr int RP = RC * PZ;
    constexpr int QU = 2 * R + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];

    const int L = xcd_remap(blockIdx.x, a.nblocks);
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
    const int D = A + 2;
    const int KD = a.kd;
    const int SLOT = a.slot_chunks * 16;   // bytes
    const T* __restrict__ gin = a.in + b * a.ng;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wave_base = wave * 64;       // first chunk of the wave per k

    // DMA chunks of this thread: c = k * NT + tid, row c / CPR of the slot
    int doff[XD_KDMAX];
    bool dzero[XD_KDMAX];
#pragma unroll
    for (int k = 0; k < XD_KDMAX; k++) {
        doff[k] = 0;
        dzero[k] = false;
        if (k < KD) {
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

    auto slot_ptr = [&](int i) -> unsigned char* {
        return smem_raw + (i % D) * SLOT;
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
    auto issue = [&](int i) {
        const long po = plane_off(i);
        const T* src = gin + (po < 0 ? 0 : po);
        unsigned char* s = slot_ptr(i);
#pragma unroll
        for (int k = 0; k < XD_KDMAX; k++)
            if (k < KD)
                xd_dma16(src + doff[k],
                         (unsigned)(size_t)(xd_lds_t*)(s + (k * NT + wave_base) * 16));
    };
    auto zero_fill = [&](int i) {
        const bool zp = plane_off(i) < 0;
        unsigned char* s = slot_ptr(i);
#pragma unroll
        for (int k = 0; k < XD_KDMAX; k++)
            if (k < KD && (zp || dzero[k]))
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

    for (int i = 0; i < A && i < NL; i++) issue(i);

    for (int ig = 0; ig < NL; ig += QU) {
#pragma unroll
        for (int u = 0; u < QU; u++) {
            const int i = ig + u;
            if (i < NL) {
                if (i + A < NL) issue(i + A);
                // DMAs issued after landing i: planes i+1 .. min(i+A, NL-1)
                xd_wait(KD * min(A, NL - 1 - i));
                asm volatile("" ::: "memory");
                zero_fill(i);
                __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
                // the waits are explicit; keep the compiler from moving
                // LDS accesses across the barrier
                asm volatile("" ::: "memory");
                __builtin_amdgcn_s_barrier();
                asm volatile("" ::: "memory");

                const T* s = reinterpret_cast<const T*>(slot_ptr(i));
                const int p = xs - R + i;
                const bool own = p >= xs && p < xe;
                // column chunks of rows lrow0 - R .. lrow0 + PY - 1 + R
                C col[PY + 2 * R];
                if (own && a.hasy) {
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
                    // x terms into outputs p - d (d = -R..R, d != 0)
                    if (a.hasx) {
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
                        if (a.hasy) {
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
                        if (a.hasz) {
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
    if (t.dma < 1 || t.dma > 4 || n[2] % PZ != 0) return 0;
    const int CPR = n[2] / PZ;
    const int NT = CPR * t.nseg;
    const int TH = t.nseg * t.py + 2 * R;
    if (NT > XR_MAXT || NT < 64) return 0;
    const int kd = (TH * CPR + NT - 1) / NT;
    if (kd > XD_KDMAX || kd * t.dma > 32) return 0;
    const size_t bytes = (size_t)(t.dma + 2) * kd * NT * 16;
    if (bytes > 160 * 1024) return 0;
    if (kd_out) *kd_out = kd;
    return bytes;
}

template <typename T, int R, int PY>
static int launch_xd_t(const AxisStencil& st, const XmTile& tl,
                       const int n[3], const BC& bc, const void* in,
                       void* out, long nin, hipStream_t stream,
                       int* nblocks_out)
{
    constexpr int PZ = PzOf<T>::value;
    int kd = 0;
    const size_t shm = xd_lds_bytes(tl, R, (int)sizeof(T), n, &kd);
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
    a.kd = kd;
    const int NT = (n[2] / PZ) * tl.nseg;
    a.slot_chunks = kd * NT;
    a.hasx = st.has[0];
    a.hasy = st.has[1];
    a.hasz = st.has[2];
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
            (const void*)k_xdma<T, R, PY, PZ>,
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
    hipLaunchKernelGGL((k_xdma<T, R, PY, PZ>), dim3((unsigned)nb),
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
#define XD_CASE(RR, PP)                                                      \
    if (st.R == RR && tl.py == PP)                                           \
        return launch_xd_t<T, RR, PP>(st, tl, n, bc, in, out, nin, stream,   \
                                      nbo);
    XD_CASE(1, 1) XD_CASE(1, 2) XD_CASE(1, 4)
    XD_CASE(2, 1) XD_CASE(2, 2) XD_CASE(2, 4)
    XD_CASE(3, 1) XD_CASE(3, 2) XD_CASE(3, 4)
    XD_CASE(4, 1) XD_CASE(4, 2) XD_CASE(4, 4)
#undef XD_CASE
    set_error("x-dma: unsupported radius %d / rows %d", st.R, tl.py);
    return ENOTSUP_;
}

}  // namespace gfd
