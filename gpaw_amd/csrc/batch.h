// Grid tables shared by the batched multigrid kernels (batch.hip,
// bxmarch.hip): every grid of a batch level lives in one concatenated level
// array and every stencil launch walks the tiles of all grids.
#pragma once

#include <hip/hip_runtime.h>

#include "gfd_common.h"

namespace gfd {

constexpr int BMAXR = 3;

// One grid of one level in a batched stencil launch.
struct BSt {
    long off;          // element offset in the level arrays
    long tile0;        // first tile of the grid in the launch
    int n[3];
    int periodic[3];
    int tz, ty;        // tiles along z and y
    int xc;            // x planes per tile (tiles = ceil(n0 / xc) ty tz)
    int mz, mseg;      // x-march tiles: z width, row segments of the block
    int has[3];
    double c0, w;      // w: Jacobi weight
    double step;       // step of EPI_UPDATE / EPI_UPDW
    double cp[3][BMAXR], cm[3][BMAXR];
};

struct BSArgs {
    const BSt* g;
    int G;
    long ntiles;
    const void* in;
    void* out;
    const void* aux;
    double* partial;   // one double per tile (EPI_RESID)
    int negate;        // EPI_UPDATE / EPI_UPDW: store the negated result
};

__device__ inline int bfind_tile(const BSt* g, int G, long t)
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

// index along an axis with periodic wrap; false if outside a zero boundary
__device__ inline bool bwrap(int& i, int n, int periodic)
{
    if (i >= 0 && i < n) return true;
    if (!periodic) return false;
    i %= n;
    if (i < 0) i += n;
    return true;
}

// One grid pair of a batched K = 2 transfer.
struct BXf {
    long in_off, out_off;
    long p0;           // first output point of the grid in the launch
    int ni[3], no[3];
    int off[3], skip0[3];
    int periodic[3];
    double alpha;      // XEPI_SCALE2: out2 = alpha out (per grid)
    // x-march tiles over the output grid (bxfer_march.hip): first tile,
    // tiles along z and y, output planes per tile, z width, row segments
    long tile0;
    int tz, ty, xc, mz, mseg;
};

struct BXArgs {
    const BXf* g;
    int G;
    long total;
    const void* in;
    void* out;
    const void* base;  // XEPI_SUBFROM: out = base - T(in)
    void* out2;        // XEPI_SCALE2
    int negate;        // XEPI_SCALE2: out = -T(in)
    long ntiles;       // x-march launch: tiles of all grids
};

// Buffer resources: raw loads and stores through a resource whose
// base is the start of a plane and whose size is the plane's bytes (0 for
// a zero plane); a byte offset at or past the size reads 0 and drops a
// store, which is how zero-boundary neighbours, points outside the grid
// and lanes without an output are masked without branches.
constexpr unsigned BXM_OOB = 0x80000000u;

__device__ inline __amdgpu_buffer_rsrc_t brsrc(const void* base,
                                               unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base),
                                             (short)0, (int)bytes,
                                             0x00020000);
}

template <typename T>
__device__ inline T bload(__amdgpu_buffer_rsrc_t r, unsigned voff)
{
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(
            T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, 0, 0));
    } else if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(
            T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, 0, 0));
    } else {
        static_assert(sizeof(T) == 16, "4, 8 or 16-byte elements");
        return __builtin_bit_cast(
            T, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, 0, 0));
    }
}

template <typename T>
__device__ inline void bstore(T v, __amdgpu_buffer_rsrc_t r, unsigned voff)
{
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(
            __builtin_bit_cast(unsigned, v), r, (int)voff, 0, 0);
    } else {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r,
                                              (int)voff, 0, 0);
    }
}

// batched-only epilogue of the x-march kernel: a Jacobi sweep (FP32 data,
// FP32 rhs in aux) whose result e is subtracted from the FP64 array `out`
// at the output points (out -= e), in place of storing e
constexpr int BEPI_RELAXSUB = 16;

// x-march batched stencil (bxmarch.hip): 1-D blocks of `nt` threads, each
// grid's tile TZ = mz columns by mseg * py rows, pf planes (1, 2) of
// loads in flight; `lds` dynamic bytes.
// dtype / epilogue pairs as for the register-tile kernel of batch.hip.
constexpr int BXM_HMAX = 2;     // halo elements per thread per plane
constexpr int BXM_MAXT = 512;
// 0: timed choice of the two kernels, 1: register tiles, 2: x-march
void batch_set_kernel(int kind);
// batched transfers tuned afterwards: 0 timed choice, 1 one point per
// thread, 2 x-march (timed tile), or a forced x-march tile nt | py << 12 |
// zmax << 16 | xc << 28 (nt >= 64)
void batch_set_xfer(long code);
// x-march tile used by every level tuned afterwards (nt 0: timed choice)
void batch_force_tile(int nt, int py, int zmax, int xc, int pf);
int launch_bxmarch(const BSArgs& a, int R, int py, int pf, int nt,
                   size_t lds, int dtype, int epi, hipStream_t s);

// x-march batched K = 2 transfers (bxfer_march.hip): one block of `nt`
// threads per tile of each grid's output (table fields tile0 .. mseg), py
// output rows per thread; dtype 0 FP64, 2 FP32, 4 FP32 from FP64 input
// (restriction); epilogue xe of transfer.h.
int launch_bxfer_march(const BXArgs& a, bool interp, int py, int nt,
                       int dtype, int xe, hipStream_t s);

}  // namespace gfd
