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

// x-march batched stencil (bxmarch.hip): 1-D blocks of `nt` threads, each
// grid's tile TZ = mz columns by mseg * py rows, pf planes (1, 2) of
// loads in flight; `lds` dynamic bytes.
// dtype / epilogue pairs as for the register-tile kernel of batch.hip.
constexpr int BXM_HMAX = 2;     // halo elements per thread per plane
constexpr int BXM_MAXT = 512;
// 0: timed choice of the two kernels, 1: register tiles, 2: x-march
void batch_set_kernel(int kind);
// x-march tile used by every level tuned afterwards (nt 0: timed choice)
void batch_force_tile(int nt, int py, int zmax, int xc, int pf);
int launch_bxmarch(const BSArgs& a, int R, int py, int pf, int nt,
                   size_t lds, int dtype, int epi, hipStream_t s);

}  // namespace gfd
