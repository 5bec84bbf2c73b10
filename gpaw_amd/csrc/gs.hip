// Lexicographic Gauss-Seidel relaxation (relax method 1) on the GPU, for
// the plain operators (c/bmgs/relax.cpp:12-40) and the weighted ones
// (c/bmgs/wrelax.cpp:26-58), bit for bit.
//
// The reference visits the points in C order and reads, for every
// neighbour, the value in its padded buffer: updated already when the
// neighbour is inside the grid and earlier in that order, the value from
// the start of the sweep otherwise; halo points (wrapped periodic
// neighbours) hold start-of-sweep copies, zero-boundary halos hold 0
// (c/operators.cpp:74-86).
//
// The same values are read here in wavefront order: a point (i0, i1, i2)
// is updated at step t = a0 i0 + a1 i1 + i2, with a0, a1 >= 1 chosen on the
// host so that every offset that precedes the centre in C order lowers t
// and every offset that follows it raises t.  Points of one step never
// read each other, so each step is one parallel pass.  Neighbours inside
// the grid are read from the array being updated; wrapped neighbours from
// a copy made at the start of the sweep.  Sums are in the reference's
// point order with FP contraction off, so the results are those of the
// reference's CPU loop.
//
// Grids whose (i0, i1) plane fits one workgroup's threads run every sweep
// of a call in one workgroup (one barrier per step); larger grids launch
// one kernel per step.
#include <algorithm>
#include <vector>

#include "gfd_common.h"
#include "gs.h"

namespace gfd {

constexpr int GS_BLOCK = 1024;
// largest (i0, i1) plane for the one-workgroup kernel: each thread then
// handles at most this / GS_BLOCK columns per step
constexpr long GS_ONE_BLOCK_MAX = 4L * GS_BLOCK;

__device__ inline bool gs_wrap(int& j, int n, int periodic, bool& wrapped)
{
    if (j >= 0 && j < n) return true;
    if (!periodic) return false;
    wrapped = true;
    j = j < 0 ? j + n : j - n;
    return true;
}

// The neighbour value the reference reads for point offset c (3 ints at
// off + 3 c): cur inside the grid, the start-of-sweep copy when wrapped,
// 0 outside a zero boundary.
__device__ inline double gs_nb(const GsArgs& a, const double* cur,
                               const double* snap, int i0, int i1, int i2,
                               int c)
{
    int j0 = i0 + a.off[3 * c], j1 = i1 + a.off[3 * c + 1],
        j2 = i2 + a.off[3 * c + 2];
    bool wrapped = false;
    if (!gs_wrap(j0, a.n[0], a.periodic[0], wrapped) ||
        !gs_wrap(j1, a.n[1], a.periodic[1], wrapped) ||
        !gs_wrap(j2, a.n[2], a.periodic[2], wrapped))
        return 0.0;
    const long j = ((long)j0 * a.n[1] + j1) * a.n[2] + j2;
    return wrapped ? snap[j] : cur[j];
}

// New value of one point (c/bmgs/relax.cpp:30-36, wrelax.cpp:40-52).
__device__ inline double gs_point(const GsArgs& a, const double* cur,
                                  const double* snap, const double* src,
                                  int i0, int i1, int i2, long idx)
{
#pragma clang fp contract(off)
    if (a.nw == 0) {
        double x = 0.0;
        for (int c = 1; c < a.P; c++)
            x += gs_nb(a, cur, snap, i0, i1, i2, c) * a.coef[c];
        return (src[idx] - x) * a.inv_c0;
    }
    double x = 0.0, coef = 0.0;
    for (int w = 0; w < a.nw; w++) {
        const double weight = a.wts[w][idx];
        double tmp = 0.0;
        for (int c = a.start[w] + 1; c < a.start[w + 1]; c++)
            tmp += gs_nb(a, cur, snap, i0, i1, i2, c) * a.coef[c];
        tmp *= weight;
        x += tmp;
        coef += weight * a.coef[a.start[w]];
    }
    return (src[idx] - x) / coef;
}

// One wavefront step t over the whole grid: thread (i0, i1) updates the
// point with i2 = t - a0 i0 - a1 i1 if it lies in the grid.
__global__ void __launch_bounds__(256)
k_gs_step(GsArgs a, double* __restrict__ f, const double* __restrict__ snap,
          const double* __restrict__ src, int t)
{
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= (long)a.n[0] * a.n[1]) return;
    const int i0 = (int)(id / a.n[1]), i1 = (int)(id % a.n[1]);
    const int i2 = t - a.a0 * i0 - a.a1 * i1;
    if (i2 < 0 || i2 >= a.n[2]) return;
    const long idx = id * a.n[2] + i2;
    f[idx] = gs_point(a, f, snap, src, i0, i1, i2, idx);
}

// Every sweep in one workgroup: copy to the snapshot, then the steps in
// order with a workgroup barrier (which orders the global stores of one
// step before the loads of the next for all threads of the group).
__global__ void __launch_bounds__(GS_BLOCK)
k_gs_block(GsArgs a, double* __restrict__ f, double* __restrict__ snap,
           const double* __restrict__ src, int nsweeps, int nsteps)
{
    const long ng = (long)a.n[0] * a.n[1] * a.n[2];
    const long ncol = (long)a.n[0] * a.n[1];
    for (int s = 0; s < nsweeps; s++) {
        for (long i = threadIdx.x; i < ng; i += GS_BLOCK) snap[i] = f[i];
        __syncthreads();
        for (int t = 0; t < nsteps; t++) {
            for (long id = threadIdx.x; id < ncol; id += GS_BLOCK) {
                const int i0 = (int)(id / a.n[1]), i1 = (int)(id % a.n[1]);
                const int i2 = t - a.a0 * i0 - a.a1 * i1;
                if (i2 < 0 || i2 >= a.n[2]) continue;
                const long idx = id * a.n[2] + i2;
                f[idx] = gs_point(a, f, snap, src, i0, i1, i2, idx);
            }
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(256)
k_gs_copy(const double* __restrict__ a, double* __restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long)gridDim.x * 256)
        b[i] = a[i];
}

// Step weights: the smallest a1, then a0, for which every point offset
// that precedes the centre in C order has a0 o0 + a1 o1 + o2 < 0 and every
// one that follows it has a0 o0 + a1 o1 + o2 > 0.
int gs_order(const int* off, int npoints, int* a0, int* a1)
{
    for (int b1 = 1; b1 <= 1024; b1++) {
        bool ok1 = true;
        for (int c = 0; c < npoints && ok1; c++) {
            const int* o = off + 3 * c;
            if (o[0] != 0 || (o[1] == 0 && o[2] == 0)) continue;
            // in C order (0, o1, o2) precedes the centre when o1 < 0, or
            // o1 = 0 and o2 < 0
            const bool before = o[1] < 0 || (o[1] == 0 && o[2] < 0);
            const long v = (long)b1 * o[1] + o[2];
            ok1 = before ? v < 0 : v > 0;
        }
        if (!ok1) continue;
        for (int b0 = 1; b0 <= (1 << 20); b0++) {
            bool ok = true;
            for (int c = 0; c < npoints && ok; c++) {
                const int* o = off + 3 * c;
                if (o[0] == 0) continue;
                const long v = (long)b0 * o[0] + (long)b1 * o[1] + o[2];
                ok = o[0] < 0 ? v < 0 : v > 0;
            }
            if (ok) {
                *a0 = b0;
                *a1 = b1;
                return OK;
            }
        }
    }
    set_error("Gauss-Seidel: no wavefront order for this stencil");
    return ENOTSUP_;
}

}  // namespace gfd

// Host-only view of the wavefront weights (declared in gpaw_fd.h).
extern "C" int gfd_gs_order(const int* offsets, int npoints, int* a0,
                            int* a1)
{
    if (!offsets || npoints <= 0 || !a0 || !a1) {
        gfd::set_error("gfd_gs_order: bad arguments");
        return gfd::EINVAL_;
    }
    return gfd::gs_order(offsets, npoints, a0, a1);
}

namespace gfd {

int gs_relax(GsArgs a, double* f, double* snap, const double* src,
             int nsweeps, hipStream_t s)
{
    if (nsweeps <= 0) return OK;
    const long ng = (long)a.n[0] * a.n[1] * a.n[2];
    const long ncol = (long)a.n[0] * a.n[1];
    const long steps =
        (long)a.a0 * (a.n[0] - 1) + (long)a.a1 * (a.n[1] - 1) + a.n[2];
    if (steps > (1L << 30)) {
        set_error("Gauss-Seidel: grid too large for the wavefront order");
        return EINVAL_;
    }
    if (ncol <= GS_ONE_BLOCK_MAX) {
        hipLaunchKernelGGL(k_gs_block, dim3(1), dim3(GS_BLOCK), 0, s, a, f,
                           snap, src, nsweeps, (int)steps);
        GFD_LAUNCH_CHECK();
        return OK;
    }
    const unsigned gcopy =
        (unsigned)std::min<long>(std::max<long>(1, (ng + 255) / 256), 8192);
    const unsigned gstep = (unsigned)((ncol + 255) / 256);
    for (int k = 0; k < nsweeps; k++) {
        hipLaunchKernelGGL(k_gs_copy, dim3(gcopy), dim3(256), 0, s,
                           (const double*)f, snap, ng);
        GFD_LAUNCH_CHECK();
        for (long t = 0; t < steps; t++) {
            hipLaunchKernelGGL(k_gs_step, dim3(gstep), dim3(256), 0, s, a, f,
                               (const double*)snap, src, (int)t);
            GFD_LAUNCH_CHECK();
        }
    }
    return OK;
}

}  // namespace gfd
