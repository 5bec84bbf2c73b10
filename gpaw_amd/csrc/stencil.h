// Stencil kernel arguments and launchers (FDOperator.apply / relax and the
// fused residual forms used by the multigrid drivers).
#pragma once

#include <vector>

#include "gfd_common.h"

namespace gfd {

constexpr int MAXR_FAST = 4;   // largest radius of the x-marching kernel
constexpr int MAXP = 256;      // largest number of stencil points

// What the kernel does with acc = sum_{p != centre} c_p in(x + o_p) and the
// centre value v = in(x):
enum Epi : int {
    EPI_PLAIN = 0,   // out = c0 v + acc
    EPI_RELAX = 1,   // out = (1 - w) v + w (aux - acc) / c0        (Jacobi)
    EPI_RESID = 2,   // out = (c0 v + acc) - aux          [+ sum out^2]
    EPI_UPDATE = 3,  // out = +-(v - s ((c0 v + acc) - aux))
    EPI_POT = 4,     // out = (c0 v + acc) + pot v [- shift[b] v], pot real
                     // per grid point, shift one real per grid of the batch
    EPI_UPDW = 5,    // EPI_UPDATE on FP32 data with aux and out in FP64
    EPI_RESN = 6,    // EPI_RESID on FP64 data, out narrowed to FP32
                     // (shared by the grids of a batch): H psi
};

struct EpiArgs {
    const void* aux;   // src (RELAX) or subtrahend (RESID/UPDATE), may be 0
    double c0;
    double w;          // RELAX weight
    double s;          // UPDATE step
    int negate;        // UPDATE: store -(...)
    double* partial;   // RESID: per-block sum of |out|^2, or 0
    const double* pot; // POT: real potential of one grid
    const double* shift = nullptr;  // POT: per-grid eigenvalue (residual)
};

// Axis-only (central-difference) stencil in per-axis form.
struct AxisStencil {
    int R;                    // radius (1..MAXR_FAST)
    double c0;                // centre coefficient
    double cp[3][MAXR_FAST];  // [axis][k-1]: coefficient of offset +k
    double cm[3][MAXR_FAST];  // [axis][k-1]: coefficient of offset -k
    int has[3];               // axis carries nonzero coefficients
};

// General stencil: coefficients and offsets live in device memory.
struct GeneralStencil {
    int P;
    int centre;               // index of the (first) zero offset, -1 if none
    double c0;
    const double* coef;       // [P] (centre term excluded from the sum)
    const int* off;           // [P][3]
};

// Tile choice of the x-marching kernel (exposed for the autotuner).
struct XmTile {
    int tz;      // tile width along z (threads per row)
    int nseg;    // row segments (threads along y)
    int py;      // rows per thread
    int xchunk;  // planes per block along x
    int dma = 0; // full-row kernel: planes in flight by LDS DMA (k_xdma),
                 // 0 for register prefetch (k_xrow)
    int order = 0;  // k_xdma block order (rows_dma.h)
    int wide = 0;   // k_xrow on 16-byte data: 2 = two points per thread
};
// Launch one stencil pass over `nin` grids of shape n[3].
// dtype: 0 = f64, 1 = c128, 2 = f32, 3 = c64 (row kernel and general
// kernel; the tiled x-march kernel takes 0-2).
int launch_axis_stencil(const AxisStencil& st, int dtype, const int n[3],
                        const BC& bc, const void* in, void* out, long nin,
                        int epi, const EpiArgs& ea, hipStream_t stream,
                        int* nblocks_out);
int launch_general_stencil(const GeneralStencil& st, int dtype,
                           const int n[3], const BC& bc, const void* in,
                           void* out, long nin, int epi, const EpiArgs& ea,
                           hipStream_t stream, int* nblocks_out);

// Full-row x-marching kernel (16-byte vector accesses; n2 % (16 / element
// size) == 0): the first choice for axis-only stencils.
int launch_row_stencil(const AxisStencil& st, int dtype, const int n[3],
                       const BC& bc, const void* in, void* out, long nin,
                       int epi, const EpiArgs& ea, hipStream_t stream,
                       int* nblocks_out);
bool row_stencil_fits(const AxisStencil& st, int dtype, const int n[3],
                      long nin);
long row_stencil_max_blocks(const AxisStencil& st, int dtype, const int n[3],
                            long nin);
XmTile xr_default_tile(int R, int dtype, const int n[3], long nin);
// The tiles the tuner times for a radius, data type and grid shape besides
// the heuristic one (k_xrow tiles, then with dma the k_xdma tiles).
std::vector<XmTile> xr_candidates(int R, int dtype, const int n[3], bool dma);
// Timing autotuner of the full-row kernel's tile (on by default; off also
// forgets the tiles found).
void xr_set_autotune(bool on);
// While on (this thread only): no timing runs; a grid the tuner has not
// seen yet gets the heuristic tile.  For callers that must not allocate.
void xr_no_tune_here(bool on);
// Let the tuner try the LDS-DMA kernel (k_xdma) for plain real passes.
void xr_set_dma(bool on);

// Non-temporal hints of the full-row kernel: -1 automatic, else bit 0
// output stores, bit 1 own-row loads.
void xr_set_nt(int flags);
void xr_set_override(const XmTile* t);

// Whether the x-marching kernel takes this stencil / grid (else general).
bool axis_stencil_fits(const AxisStencil& st, int dtype, const int n[3],
                       long nin);

// Block-count upper bounds for the residual-norm partial buffer.
long axis_stencil_max_blocks(const AxisStencil& st, int dtype, const int n[3],
                             long nin);
long general_stencil_max_blocks(const int n[3], long nin);

// Deterministic final reduction of `np` partial sums into *result (device).
int launch_sum_partials(const double* partial, long np, double* result,
                        hipStream_t stream);

XmTile xm_default_tile(int R, int dtype, const int n[3], long nin);
void xm_set_override(const XmTile* t);  // nullptr clears

}  // namespace gfd
