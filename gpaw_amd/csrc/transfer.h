// Restriction / interpolation (Transformer.apply) on gfx950.
#pragma once

#include "gfd_common.h"

namespace gfd {

// What the last pass does with the transferred value T:
enum XEpi : int {
    XEPI_PLAIN = 0,  // out = T
    XEPI_NEG = 1,    // out = -T
    XEPI_SCALE2 = 2, // out = T, out2 = alpha T          (alpha negated w/ neg)
    XEPI_SUBFROM = 3 // out = base - T
};

struct XEpiArgs {
    int epi = XEPI_PLAIN;
    int negate = 0;          // applies to T before the epilogue (SCALE2)
    double alpha = 0.0;
    void* out2 = nullptr;
    const void* base = nullptr;
    // FP32 data (dtype 2 / 3) only, K = 2: the input is the FP64
    // counterpart (f64 / c128), narrowed as it is read
    int in_wide = 0;
};

// Per-axis description of one 1-D transfer pass.
struct XferAxis {
    int nin;     // input points along the axis
    int nout;    // output points along the axis
    int off;     // interior index of the first centre (see transfer.hip)
    int skip0;   // interpolation: first even output dropped
};

struct XferPlan {
    int interpolate;
    int K;           // 2, 4, 6, 8
    int nin[3];      // input grid
    int nout[3];     // output grid
    XferAxis ax[3];
};

// K = 2 transfers run as one fused kernel unless forced to the 3 passes.
void xfer_force_passes(bool on);
// Use the x-marching fused kernels (every K; default on).  Off: K = 2 by
// the one-point-per-thread fused kernels, K > 2 by three 1-D passes.
void xfer_set_march(bool on);
// Whether K = 2 transfers are forced onto the separable passes (they then
// need the passes' work space).
bool xfer_passes_forced();
// Tiles per resident block of the marching kernels, in units of four
// (default 1: about four).
void xfer_set_march_scale(int s);

// Timed choice between the x-marching kernel and the other kernels per
// transfer shape (on by default; off also forgets the choices).
void xfer_set_autotune(bool on);
// Whether this shape has no choice yet (and the tuner is on).
bool xfer_needs_tuning(const XferPlan& p, int dtype, long nin, int in_wide);
// Time both kernel families with PLAIN passes of `in` into tune_out
// (output-sized scratch) and record the faster.  Not under graph capture.
int xfer_tune(const XferPlan& p, int dtype, const BC& bc, const void* in,
              void* tune_out, long nin, void* scratch, int in_wide,
              hipStream_t s);

// The x-marching fused transfer (xfer_march.hip); ENOTSUP_ for a shape
// its tiles cannot take.
int launch_transfer_march(const XferPlan& p, int dtype, const BC& bc,
                          const void* in, void* out, long nin,
                          const XEpiArgs& ea, hipStream_t stream);

// Scratch (in elements of the data type) needed for `nin` grids.
long xfer_scratch_elems(const XferPlan& p, long nin);

// dtype: 0 f64, 1 c128, 2 f32, 3 c64 (c64: K = 2 only)
int launch_transfer(const XferPlan& p, int dtype, const BC& bc,
                    const void* in, void* out, long nin, void* scratch,
                    const XEpiArgs& ea, hipStream_t stream);

}  // namespace gfd
