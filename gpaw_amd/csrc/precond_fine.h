// Fused fine level of the kinetic preconditioner (real or complex; the
// FP32 chain or the FP64 one): the interpolation d0' = I1(-d1) and the
// update d0 = -(d0' - step (K0 d0' - res)) in one pass, d0' never written
// to memory.
#pragma once

#include "stencil.h"
#include "transfer.h"

namespace gfd {

// Launch the fused pass on nin bands: d1 (f32 / c64, or f64 / c128 with
// fp64, the interpolation's coarse input, plan ip), res and out (f64 / c128
// on the fine grid n of the axis
// stencil ax).  ENOTSUP_ when the shape or stencil does not fit the kernel
// (the caller then runs the two passes); OK when launched.
int launch_precond_fine(const AxisStencil& ax, double c0, const int n[3],
                        const int fine_periodic[3], const XferPlan& ip,
                        const int coarse_periodic[3], const double* phases,
                        const void* d1, const void* res, void* out, long nin,
                        double step, int real, int fp64, hipStream_t s);

// 0: off (always two passes), 1: on where the shape fits (default)
void precond_fine_enable(bool on);
// non-temporal stores of the result (default off)
void precond_fine_nt(bool on);
// force the tile (rows per thread py, thread rows nys, x chunk xc; 0:
// automatic)
void precond_fine_force(int py, int nys, int xc);

}  // namespace gfd
