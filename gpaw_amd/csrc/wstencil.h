// Weighted operators: the axis-only form run by the weighted x-march
// kernel (wxmarch.hip).
#pragma once

#include "stencil.h"

namespace gfd {

constexpr int WX_NW = 4;  // stencils of the fast path (dielectric: 4)

// Every stencil of the operator as per-axis coefficients around its
// centre, with the weight grids (device pointers, ng doubles each).
struct WAxis {
    int nw = 0;
    int R = 0;
    int has[3] = {0, 0, 0};
    int hasw[WX_NW][3];  // stencil w has coefficients along the axis
    double c0[WX_NW];
    double cp[WX_NW][3][MAXR_FAST];  // [stencil][axis][k - 1]: offset +k
    double cm[WX_NW][3][MAXR_FAST];  // offset -k
    const double* wts[WX_NW];
};

// out = sum_w W_w (stencil_w in) over nin grids, or (relax, real data, one
// grid) one weighted Jacobi sweep in -> out with right-hand side src and
// weight w.  ENOTSUP_ when the operator or grid does not fit the kernel.
int launch_wxmarch(const WAxis& st, int real, const int n[3], const BC& bc,
                   const void* in, void* out, const double* src, long nin,
                   double w, bool relax, hipStream_t s);
// Route weighted operators through the fast kernel (default on).
void wx_set_enabled(bool on);
// Tile overrides of the weighted x-march kernel (0: automatic): rows per
// thread (1, 2, 4) and x planes per block.
void wx_set_tile(int py, int xchunk);

}  // namespace gfd
