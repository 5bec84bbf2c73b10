// Lexicographic Gauss-Seidel sweeps (relax method 1), plain and weighted
// operators; kernels and the wavefront order in gs.hip.
#pragma once

#include "gfd_common.h"

namespace gfd {

struct GsArgs {
    int n[3];
    int periodic[3];
    int a0, a1;            // step t = a0 i0 + a1 i1 + i2
    int nw;                // 0: plain operator; > 0: weighted, nw stencils
    int P;                 // plain: points (the first is the centre)
    const int* start;      // weighted: first point of stencil w, [nw + 1]
    const double* coef;    // device, per point
    const int* off;        // device, per point (o0, o1, o2)
    const double* const* wts;  // weighted: nw device grids
    double inv_c0;         // plain: 1 / coef[0], computed on the host
};

// (a0, a1) for the point offsets off[3 * npoints] (host array).
int gs_order(const int* off, int npoints, int* a0, int* a1);

// nsweeps Gauss-Seidel sweeps on f in place; snap: work array of the grid's
// size (start-of-sweep values).
int gs_relax(GsArgs a, double* f, double* snap, const double* src,
             int nsweeps, hipStream_t s);

}  // namespace gfd
