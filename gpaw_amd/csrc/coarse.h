// LDS-resident coarse V-cycle: one workgroup runs every level of a grid's
// hierarchy from level lc down to its coarsest, with all level arrays in
// LDS, in a single launch (one workgroup per grid for a batch).
#pragma once

#include "gfd_common.h"

namespace gfd {

constexpr int CMAXLEV = 8;
constexpr int CMAXR = 3;

struct CLevel {
    int n[3];
    int periodic[3];
    int R;
    int has[3];
    double c0;
    double cp[3][CMAXR], cm[3][CMAXR];
    double w;
    int pre, post;           // sweeps (post only on the coarsest level)
    int a, b, r;             // LDS element offsets: e ping-pong and rhs
                             // (first level: staged from global memory,
                             // or r < 0: read there in every pass)
    // transfer to the next coarser level
    int roff[3];             // restriction: first centre per axis
    int ioff[3], iskip0[3];  // interpolation back to this level
};

struct CGrid {
    int nlev;                // levels handled (lc .. lc + nlev - 1)
    int lds_elems;           // LDS elements used
    long rhs_off, out_off;   // element offsets of the grid's rhs / result
    CLevel lev[CMAXLEV];
};

// LDS bytes available to one workgroup for the coarse solve
constexpr size_t COARSE_LDS_MAX = 150 * 1024;

// dtype 0: f64, 2: f32.  rhs: level-lc right-hand sides, out: level-lc
// corrections (zero initial guess), grid g at rhs_off / out_off.
int launch_coarse(const CGrid* d_grids, int G, size_t lds_bytes, int dtype,
                  const void* rhs, void* out, hipStream_t s);

}  // namespace gfd
