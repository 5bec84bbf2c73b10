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
    // coarsest level only: cheb = 1 runs `post` steps of Chebyshev
    // iteration on D^-1 A with eigenvalue bounds [lam_lo, lam_hi] in
    // place of Jacobi sweeps; dbuf: LDS offset of its update array
    int cheb, dbuf;
    double lam_lo, lam_hi;
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

// Chebyshev bounds of level L: the extreme nonzero eigenvalues of
// D^-1 A for a radius-1 axis stencil (Fourier modes on periodic axes, sine
// modes on zero-boundary axes, which are its exact eigenvectors), widened
// by a safety margin.  False if they do not apply (radius > 1, or an
// operator that is not a symmetric negative-definite-like stencil).
bool cheb_bounds(const CLevel& L, double* lo, double* hi);

// Chebyshev coarsest-level solve in the mixed-precision schedules (on by
// default; the reference schedule always smooths with Jacobi)
void coarse_set_cheb(bool on);
bool coarse_cheb_enabled();

// dtype 0: f64, 2: f32.  rhs: level-lc right-hand sides, out: level-lc
// corrections (zero initial guess), grid g at rhs_off / out_off.
int launch_coarse(const CGrid* d_grids, int G, size_t lds_bytes, int dtype,
                  const void* rhs, void* out, hipStream_t s);

}  // namespace gfd
