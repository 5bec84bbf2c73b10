// Shared definitions for the gfx950 FD grid-path library.
//
// Data layout (HBM): every grid is C-order (n0, n1, n2) with n2 contiguous;
// batches of grids are stored back to back (band-major), complex values are
// interleaved (re, im) doubles (hip double2).  No padded copies are ever
// made: the boundary is applied by index wrap in the loaders, with the
// Bloch factor phase[axis][0] above the upper end and phase[axis][1] below
// the lower end of a periodic axis (c/bc.cpp:242-257 semantics), and zero
// outside a zero-boundary axis.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace gfd {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
const char* last_error();

enum Status : int {
    OK = 0,
    EINVAL_ = 1,    // bad argument
    EHIP = 2,       // HIP runtime error
    ENOMEM_ = 3,    // device allocation failed
    ENOTSUP_ = 4,   // combination not supported
    ECONVERGE = 5,  // iterative solver did not converge
};

#define GFD_CHECK_HIP(expr)                                                  \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            ::gfd::set_error("%s failed: %s (%s:%d)", #expr,                 \
                             hipGetErrorString(e_), __FILE__, __LINE__);     \
            return ::gfd::EHIP;                                              \
        }                                                                    \
    } while (0)

#define GFD_LAUNCH_CHECK()                                                   \
    do {                                                                     \
        hipError_t e_ = hipGetLastError();                                   \
        if (e_ != hipSuccess) {                                              \
            ::gfd::set_error("kernel launch failed: %s (%s:%d)",             \
                             hipGetErrorString(e_), __FILE__, __LINE__);     \
            return ::gfd::EHIP;                                              \
        }                                                                    \
    } while (0)

#define GFD_TRY(expr)                                                        \
    do {                                                                     \
        int s_ = (expr);                                                     \
        if (s_ != ::gfd::OK) return s_;                                      \
    } while (0)

// ---------------------------------------------------------------- scalars
typedef double2 cplx;

__host__ __device__ inline double2 operator+(double2 a, double2 b)
{
    return make_double2(a.x + b.x, a.y + b.y);
}
__host__ __device__ inline double2 operator-(double2 a, double2 b)
{
    return make_double2(a.x - b.x, a.y - b.y);
}
__host__ __device__ inline double2 operator*(double s, double2 a)
{
    return make_double2(s * a.x, s * a.y);
}
__host__ __device__ inline double2 operator-(double2 a)
{
    return make_double2(-a.x, -a.y);
}
__host__ __device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__host__ __device__ inline float2 operator+(float2 a, float2 b)
{
    return make_float2(a.x + b.x, a.y + b.y);
}
__host__ __device__ inline float2 operator-(float2 a, float2 b)
{
    return make_float2(a.x - b.x, a.y - b.y);
}
__host__ __device__ inline float2 operator*(float s, float2 a)
{
    return make_float2(s * a.x, s * a.y);
}
__host__ __device__ inline float2 operator-(float2 a)
{
    return make_float2(-a.x, -a.y);
}
__host__ __device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Scalar-generic helpers used by the templated kernels (T = double|double2
// for the FP64 paths, float|float2 for the mixed-precision smoothers and
// preconditioner).
template <typename T> struct ScalarOf;
template <> struct ScalarOf<double> { typedef double type; };
template <> struct ScalarOf<double2> { typedef double type; };
template <> struct ScalarOf<float> { typedef float type; };
template <> struct ScalarOf<float2> { typedef float type; };

template <typename T>
__device__ inline T zero_of();
template <> __device__ inline double zero_of<double>() { return 0.0; }
template <> __device__ inline float zero_of<float>() { return 0.0f; }
template <> __device__ inline double2 zero_of<double2>()
{
    return make_double2(0.0, 0.0);
}
template <> __device__ inline float2 zero_of<float2>()
{
    return make_float2(0.0f, 0.0f);
}

// acc + c * v with c real
__device__ inline double fmac(double acc, double c, double v)
{
    return fma(c, v, acc);
}
__device__ inline float fmac(float acc, float c, float v)
{
    return fmaf(c, v, acc);
}
__device__ inline double2 fmac(double2 acc, double c, double2 v)
{
    return make_double2(fma(c, v.x, acc.x), fma(c, v.y, acc.y));
}
__device__ inline float2 fmac(float2 acc, float c, float2 v)
{
    return make_float2(fmaf(c, v.x, acc.x), fmaf(c, v.y, acc.y));
}

// value times a boundary factor (1, 0 or a Bloch phase)
__device__ inline double apply_phase(double v, double2 ph) { return ph.x * v; }
__device__ inline float apply_phase(float v, double2 ph)
{
    return (float)ph.x * v;
}
__device__ inline double2 apply_phase(double2 v, double2 ph)
{
    return cmul(ph, v);
}
__device__ inline float2 apply_phase(float2 v, double2 ph)
{
    return cmul(make_float2((float)ph.x, (float)ph.y), v);
}

// ---------------------------------------------------------------- grids
// Boundary of one grid: per axis a periodic flag and the two Bloch phases.
struct BC {
    int periodic[3];
    double2 phase[3][2];  // [axis][0]: above the upper end, [1]: below lower
};

// Flat point index gid of a batch of n0 x n1 x n2 grids (grid points each,
// total points in the launch) -> (item b, i0, i1, i2).  32-bit divisions
// when the launch has fewer than 2^31 points (a 64-bit division expands to
// about a hundred instructions, more than the point's own work in the
// one-point-per-thread kernels); 64-bit beyond.
__device__ inline void flat_split(long gid, long total, long grid, int n1,
                                  int n2, long& b, int& i0, int& i1, int& i2)
{
    if (total <= 0x7fffffffL) {
        const unsigned g = (unsigned)gid, gs = (unsigned)grid;
        const unsigned bb = g / gs;
        const unsigned r = g - bb * gs;
        const unsigned q = r / (unsigned)n2;
        i2 = (int)(r - q * (unsigned)n2);
        const unsigned q1 = q / (unsigned)n1;
        i1 = (int)(q - q1 * (unsigned)n1);
        i0 = (int)q1;
        b = bb;
    } else {
        b = gid / grid;
        long r = gid - b * grid;
        i2 = (int)(r % n2);
        r /= n2;
        i1 = (int)(r % n1);
        i0 = (int)(r / n1);
    }
}

// Map a possibly out-of-range index along one axis to (index, factor).
// Returns false when the value is zero (zero boundary).  `i` must lie in
// [-n, 2n) -- the stencil range never exceeds the grid size.
__device__ inline bool wrap_axis(int& i, int n, int periodic, double2 ph_hi,
                                 double2 ph_lo, double2& f, bool& phased)
{
    if (i >= 0 && i < n) return true;
    if (!periodic) return false;
    if (i < 0) {
        i += n;
        f = phased ? cmul(ph_lo, f) : ph_lo;
    } else {
        i -= n;
        f = phased ? cmul(ph_hi, f) : ph_hi;
    }
    phased = true;
    return true;
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace gfd
