// Device helpers shared by the stencil kernels (wrap, epilogues,
// reductions, XCD-aware block order).
#pragma once

#include "stencil.h"

namespace gfd {

template <typename T> struct IsCplx { static constexpr bool value = false; };
template <> struct IsCplx<double2> { static constexpr bool value = true; };
template <> struct IsCplx<float2> { static constexpr bool value = true; };

// Bring i into [0, n); multiply the factor by the Bloch phase once per wrap.
__device__ inline bool wrap_full(int& i, int n, int periodic, double2 hi,
                                 double2 lo, double2& f, bool& phased)
{
    if (i >= 0 && i < n) return true;
    if (!periodic) return false;
    while (i < 0) {
        i += n;
        f = phased ? cmul(lo, f) : lo;
        phased = true;
    }
    while (i >= n) {
        i -= n;
        f = phased ? cmul(hi, f) : hi;
        phased = true;
    }
    return true;
}

template <typename T>
__device__ inline T scale_by(T v, double2 f, bool phased);
template <> __device__ inline double scale_by<double>(double v, double2, bool)
{
    return v;
}
template <> __device__ inline float scale_by<float>(float v, double2, bool)
{
    return v;
}
template <>
__device__ inline double2 scale_by<double2>(double2 v, double2 f, bool phased)
{
    return phased ? cmul(f, v) : v;
}
template <>
__device__ inline float2 scale_by<float2>(float2 v, double2 f, bool phased)
{
    return phased ? cmul(make_float2((float)f.x, (float)f.y), v) : v;
}

// FP64 counterpart of an FP32 data type (EPI_UPDW's aux and out)
template <typename T> struct WideOf { typedef T type; };
template <> struct WideOf<float> { typedef double type; };
template <> struct WideOf<float2> { typedef double2 type; };

__device__ inline float to_narrow(double v) { return (float)v; }
__device__ inline float2 to_narrow(double2 v)
{
    return make_float2((float)v.x, (float)v.y);
}
__device__ inline double to_wide(float v) { return (double)v; }
__device__ inline double2 to_wide(float2 v)
{
    return make_double2((double)v.x, (double)v.y);
}
__device__ inline double sq(double v) { return v * v; }
__device__ inline double sq(float v) { return (double)v * (double)v; }
__device__ inline double sq(double2 v) { return v.x * v.x + v.y * v.y; }
__device__ inline double sq(float2 v)
{
    return (double)v.x * v.x + (double)v.y * v.y;
}

// Epilogue (see stencil.h) on one point: acc excludes the centre term, v
// is the centre value, av the auxiliary operand (src or subtrahend).
template <typename T, int EPI>
__device__ inline T epi_apply(T acc, T v, T av, double c0, double w,
                              double s, int negate, double& nrm)
{
    typedef typename ScalarOf<T>::type S;
    if constexpr (EPI == EPI_PLAIN) {
        return fmac(acc, (S)c0, v);
    } else if constexpr (EPI == EPI_RELAX) {
        if constexpr (IsCplx<T>::value) {
            return v;  // relax is real only (rejected on the host)
        } else {
            return ((S)1.0 - (S)w) * v + (S)w * (av - acc) / (S)c0;
        }
    } else if constexpr (EPI == EPI_RESID) {
        T r = fmac(acc, (S)c0, v);
        r = r - av;
        nrm += sq(r);
        return r;
    } else {
        T r = fmac(acc, (S)c0, v);
        r = r - av;
        T d = fmac(v, -(S)s, r);
        return negate ? -d : d;
    }
}

template <typename T, int EPI>
__device__ inline T epilogue(T acc, T v, const T* __restrict__ aux, long idx,
                             double c0, double w, double s, int negate,
                             double& nrm)
{
    T av = zero_of<T>();
    if constexpr (EPI != EPI_PLAIN) av = aux[idx];
    return epi_apply<T, EPI>(acc, v, av, c0, w, s, negate, nrm);
}

// Block reduction of one double per thread into partial[slot].
__device__ inline void block_sum_store(double v, double* partial, int slot)
{
    __shared__ double red[16];
    const int nt = blockDim.x * blockDim.y;
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = t & 63, wv = t >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        for (int i = 0; i < (nt + 63) / 64; i++) s += red[i];
        partial[slot] = s;
    }
}

// Blocks are handed to the 8 XCDs round robin; remap so that consecutive
// logical tiles (neighbours of one grid) share an XCD and its L2.
__device__ inline int xcd_remap(int h, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int x = h & 7, k = h >> 3;
    return x * q + min(x, r) + k;
}

}  // namespace gfd
