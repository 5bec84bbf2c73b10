#pragma once

#include "gfd_common.h"
#include "stencil.h"

namespace gfd {

int vec_zero(double* a, long n, hipStream_t s);
int vec_copy(const double* a, double* b, long n, hipStream_t s);
int vec_zero_bytes(void* a, long nbytes, hipStream_t s);  // nbytes % 4 == 0
int vec_d2f(const double* a, float* b, long n, hipStream_t s);
int vec_sub_f(double* a, const float* e, long n, hipStream_t s);
int vec_sub_d(double* a, const double* e, long n, hipStream_t s);  // a -= e
// out = w f / c0 (dtype 0 f64, 2 f32): a Jacobi sweep from zero
int vec_relax0(void* out, const void* f, long n, double w, double c0,
               int dtype, hipStream_t s);
int vec_add_scalar(double* a, long n, double v, hipStream_t s);
int vec_copy16(const void* a, void* b, long nbytes, hipStream_t s);
long vec_dot_partials(long n);
// *dev_result = sum_i a_i b_i (b == nullptr: sum_i a_i), deterministic order
int vec_dot(const double* a, const double* b, long n, double* partial,
            double* dev_result, hipStream_t s);

}  // namespace gfd
