// Handle objects behind the C ABI's opaque pointers (gfd_operator,
// gfd_transformer) and the grow-only device buffer they own; shared by the
// single-grid drivers (api.hip) and the batched ones (batch.hip).
#pragma once

#include <vector>

#include "gfd_common.h"
#include "stencil.h"
#include "transfer.h"

namespace gfd {

// 0 auto, 1 general kernel, 2 tiled x-march kernel, 3 full-row kernel
extern int g_force_path;

// Grow-only device buffer.  Growing waits for the stream first, so a buffer
// still read by queued kernels is never freed under them.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want, hipStream_t s)
    {
        if (want <= bytes) return OK;
        if (p) {
            GFD_CHECK_HIP(hipStreamSynchronize(s));
            GFD_CHECK_HIP(hipFree(p));
            p = nullptr;
            bytes = 0;
        }
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            set_error("hipMalloc of %zu bytes failed", want);
            return ENOMEM_;
        }
        bytes = want;
        return OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

}  // namespace gfd

using gfd::AxisStencil;
using gfd::DevBuf;
using gfd::GeneralStencil;
using gfd::XferPlan;
using gfd::g_force_path;

struct gfd_operator {
    int P = 0;
    std::vector<double> coef;
    std::vector<int> off;  // P x 3
    int n[3] = {0, 0, 0};
    int range = 0;
    int periodic[3] = {0, 0, 0};
    int real = 1;
    double diag = 0.0;      // get_diagonal_element (last zero offset)
    int centre = -1;        // first zero offset
    bool axis_ok = false;
    AxisStencil ax{};
    double* d_coef = nullptr;
    int* d_off = nullptr;
    DevBuf scratch;         // relax ping-pong / in-place copies
    DevBuf partial;         // residual-norm partial sums
    int gs_a0 = 0, gs_a1 = 0;  // Gauss-Seidel wavefront order (0: unset)
    DevBuf scalar;          // one device double
    long ng() const { return (long)n[0] * n[1] * n[2]; }
    int dtype() const { return real ? 0 : 1; }
    GeneralStencil gen(double c0) const
    {
        GeneralStencil g;
        g.P = P;
        g.centre = centre;
        g.c0 = c0;
        g.coef = d_coef;
        g.off = d_off;
        return g;
    }
    bool use_fast() const { return axis_ok && g_force_path != 1; }
};

struct gfd_transformer {
    XferPlan plan{};
    int periodic[3] = {0, 0, 0};
    int real = 1;
    DevBuf scratch;
    DevBuf tune_out;  // output-sized target of the kernel-choice timing
    long ngin() const
    {
        return (long)plan.nin[0] * plan.nin[1] * plan.nin[2];
    }
    long ngout() const
    {
        return (long)plan.nout[0] * plan.nout[1] * plan.nout[2];
    }
    int dtype() const { return real ? 0 : 1; }
};

