// FFT / DST Poisson solve (FastPoissonSolver) on hipFFT.
//
// The FD operator is diagonal in the Fourier basis of a periodic axis and in
// the DST-I basis of a zero-boundary axis.  A zero-boundary axis of N - 1
// interior points is extended oddly to a period of 2N (0, a_1 .. a_{N-1},
// 0, -a_{N-1} .. -a_1), on which the periodic stencil reproduces the
// zero-boundary operator exactly; the solve is then one real-to-complex 3-D
// FFT of the extended grid, a multiplication by the inverse eigenvalues
// (computed by the caller from the stencil), the inverse FFT and the
// extraction of the interior -- the DST-I of the reference's `fst` by the
// same odd extension (gpaw/poisson.py:721-771).
#include "../../include/gpaw_fd.h"

#include <hipfft/hipfft.h>

#include <vector>

#include "gfd_common.h"

namespace gfd {

struct FExt {
    int n[3];    // grid points (interior for zero axes)
    int m[3];    // extended (FFT) grid
    int odd[3];  // zero-boundary axis: odd extension
};

// extended index j (0 .. m-1) -> grid index and sign (0 sign: zero)
__device__ inline int fsrc(int j, int n, int m, int odd, int& sg)
{
    if (!odd) {
        sg = 1;
        return j;
    }
    // m = 2 (n + 1): j = 0 and j = n + 1 are zero
    const int N = n + 1;
    if (j == 0 || j == N) {
        sg = 0;
        return 0;
    }
    if (j < N) {
        sg = 1;
        return j - 1;
    }
    sg = -1;
    return m - j - 1;
}

__global__ void __launch_bounds__(256)
k_fext(FExt e, const double* __restrict__ rho, double* __restrict__ ext)
{
    const long total = (long)e.m[0] * e.m[1] * e.m[2];
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < total;
         p += (long)gridDim.x * 256) {
        const int j2 = (int)(p % e.m[2]);
        const long q = p / e.m[2];
        const int j1 = (int)(q % e.m[1]);
        const int j0 = (int)(q / e.m[1]);
        int s0, s1, s2;
        const int i0 = fsrc(j0, e.n[0], e.m[0], e.odd[0], s0);
        const int i1 = fsrc(j1, e.n[1], e.m[1], e.odd[1], s1);
        const int i2 = fsrc(j2, e.n[2], e.m[2], e.odd[2], s2);
        const int sg = s0 * s1 * s2;
        ext[p] = sg == 0 ? 0.0
                         : sg * rho[((long)i0 * e.n[1] + i1) * e.n[2] + i2];
    }
}

__global__ void __launch_bounds__(256)
k_fscale(double2* __restrict__ spec, const double* __restrict__ inv_lambda,
         long total)
{
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < total;
         p += (long)gridDim.x * 256) {
        const double f = inv_lambda[p];
        const double2 v = spec[p];
        spec[p] = make_double2(f * v.x, f * v.y);
    }
}

__global__ void __launch_bounds__(256)
k_fextract(FExt e, const double* __restrict__ ext, double* __restrict__ phi)
{
    const long total = (long)e.n[0] * e.n[1] * e.n[2];
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < total;
         p += (long)gridDim.x * 256) {
        const int i2 = (int)(p % e.n[2]);
        const long q = p / e.n[2];
        const int i1 = (int)(q % e.n[1]);
        const int i0 = (int)(q / e.n[1]);
        const int j0 = i0 + e.odd[0], j1 = i1 + e.odd[1], j2 = i2 + e.odd[2];
        phi[p] = ext[((long)j0 * e.m[1] + j1) * e.m[2] + j2];
    }
}

static unsigned fgrid(long n)
{
    long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

static int fft_ok(hipfftResult r, const char* what)
{
    if (r != HIPFFT_SUCCESS) {
        set_error("%s failed (hipfft status %d)", what, (int)r);
        return EHIP;
    }
    return OK;
}

}  // namespace gfd

using namespace gfd;

struct gfd_fft_poisson {
    FExt e{};
    long nspec = 0;
    double* inv = nullptr;     // inverse eigenvalues / (m0 m1 m2)
    double* ext = nullptr;     // extended real grid
    double2* spec = nullptr;   // half spectrum
    hipfftHandle fwd = 0, bwd = 0;
    bool have_fwd = false, have_bwd = false;
};

extern "C" void gfd_fft_poisson_free(gfd_fft_poisson* f)
{
    if (!f) return;
    if (f->have_fwd) (void)hipfftDestroy(f->fwd);
    if (f->have_bwd) (void)hipfftDestroy(f->bwd);
    if (f->inv) (void)hipFree(f->inv);
    if (f->ext) (void)hipFree(f->ext);
    if (f->spec) (void)hipFree(f->spec);
    delete f;
}

extern "C" int gfd_fft_poisson_new(const long n[3], const int zero_bc[3],
                                   const double* inv_lambda,
                                   gfd_fft_poisson** out)
{
    if (!out || !n || !zero_bc || !inv_lambda) {
        set_error("gfd_fft_poisson_new: bad arguments");
        return EINVAL_;
    }
    *out = nullptr;
    auto* f = new gfd_fft_poisson;
    long mtot = 1;
    for (int c = 0; c < 3; c++) {
        if (n[c] < 1 || n[c] > (1L << 20)) {
            set_error("gfd_fft_poisson_new: bad grid size");
            delete f;
            return EINVAL_;
        }
        f->e.n[c] = (int)n[c];
        f->e.odd[c] = zero_bc[c] ? 1 : 0;
        f->e.m[c] = zero_bc[c] ? 2 * ((int)n[c] + 1) : (int)n[c];
        mtot *= f->e.m[c];
    }
    f->nspec = (long)f->e.m[0] * f->e.m[1] * (f->e.m[2] / 2 + 1);
    const double scale = 1.0 / (double)mtot;
    std::vector<double> h(inv_lambda, inv_lambda + f->nspec);
    for (double& v : h) v *= scale;
    int st = OK;
    if (hipMalloc(&f->inv, sizeof(double) * f->nspec) != hipSuccess ||
        hipMalloc(&f->ext, sizeof(double) * mtot) != hipSuccess ||
        hipMalloc(&f->spec, sizeof(double2) * f->nspec) != hipSuccess) {
        (void)hipGetLastError();
        set_error("gfd_fft_poisson_new: out of device memory");
        gfd_fft_poisson_free(f);
        return ENOMEM_;
    }
    if (hipMemcpy(f->inv, h.data(), sizeof(double) * f->nspec,
                  hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        set_error("gfd_fft_poisson_new: copy failed");
        gfd_fft_poisson_free(f);
        return EHIP;
    }
    st = fft_ok(hipfftPlan3d(&f->fwd, f->e.m[0], f->e.m[1], f->e.m[2],
                             HIPFFT_D2Z),
                "hipfftPlan3d(D2Z)");
    if (st == OK) {
        f->have_fwd = true;
        st = fft_ok(hipfftPlan3d(&f->bwd, f->e.m[0], f->e.m[1], f->e.m[2],
                                 HIPFFT_Z2D),
                    "hipfftPlan3d(Z2D)");
        if (st == OK) f->have_bwd = true;
    }
    if (st != OK) {
        gfd_fft_poisson_free(f);
        return st;
    }
    *out = f;
    return OK;
}

extern "C" long gfd_fft_poisson_spectrum_shape(const long n[3],
                                               const int zero_bc[3],
                                               long m_out[3])
{
    if (!n || !zero_bc || !m_out) return -1;
    for (int c = 0; c < 3; c++)
        m_out[c] = zero_bc[c] ? 2 * (n[c] + 1) : n[c];
    return m_out[0] * m_out[1] * (m_out[2] / 2 + 1);
}

extern "C" int gfd_fft_poisson_solve(gfd_fft_poisson* f, const double* rho,
                                     double* phi, void* stream)
{
    if (!f || !rho || !phi) {
        set_error("gfd_fft_poisson_solve: bad arguments");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    const long mtot = (long)f->e.m[0] * f->e.m[1] * f->e.m[2];
    const long ntot = (long)f->e.n[0] * f->e.n[1] * f->e.n[2];
    hipLaunchKernelGGL(k_fext, dim3(fgrid(mtot)), dim3(256), 0, s, f->e, rho,
                       f->ext);
    GFD_LAUNCH_CHECK();
    GFD_TRY(fft_ok(hipfftSetStream(f->fwd, s), "hipfftSetStream"));
    GFD_TRY(fft_ok(hipfftExecD2Z(f->fwd, f->ext,
                                 (hipfftDoubleComplex*)f->spec),
                   "hipfftExecD2Z"));
    hipLaunchKernelGGL(k_fscale, dim3(fgrid(f->nspec)), dim3(256), 0, s,
                       f->spec, f->inv, f->nspec);
    GFD_LAUNCH_CHECK();
    GFD_TRY(fft_ok(hipfftSetStream(f->bwd, s), "hipfftSetStream"));
    GFD_TRY(fft_ok(hipfftExecZ2D(f->bwd, (hipfftDoubleComplex*)f->spec,
                                 f->ext),
                   "hipfftExecZ2D"));
    hipLaunchKernelGGL(k_fextract, dim3(fgrid(ntot)), dim3(256), 0, s, f->e,
                       f->ext, phi);
    GFD_LAUNCH_CHECK();
    return OK;
}
