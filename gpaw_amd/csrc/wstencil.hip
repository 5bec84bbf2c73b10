// Weighted FD operators (WeightedFDOperator, the dielectric Poisson
// operator of the solvation code): out(x) = sum_w weight_w(x) *
// sum_p c_{w,p} in(x + o_{w,p}), and its weighted Jacobi relaxation
// (c/woperators.cpp, c/bmgs/wfd.cpp, c/bmgs/wrelax.cpp).
//
// One thread per grid point; the neighbours of every stencil come through
// L1/L2 (the stencils of one operator share their offsets, so each
// neighbour is fetched from memory once and re-read from cache), the
// weights are read once per point.  Wrapping and Bloch phases as in the
// other stencil kernels (wrap_full); no padded copies.
#include "../../include/gpaw_fd.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gs.h"
#include "handles.h"
#include "stencil_dev.h"
#include "wstencil.h"

namespace gfd {

constexpr int WMAXW = 64;  // weights per operator

struct WArgs {
    int nw;
    int n[3];
    long ng;
    const int* start;          // [nw + 1]: first point of each stencil
    const double* coef;        // [points]
    const int* off;            // [points][3]
    const double* const* wts;  // [nw] device pointers, ng doubles each
    BC bc;
    // flat offset of the (x, y) part of every point (points away from
    // the x and y edges) and the largest offset component
    const long* foff;
    int R;
};

template <typename T>
__device__ inline T wneighbour(const WArgs& a, const T* g, int i0, int i1,
                               int i2, int c, bool& ok)
{
    int j0 = i0 + a.off[3 * c], j1 = i1 + a.off[3 * c + 1],
        j2 = i2 + a.off[3 * c + 2];
    double2 f = make_double2(1.0, 0.0);
    bool ph = false;
    ok = wrap_full(j0, a.n[0], a.bc.periodic[0], a.bc.phase[0][0],
                   a.bc.phase[0][1], f, ph) &&
         wrap_full(j1, a.n[1], a.bc.periodic[1], a.bc.phase[1][0],
                   a.bc.phase[1][1], f, ph) &&
         wrap_full(j2, a.n[2], a.bc.periodic[2], a.bc.phase[2][0],
                   a.bc.phase[2][1], f, ph);
    if (!ok) return zero_of<T>();
    const T v = g[((long)j0 * a.n[1] + j1) * a.n[2] + j2];
    return scale_by<T>(v, f, ph);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_wapply(WArgs a, const T* __restrict__ in, T* __restrict__ out, long nin)
{
    typedef typename ScalarOf<T>::type S;
    const long total = a.ng * nin;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < total;
         p += (long)gridDim.x * 256) {
        const long b = p / a.ng, idx = p - b * a.ng;
        const int i2 = (int)(idx % a.n[2]);
        const long q = idx / a.n[2];
        const int i1 = (int)(q % a.n[1]), i0 = (int)(q / a.n[1]);
        const T* g = in + b * a.ng;
        T x = zero_of<T>();
        const int R = a.R;
        if (i0 >= R && i0 < a.n[0] - R && i1 >= R && i1 < a.n[1] - R) {
            // away from the x and y edges (uniform across most waves): flat
            // (x, y) offsets, z wrapped by selects
            const long row = idx - i2;
            for (int w = 0; w < a.nw; w++) {
                T t = zero_of<T>();
                for (int c = a.start[w]; c < a.start[w + 1]; c++) {
                    int j2 = i2 + a.off[3 * c + 2];
                    double2 f = make_double2(1.0, 0.0);
                    bool ph = false;
                    if (j2 < 0 || j2 >= a.n[2]) {
                        if (!a.bc.periodic[2]) continue;
                        ph = true;
                        f = j2 < 0 ? a.bc.phase[2][1] : a.bc.phase[2][0];
                        j2 = j2 < 0 ? j2 + a.n[2] : j2 - a.n[2];
                    }
                    const T v = scale_by<T>(g[row + a.foff[c] + j2], f, ph);
                    t = fmac(t, (S)a.coef[c], v);
                }
                x = fmac(x, (S)a.wts[w][idx], t);
            }
        } else {
            for (int w = 0; w < a.nw; w++) {
                T t = zero_of<T>();
                for (int c = a.start[w]; c < a.start[w + 1]; c++) {
                    bool ok;
                    const T v = wneighbour(a, g, i0, i1, i2, c, ok);
                    if (ok) t = fmac(t, (S)a.coef[c], v);
                }
                x = fmac(x, (S)a.wts[w][idx], t);
            }
        }
        out[p] = x;
    }
}

// Weighted Jacobi sweep: nxt = (1 - w) cur + w (src - x) / coef with
// x = sum_w weight_w sum_{p >= 1} c_{w,p} cur(x + o), coef = sum_w weight_w
// c_{w,0} (the first point of every stencil is the centre)
__global__ void __launch_bounds__(256)
k_wrelax(WArgs a, const double* __restrict__ cur, double* __restrict__ nxt,
         const double* __restrict__ src, double w)
{
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < a.ng;
         idx += (long)gridDim.x * 256) {
        const int i2 = (int)(idx % a.n[2]);
        const long q = idx / a.n[2];
        const int i1 = (int)(q % a.n[1]), i0 = (int)(q / a.n[1]);
        double x = 0.0, coef = 0.0;
        const int R = a.R;
        const bool inner = i0 >= R && i0 < a.n[0] - R && i1 >= R &&
                           i1 < a.n[1] - R && i2 >= R && i2 < a.n[2] - R;
        for (int v = 0; v < a.nw; v++) {
            const double wt = a.wts[v][idx];
            double t = 0.0;
            for (int c = a.start[v] + 1; c < a.start[v + 1]; c++) {
                if (inner) {
                    t = fma(a.coef[c],
                            cur[idx + a.foff[c] + a.off[3 * c + 2]], t);
                } else {
                    bool ok;
                    const double nb = wneighbour(a, cur, i0, i1, i2, c, ok);
                    if (ok) t = fma(a.coef[c], nb, t);
                }
            }
            x = fma(wt, t, x);
            coef = fma(wt, a.coef[a.start[v]], coef);
        }
        nxt[idx] = (1.0 - w) * cur[idx] + w * (src[idx] - x) / coef;
    }
}

// per-block max over points of |sum_w weight_w c_{w,0}|
__global__ void __launch_bounds__(256)
k_wdiag(WArgs a, double* __restrict__ partial)
{
    double m = 0.0;
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < a.ng;
         idx += (long)gridDim.x * 256) {
        double coef = 0.0;
        for (int v = 0; v < a.nw; v++)
            coef = fma(a.wts[v][idx], a.coef[a.start[v]], coef);
        m = fmax(m, fabs(coef));
    }
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

static unsigned wgrid(long n)
{
    long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace gfd

using namespace gfd;

struct gfd_woperator {
    int nw = 0;
    int n[3] = {0, 0, 0};
    int periodic[3] = {0, 0, 0};
    int real = 1;
    int* d_start = nullptr;
    double* d_coef = nullptr;
    int* d_off = nullptr;
    const double** d_wts = nullptr;
    int R = 0;
    long* d_foff = nullptr;
    std::vector<int> h_off;   // point offsets (o0, o1, o2), host copy
    int gs_a0 = 0, gs_a1 = 0; // Gauss-Seidel wavefront order (0: unset)
    DevBuf scratch;
    DevBuf partial;
    bool axis_ok = false;  // every stencil axis-only: `ax` is set
    WAxis ax;
    long ng() const { return (long)n[0] * n[1] * n[2]; }
    WArgs args(const double* phases) const
    {
        WArgs a;
        a.nw = nw;
        for (int c = 0; c < 3; c++) a.n[c] = n[c];
        a.ng = ng();
        a.start = d_start;
        a.coef = d_coef;
        a.off = d_off;
        a.wts = d_wts;
        a.R = R;
        a.foff = d_foff;
        for (int c = 0; c < 3; c++) {
            a.bc.periodic[c] = periodic[c];
            for (int d = 0; d < 2; d++)
                a.bc.phase[c][d] =
                    phases ? make_double2(phases[4 * c + 2 * d],
                                          phases[4 * c + 2 * d + 1])
                           : make_double2(1.0, 0.0);
        }
        return a;
    }
};

extern "C" void gfd_woperator_free(gfd_woperator* op)
{
    if (!op) return;
    if (op->d_start) (void)hipFree(op->d_start);
    if (op->d_coef) (void)hipFree(op->d_coef);
    if (op->d_off) (void)hipFree(op->d_off);
    if (op->d_wts) (void)hipFree((void*)op->d_wts);
    if (op->d_foff) (void)hipFree(op->d_foff);
    op->scratch.release();
    op->partial.release();
    delete op;
}

extern "C" int gfd_woperator_new(int nweights, const int* ncoefs,
                                 const double* coefs, const long* offsets,
                                 const long size[3], int range,
                                 const long neighbors[6], int real,
                                 const double* const* weights,
                                 gfd_woperator** out)
{
    if (!out || nweights <= 0 || nweights > WMAXW || !ncoefs || !coefs ||
        !offsets || !size || !neighbors || !weights || range < 0) {
        set_error("gfd_woperator_new: bad arguments");
        return EINVAL_;
    }
    *out = nullptr;
    auto* op = new gfd_woperator;
    op->nw = nweights;
    op->real = real ? 1 : 0;
    for (int c = 0; c < 3; c++) {
        if (size[c] <= 0 || size[c] > (1L << 20)) {
            gfd_woperator_free(op);
            set_error("gfd_woperator_new: bad grid size");
            return EINVAL_;
        }
        op->n[c] = (int)size[c];
        const bool lo = neighbors[2 * c] >= 0, hi = neighbors[2 * c + 1] >= 0;
        if (lo != hi) {
            gfd_woperator_free(op);
            set_error("axis %d: mixed periodic / zero sides are not "
                      "supported on a single domain", c);
            return ENOTSUP_;
        }
        op->periodic[c] = lo ? 1 : 0;
    }
    // decode the padded-stride offsets (as gfd_operator_new); the first
    // point of every stencil must be the centre
    const long M1 = size[1] + 2 * range, M2 = size[2] + 2 * range;
    const long shift = range * (M1 * M2 + M2 + 1);
    std::vector<int> start(nweights + 1, 0), off;
    long total = 0;
    for (int w = 0; w < nweights; w++) {
        if (ncoefs[w] <= 0 || ncoefs[w] > MAXP) {
            gfd_woperator_free(op);
            set_error("gfd_woperator_new: bad point count");
            return EINVAL_;
        }
        start[w] = (int)total;
        total += ncoefs[w];
    }
    start[nweights] = (int)total;
    off.resize(3 * total);
    for (long p = 0; p < total; p++) {
        const long t = offsets[p] + shift;
        const long o2 = t % M2 - range;
        const long o1 = (t / M2) % M1 - range;
        const long o0 = t / (M1 * M2) - range;
        if (t < 0 || o0 < -range || o0 > range) {
            gfd_woperator_free(op);
            set_error("gfd_woperator_new: offset %ld outside range %d",
                      offsets[p], range);
            return EINVAL_;
        }
        const long oc[3] = {o0, o1, o2};
        for (int c = 0; c < 3; c++) {
            if (oc[c] > size[c] || -oc[c] > size[c]) {
                gfd_woperator_free(op);
                set_error("Stencil longer than domain");
                return EINVAL_;
            }
            off[3 * p + c] = (int)oc[c];
        }
    }
    for (int w = 0; w < nweights; w++) {
        const int* o = &off[3 * start[w]];
        if (o[0] || o[1] || o[2]) {
            gfd_woperator_free(op);
            set_error("gfd_woperator_new: stencil %d does not start at the "
                      "centre", w);
            return EINVAL_;
        }
    }
    // flat offsets for the interior path
    std::vector<long> fo(total);
    int R = 0;
    for (long p = 0; p < total; p++) {
        const int* o = &off[3 * p];
        fo[p] = ((long)o[0] * size[1] + o[1]) * size[2];
        for (int c = 0; c < 3; c++) R = std::max(R, std::abs(o[c]));
    }
    op->R = R;
    op->h_off = off;
    // per-axis form of every stencil for the weighted x-march kernel
    if (nweights <= WX_NW && R <= MAXR_FAST) {
        WAxis& ax = op->ax;
        ax.nw = nweights;
        ax.R = R;
        bool ok = true;
        for (int w = 0; w < nweights && ok; w++) {
            ax.c0[w] = 0.0;
            for (int a = 0; a < 3; a++) {
                ax.hasw[w][a] = 0;
                for (int k = 0; k < MAXR_FAST; k++)
                    ax.cp[w][a][k] = ax.cm[w][a][k] = 0.0;
            }
            ax.wts[w] = weights[w];
            for (int p = start[w]; p < start[w + 1] && ok; p++) {
                const int* o = &off[3 * p];
                const int nz = (o[0] != 0) + (o[1] != 0) + (o[2] != 0);
                if (nz == 0) {
                    ax.c0[w] += coefs[p];
                } else if (nz == 1) {
                    const int a = o[0] ? 0 : (o[1] ? 1 : 2);
                    const int k = o[a];
                    if (k > 0)
                        ax.cp[w][a][k - 1] += coefs[p];
                    else
                        ax.cm[w][a][-k - 1] += coefs[p];
                    if (coefs[p] != 0.0) ax.has[a] = ax.hasw[w][a] = 1;
                } else {
                    ok = false;
                }
            }
        }
        op->axis_ok = ok;
    }
    if (hipMalloc(&op->d_foff, sizeof(long) * total) != hipSuccess ||
        hipMemcpy(op->d_foff, fo.data(), sizeof(long) * total,
                  hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        gfd_woperator_free(op);
        set_error("gfd_woperator_new: out of device memory");
        return ENOMEM_;
    }
    if (hipMalloc(&op->d_start, sizeof(int) * (nweights + 1)) != hipSuccess ||
        hipMalloc(&op->d_coef, sizeof(double) * total) != hipSuccess ||
        hipMalloc(&op->d_off, sizeof(int) * 3 * total) != hipSuccess ||
        hipMalloc((void**)&op->d_wts, sizeof(double*) * nweights) !=
            hipSuccess) {
        (void)hipGetLastError();
        gfd_woperator_free(op);
        set_error("gfd_woperator_new: out of device memory");
        return ENOMEM_;
    }
    if (hipMemcpy(op->d_start, start.data(), sizeof(int) * (nweights + 1),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(op->d_coef, coefs, sizeof(double) * total,
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(op->d_off, off.data(), sizeof(int) * 3 * total,
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy((void*)op->d_wts, weights, sizeof(double*) * nweights,
                  hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        gfd_woperator_free(op);
        set_error("gfd_woperator_new: copy failed");
        return EHIP;
    }
    *out = op;
    return OK;
}

extern "C" int gfd_woperator_apply(gfd_woperator* op, const void* in,
                                   void* out, long nin, const double* phases,
                                   void* stream)
{
    if (!op || nin < 0 || (nin > 0 && (!in || !out)) || in == out) {
        set_error("gfd_woperator_apply: bad arguments");
        return EINVAL_;
    }
    if (nin == 0) return OK;
    hipStream_t s = (hipStream_t)stream;
    const WArgs a = op->args(op->real ? nullptr : phases);
    if (op->axis_ok) {
        const int st = launch_wxmarch(op->ax, op->real, op->n, a.bc, in, out,
                                      nullptr, nin, 0.0, false, s);
        if (st != ENOTSUP_) return st;
        set_error("");
    }
    const unsigned g = wgrid(op->ng() * nin);
    if (op->real)
        hipLaunchKernelGGL(k_wapply<double>, dim3(g), dim3(256), 0, s, a,
                           (const double*)in, (double*)out, nin);
    else
        hipLaunchKernelGGL(k_wapply<double2>, dim3(g), dim3(256), 0, s, a,
                           (const double2*)in, (double2*)out, nin);
    GFD_LAUNCH_CHECK();
    return OK;
}

extern "C" int gfd_woperator_relax(gfd_woperator* op, int method, double* f,
                                   const double* src, int nrelax, double w,
                                   void* stream)
{
    if (!op || !f || !src || nrelax < 0 || f == src) {
        set_error("gfd_woperator_relax: bad arguments");
        return EINVAL_;
    }
    if (!op->real) {
        set_error("gfd_woperator_relax: real data only");
        return EINVAL_;
    }
    if (method != 1 && method != 2) {
        set_error("gfd_woperator_relax: unknown method %d", method);
        return EINVAL_;
    }
    if (nrelax == 0) return OK;
    hipStream_t s = (hipStream_t)stream;
    const long ng = op->ng();
    GFD_TRY(op->scratch.ensure(sizeof(double) * ng, s));
    double* tmp = (double*)op->scratch.p;
    if (method == 1) {
        // lexicographic Gauss-Seidel (c/bmgs/wrelax.cpp:26-58), tmp holds
        // the start-of-sweep values
        if (op->gs_a0 == 0)
            GFD_TRY(gs_order(op->h_off.data(), (int)(op->h_off.size() / 3),
                             &op->gs_a0, &op->gs_a1));
        GsArgs g{};
        for (int c = 0; c < 3; c++) {
            g.n[c] = op->n[c];
            g.periodic[c] = op->periodic[c];
        }
        g.a0 = op->gs_a0;
        g.a1 = op->gs_a1;
        g.nw = op->nw;
        g.start = op->d_start;
        g.coef = op->d_coef;
        g.off = op->d_off;
        g.wts = op->d_wts;
        return gs_relax(g, f, tmp, src, nrelax, s);
    }
    const WArgs a = op->args(nullptr);
    const unsigned g = wgrid(ng);
    // sweeps alternate f -> tmp -> f; an odd count ends with a copy back
    double* cur = f;
    double* nxt = tmp;
    bool fast = op->axis_ok;
    for (int k = 0; k < nrelax; k++) {
        if (fast) {
            const int st = launch_wxmarch(op->ax, 1, op->n, a.bc, cur, nxt,
                                          src, 1, w, true, s);
            if (st == ENOTSUP_) {
                fast = false;
                set_error("");
            } else {
                GFD_TRY(st);
            }
        }
        if (!fast) {
            hipLaunchKernelGGL(k_wrelax, dim3(g), dim3(256), 0, s, a,
                               (const double*)cur, nxt, src, w);
            GFD_LAUNCH_CHECK();
        }
        double* t = cur;
        cur = nxt;
        nxt = t;
    }
    if (cur != f)
        GFD_CHECK_HIP(hipMemcpyAsync(f, cur, sizeof(double) * ng,
                                     hipMemcpyDeviceToDevice, s));
    return OK;
}

extern "C" int gfd_woperator_get_diagonal_element(gfd_woperator* op,
                                                  double* out, void* stream)
{
    if (!op || !out) {
        set_error("gfd_woperator_get_diagonal_element: bad arguments");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = std::min(wgrid(op->ng()), 1024u);
    GFD_TRY(op->partial.ensure(sizeof(double) * g, s));
    hipLaunchKernelGGL(k_wdiag, dim3(g), dim3(256), 0, s, op->args(nullptr),
                       (double*)op->partial.p);
    GFD_LAUNCH_CHECK();
    std::vector<double> h(g);
    GFD_CHECK_HIP(hipMemcpyAsync(h.data(), op->partial.p, sizeof(double) * g,
                                 hipMemcpyDeviceToHost, s));
    GFD_CHECK_HIP(hipStreamSynchronize(s));
    double d = 0.0;
    for (double v : h) d = std::max(d, v);
    *out = d;
    return OK;
}
