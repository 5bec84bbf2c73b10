// C ABI of libgpaw_fd.so: handles, argument checking and the native
// multigrid drivers (Poisson V-cycle, kinetic preconditioner chain).
#include "../../include/gpaw_fd.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gfd_common.h"
#include "stencil.h"
#include "transfer.h"
#include "vector.h"
#include "bands.h"
#include "handles.h"
#include "coarse.h"
#include "gs.h"
#include "wstencil.h"
#include "precond_fine.h"
#include "batch.h"

namespace gfd {

static thread_local std::string g_err;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

const char* last_error() { return g_err.c_str(); }

// 0 auto, 1 general kernel, 2 tiled x-march kernel, 3 full-row kernel
int g_force_path = 0;

static BC make_bc(const int periodic[3], const double* phases)
{
    BC bc;
    for (int c = 0; c < 3; c++) {
        bc.periodic[c] = periodic[c];
        for (int d = 0; d < 2; d++)
            bc.phase[c][d] = phases ? make_double2(phases[4 * c + 2 * d],
                                                   phases[4 * c + 2 * d + 1])
                                    : make_double2(1.0, 0.0);
    }
    return bc;
}

static int periodic_from_neighbors(const long nb[6], int periodic[3])
{
    for (int c = 0; c < 3; c++) {
        const bool lo = nb[2 * c] >= 0, hi = nb[2 * c + 1] >= 0;
        if (lo != hi) {
            set_error("axis %d: mixed periodic / zero sides are not supported "
                      "on a single domain",
                      c);
            return ENOTSUP_;
        }
        periodic[c] = lo ? 1 : 0;
    }
    return OK;
}

}  // namespace gfd

using namespace gfd;

// ====================================================================== ops

extern "C" const char* gfd_last_error(void) { return gfd::last_error(); }

extern "C" int gfd_abi_version(void) { return 1; }

extern "C" int gfd_operator_new(int ncoefs, const double* coefs,
                                const long* offsets, const long size[3],
                                int range, const long neighbors[6], int real,
                                int cfd, gfd_operator** out)
{
    (void)cfd;
    if (!out || ncoefs <= 0 || ncoefs > MAXP || !coefs || !offsets || !size ||
        !neighbors || range < 0) {
        set_error("gfd_operator_new: bad arguments");
        return EINVAL_;
    }
    *out = nullptr;
    gfd_operator* op = new gfd_operator();
    op->P = ncoefs;
    op->range = range;
    op->real = real ? 1 : 0;
    for (int c = 0; c < 3; c++) {
        if (size[c] <= 0 || size[c] > (1L << 20)) {
            delete op;
            set_error("gfd_operator_new: bad grid size");
            return EINVAL_;
        }
        op->n[c] = (int)size[c];
    }
    int st = periodic_from_neighbors(neighbors, op->periodic);
    if (st != OK) {
        delete op;
        return st;
    }
    // decode the padded-stride offsets (gpaw/fd_operators.py:61-63)
    const long M1 = size[1] + 2 * range, M2 = size[2] + 2 * range;
    const long shift = range * (M1 * M2 + M2 + 1);
    op->coef.assign(coefs, coefs + ncoefs);
    op->off.resize(3 * ncoefs);
    for (int p = 0; p < ncoefs; p++) {
        const long t = offsets[p] + shift;
        if (t < 0) {
            delete op;
            set_error("gfd_operator_new: offset %ld outside range %d",
                      offsets[p], range);
            return EINVAL_;
        }
        const long o2 = t % M2 - range;
        const long o1 = (t / M2) % M1 - range;
        const long o0 = t / (M1 * M2) - range;
        if (o0 < -range || o0 > range) {
            delete op;
            set_error("gfd_operator_new: offset %ld outside range %d",
                      offsets[p], range);
            return EINVAL_;
        }
        const long oc[3] = {o0, o1, o2};
        for (int c = 0; c < 3; c++)
            if (oc[c] > size[c] || -oc[c] > size[c]) {
                delete op;
                set_error("Stencil longer than domain");
                return EINVAL_;
            }
        op->off[3 * p] = (int)o0;
        op->off[3 * p + 1] = (int)o1;
        op->off[3 * p + 2] = (int)o2;
        if (o0 == 0 && o1 == 0 && o2 == 0) {
            op->diag = coefs[p];
            if (op->centre < 0) op->centre = p;
        }
    }
    // per-axis form for the x-march kernel
    AxisStencil& ax = op->ax;
    std::memset(&ax, 0, sizeof ax);
    bool ok = true;
    int R = 0, ncentre = 0;
    std::vector<int> seen(3 * (2 * MAXR_FAST + 1), 0);
    for (int p = 0; p < ncoefs && ok; p++) {
        const int* o = &op->off[3 * p];
        const int nz = (o[0] != 0) + (o[1] != 0) + (o[2] != 0);
        if (nz == 0) {
            ax.c0 = coefs[p];
            ncentre++;
            continue;
        }
        if (nz > 1) {
            ok = false;
            break;
        }
        const int a = o[0] ? 0 : (o[1] ? 1 : 2);
        const int k = o[a];
        if (k > MAXR_FAST || -k > MAXR_FAST) {
            ok = false;
            break;
        }
        int& sflag = seen[a * (2 * MAXR_FAST + 1) + k + MAXR_FAST];
        if (sflag) {
            ok = false;  // duplicate offset: keep separate terms (general)
            break;
        }
        sflag = 1;
        if (k > 0) ax.cp[a][k - 1] = coefs[p];
        else ax.cm[a][-k - 1] = coefs[p];
        ax.has[a] = 1;
        R = std::max(R, std::abs(k));
    }
    if (ncentre > 1 || (ncentre == 1 && op->centre != 0 && false)) ok = false;
    if (R == 0) ok = false;
    for (int c = 0; c < 3; c++)
        if (op->n[c] < R) ok = false;
    ax.R = R;
    op->axis_ok = ok;
    // device copies for the general kernel
    if (hipMalloc((void**)&op->d_coef, sizeof(double) * ncoefs) != hipSuccess ||
        hipMalloc((void**)&op->d_off, sizeof(int) * 3 * ncoefs) != hipSuccess) {
        (void)hipGetLastError();
        if (op->d_coef) (void)hipFree(op->d_coef);
        delete op;
        set_error("gfd_operator_new: device allocation failed (no GPU?)");
        return ENOMEM_;
    }
    if (hipMemcpy(op->d_coef, op->coef.data(), sizeof(double) * ncoefs,
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(op->d_off, op->off.data(), sizeof(int) * 3 * ncoefs,
                  hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(op->d_coef);
        (void)hipFree(op->d_off);
        delete op;
        set_error("gfd_operator_new: copy of coefficients failed");
        return EHIP;
    }
    *out = op;
    return OK;
}

extern "C" void gfd_operator_free(gfd_operator* op)
{
    if (!op) return;
    if (op->d_coef) (void)hipFree(op->d_coef);
    if (op->d_off) (void)hipFree(op->d_off);
    op->scratch.release();
    op->partial.release();
    op->scalar.release();
    delete op;
}

extern "C" double gfd_operator_get_diagonal_element(const gfd_operator* op)
{
    return op ? op->diag : 0.0;
}

static bool aligned16(const void* p)
{
    return ((uintptr_t)p & 15) == 0;
}

// Kernel family for one pass: 3 full-row, 2 tiled x-march, 1 general.
static int op_path(gfd_operator* op, int dtype, long nin, const void* in,
                   const void* out, const void* aux, int epi)
{
    if (!op->use_fast()) return 1;
    const bool row = (g_force_path == 0 || g_force_path == 3) &&
                     aligned16(in) && aligned16(out) &&
                     (aux == nullptr || aligned16(aux)) &&
                     row_stencil_fits(op->ax, dtype, op->n, nin);
    if (row) return 3;
    // epilogues the tiled x-march kernel instantiates per data type
    const bool xm_has =
        (dtype == 0 && epi <= EPI_UPDATE) ||
        (dtype == 1 && (epi == EPI_PLAIN || epi == EPI_RESID ||
                        epi == EPI_UPDATE)) ||
        (dtype == 2 && epi <= EPI_RESID);
    if (xm_has && (g_force_path == 0 || g_force_path == 2) &&
        axis_stencil_fits(op->ax, dtype, op->n, nin))
        return 2;
    return 1;
}

// One stencil pass with an epilogue (the building block of every driver).
static int op_pass(gfd_operator* op, int dtype, const void* in, void* out,
                   long nin, const double* phases, int epi, EpiArgs ea,
                   hipStream_t s, int* nblocks)
{
    const BC bc = make_bc(op->periodic, phases);
    const int path = op_path(op, dtype, nin, in, out, ea.aux, epi);
    if (path == 3)
        return launch_row_stencil(op->ax, dtype, op->n, bc, in, out, nin, epi,
                                  ea, s, nblocks);
    if (path == 2)
        return launch_axis_stencil(op->ax, dtype, op->n, bc, in, out, nin, epi,
                                   ea, s, nblocks);
    return launch_general_stencil(op->gen(ea.c0), dtype, op->n, bc, in, out,
                                  nin, epi, ea, s, nblocks);
}

// Upper bound of the blocks a residual pass may launch (partial sums).
static long op_max_blocks(gfd_operator* op, int dtype, long nin)
{
    long m = general_stencil_max_blocks(op->n, nin);
    if (op->use_fast()) {
        m = std::max(m, axis_stencil_max_blocks(op->ax, dtype, op->n, nin));
        m = std::max(m, row_stencil_max_blocks(op->ax, dtype, op->n, nin));
    }
    return m;
}

static double centre_coef(const gfd_operator* op)
{
    return op->centre >= 0 ? op->coef[op->centre] : 0.0;
}

extern "C" int gfd_operator_apply(gfd_operator* op, const void* in, void* out,
                                  long nin, const double* phases, void* stream)
{
    if (!op || nin < 0 || (nin > 0 && (!in || !out))) {
        set_error("gfd_operator_apply: bad arguments");
        return EINVAL_;
    }
    if (nin == 0) return OK;
    hipStream_t s = (hipStream_t)stream;
    const size_t es = op->real ? 8 : 16;
    const void* src = in;
    if (in == out) {  // the reference copies into a padded buffer first
        GFD_TRY(op->scratch.ensure(es * op->ng() * nin, s));
        GFD_CHECK_HIP(hipMemcpyAsync(op->scratch.p, in, es * op->ng() * nin,
                                     hipMemcpyDeviceToDevice, s));
        src = op->scratch.p;
    }
    EpiArgs ea{};
    ea.c0 = centre_coef(op);
    return op_pass(op, op->dtype(), src, out, nin, op->real ? nullptr : phases,
                   EPI_PLAIN, ea, s, nullptr);
}

// nrelax Jacobi sweeps ping-ponging between *cur and *alt; on return *cur
// holds the result (the pointers are swapped once per sweep).
static int relax_pp(gfd_operator* op, double** cur, double** alt,
                    const double* src, int nrelax, double w, hipStream_t s)
{
    EpiArgs ea{};
    ea.aux = src;
    ea.c0 = op->coef[0];
    ea.w = w;
    for (int k = 0; k < nrelax; k++) {
        GFD_TRY(op_pass(op, 0, *cur, *alt, 1, nullptr, EPI_RELAX, ea, s,
                        nullptr));
        std::swap(*cur, *alt);
    }
    return OK;
}

static int check_relax(const gfd_operator* op, int method)
{
    if (!op->real) {
        set_error("relax: real data only");
        return EINVAL_;
    }
    if (method != 1 && method != 2) {
        set_error("relax: unknown method %d", method);
        return EINVAL_;
    }
    if (op->centre != 0) {
        set_error("relax: the first stencil point must be the centre");
        return EINVAL_;
    }
    return OK;
}

// nsweeps Gauss-Seidel sweeps on f in place (c/bmgs/relax.cpp:12-40);
// snap: a work grid of the operator's size.
static int op_gauss_seidel(gfd_operator* op, double* f, double* snap,
                           const double* src, int nsweeps, hipStream_t s)
{
    if (op->gs_a0 == 0)
        GFD_TRY(gs_order(op->off.data(), op->P, &op->gs_a0, &op->gs_a1));
    GsArgs a{};
    for (int c = 0; c < 3; c++) {
        a.n[c] = op->n[c];
        a.periodic[c] = op->periodic[c];
    }
    a.a0 = op->gs_a0;
    a.a1 = op->gs_a1;
    a.nw = 0;
    a.P = op->P;
    a.coef = op->d_coef;
    a.off = op->d_off;
    a.inv_c0 = 1.0 / op->coef[0];
    return gs_relax(a, f, snap, src, nsweeps, s);
}

extern "C" int gfd_operator_relax(gfd_operator* op, int method, double* f,
                                  const double* src, int nrelax, double w,
                                  void* stream)
{
    if (!op || !f || !src || nrelax < 0) {
        set_error("gfd_operator_relax: bad arguments");
        return EINVAL_;
    }
    GFD_TRY(check_relax(op, method));
    if (nrelax == 0) return OK;
    hipStream_t s = (hipStream_t)stream;
    const long ng = op->ng();
    if (method == 1) {
        GFD_TRY(op->scratch.ensure(sizeof(double) * ng, s));
        return op_gauss_seidel(op, f, (double*)op->scratch.p, src, nrelax, s);
    }
    GFD_TRY(op->scratch.ensure(2 * sizeof(double) * ng, s));
    double* t1 = (double*)op->scratch.p;
    double* t2 = t1 + ng;
    // f -> t1 -> t2 -> ... -> f: the last sweep writes f directly
    double* cur = f;
    EpiArgs ea{};
    ea.aux = src;
    ea.c0 = op->coef[0];
    ea.w = w;
    for (int k = 0; k < nrelax; k++) {
        double* dst;
        if (k == nrelax - 1 && cur != f) dst = f;
        else dst = (cur == t1) ? t2 : t1;
        GFD_TRY(op_pass(op, 0, cur, dst, 1, nullptr, EPI_RELAX, ea, s, nullptr));
        cur = dst;
    }
    if (cur != f)
        GFD_CHECK_HIP(hipMemcpyAsync(f, cur, sizeof(double) * ng,
                                     hipMemcpyDeviceToDevice, s));
    return OK;
}

// partial: caller-owned partial-sum buffer of at least
// op_max_blocks(op, dtype, nin) + 1 doubles, or NULL for the operator's own
// (grown here; a caller that captures the pass into a graph passes its own,
// sized beforehand, so no later call can reallocate it under the graph).
static int residual_norm(gfd_operator* op, int dtype, const void* in,
                         const void* b, void* r, long nin,
                         const double* phases, double* dev_norm,
                         hipStream_t s, bool narrow = false,
                         double* partial = nullptr)
{
    EpiArgs ea{};
    ea.aux = b;
    ea.c0 = centre_coef(op);
    if (dev_norm) {
        if (!partial) {
            const long nbmax = op_max_blocks(op, dtype, nin);
            GFD_TRY(op->partial.ensure(sizeof(double) * (nbmax + 1), s));
            partial = (double*)op->partial.p;
        }
        ea.partial = partial;
    }
    int nb = 0;
    GFD_TRY(op_pass(op, dtype, in, r, nin, phases,
                    narrow ? EPI_RESN : EPI_RESID, ea, s, &nb));
    if (dev_norm) GFD_TRY(launch_sum_partials(ea.partial, nb, dev_norm, s));
    return OK;
}

extern "C" int gfd_operator_residual(gfd_operator* op, const void* in,
                                     const void* b, void* r, long nin,
                                     const double* phases, double* norm2,
                                     void* stream)
{
    if (!op || nin < 0 || (nin > 0 && (!in || !b || !r)) || in == r) {
        set_error("gfd_operator_residual: bad arguments");
        return EINVAL_;
    }
    if (nin == 0) {
        if (norm2) *norm2 = 0.0;
        return OK;
    }
    hipStream_t s = (hipStream_t)stream;
    double* dn = nullptr;
    if (norm2) {
        GFD_TRY(op->scalar.ensure(sizeof(double), s));
        dn = (double*)op->scalar.p;
    }
    GFD_TRY(residual_norm(op, op->dtype(), in, b, r, nin,
                          op->real ? nullptr : phases, dn, s));
    if (norm2) {
        GFD_CHECK_HIP(hipMemcpyAsync(norm2, dn, sizeof(double),
                                     hipMemcpyDeviceToHost, s));
        GFD_CHECK_HIP(hipStreamSynchronize(s));
    }
    return OK;
}

// ------------------------------------------- residual without allocation
//
// gfd_operator_residual's norm needs a partial-sum buffer (one double per
// launched block) and a result slot; the operator grows its own with
// hipMalloc.  The variant below takes them from one fixed device array
// instead, so it never allocates: layout [0] the norm, [NA_PARTIAL ..] the
// block partials.  Every launched block writes its own slot and the final
// sum writes [0], so nothing has to be cleared between calls.  All calls on
// one GPU share that GPU's array: calls that overlap in time race.
constexpr long NA_DOUBLES = 1L << 17;  // 1 MiB per GPU
constexpr long NA_PARTIAL = 32;        // partials start 256 bytes in
static __device__ __align__(256) double scratch[NA_DOUBLES];

constexpr int NA_MAXDEV = 64;
static std::mutex g_na_mu;
static double* g_na_ptr[NA_MAXDEV];  // device address of scratch, per GPU

// The current device's address of `scratch`, looked up once per device.
static int noalloc_scratch(double** out)
{
    int dev = -1;
    GFD_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAXDEV) {
        set_error("residual_noalloc: device %d beyond the %d supported", dev,
                  NA_MAXDEV);
        return EINVAL_;
    }
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        GFD_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_na_ptr[dev] = (double*)p;
    }
    *out = g_na_ptr[dev];
    return OK;
}

extern "C" long gfd_operator_residual_noalloc_max_blocks(void)
{
    return NA_DOUBLES - NA_PARTIAL;
}

extern "C" int gfd_operator_residual_noalloc(gfd_operator* op,
                                             const void* in, const void* b,
                                             void* r, long nin,
                                             const double* phases,
                                             double* norm2, void* stream)
{
    if (!op || nin < 0 || (nin > 0 && (!in || !b || !r)) || in == r) {
        set_error("gfd_operator_residual_noalloc: bad arguments");
        return EINVAL_;
    }
    if (nin == 0) {
        if (norm2) *norm2 = 0.0;
        return OK;
    }
    const int dtype = op->dtype();
    const long nbmax = op_max_blocks(op, dtype, nin);
    if (norm2 && nbmax > NA_DOUBLES - NA_PARTIAL) {
        set_error("gfd_operator_residual_noalloc: %ld grids of %d x %d x %d "
                  "need %ld partial sums, above the fixed %ld; split the "
                  "batch", nin, op->n[0], op->n[1], op->n[2], nbmax,
                  NA_DOUBLES - NA_PARTIAL);
        return EINVAL_;
    }
    double* sp = nullptr;
    if (norm2) GFD_TRY(noalloc_scratch(&sp));
    hipStream_t s = (hipStream_t)stream;
    xr_no_tune_here(true);
    const int st =
        residual_norm(op, dtype, in, b, r, nin, op->real ? nullptr : phases,
                      sp, s, false, sp ? sp + NA_PARTIAL : nullptr);
    xr_no_tune_here(false);
    GFD_TRY(st);
    if (norm2) {
        GFD_CHECK_HIP(hipMemcpyAsync(norm2, sp, sizeof(double),
                                     hipMemcpyDeviceToHost, s));
        GFD_CHECK_HIP(hipStreamSynchronize(s));
    }
    return OK;
}

// ================================================================ transfers

extern "C" int gfd_transformer_new(const long size_in[3],
                                   const long size_out[3], int k,
                                   const long pad[6], const long npad[6],
                                   const long skip[6],
                                   const long neighbors[6], int real,
                                   int interpolate, gfd_transformer** out)
{
    if (!out || !size_in || !size_out || !pad || !npad || !skip ||
        !neighbors || (k != 2 && k != 4 && k != 6 && k != 8)) {
        set_error("gfd_transformer_new: bad arguments (k must be 2,4,6,8)");
        return EINVAL_;
    }
    *out = nullptr;
    gfd_transformer* t = new gfd_transformer();
    t->real = real ? 1 : 0;
    int st = periodic_from_neighbors(neighbors, t->periodic);
    if (st != OK) {
        delete t;
        return st;
    }
    XferPlan& p = t->plan;
    p.interpolate = interpolate ? 1 : 0;
    p.K = k;
    for (int c = 0; c < 3; c++) {
        p.nin[c] = (int)size_in[c];
        p.nout[c] = (int)size_out[c];
        const long p0 = pad[2 * c], p1 = pad[2 * c + 1];
        if (p0 < 0 || p1 < 0 || size_in[c] <= 0 || size_out[c] <= 0) {
            delete t;
            set_error("gfd_transformer_new: negative padding or empty grid");
            return EINVAL_;
        }
        // a periodic halo is a pure wrap only when each side receives what
        // the other side sends (c/bc.cpp:66-78)
        if (t->periodic[c] && (npad[2 * c] != p1 || npad[2 * c + 1] != p0)) {
            delete t;
            set_error("gfd_transformer_new: axis %d paddings are not a "
                      "single-domain wrap",
                      c);
            return ENOTSUP_;
        }
        const long M = size_in[c] + p0 + p1;
        XferAxis& a = p.ax[c];
        a.nin = (int)size_in[c];
        a.nout = (int)size_out[c];
        if (!interpolate) {
            const long nout = (M - (2 * k - 3)) / 2;
            if (nout != size_out[c]) {
                delete t;
                set_error("gfd_transformer_new: axis %d restricts %ld -> %ld, "
                          "not %ld",
                          c, size_in[c], nout, size_out[c]);
                return EINVAL_;
            }
            a.off = (int)((k - 1) - p0);
            a.skip0 = 0;
        } else {
            const long s0 = skip[2 * c], s1 = skip[2 * c + 1];
            const long nloop = M - (k - 1) + s1;
            const long nout = 2 * nloop - s0 - s1;
            if (nout != size_out[c]) {
                delete t;
                set_error("gfd_transformer_new: axis %d interpolates %ld -> "
                          "%ld, not %ld",
                          c, size_in[c], nout, size_out[c]);
                return EINVAL_;
            }
            a.off = (int)((k / 2 - 1) - p0);
            a.skip0 = (int)s0;
        }
    }
    *out = t;
    return OK;
}

extern "C" void gfd_transformer_free(gfd_transformer* t)
{
    if (!t) return;
    t->scratch.release();
    delete t;
}

static size_t xfer_scratch_bytes(const gfd_transformer* t, int dtype,
                                 long nin)
{
    const size_t es = dtype == 1 ? 16 : (dtype == 2 ? 4 : 8);  // 3: c64
    return t->plan.K == 2 && !xfer_passes_forced()
               ? 0
               : es * xfer_scratch_elems(t->plan, nin) + 16;
}

// Time the transfer kernel families for this shape once (not under graph
// capture), into the transformer's output-sized tuning buffer.
static int xfer_maybe_tune(gfd_transformer* t, int dtype, const BC& bc,
                           const void* in, long nin, void* work, int in_wide,
                           hipStream_t s)
{
    if (!xfer_needs_tuning(t->plan, dtype, nin, in_wide)) return OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return OK;
    }
    if (cs != hipStreamCaptureStatusNone) return OK;
    const size_t es = dtype == 1 ? 16 : (dtype == 2 ? 4 : 8);
    GFD_TRY(t->tune_out.ensure(es * t->ngout() * nin, s));
    return xfer_tune(t->plan, dtype, bc, in, t->tune_out.p, nin, work,
                     in_wide, s);
}

// work: caller-owned work space of xfer_scratch_bytes() (graph-captured
// drivers), or NULL for the transformer's own.
static int xfer_dt(gfd_transformer* t, int dtype, const void* in, void* out,
                   long nin, const double* phases, const XEpiArgs& ea,
                   hipStream_t s, void* work = nullptr)
{
    if (!work && (t->plan.K != 2 || xfer_passes_forced())) {
        GFD_TRY(t->scratch.ensure(xfer_scratch_bytes(t, dtype, nin), s));
        work = t->scratch.p;
    }
    const BC bc = make_bc(t->periodic, t->real ? nullptr : phases);
    GFD_TRY(xfer_maybe_tune(t, dtype, bc, in, nin, work, ea.in_wide, s));
    return launch_transfer(t->plan, dtype, bc, in, out, nin, work, ea, s);
}

static int xfer(gfd_transformer* t, const void* in, void* out, long nin,
                const double* phases, const XEpiArgs& ea, hipStream_t s)
{
    return xfer_dt(t, t->dtype(), in, out, nin, phases, ea, s);
}

extern "C" int gfd_transformer_apply(gfd_transformer* t, const void* in,
                                     void* out, long nin,
                                     const double* phases, void* stream)
{
    if (!t || nin < 0 || (nin > 0 && (!in || !out))) {
        set_error("gfd_transformer_apply: bad arguments");
        return EINVAL_;
    }
    if (nin == 0) return OK;
    XEpiArgs ea;
    return xfer(t, in, out, nin, phases, ea, (hipStream_t)stream);
}

// ============================================================ preconditioner

static bool same3(const int a[3], const int b[3])
{
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// fp32: the work arrays q0, r1, .., q2 are FP32 (f32 / c64) and every
// pass computes in FP32; residuals and d0 stay FP64 (the first
// restriction narrows, the last stencil pass widens).
static int fd_precond(gfd_transformer* rest0, gfd_transformer* rest1,
                      gfd_transformer* intp1, gfd_transformer* intp2,
                      gfd_operator* kin0, gfd_operator* kin1,
                      gfd_operator* kin2, void* d0, void* q0, void* r1,
                      void* d1, void* q1, void* r2, void* d2, void* q2,
                      const void* residuals, long nin, double step,
                      const double* phases, void* stream, bool fp32)
{
    if (!rest0 || !rest1 || !intp1 || !intp2 || !kin0 || !kin1 || !kin2 ||
        nin < 0) {
        set_error("gfd_fd_precond: bad arguments");
        return EINVAL_;
    }
    if (nin == 0) return OK;
    if (!d0 || !q0 || !r1 || !d1 || !q1 || !r2 || !d2 || !q2 || !residuals) {
        set_error("gfd_fd_precond: null work array");
        return EINVAL_;
    }
    const int real = kin0->real;
    if (kin1->real != real || kin2->real != real || rest0->real != real ||
        rest1->real != real || intp1->real != real || intp2->real != real) {
        set_error("gfd_fd_precond: mixed real/complex objects");
        return EINVAL_;
    }
    if (rest0->plan.interpolate || rest1->plan.interpolate ||
        !intp1->plan.interpolate || !intp2->plan.interpolate ||
        !same3(rest0->plan.nin, kin0->n) || !same3(rest0->plan.nout, kin1->n) ||
        !same3(rest1->plan.nin, kin1->n) || !same3(rest1->plan.nout, kin2->n) ||
        !same3(intp2->plan.nin, kin2->n) || !same3(intp2->plan.nout, kin1->n) ||
        !same3(intp1->plan.nin, kin1->n) || !same3(intp1->plan.nout, kin0->n)) {
        set_error("Input grid shape does not match with arrays d0 and q0.");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    const double* ph = real ? nullptr : phases;
    const int dt = fp32 ? (real ? 2 : 3) : (real ? 0 : 1);
    // r1 = R0(-res); d1 = 4 step r1
    {
        XEpiArgs e;
        e.epi = XEPI_SCALE2;
        e.negate = 1;
        e.alpha = 4 * step;
        e.out2 = d1;
        e.in_wide = fp32 ? 1 : 0;
        GFD_TRY(xfer_dt(rest0, dt, residuals, r1, nin, ph, e, s));
    }
    // q1 = K1 d1 - r1
    {
        EpiArgs e{};
        e.aux = r1;
        e.c0 = centre_coef(kin1);
        GFD_TRY(op_pass(kin1, dt, d1, q1, nin, ph, EPI_RESID, e, s, nullptr));
    }
    // r2 = R1(q1); d2 = 16 step r2
    {
        XEpiArgs e;
        e.epi = XEPI_SCALE2;
        e.alpha = 16 * step;
        e.out2 = d2;
        GFD_TRY(xfer_dt(rest1, dt, q1, r2, nin, ph, e, s));
    }
    // q2 = K2 d2 - r2; d2 -= 16 step q2        (new d2 held in q2)
    {
        EpiArgs e{};
        e.aux = r2;
        e.c0 = centre_coef(kin2);
        e.s = 16 * step;
        GFD_TRY(op_pass(kin2, dt, d2, q2, nin, ph, EPI_UPDATE, e, s, nullptr));
    }
    // q1 = I2(d2); d1 -= q1
    {
        XEpiArgs e;
        e.epi = XEPI_SUBFROM;
        e.base = d1;
        GFD_TRY(xfer_dt(intp2, dt, q2, d1, nin, ph, e, s));
    }
    // q1 = K1 d1 - r1; d1 -= 4 step q1          (new d1 held in q1)
    {
        EpiArgs e{};
        e.aux = r1;
        e.c0 = centre_coef(kin1);
        e.s = 4 * step;
        GFD_TRY(op_pass(kin1, dt, d1, q1, nin, ph, EPI_UPDATE, e, s, nullptr));
    }
    // both at once where the fine shape fits the fused kernel: d0' is then
    // never stored
    if (kin0->use_fast()) {
        const int st = launch_precond_fine(
            kin0->ax, centre_coef(kin0), kin0->n, kin0->periodic,
            intp1->plan, intp1->periodic, ph, q1, residuals, d0, nin, step,
            real, fp32 ? 0 : 1, s);
        if (st != ENOTSUP_) return st;
    }
    // d0' = I1(-d1)                              (held in q0)
    {
        XEpiArgs e;
        e.epi = XEPI_NEG;
        GFD_TRY(xfer_dt(intp1, dt, q1, q0, nin, ph, e, s));
    }
    // q0 = K0 d0' - res; d0 = -(d0' - step q0)
    {
        EpiArgs e{};
        e.aux = residuals;
        e.c0 = centre_coef(kin0);
        e.s = step;
        e.negate = 1;
        GFD_TRY(op_pass(kin0, dt, q0, d0, nin, ph,
                        fp32 ? EPI_UPDW : EPI_UPDATE, e, s, nullptr));
    }
    return OK;
}

extern "C" int gfd_fd_precond(gfd_transformer* rest0, gfd_transformer* rest1,
                              gfd_transformer* intp1, gfd_transformer* intp2,
                              gfd_operator* kin0, gfd_operator* kin1,
                              gfd_operator* kin2, void* d0, void* q0, void* r1,
                              void* d1, void* q1, void* r2, void* d2, void* q2,
                              const void* residuals, long nin, double step,
                              const double* phases, void* stream)
{
    return fd_precond(rest0, rest1, intp1, intp2, kin0, kin1, kin2, d0, q0,
                      r1, d1, q1, r2, d2, q2, residuals, nin, step, phases,
                      stream, false);
}

extern "C" int gfd_fd_precond32(gfd_transformer* rest0,
                                gfd_transformer* rest1,
                                gfd_transformer* intp1,
                                gfd_transformer* intp2, gfd_operator* kin0,
                                gfd_operator* kin1, gfd_operator* kin2,
                                void* d0, void* q0, void* r1, void* d1,
                                void* q1, void* r2, void* d2, void* q2,
                                const void* residuals, long nin, double step,
                                const double* phases, void* stream)
{
    if (rest0 && rest1 && intp1 && intp2 &&
        (rest0->plan.K != 2 || rest1->plan.K != 2 || intp1->plan.K != 2 ||
         intp2->plan.K != 2)) {
        set_error("gfd_fd_precond32: nn=1 transformers only");
        return EINVAL_;
    }
    return fd_precond(rest0, rest1, intp1, intp2, kin0, kin1, kin2, d0, q0,
                      r1, d1, q1, r2, d2, q2, residuals, nin, step, phases,
                      stream, true);
}

// ================================================================= Poisson

// Level buffers of one precision (dtype 0: f64, 2: f32).
struct Levels {
    int dtype = 0;
    int L = -1;  // coarsest level index held
    DevBuf pool;
    std::vector<void*> phi, tmp, rho, res;
};

// Coarse levels of a schedule run by the LDS-resident workgroup.
struct CoarseSolve {
    bool valid = false;
    int lc = -1;      // first level in LDS (-1: none)
    size_t lds = 0;
    CGrid host{};
    DevBuf dev;
};

// LDS-resident coarse solve on (default) / off, and the largest first
// level (points) it takes
static bool g_coarse_lds = true;
static long g_coarse_max_pts = 4096;

struct gfd_poisson {
    int L = 0;  // number of coarsenings
    int method = 2;  // FP64 smoother: 2 weighted Jacobi, 1 Gauss-Seidel
    std::vector<gfd_operator*> ops;
    std::vector<gfd_transformer*> rst, itp;
    std::vector<int> pre, post;
    std::vector<double> w;
    double dv = 1.0;
    Levels lv;     // f64 levels (phi[0], rho[0]: owned copies for graphs)
    Levels lv32;   // f32 levels of the mixed-precision inner cycles
    Levels lv64r;  // f64 correction levels: refinement with FP64 inner
                   // cycles on the mixed solve's schedule
    DevBuf scalar; // norms
    // partial sums of the norms and work space of K > 2 transfers: owned
    // here (not borrowed from the level operators / transformers, which
    // other callers may regrow) because captured graphs hold their
    // addresses; sized once in gfd_poisson_new
    DevBuf partial, xscratch;
    // graph replay of one V-cycle (+ residual norm) on an owned stream
    bool use_graph = true;
    hipStream_t gs = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t g64 = nullptr;
    hipGraphExec_t g32 = nullptr;
    int g32_key[6] = {-1, -1, -1, -1, -1, -1};
    double g32_w[2] = {0, 0};
    hipGraphExec_t g64r = nullptr;
    int g64r_key[6] = {-1, -1, -1, -1, -1, -1};
    double g64r_w[2] = {0, 0};
    CoarseSolve cs64, cs32, cs64r;
};

static int alloc_levels(gfd_poisson* p, Levels& lv, int dtype, int L)
{
    if (lv.L >= L && lv.dtype == dtype) return OK;
    const size_t es = dtype == 2 ? 4 : 8;
    long total = 0;
    for (int l = 0; l <= L; l++) total += 4 * ((p->ops[l]->ng() + 3) & ~3L);
    lv.pool.release();
    GFD_TRY(lv.pool.ensure(es * total, nullptr));
    lv.phi.clear();
    lv.tmp.clear();
    lv.rho.clear();
    lv.res.clear();
    char* q = (char*)lv.pool.p;
    for (int l = 0; l <= L; l++) {
        const size_t nb = es * ((p->ops[l]->ng() + 3) & ~3L);
        lv.phi.push_back(q);
        lv.tmp.push_back(q + nb);
        lv.rho.push_back(q + 2 * nb);
        lv.res.push_back(q + 3 * nb);
        q += 4 * nb;
    }
    lv.dtype = dtype;
    lv.L = L;
    return OK;
}

extern "C" int gfd_poisson_new(int nlevels, gfd_operator* const* ops,
                               gfd_transformer* const* restrictors,
                               gfd_transformer* const* interpolators,
                               const int* presmooths, const int* postsmooths,
                               const double* weights, int relax_method,
                               double dv, gfd_poisson** out)
{
    if (!out || nlevels < 0 || !ops || (nlevels > 0 && (!restrictors ||
                                                        !interpolators)) ||
        !presmooths || !postsmooths || !weights) {
        set_error("gfd_poisson_new: bad arguments");
        return EINVAL_;
    }
    *out = nullptr;
    for (int l = 0; l <= nlevels; l++) {
        if (!ops[l]) {
            set_error("gfd_poisson_new: null operator");
            return EINVAL_;
        }
        int st = check_relax(ops[l], relax_method);
        if (st != OK) return st;
    }
    for (int l = 0; l < nlevels; l++) {
        gfd_transformer* r = restrictors[l];
        gfd_transformer* i = interpolators[l];
        if (!r || !i || r->plan.interpolate || !i->plan.interpolate ||
            !r->real || !i->real || !same3(r->plan.nin, ops[l]->n) ||
            !same3(r->plan.nout, ops[l + 1]->n) ||
            !same3(i->plan.nin, ops[l + 1]->n) ||
            !same3(i->plan.nout, ops[l]->n)) {
            set_error("gfd_poisson_new: level %d transformers do not match", l);
            return EINVAL_;
        }
    }
    gfd_poisson* p = new gfd_poisson();
    p->L = nlevels;
    p->ops.assign(ops, ops + nlevels + 1);
    if (nlevels > 0) {
        p->rst.assign(restrictors, restrictors + nlevels);
        p->itp.assign(interpolators, interpolators + nlevels);
    }
    p->pre.assign(presmooths, presmooths + nlevels + 1);
    p->post.assign(postsmooths, postsmooths + nlevels + 1);
    p->w.assign(weights, weights + nlevels + 1);
    p->dv = dv;
    p->method = relax_method;
    int st = alloc_levels(p, p->lv, 0, nlevels);
    // every norm of the drivers is over level 0 (one grid); transfers of
    // either precision use at most the FP64 scratch size
    const long ng0 = p->ops[0]->ng();
    const long np = std::max(op_max_blocks(p->ops[0], 0, 1),
                             vec_dot_partials(ng0)) + 1;
    if (st == OK) st = p->partial.ensure(sizeof(double) * np, nullptr);
    size_t xb = 16;
    for (int l = 0; l < nlevels; l++)
        xb = std::max({xb, xfer_scratch_bytes(p->rst[l], 0, 1),
                       xfer_scratch_bytes(p->itp[l], 0, 1)});
    if (st == OK) st = p->xscratch.ensure(xb, nullptr);
    if (st == OK) st = p->scalar.ensure(4 * sizeof(double), nullptr);
    if (st != OK) {
        gfd_poisson_free(p);
        return st;
    }
    *out = p;
    return OK;
}

static void drop_graphs(gfd_poisson* p)
{
    if (p->g64) (void)hipGraphExecDestroy(p->g64);
    if (p->g32) (void)hipGraphExecDestroy(p->g32);
    if (p->g64r) (void)hipGraphExecDestroy(p->g64r);
    p->g64 = nullptr;
    p->g32 = nullptr;
    p->g64r = nullptr;
}

extern "C" void gfd_poisson_free(gfd_poisson* p)
{
    if (!p) return;
    drop_graphs(p);
    if (p->ev_in) (void)hipEventDestroy(p->ev_in);
    if (p->ev_out) (void)hipEventDestroy(p->ev_out);
    if (p->gs) (void)hipStreamDestroy(p->gs);
    p->lv.pool.release();
    p->lv32.pool.release();
    p->lv64r.pool.release();
    p->scalar.release();
    p->partial.release();
    p->xscratch.release();
    delete p;
}

// nrelax smoothing sweeps: Jacobi of any precision ping-ponging between
// *cur / *alt, or (method 1, FP64) Gauss-Seidel in place on *cur with *alt
// as its start-of-sweep copy
static int relax_pp_t(gfd_operator* op, int dtype, void** cur, void** alt,
                      const void* src, int nrelax, double w, hipStream_t s,
                      int method = 2)
{
    if (method == 1 && dtype == 0)
        return op_gauss_seidel(op, (double*)*cur, (double*)*alt,
                               (const double*)src, nrelax, s);
    EpiArgs ea{};
    ea.aux = src;
    ea.c0 = op->coef[0];
    ea.w = w;
    for (int k = 0; k < nrelax; k++) {
        GFD_TRY(op_pass(op, dtype, *cur, *alt, 1, nullptr, EPI_RELAX, ea, s,
                        nullptr));
        std::swap(*cur, *alt);
    }
    return OK;
}

// Smoothing schedule of one V-cycle.
struct Schedule {
    int coarsest;   // level treated as the coarsest
    const int* pre;
    const int* post;
    const double* w;
    int pre_c, post_c;       // schedule override (>= 0) for l < coarsest
    int coarse_sweeps;       // sweeps at the coarsest level (>= 0: override)
    double w0, wc;           // weights override (> 0) for level 0 / others
};

static int sched_pre(const Schedule& sc, int l)
{
    return sc.pre_c >= 0 ? sc.pre_c : sc.pre[l];
}
static int sched_post(const Schedule& sc, int l)
{
    if (l == sc.coarsest && sc.coarse_sweeps >= 0) return sc.coarse_sweeps;
    return sc.post_c >= 0 ? sc.post_c : sc.post[l];
}
static double sched_w(const Schedule& sc, int l)
{
    if (l == 0 && sc.w0 > 0) return sc.w0;
    if (l > 0 && sc.wc > 0) return sc.wc;
    return sc.w[l];
}

// Plan the LDS-resident solve of levels [lc, coarsest] for a schedule: the
// first lc >= 1 whose level arrays fit one workgroup's LDS (rhs of level lc
// stays in global memory).  *changed reports a new plan (graphs that
// captured the old one must be dropped).  Not callable under capture.
// cheb: Chebyshev iteration on the coarsest level (mixed schedules; the
// reference schedule smooths every level with Jacobi)
static int plan_coarse(gfd_poisson* p, const Schedule& sc, int dtype,
                       CoarseSolve& c, hipStream_t s, bool* changed,
                       bool cheb = false)
{
    const size_t es = dtype == 2 ? 4 : 8;
    int lc = -1;
    size_t bytes = 0;
    // the LDS-resident solve smooths with Jacobi: not for the FP64
    // Gauss-Seidel schedule
    const bool lds_ok = g_coarse_lds && !(dtype == 0 && p->method == 1);
    for (int l0 = 1; lds_ok && l0 <= sc.coarsest; l0++) {
        if (sc.coarsest - l0 + 1 > CMAXLEV) continue;
        if (p->ops[l0]->ng() > g_coarse_max_pts) continue;
        size_t b = 0;
        bool ok = true;
        for (int l = l0; l <= sc.coarsest && ok; l++) {
            const gfd_operator* op = p->ops[l];
            ok = op->axis_ok && op->real && op->ax.R <= CMAXR;
            if (l < sc.coarsest)
                ok = ok && p->rst[l]->plan.K == 2 && p->itp[l]->plan.K == 2;
            b += es * op->ng() * (l == l0 ? 2 : 3);
        }
        if (ok && b <= COARSE_LDS_MAX) {
            lc = l0;
            bytes = b;
            break;
        }
    }
    // stage the first level's rhs in LDS too when it fits
    const bool r_lds =
        lc >= 0 && bytes + es * p->ops[lc]->ng() <= COARSE_LDS_MAX;
    if (r_lds) bytes += es * p->ops[lc]->ng();
    CGrid g{};
    if (lc >= 0) {
        g.nlev = sc.coarsest - lc + 1;
        int off = 0;
        for (int i = 0; i < g.nlev; i++) {
            const int l = lc + i;
            const gfd_operator* op = p->ops[l];
            CLevel& L = g.lev[i];
            for (int d = 0; d < 3; d++) {
                L.n[d] = op->n[d];
                L.periodic[d] = op->periodic[d];
                L.has[d] = op->ax.has[d];
                for (int k = 0; k < CMAXR; k++) {
                    L.cp[d][k] = k < op->ax.R ? op->ax.cp[d][k] : 0.0;
                    L.cm[d][k] = k < op->ax.R ? op->ax.cm[d][k] : 0.0;
                }
            }
            L.R = op->ax.R;
            L.c0 = centre_coef(op);
            L.w = sched_w(sc, l);
            L.pre = sched_pre(sc, l);
            L.post = sched_post(sc, l);
            const int n = (int)op->ng();
            L.a = off;
            off += n;
            L.b = off;
            off += n;
            if (i == 0 && !r_lds) {
                L.r = -1;
            } else {
                L.r = off;
                off += n;
            }
            if (l < sc.coarsest)
                for (int d = 0; d < 3; d++) {
                    L.roff[d] = p->rst[l]->plan.ax[d].off;
                    L.ioff[d] = p->itp[l]->plan.ax[d].off;
                    L.iskip0[d] = p->itp[l]->plan.ax[d].skip0;
                }
        }
        CLevel& C = g.lev[g.nlev - 1];
        const int nc = C.n[0] * C.n[1] * C.n[2];
        double lo, hi;
        C.cheb = 0;
        C.dbuf = -1;
        if (cheb && coarse_cheb_enabled() &&
            bytes + es * nc <= COARSE_LDS_MAX && cheb_bounds(C, &lo, &hi)) {
            C.cheb = 1;
            C.dbuf = off;
            C.lam_lo = lo;
            C.lam_hi = hi;
            off += nc;
            bytes += es * nc;
        }
        g.lds_elems = off;
    }
    const bool same = c.valid && c.lc == lc &&
                      (lc < 0 || std::memcmp(&c.host, &g, sizeof g) == 0);
    if (changed) *changed = !same;
    if (same) return OK;
    if (lc >= 0) {
        GFD_TRY(c.dev.ensure(sizeof(CGrid), s));
        GFD_CHECK_HIP(hipMemcpyAsync(c.dev.p, &g, sizeof(CGrid),
                                     hipMemcpyHostToDevice, s));
        GFD_CHECK_HIP(hipStreamSynchronize(s));
    }
    c.host = g;
    c.lc = lc;
    c.lds = bytes;
    c.valid = true;
    return OK;
}

// gpaw/poisson.py:500-526 for one level; *cur / *alt hold phi_l.
// zero: phi_l starts at 0 (its contents are not read).  On the mixed
// schedules' levels the first Jacobi sweep from 0 is then the scaled
// right-hand side (one pass without the stencil); the reference schedule
// zeroes phi_l and sweeps as the reference does.
static int vcycle(gfd_poisson* p, Levels& lv, const Schedule& sc, int l,
                  void** cur, void** alt, const void* rho, hipStream_t s,
                  bool zero = false)
{
    gfd_operator* op = p->ops[l];
    const int dt = lv.dtype;
    const size_t es = dt == 2 ? 4 : 8;
    const bool mixed = &lv != &p->lv;
    // sweeps of the first smoothing stage from phi_l = 0
    auto relax_from = [&](int n, int& done) -> int {
        done = 0;
        if (!zero) return OK;
        zero = false;
        const bool jacobi = !(p->method == 1 && dt == 0);
        if (mixed && jacobi && n > 0) {
            done = 1;
            return vec_relax0(*cur, rho, op->ng(), sched_w(sc, l),
                              op->coef[0], dt, s);
        }
        return vec_zero_bytes(*cur, es * op->ng(), s);
    };
    if (l < sc.coarsest) {
        int done = 0;
        GFD_TRY(relax_from(sched_pre(sc, l), done));
        GFD_TRY(relax_pp_t(op, dt, cur, alt, rho, sched_pre(sc, l) - done,
                           sched_w(sc, l), s, p->method));
        EpiArgs e{};
        e.aux = rho;
        e.c0 = centre_coef(op);
        GFD_TRY(op_pass(op, dt, *cur, lv.res[l], 1, nullptr, EPI_RESID, e, s,
                        nullptr));
        XEpiArgs xe;
        GFD_TRY(xfer_dt(p->rst[l], dt, lv.res[l], lv.rho[l + 1], 1, nullptr,
                        xe, s, p->xscratch.p));
        void* c1 = lv.phi[l + 1];
        void* a1 = lv.tmp[l + 1];
        const CoarseSolve& cs =
            &lv == &p->lv64r ? p->cs64r : (dt == 2 ? p->cs32 : p->cs64);
        if (cs.valid && cs.lc == l + 1) {
            // levels l+1 .. coarsest in one LDS-resident workgroup
            GFD_TRY(launch_coarse((const CGrid*)cs.dev.p, 1, cs.lds, dt,
                                  lv.rho[l + 1], c1, s));
        } else {
            GFD_TRY(vcycle(p, lv, sc, l + 1, &c1, &a1, lv.rho[l + 1], s,
                           true));
        }
        XEpiArgs ce;
        ce.epi = XEPI_SUBFROM;
        ce.base = *cur;
        GFD_TRY(xfer_dt(p->itp[l], dt, c1, *alt, 1, nullptr, ce, s,
                        p->xscratch.p));
        std::swap(*cur, *alt);
    }
    int done = 0;
    GFD_TRY(relax_from(sched_post(sc, l), done));
    GFD_TRY(relax_pp_t(op, dt, cur, alt, rho, sched_post(sc, l) - done,
                       sched_w(sc, l), s, p->method));
    return OK;
}

static Schedule ref_schedule(gfd_poisson* p)
{
    Schedule sc;
    sc.coarsest = p->L;
    sc.pre = p->pre.data();
    sc.post = p->post.data();
    sc.w = p->w.data();
    sc.pre_c = sc.post_c = sc.coarse_sweeps = -1;
    sc.w0 = sc.wc = -1.0;
    return sc;
}

// One V-cycle on (phi, rho) + residual norm into *dn (device).
static int vcycle64_norm(gfd_poisson* p, double* phi, const double* rho,
                         double* dn, hipStream_t s)
{
    const Schedule sc = ref_schedule(p);
    void* cur = phi;
    void* alt = p->lv.tmp[0];
    GFD_TRY(vcycle(p, p->lv, sc, 0, &cur, &alt, rho, s));
    if (cur != (void*)phi)
        GFD_TRY(vec_copy((const double*)cur, phi, p->ops[0]->ng(), s));
    GFD_TRY(residual_norm(p->ops[0], 0, phi, rho, p->lv.res[0], 1, nullptr,
                          dn, s, false, (double*)p->partial.p));
    return OK;
}

static int ensure_stream(gfd_poisson* p)
{
    if (!p->gs) {
        GFD_CHECK_HIP(hipStreamCreateWithFlags(&p->gs, hipStreamNonBlocking));
        GFD_CHECK_HIP(hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming));
        GFD_CHECK_HIP(hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming));
    }
    return OK;
}

// Capture fn(stream) into an executable graph.  Buffers that the captured
// work may grow are sized first: nothing allocates during a capture.
// One eager PLAIN pass per level on the scratch arrays lets the stencil
// tile tuner time its candidates before a graph freezes the tiles (the
// tuner never runs under capture); tmp and res are rewritten before they
// are read in every V-cycle.
static int pretune_levels(gfd_poisson* p, Levels& lv, hipStream_t s)
{
    for (int l = 0; l <= lv.L && l <= p->L; l++) {
        EpiArgs e{};
        e.c0 = centre_coef(p->ops[l]);
        GFD_TRY(op_pass(p->ops[l], lv.dtype, lv.tmp[l], lv.res[l], 1,
                        nullptr, EPI_PLAIN, e, s, nullptr));
        if (l < lv.L && l < p->L) {
            // the transfer kernel choices of this level pair
            const BC br = make_bc(p->rst[l]->periodic, nullptr);
            GFD_TRY(xfer_maybe_tune(p->rst[l], lv.dtype, br, lv.res[l], 1,
                                    p->xscratch.p, 0, s));
            const BC bi = make_bc(p->itp[l]->periodic, nullptr);
            GFD_TRY(xfer_maybe_tune(p->itp[l], lv.dtype, bi, lv.phi[l + 1],
                                    1, p->xscratch.p, 0, s));
        }
    }
    return OK;
}

template <typename F>
static int capture(gfd_poisson* p, hipGraphExec_t* exec, F fn)
{
    GFD_TRY(pretune_levels(p, p->lv, p->gs));
    if (p->lv32.L >= 0) GFD_TRY(pretune_levels(p, p->lv32, p->gs));
    if (p->lv64r.L >= 0) GFD_TRY(pretune_levels(p, p->lv64r, p->gs));
    hipGraph_t g = nullptr;
    GFD_CHECK_HIP(hipStreamBeginCapture(p->gs, hipStreamCaptureModeThreadLocal));
    int st = fn(p->gs);
    hipError_t e = hipStreamEndCapture(p->gs, &g);
    if (st != OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        set_error("graph capture failed: %s", hipGetErrorString(e));
        return EHIP;
    }
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *exec = nullptr;
        set_error("graph instantiation failed: %s", hipGetErrorString(e));
        return EHIP;
    }
    return OK;
}

static int read_scalar(gfd_poisson* p, hipStream_t s, double* h)
{
    GFD_CHECK_HIP(hipMemcpyAsync(h, p->scalar.p, sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    GFD_CHECK_HIP(hipStreamSynchronize(s));
    return OK;
}

static int poisson_iterate(gfd_poisson* p, double* phi, const double* rho,
                           double* error, hipStream_t s)
{
    GFD_TRY(plan_coarse(p, ref_schedule(p), 0, p->cs64, s, nullptr));
    GFD_TRY(p->scalar.ensure(4 * sizeof(double), s));
    GFD_TRY(vcycle64_norm(p, phi, rho, (double*)p->scalar.p, s));
    double h = 0.0;
    GFD_TRY(read_scalar(p, s, &h));
    *error = h * p->dv;
    return OK;
}

extern "C" int gfd_poisson_iterate(gfd_poisson* p, double* phi,
                                   const double* rho, double* error,
                                   void* stream)
{
    if (!p || !phi || !rho || !error) {
        set_error("gfd_poisson_iterate: bad arguments");
        return EINVAL_;
    }
    return poisson_iterate(p, phi, rho, error, (hipStream_t)stream);
}

extern "C" int gfd_add_scalar(double* a, long n, double v, void* stream);

static int remove_mean(gfd_poisson* p, double* phi, hipStream_t s)
{
    const long ng = p->ops[0]->ng();
    double* dn = (double*)p->scalar.p;
    GFD_TRY(vec_dot(phi, nullptr, ng, (double*)p->partial.p, dn, s));
    double sum = 0.0;
    GFD_TRY(read_scalar(p, s, &sum));
    const double ave = sum / (double)ng;
    return vec_add_scalar(phi, ng, -ave, s);
}

// Run the solve on the owned stream: phi / rho are copied into the owned
// level-0 buffers, the captured graph is replayed, phi is copied back.
static int enter_owned(gfd_poisson* p, double* phi, const double* rho,
                       hipStream_t s)
{
    GFD_TRY(ensure_stream(p));
    const long ng = p->ops[0]->ng();
    GFD_TRY(vec_copy(phi, (double*)p->lv.phi[0], ng, s));
    GFD_TRY(vec_copy(rho, (double*)p->lv.rho[0], ng, s));
    GFD_CHECK_HIP(hipEventRecord(p->ev_in, s));
    GFD_CHECK_HIP(hipStreamWaitEvent(p->gs, p->ev_in, 0));
    return OK;
}

static int leave_owned(gfd_poisson* p, double* phi, hipStream_t s)
{
    GFD_CHECK_HIP(hipEventRecord(p->ev_out, p->gs));
    GFD_CHECK_HIP(hipStreamWaitEvent(s, p->ev_out, 0));
    return vec_copy((const double*)p->lv.phi[0], phi, p->ops[0]->ng(), s);
}

extern "C" int gfd_poisson_solve(gfd_poisson* p, double* phi,
                                 const double* rho, double eps, int maxiter,
                                 int remove_mean_flag, int* niter,
                                 double* error, void* stream)
{
    if (!p || !phi || !rho || !niter || maxiter < 1) {
        set_error("gfd_poisson_solve: bad arguments");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    GFD_TRY(p->scalar.ensure(4 * sizeof(double), s));
    double* dn = (double*)p->scalar.p;
    int it = 1;
    double err = 0.0;
    {
        bool changed = false;
        GFD_TRY(plan_coarse(p, ref_schedule(p), 0, p->cs64, s, &changed));
        if (changed && p->g64) {
            (void)hipGraphExecDestroy(p->g64);
            p->g64 = nullptr;
        }
    }
    if (p->use_graph) {
        GFD_TRY(enter_owned(p, phi, rho, s));
        double* phi0 = (double*)p->lv.phi[0];
        const double* rho0 = (const double*)p->lv.rho[0];
        if (!p->g64)
            GFD_TRY(capture(p, &p->g64, [&](hipStream_t cs) {
                return vcycle64_norm(p, phi0, rho0, dn, cs);
            }));
        double h = 0.0;
        GFD_CHECK_HIP(hipGraphLaunch(p->g64, p->gs));
        GFD_TRY(read_scalar(p, p->gs, &h));
        err = h * p->dv;
        while (err > eps && it < maxiter) {
            it++;
            GFD_CHECK_HIP(hipGraphLaunch(p->g64, p->gs));
            GFD_TRY(read_scalar(p, p->gs, &h));
            err = h * p->dv;
        }
        GFD_TRY(leave_owned(p, phi, s));
    } else {
        GFD_TRY(poisson_iterate(p, phi, rho, &err, s));
        while (err > eps && it < maxiter) {
            it++;
            GFD_TRY(poisson_iterate(p, phi, rho, &err, s));
        }
    }
    *niter = it;
    if (error) *error = err;
    if (it == maxiter) {
        set_error("Poisson solver did not converge in %d iterations!",
                  maxiter);
        return ECONVERGE;
    }
    if (remove_mean_flag) GFD_TRY(remove_mean(p, phi, s));
    return OK;
}

extern "C" int gfd_poisson_set_graph(gfd_poisson* p, int on)
{
    if (!p) {
        set_error("gfd_poisson_set_graph: null handle");
        return EINVAL_;
    }
    p->use_graph = on != 0;
    return OK;
}

// One refinement step of the mixed-precision solve on the owned buffers:
// e = 0; `inner` f32 V-cycles on L e = r32; phi -= e; r32 = (float)(L phi
// - rho) with the FP64 squared norm into *dn.
// One refinement step: e from `inner` V-cycles on L e = r (FP32 levels, or
// FP64 with fp64), phi -= e, then the new FP64 residual r and its norm.
static int mixed_step(gfd_poisson* p, const Schedule& sc, int inner,
                      double* dn, hipStream_t s, bool fp64)
{
    const long ng = p->ops[0]->ng();
    double* phi0 = (double*)p->lv.phi[0];
    const double* rho0 = (const double*)p->lv.rho[0];
    Levels& lv = fp64 ? p->lv64r : p->lv32;
    void* r0 = lv.rho[0];
    void* cur = lv.phi[0];
    void* alt = lv.tmp[0];
    GFD_TRY(vec_zero_bytes(cur, (fp64 ? 8 : 4) * ng, s));
    for (int c = 0; c < inner; c++)
        GFD_TRY(vcycle(p, lv, sc, 0, &cur, &alt, r0, s));
    if (fp64)
        GFD_TRY(vec_sub_d(phi0, (const double*)cur, ng, s));
    else
        GFD_TRY(vec_sub_f(phi0, (const float*)cur, ng, s));
    return residual_norm(p->ops[0], 0, phi0, rho0, r0, 1, nullptr, dn, s,
                         !fp64, (double*)p->partial.p);
}

extern "C" int gfd_poisson_solve_mixed(gfd_poisson* p, double* phi,
                                       const double* rho, double tol_rel,
                                       int levels, int pre, int post,
                                       int coarse_sweeps, double w0,
                                       double wc, int inner, int max_outer,
                                       int inner_fp64, int* n_outer,
                                       double* relres, void* stream)
{
    if (!p || !phi || !rho || !n_outer || max_outer < 1 || inner < 1 ||
        levels < 1 || pre < 0 || post < 0) {
        set_error("gfd_poisson_solve_mixed: bad arguments");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    const bool fp64 = inner_fp64 != 0;
    const int L32 = std::min(levels - 1, p->L);
    Levels& lvi = fp64 ? p->lv64r : p->lv32;
    CoarseSolve& csi = fp64 ? p->cs64r : p->cs32;
    hipGraphExec_t& gi = fp64 ? p->g64r : p->g32;
    int* gkey = fp64 ? p->g64r_key : p->g32_key;
    double* gw = fp64 ? p->g64r_w : p->g32_w;
    GFD_TRY(alloc_levels(p, lvi, fp64 ? 0 : 2, L32));
    GFD_TRY(p->scalar.ensure(4 * sizeof(double), s));
    double* dn = (double*)p->scalar.p;
    Schedule sc = ref_schedule(p);
    sc.coarsest = L32;
    sc.pre_c = pre;
    sc.post_c = post;
    sc.coarse_sweeps = coarse_sweeps;
    sc.w0 = w0;
    sc.wc = wc;
    bool cs_changed = false;
    GFD_TRY(plan_coarse(p, sc, fp64 ? 0 : 2, csi, s, &cs_changed, true));
    const int key[6] = {L32, pre, post, coarse_sweeps, inner, 0};
    bool same = gi && !cs_changed && gw[0] == w0 && gw[1] == wc;
    for (int i = 0; i < 6; i++) same = same && gkey[i] == key[i];
    if (!same && gi) {
        (void)hipGraphExecDestroy(gi);
        gi = nullptr;
    }
    GFD_TRY(enter_owned(p, phi, rho, s));
    const long ng = p->ops[0]->ng();
    double* phi0 = (double*)p->lv.phi[0];
    const double* rho0 = (const double*)p->lv.rho[0];
    // |rho|^2 and the first residual
    double rn2 = 0.0, err2 = 0.0;
    GFD_TRY(vec_dot(rho0, rho0, ng, (double*)p->partial.p, dn, p->gs));
    GFD_TRY(read_scalar(p, p->gs, &rn2));
    GFD_TRY(residual_norm(p->ops[0], 0, phi0, rho0, lvi.rho[0], 1, nullptr,
                          dn, p->gs, !fp64, (double*)p->partial.p));
    GFD_TRY(read_scalar(p, p->gs, &err2));
    const double tol2 = tol_rel * tol_rel * rn2;
    int outer = 0;
    while (err2 > tol2 && outer < max_outer) {
        if (p->use_graph) {
            if (!gi) {
                GFD_TRY(capture(p, &gi, [&](hipStream_t cs) {
                    return mixed_step(p, sc, inner, dn, cs, fp64);
                }));
                for (int i = 0; i < 6; i++) gkey[i] = key[i];
                gw[0] = w0;
                gw[1] = wc;
            }
            GFD_CHECK_HIP(hipGraphLaunch(gi, p->gs));
        } else {
            GFD_TRY(mixed_step(p, sc, inner, dn, p->gs, fp64));
        }
        GFD_TRY(read_scalar(p, p->gs, &err2));
        outer++;
    }
    GFD_TRY(leave_owned(p, phi, s));
    *n_outer = outer;
    if (relres) *relres = rn2 > 0 ? std::sqrt(err2 / rn2) : std::sqrt(err2);
    if (err2 > tol2) {
        set_error("mixed-precision Poisson solve did not reach %g in %d "
                  "refinement steps",
                  tol_rel, max_outer);
        return ECONVERGE;
    }
    return OK;
}

// =================================================================== bands

// Work spaces of the band calls and of gfd_dot, one per stream, so that
// calls on different streams never share partial sums.  They live until the
// process ends (the HIP runtime may be gone at static teardown).
static std::mutex g_ws_mu;

static BandWork& band_work(hipStream_t s)
{
    static std::map<hipStream_t, BandWork*> m;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    BandWork*& w = m[s];
    if (!w) w = new BandWork();
    return *w;
}

static DevBuf& dot_work(hipStream_t s)
{
    static std::map<hipStream_t, DevBuf*> m;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    DevBuf*& w = m[s];
    if (!w) w = new DevBuf();
    return *w;
}

extern "C" int gfd_h_apply(gfd_operator* kin, const double* vpot,
                           const void* psi, void* hpsi, long nbands,
                           const double* phases, void* stream)
{
    if (!kin || !vpot || nbands < 0 || (nbands > 0 && (!psi || !hpsi)) ||
        psi == hpsi) {
        set_error("gfd_h_apply: bad arguments");
        return EINVAL_;
    }
    if (nbands == 0) return OK;
    EpiArgs ea{};
    ea.c0 = centre_coef(kin);
    ea.pot = vpot;
    return op_pass(kin, kin->dtype(), psi, hpsi, nbands,
                   kin->real ? nullptr : phases, EPI_POT, ea,
                   (hipStream_t)stream, nullptr);
}

extern "C" int gfd_h_residual(gfd_operator* kin, const double* vpot,
                              const double* eps, const void* psi, void* r,
                              long nbands, const double* phases,
                              void* stream)
{
    if (!kin || !vpot || !eps || nbands < 0 ||
        (nbands > 0 && (!psi || !r)) || psi == r) {
        set_error("gfd_h_residual: bad arguments");
        return EINVAL_;
    }
    if (nbands == 0) return OK;
    EpiArgs ea{};
    ea.c0 = centre_coef(kin);
    ea.pot = vpot;
    ea.shift = eps;
    return op_pass(kin, kin->dtype(), psi, r, nbands,
                   kin->real ? nullptr : phases, EPI_POT, ea,
                   (hipStream_t)stream, nullptr);
}

extern "C" int gfd_band_dots(int is_complex, const void* a, const void* b,
                             int nbands, long npts, long ld, double dv,
                             double* out, void* stream)
{
    if (!a || !b || !out || nbands < 0 || npts < 0 || ld < npts) {
        set_error("gfd_band_dots: bad arguments");
        return EINVAL_;
    }
    return band_dots(is_complex ? 1 : 0, a, b, nbands, npts, ld, dv, out,
                     (hipStream_t)stream);
}

extern "C" int gfd_rmm_lambda(int is_complex, const void* dR, const void* R,
                              int nbands, long npts, long ld, double* lam,
                              void* stream)
{
    if (!dR || !R || !lam || nbands < 0 || npts < 0 || ld < npts) {
        set_error("gfd_rmm_lambda: bad arguments");
        return EINVAL_;
    }
    return rmm_lambda(is_complex ? 1 : 0, dR, R, nbands, npts, ld, lam,
                      (hipStream_t)stream);
}

extern "C" int gfd_band_update(int is_complex, int mode, const double* alpha,
                               double scale, const void* x, void* y,
                               const void* v, void* w, int nbands, long npts,
                               long ld, void* stream)
{
    if (!alpha || !x || !y || (mode != 0 && mode != 1) ||
        (mode == 1 && (!v || !w || w == y)) || nbands < 0 || npts < 0 ||
        ld < npts) {
        set_error("gfd_band_update: bad arguments");
        return EINVAL_;
    }
    return band_update(is_complex ? 1 : 0, mode, alpha, scale, x, y, v, w,
                       nbands, npts, ld, (hipStream_t)stream);
}

extern "C" int gfd_band_overlap(const void* psi, int nbands, long npts,
                                long ld, double dv, void* S, int lower_only,
                                void* stream)
{
    if (!psi || !S || nbands < 0 || npts < 0 || ld < npts) {
        set_error("gfd_band_overlap: bad arguments");
        return EINVAL_;
    }
    return gram(band_work((hipStream_t)stream), (const double2*)psi, ld,
                (const double2*)psi, ld,
                nbands, nbands, npts, dv, 0, lower_only, (double2*)S, nbands,
                (hipStream_t)stream);
}

extern "C" int gfd_band_matrix(const void* psi, const void* hpsi, int nbands,
                               long npts, long ld, double dv, void* H,
                               void* stream)
{
    if (!psi || !hpsi || !H || nbands < 0 || npts < 0 || ld < npts) {
        set_error("gfd_band_matrix: bad arguments");
        return EINVAL_;
    }
    return gram(band_work((hipStream_t)stream), (const double2*)psi, ld,
                (const double2*)hpsi, ld,
                nbands, nbands, npts, dv, 1, 0, (double2*)H, nbands,
                (hipStream_t)stream);
}

extern "C" int gfd_band_cholesky_orth(void* S, int nbands, const void* psi,
                                      long npts, long ld, void* out,
                                      void* stream)
{
    if (!S || !psi || !out || nbands <= 0 || npts <= 0 || ld < npts ||
        psi == out) {
        set_error("gfd_band_cholesky_orth: bad arguments");
        return EINVAL_;
    }
    return cholesky_orth(band_work((hipStream_t)stream), (double2*)S, nbands,
                         (const double2*)psi, ld,
                         npts, (double2*)out, ld, (hipStream_t)stream);
}

extern "C" int gfd_band_subspace_diag(void* H, int nbands, const void* psi,
                                      long npts, long ld, void* out,
                                      double* eps, void* stream)
{
    if (!H || !psi || !out || !eps || nbands <= 0 || npts <= 0 ||
        ld < npts || psi == out) {
        set_error("gfd_band_subspace_diag: bad arguments");
        return EINVAL_;
    }
    return subspace_diag(band_work((hipStream_t)stream), (double2*)H, nbands,
                         (const double2*)psi, ld, npts, (double2*)out, ld,
                         eps, (hipStream_t)stream);
}

// ==================================================================== misc

extern "C" int gfd_dot(const double* a, const double* b, long n,
                       double* result, void* stream)
{
    if (!a || !result || n < 0) {
        set_error("gfd_dot: bad arguments");
        return EINVAL_;
    }
    hipStream_t s = (hipStream_t)stream;
    DevBuf& buf = dot_work(s);
    GFD_TRY(buf.ensure(sizeof(double) * (vec_dot_partials(n) + 2), s));
    double* part = (double*)buf.p;
    double* dres = part + vec_dot_partials(n) + 1;
    GFD_TRY(vec_dot(a, b, n, part, dres, s));
    GFD_CHECK_HIP(hipMemcpyAsync(result, dres, sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    GFD_CHECK_HIP(hipStreamSynchronize(s));
    return OK;
}

extern "C" int gfd_add_scalar(double* a, long n, double v, void* stream)
{
    if (!a || n < 0) {
        set_error("gfd_add_scalar: bad arguments");
        return EINVAL_;
    }
    return vec_add_scalar(a, n, v, (hipStream_t)stream);
}

extern "C" int gfd_stream_copy(const void* src, void* dst, long nbytes,
                               void* stream)
{
    if (!src || !dst || nbytes < 0 || (nbytes & 15) ||
        ((uintptr_t)src & 15) || ((uintptr_t)dst & 15)) {
        set_error("gfd_stream_copy: 16-byte aligned pointers and sizes");
        return EINVAL_;
    }
    return vec_copy16(src, dst, nbytes, (hipStream_t)stream);
}

static long tile_code(const XmTile& t)
{
    return (long)t.tz | (long)t.nseg << 12 | (long)t.py << 24 |
           (long)t.xchunk << 32 | (long)t.dma << 48 | (long)t.order << 52 |
           (long)t.wide << 56;
}

extern "C" int gfd_row_tile_candidates(int radius, int dtype,
                                       const long n[3], int with_dma,
                                       long* codes, int max_codes)
{
    if (radius < 1 || radius > MAXR_FAST || dtype < 0 || dtype > 3 || !n ||
        n[0] < 1 || n[1] < 1 || n[2] < 1 || max_codes < 0 ||
        (max_codes > 0 && !codes)) {
        set_error("gfd_row_tile_candidates: bad arguments");
        return -1;
    }
    const int nn[3] = {(int)n[0], (int)n[1], (int)n[2]};
    const std::vector<XmTile> c = xr_candidates(radius, dtype, nn,
                                                with_dma != 0);
    for (int i = 0; i < (int)c.size() && i < max_codes; i++)
        codes[i] = tile_code(c[i]);
    return (int)c.size();
}

extern "C" int gfd_set_tuning(const char* key, long value)
{
    if (!key) {
        set_error("gfd_set_tuning: null key");
        return EINVAL_;
    }
    if (std::strcmp(key, "force_general") == 0) {
        g_force_path = value ? 1 : 0;
        return OK;
    }
    if (std::strcmp(key, "force_path") == 0) {
        if (value < 0 || value > 3) {
            set_error("gfd_set_tuning: force_path must be 0..3");
            return EINVAL_;
        }
        g_force_path = (int)value;
        return OK;
    }
    if (std::strcmp(key, "xdma") == 0) {
        xr_set_dma(value != 0);
        return OK;
    }
    if (std::strcmp(key, "autotune") == 0) {
        xr_set_autotune(value != 0);
        xfer_set_autotune(value != 0);
        return OK;
    }
    if (std::strcmp(key, "xr_nt") == 0) {
        xr_set_nt((int)value);
        return OK;
    }
    if (std::strcmp(key, "coarse_lds") == 0) {
        g_coarse_lds = value != 0;
        return OK;
    }
    if (std::strcmp(key, "coarse_cheb") == 0) {
        coarse_set_cheb(value != 0);
        return OK;
    }
    if (std::strcmp(key, "coarse_max_points") == 0) {
        g_coarse_max_pts = value > 0 ? value : 4096;
        return OK;
    }
    if (std::strcmp(key, "xfer_passes") == 0) {
        xfer_force_passes(value != 0);
        return OK;
    }
    if (std::strcmp(key, "wx_march") == 0) {
        wx_set_enabled(value != 0);
        return OK;
    }
    if (std::strcmp(key, "wx_tile") == 0) {
        // py | xchunk << 8 (0: automatic)
        const int py = (int)(value & 0xff), xc = (int)((value >> 8) & 0xffff);
        if (py != 0 && py != 1 && py != 2 && py != 4) {
            set_error("gfd_set_tuning: wx_tile rows per thread 1, 2 or 4");
            return EINVAL_;
        }
        wx_set_tile(py, xc);
        return OK;
    }
    if (std::strcmp(key, "batch_tile") == 0) {
        // nt | py << 12 | zmax << 16 | xc << 28 | pf << 40 (0: timed)
        const int nt = (int)(value & 0xfff), py = (int)((value >> 12) & 0xf);
        const int zmax = (int)((value >> 16) & 0xfff);
        const int xc = (int)((value >> 28) & 0xfff);
        const int pf = (int)((value >> 40) & 0xf);
        if (value != 0 &&
            (nt < 64 || nt > 512 || nt % 64 || (py != 1 && py != 2 && py != 4) ||
             zmax < 1 || xc < 1 || pf < 1 || pf > 2)) {
            set_error("gfd_set_tuning: batch_tile fields out of range");
            return EINVAL_;
        }
        batch_force_tile(value ? nt : 0, py, zmax, xc, pf);
        return OK;
    }
    if (std::strcmp(key, "batch_xfer") == 0) {
        const long nt = value & 0xfff, py = (value >> 12) & 0xf;
        const long zmax = (value >> 16) & 0xfff, xc = (value >> 28) & 0xfff;
        if (value < 0 || (value > 2 && value < 64) ||
            (value >= 64 && (nt < 64 || nt > 512 || nt % 64 ||
                             (py != 1 && py != 2 &&
                                                    py != 4) ||
                             zmax < 1 || xc < 1))) {
            set_error("gfd_set_tuning: batch_xfer 0, 1, 2 or a tile code");
            return EINVAL_;
        }
        batch_set_xfer(value);
        return OK;
    }
    if (std::strcmp(key, "batch_kernel") == 0) {
        if (value < 0 || value > 2) {
            set_error("gfd_set_tuning: batch_kernel 0, 1 or 2");
            return EINVAL_;
        }
        batch_set_kernel((int)value);
        return OK;
    }
    if (std::strcmp(key, "precond_fine") == 0) {
        // 0: the FP32 chain's fine level as two passes; 1: fused where it
        // fits (default)
        precond_fine_enable(value != 0);
        precond_fine_nt(value == 2);
        return OK;
    }
    if (std::strcmp(key, "precond_fine_tile") == 0) {
        // py | nys << 8 | xc << 16 (0: automatic)
        const int py = (int)(value & 0xff), nys = (int)((value >> 8) & 0xff);
        const int xc = (int)((value >> 16) & 0xffff);
        if (value < 0 || py > 2 || nys > 256) {
            set_error("gfd_set_tuning: precond_fine_tile fields out of range");
            return EINVAL_;
        }
        precond_fine_force(py, nys, xc);
        return OK;
    }
    if (std::strcmp(key, "xfer_march_scale") == 0) {
        xfer_set_march_scale((int)value);
        return OK;
    }
    if (std::strcmp(key, "xfer_march") == 0) {
        xfer_set_march(value != 0);
        return OK;
    }
    if (std::strcmp(key, "xr_tile") == 0) {
        if (value == 0) {
            xr_set_override(nullptr);
            return OK;
        }
        XmTile t;
        t.tz = (int)(value & 0xfff);
        t.nseg = (int)((value >> 12) & 0xfff);
        t.py = (int)((value >> 24) & 0xff);
        t.xchunk = (int)((value >> 32) & 0xffff);
        t.dma = (int)((value >> 48) & 0xf);
        t.order = (int)((value >> 52) & 0xf);
        t.wide = (int)((value >> 56) & 0xf);
        if (t.tz <= 0 || t.nseg <= 0 || t.py <= 0 || t.xchunk <= 0) {
            set_error("gfd_set_tuning: bad xr_tile");
            return EINVAL_;
        }
        xr_set_override(&t);
        return OK;
    }
    if (std::strcmp(key, "xm_tile") == 0) {
        if (value == 0) {
            xm_set_override(nullptr);
            return OK;
        }
        XmTile t;
        t.tz = (int)(value & 0xfff);
        t.nseg = (int)((value >> 12) & 0xfff);
        t.py = (int)((value >> 24) & 0xff);
        t.xchunk = (int)((value >> 32) & 0xffff);
        if (t.tz <= 0 || t.nseg <= 0 || t.py <= 0 || t.xchunk <= 0) {
            set_error("gfd_set_tuning: bad xm_tile");
            return EINVAL_;
        }
        xm_set_override(&t);
        return OK;
    }
    set_error("gfd_set_tuning: unknown key '%s'", key);
    return EINVAL_;
}
