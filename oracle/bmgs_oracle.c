/* CPU oracle for the GPAW finite-difference grid path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU algorithm, used as the checker for the HIP kernels (tests/, smoke() and
 * the CPU-baseline leg of bench.py).  The product never links or calls it.
 *
 * It restates, in the same operation order (and built without FP
 * contraction), the reference's
 *   - halo fill of a padded buffer        c/bc.cpp:17-142 (bc_init),
 *                                         c/bc.cpp:146-258 (bc_unpack1,
 *                                         single-domain COPY_DATA branch)
 *   - P-term stencil                      c/bmgs/fd.cpp:9-33
 *   - Jacobi / Gauss-Seidel sweeps        c/bmgs/relax.cpp:8-66,
 *                                         c/operators.cpp:50-92
 *   - 3-pass restriction                  c/bmgs/restrict.cpp:18-63,137-156
 *   - 3-pass interpolation                c/bmgs/interpolate.cpp:10-46,70-99
 *   - batch loop + OpenMP over grids      c/operators.cpp:245-316,
 *                                         c/transformers.cpp:38-158
 * so that its results equal the reference extension's bit for bit; the
 * tests check exactly that against oracle/_ref and the committed fixtures.
 */
#include "bmgs_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b)
{
    /* (a.re b.re - a.im b.im) + i (a.re b.im + a.im b.re), two roundings per
     * part, as std::complex<double> multiplication without FMA. */
    cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

/* ------------------------------------------------------------------------ */
/* Padded-buffer halo fill (single domain).                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    long n[3];       /* interior size */
    long pad[3][2];  /* halo widths below/above */
    long npad[3][2]; /* widths sent to the neighbour below/above */
    long nb[3][2];   /* >= 0: periodic wrap onto self; < 0: zero boundary */
    int cfd;         /* 1: axis-only stencil, edges/corners left at zero */
    int nd;          /* doubles per point: 1 real, 2 complex */
} halo_t;

static long padded_size(const halo_t* h, int i)
{
    return h->n[i] + h->pad[i][0] + h->pad[i][1];
}

/* Copy box [src_start, src_start+size) -> [dst_start, ...) inside the padded
 * array, multiplying by `ph` for complex data (bmgs_translate[mz]). */
static void box_translate(double* a, const long M[3], const long size[3],
                          const long s0[3], const long d0[3], int nd,
                          const double* ph)
{
    for (long i0 = 0; i0 < size[0]; i0++)
        for (long i1 = 0; i1 < size[1]; i1++) {
            long s = ((s0[0] + i0) * M[1] + (s0[1] + i1)) * M[2] + s0[2];
            long d = ((d0[0] + i0) * M[1] + (d0[1] + i1)) * M[2] + d0[2];
            if (nd == 1) {
                memcpy(a + d, a + s, size[2] * sizeof(double));
            } else {
                cplx p = {ph[0], ph[1]};
                for (long i2 = 0; i2 < size[2]; i2++) {
                    cplx v = {a[2 * (s + i2)], a[2 * (s + i2) + 1]};
                    cplx r = cmul(p, v);
                    a[2 * (d + i2)] = r.re;
                    a[2 * (d + i2) + 1] = r.im;
                }
            }
        }
}

static void halo_fill(const halo_t* h, const double* in, double* buf,
                      const double* phases)
{
    const int nd = h->nd;
    long M[3];
    for (int i = 0; i < 3; i++)
        M[i] = padded_size(h, i);
    memset(buf, 0, (size_t)(M[0] * M[1] * M[2] * nd) * sizeof(double));

    /* interior */
    for (long i0 = 0; i0 < h->n[0]; i0++)
        for (long i1 = 0; i1 < h->n[1]; i1++) {
            long d = ((i0 + h->pad[0][0]) * M[1] + (i1 + h->pad[1][0])) * M[2] +
                     h->pad[2][0];
            long s = (i0 * h->n[1] + i1) * h->n[2];
            memcpy(buf + d * nd, in + s * nd,
                   (size_t)(h->n[2] * nd) * sizeof(double));
        }

    /* faces, axis by axis; with cfd == 0 each later axis spans the full
     * padded extent of the earlier ones, which fills edges and corners. */
    long start[3], size[3];
    for (int i = 0; i < 3; i++) {
        start[i] = h->pad[i][0];
        size[i] = h->n[i];
    }
    for (int i = 0; i < 3; i++) {
        for (int d = 0; d < 2; d++) {
            long ds = h->npad[i][d];
            if (h->nb[i][d] < 0 || ds <= 0 || h->pad[i][1 - d] <= 0)
                continue;
            long s0[3], d0[3], sz[3];
            for (int j = 0; j < 3; j++) {
                s0[j] = start[j];
                d0[j] = start[j];
                sz[j] = size[j];
            }
            if (d == 0) {
                s0[i] = h->pad[i][0];
                d0[i] = h->pad[i][0] + h->n[i];
            } else {
                s0[i] = h->pad[i][0] + h->n[i] - ds;
                d0[i] = 0;
            }
            sz[i] = ds;
            box_translate(buf, M, sz, s0, d0, nd,
                          phases ? phases + 4 * i + 2 * d : NULL);
        }
        if (!h->cfd) {
            start[i] = 0;
            size[i] = M[i];
        }
    }
}

/* Flattened padded-buffer offsets of the stencil points. */
static long* flat_offsets(int ncoefs, const long* off3, const long M[3])
{
    long* o = (long*)malloc(sizeof(long) * (size_t)ncoefs);
    for (int c = 0; c < ncoefs; c++)
        o[c] = (off3[3 * c] * M[1] + off3[3 * c + 1]) * M[2] + off3[3 * c + 2];
    return o;
}

static int check_range(int ncoefs, const long* off3, int range)
{
    for (int c = 0; c < 3 * ncoefs; c++)
        if (off3[c] > range || off3[c] < -range)
            return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Stencil apply                                                            */
/* ------------------------------------------------------------------------ */

static void fd_one(int ncoefs, const double* coefs, const long* off,
                   const halo_t* h, const long M[3], const double* buf,
                   double* out)
{
    const long r0 = h->pad[0][0], r1 = h->pad[1][0], r2 = h->pad[2][0];
    for (long i0 = 0; i0 < h->n[0]; i0++)
        for (long i1 = 0; i1 < h->n[1]; i1++)
            for (long i2 = 0; i2 < h->n[2]; i2++) {
                long i = ((i0 + r0) * M[1] + (i1 + r1)) * M[2] + (i2 + r2);
                long j = (i0 * h->n[1] + i1) * h->n[2] + i2;
                if (h->nd == 1) {
                    double x = 0.0;
                    for (int c = 0; c < ncoefs; c++)
                        x += buf[i + off[c]] * coefs[c];
                    out[j] = x;
                } else {
                    double xr = 0.0, xi = 0.0;
                    for (int c = 0; c < ncoefs; c++) {
                        xr += buf[2 * (i + off[c])] * coefs[c];
                        xi += buf[2 * (i + off[c]) + 1] * coefs[c];
                    }
                    out[2 * j] = xr;
                    out[2 * j + 1] = xi;
                }
            }
}

int oracle_fd_apply(int ncoefs, const double* coefs, const long* off3,
                    const long n[3], int range, const long nb[6], int real,
                    int cfd, const double* phases, const double* in,
                    double* out, long nin, int nthreads)
{
    if (ncoefs <= 0 || range < 0 || check_range(ncoefs, off3, range))
        return 1;
    if (!real && phases == NULL)
        return 2;
    halo_t h;
    h.cfd = cfd;
    h.nd = real ? 1 : 2;
    for (int i = 0; i < 3; i++) {
        h.n[i] = n[i];
        h.pad[i][0] = h.pad[i][1] = range;
        h.npad[i][0] = h.npad[i][1] = range;
        h.nb[i][0] = nb[2 * i];
        h.nb[i][1] = nb[2 * i + 1];
    }
    long M[3] = {padded_size(&h, 0), padded_size(&h, 1), padded_size(&h, 2)};
    long* off = flat_offsets(ncoefs, off3, M);
    const long ng = n[0] * n[1] * n[2] * h.nd;
    const size_t ng2 = (size_t)(M[0] * M[1] * M[2] * h.nd);
    if (nthreads < 1)
        nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        double* buf = (double*)malloc(ng2 * sizeof(double));
#pragma omp for schedule(static)
        for (long m = 0; m < nin; m++) {
            halo_fill(&h, in + m * ng, buf, real ? NULL : phases);
            fd_one(ncoefs, coefs, off, &h, M, buf, out + m * ng);
        }
        free(buf);
    }
    free(off);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Relaxation                                                               */
/* ------------------------------------------------------------------------ */

int oracle_fd_relax(int method, int ncoefs, const double* coefs,
                    const long* off3, const long n[3], int range,
                    const long nb[6], int cfd, double* f, const double* src,
                    int nrelax, double w)
{
    if (ncoefs <= 0 || range < 0 || check_range(ncoefs, off3, range))
        return 1;
    if (off3[0] != 0 || off3[1] != 0 || off3[2] != 0)
        return 3; /* the first coefficient must be the centre */
    halo_t h;
    h.cfd = cfd;
    h.nd = 1;
    for (int i = 0; i < 3; i++) {
        h.n[i] = n[i];
        h.pad[i][0] = h.pad[i][1] = range;
        h.npad[i][0] = h.npad[i][1] = range;
        h.nb[i][0] = nb[2 * i];
        h.nb[i][1] = nb[2 * i + 1];
    }
    long M[3] = {padded_size(&h, 0), padded_size(&h, 1), padded_size(&h, 2)};
    long* off = flat_offsets(ncoefs, off3, M);
    double* buf = (double*)malloc(sizeof(double) * (size_t)(M[0] * M[1] * M[2]));
    const long r0 = range, r1 = range, r2 = range;
    for (int sweep = 0; sweep < nrelax; sweep++) {
        halo_fill(&h, f, buf, NULL);
        if (method == 1) {
            const double coef = 1.0 / coefs[0];
            for (long i0 = 0; i0 < n[0]; i0++)
                for (long i1 = 0; i1 < n[1]; i1++)
                    for (long i2 = 0; i2 < n[2]; i2++) {
                        long i = (i0 * n[1] + i1) * n[2] + i2;
                        long j = ((i0 + r0) * M[1] + (i1 + r1)) * M[2] + (i2 + r2);
                        double x = 0.0;
                        for (int c = 1; c < ncoefs; c++)
                            x += buf[j + off[c]] * coefs[c];
                        x = (src[i] - x) * coef;
                        f[i] = x;
                        buf[j] = x;
                    }
        } else {
            for (long i0 = 0; i0 < n[0]; i0++)
                for (long i1 = 0; i1 < n[1]; i1++)
                    for (long i2 = 0; i2 < n[2]; i2++) {
                        long i = (i0 * n[1] + i1) * n[2] + i2;
                        long j = ((i0 + r0) * M[1] + (i1 + r1)) * M[2] + (i2 + r2);
                        double x = 0.0;
                        for (int c = 1; c < ncoefs; c++)
                            x += buf[j + off[c]] * coefs[c];
                        f[i] = (1.0 - w) * f[i] + w * (src[i] - x) / coefs[0];
                    }
        }
    }
    free(buf);
    free(off);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Weighted operators (c/bmgs/wfd.cpp, c/bmgs/wrelax.cpp)                   */
/* ------------------------------------------------------------------------ */

static int wsetup(int nweights, const int* ncoefs, const long* off3,
                  const long n[3], int range, const long nb[6], int cfd,
                  int nd, halo_t* h, long M[3], long** off)
{
    if (nweights <= 0 || range < 0)
        return 1;
    long tot = 0;
    for (int w = 0; w < nweights; w++) {
        if (ncoefs[w] <= 0 || check_range(ncoefs[w], off3 + 3 * tot, range))
            return 1;
        /* the first point of every stencil is the centre */
        if (off3[3 * tot] != 0 || off3[3 * tot + 1] != 0 ||
            off3[3 * tot + 2] != 0)
            return 3;
        tot += ncoefs[w];
    }
    h->cfd = cfd;
    h->nd = nd;
    for (int i = 0; i < 3; i++) {
        h->n[i] = n[i];
        h->pad[i][0] = h->pad[i][1] = range;
        h->npad[i][0] = h->npad[i][1] = range;
        h->nb[i][0] = nb[2 * i];
        h->nb[i][1] = nb[2 * i + 1];
    }
    for (int i = 0; i < 3; i++)
        M[i] = padded_size(h, i);
    *off = flat_offsets((int)tot, off3, M);
    return 0;
}

int oracle_wfd_apply(int nweights, const int* ncoefs, const double* coefs,
                     const long* off3, const double* weights, const long n[3],
                     int range, const long nb[6], int real, int cfd,
                     const double* phases, const double* in, double* out,
                     long nin)
{
    halo_t h;
    long M[3];
    long* off;
    if (!real && phases == NULL)
        return 2;
    int st = wsetup(nweights, ncoefs, off3, n, range, nb, cfd, real ? 1 : 2,
                    &h, M, &off);
    if (st)
        return st;
    const long np = n[0] * n[1] * n[2];
    const long ng = np * h.nd;
    double* buf =
        (double*)malloc(sizeof(double) * (size_t)(M[0] * M[1] * M[2] * h.nd));
    for (long m = 0; m < nin; m++) {
        halo_fill(&h, in + m * ng, buf, real ? NULL : phases);
        double* o = out + m * ng;
        for (long i0 = 0; i0 < n[0]; i0++)
            for (long i1 = 0; i1 < n[1]; i1++)
                for (long i2 = 0; i2 < n[2]; i2++) {
                    const long i =
                        ((i0 + range) * M[1] + (i1 + range)) * M[2] + i2 + range;
                    const long j = (i0 * n[1] + i1) * n[2] + i2;
                    double xr = 0.0, xi = 0.0;
                    long c0 = 0;
                    for (int w = 0; w < nweights; w++) {
                        double tr = 0.0, ti = 0.0;
                        for (int c = 0; c < ncoefs[w]; c++) {
                            const long k = i + off[c0 + c];
                            if (real) {
                                tr += buf[k] * coefs[c0 + c];
                            } else {
                                tr += buf[2 * k] * coefs[c0 + c];
                                ti += buf[2 * k + 1] * coefs[c0 + c];
                            }
                        }
                        const double wt = weights[w * np + j];
                        tr *= wt;
                        ti *= wt;
                        xr += tr;
                        xi += ti;
                        c0 += ncoefs[w];
                    }
                    if (real) {
                        o[j] = xr;
                    } else {
                        o[2 * j] = xr;
                        o[2 * j + 1] = xi;
                    }
                }
    }
    free(buf);
    free(off);
    return 0;
}

int oracle_wfd_relax(int method, int nweights, const int* ncoefs,
                     const double* coefs, const long* off3,
                     const double* weights, const long n[3], int range,
                     const long nb[6], int cfd, double* f, const double* src,
                     int nrelax, double w)
{
    halo_t h;
    long M[3];
    long* off;
    int st = wsetup(nweights, ncoefs, off3, n, range, nb, cfd, 1, &h, M, &off);
    if (st)
        return st;
    const long np = n[0] * n[1] * n[2];
    double* buf = (double*)malloc(sizeof(double) * (size_t)(M[0] * M[1] * M[2]));
    for (int sweep = 0; sweep < nrelax; sweep++) {
        halo_fill(&h, f, buf, NULL);
        for (long i0 = 0; i0 < n[0]; i0++)
            for (long i1 = 0; i1 < n[1]; i1++)
                for (long i2 = 0; i2 < n[2]; i2++) {
                    const long i = (i0 * n[1] + i1) * n[2] + i2;
                    const long j =
                        ((i0 + range) * M[1] + (i1 + range)) * M[2] + i2 + range;
                    double x = 0.0, coef = 0.0;
                    long c0 = 0;
                    for (int q = 0; q < nweights; q++) {
                        const double wt = weights[q * np + i];
                        double t = 0.0;
                        for (int c = 1; c < ncoefs[q]; c++)
                            t += buf[j + off[c0 + c]] * coefs[c0 + c];
                        t *= wt;
                        x += t;
                        coef += wt * coefs[c0];
                        c0 += ncoefs[q];
                    }
                    if (method == 1) {
                        x = (src[i] - x) / coef;
                        f[i] = x;
                        buf[j] = x;
                    } else {
                        f[i] = (1.0 - w) * f[i] + w * (src[i] - x) / coef;
                    }
                }
    }
    free(buf);
    free(off);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Restriction and interpolation                                            */
/* ------------------------------------------------------------------------ */

/* One K-tap restriction pass over m rows of length n, writing transposed
 * (output element i of row j goes to b[i * m + j]). */
static void restrict_1d(int K, int nd, const double* a, long n, long m,
                        double* b)
{
    const long nout = (n - 2 * K + 3) / 2;
    a += (K - 1) * nd;
    for (long j = 0; j < m; j++) {
        for (int p = 0; p < nd; p++) {
            const double* aa = a + j * n * nd + p;
            double* bb = b + j * nd + p;
            for (long i = 0; i < nout; i++) {
#define A(k) aa[(k) * nd]
                if (K == 2)
                    bb[0] = 0.5 * (A(0) + 0.5 * (A(1) + A(-1)));
                else if (K == 4)
                    bb[0] = 0.5 * (A(0) + 0.5625 * (A(1) + A(-1)) +
                                   -0.0625 * (A(3) + A(-3)));
                else if (K == 6)
                    bb[0] = 0.5 * (A(0) + 0.58593750 * (A(1) + A(-1)) +
                                   -0.09765625 * (A(3) + A(-3)) +
                                   0.01171875 * (A(5) + A(-5)));
                else
                    bb[0] = 0.5 * (A(0) + 0.59814453125 * (A(1) + A(-1)) +
                                   -0.11962890625 * (A(3) + A(-3)) +
                                   0.02392578125 * (A(5) + A(-5)) +
                                   -0.00244140625 * (A(7) + A(-7)));
#undef A
                aa += 2 * nd;
                bb += m * nd;
            }
        }
    }
}

/* One K-tap interpolation pass: n coarse points per row (input row stride
 * `stride`), output rows of 2n - skip0 - skip1 points written transposed. */
static void interpolate_1d(int K, int nd, const double* a, long n, long m,
                           long stride, double* b, const long skip[2])
{
    a += (K / 2 - 1) * nd;
    for (long j = 0; j < m; j++) {
        for (int p = 0; p < nd; p++) {
            const double* aa = a + j * stride * nd + p;
            double* bb = b + j * nd + p;
            for (long i = 0; i < n; i++) {
#define A(k) aa[(k) * nd]
                if (i == 0 && skip[0])
                    bb -= m * nd;
                else
                    bb[0] = A(0);
                if (i == n - 1 && skip[1])
                    bb -= m * nd;
                else {
                    double v;
                    if (K == 2)
                        v = 0.5 * (A(0) + A(1));
                    else if (K == 4)
                        v = (0.5625 * (A(0) + A(1)) + -0.0625 * (A(-1) + A(2)));
                    else if (K == 6)
                        v = (0.58593750 * (A(0) + A(1)) +
                             -0.09765625 * (A(-1) + A(2)) +
                             0.01171875 * (A(-2) + A(3)));
                    else
                        v = (0.59814453125 * (A(0) + A(1)) +
                             -0.11962890625 * (A(-1) + A(2)) +
                             0.02392578125 * (A(-2) + A(3)) +
                             -0.00244140625 * (A(-3) + A(4)));
                    bb[m * nd] = v;
                }
#undef A
                aa += nd;
                bb += 2 * m * nd;
            }
        }
    }
}

int oracle_transform(int interpolate, int k, const long size_in[3],
                     const long size_out[3], const long pad[6],
                     const long npad[6], const long skip[6], const long nb[6],
                     int real, const double* phases, const double* in,
                     double* out, long nin, int nthreads)
{
    if (k != 2 && k != 4 && k != 6 && k != 8)
        return 1;
    if (!real && phases == NULL)
        return 2;
    halo_t h;
    h.cfd = 0;
    h.nd = real ? 1 : 2;
    for (int i = 0; i < 3; i++) {
        h.n[i] = size_in[i];
        h.pad[i][0] = pad[2 * i];
        h.pad[i][1] = pad[2 * i + 1];
        h.npad[i][0] = npad[2 * i];
        h.npad[i][1] = npad[2 * i + 1];
        h.nb[i][0] = nb[2 * i];
        h.nb[i][1] = nb[2 * i + 1];
        if (h.pad[i][0] < 0 || h.pad[i][1] < 0)
            return 3;
    }
    long M[3] = {padded_size(&h, 0), padded_size(&h, 1), padded_size(&h, 2)};
    const long nd = h.nd;
    const long ng = size_in[0] * size_in[1] * size_in[2] * nd;
    const long ngout = size_out[0] * size_out[1] * size_out[2] * nd;
    const size_t ng2 = (size_t)(M[0] * M[1] * M[2] * nd);
    long sk[3][2];
    for (int i = 0; i < 3; i++) {
        sk[i][0] = skip[2 * i];
        sk[i][1] = skip[2 * i + 1];
    }
    if (nthreads < 1)
        nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        double* buf = (double*)malloc(ng2 * sizeof(double));
        /* pass temporaries (the reference reuses `buf`, a half/16x
         * workspace and the output array; values are the same) */
        double* t1 = (double*)malloc(16 * ng2 * sizeof(double));
        double* t2 = (double*)malloc(16 * ng2 * sizeof(double));
#pragma omp for schedule(static)
        for (long m = 0; m < nin; m++) {
            halo_fill(&h, in + m * ng, buf, real ? NULL : phases);
            double* o = out + m * ngout;
            if (!interpolate) {
                const long e = 2 * k - 3;
                long nb3[3] = {(M[0] - e) / 2, (M[1] - e) / 2, (M[2] - e) / 2};
                restrict_1d(k, h.nd, buf, M[2], M[0] * M[1], t1);
                restrict_1d(k, h.nd, t1, M[1], M[0] * nb3[2], t2);
                restrict_1d(k, h.nd, t2, M[0], nb3[1] * nb3[2], o);
            } else {
                const long e = k - 1;
                long o2 = (M[2] - e) * 2 - sk[2][0] + sk[2][1];
                long o1 = (M[1] - e) * 2 - sk[1][0] + sk[1][1];
                interpolate_1d(k, h.nd, buf, M[2] - e + sk[2][1], M[0] * M[1],
                               M[2], t1, sk[2]);
                interpolate_1d(k, h.nd, t1, M[1] - e + sk[1][1], M[0] * o2,
                               M[1], t2, sk[1]);
                interpolate_1d(k, h.nd, t2, M[0] - e + sk[0][1], o1 * o2, M[0],
                               o, sk[0]);
            }
        }
        free(t2);
        free(t1);
        free(buf);
    }
    return 0;
}
