// Tile choice and dispatch of the full-row x-marching stencil kernel
// (k_xrow, rows_kernel.h): a tile spans whole z rows and TY rows in y and
// marches along x.
#include "rows.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>

namespace gfd {

static int xr_es(int dtype) { return dtype == 1 ? 16 : (dtype == 2 ? 4 : 8); }  // 3: c64
static int xr_pz(int dtype) { return 16 / xr_es(dtype); }

static bool xr_tile_ok(const XmTile& t, int R, int dtype, const int n[3])
{
    const int PZ = xr_pz(dtype);
    if (n[2] % PZ != 0 || t.tz != n[2]) return false;
    const int TZT = n[2] / PZ;
    const int NT = TZT * t.nseg;
    const int TY = t.nseg * t.py;
    const long lds = 2L * (TY + 2 * R) * n[2] * xr_es(dtype);
    return NT <= XR_MAXT && t.nseg * XR_HROWS >= 2 * R && n[1] % TY == 0 &&
           n[2] >= ((R + PZ - 1) / PZ) * PZ && R <= n[0] && R <= n[1] &&
           t.xchunk >= 1 && lds <= 150 * 1024;
}

static int g_xr_nt = -1;

void xr_set_nt(int flags) { g_xr_nt = flags; }

int xr_nt_flags(long bytes)
{
    if (g_xr_nt >= 0) return g_xr_nt;
    // non-temporal output stores for arrays far beyond the 256 MB
    // Infinity Cache (1-2 % on 256 x 96^3 f64); non-temporal loads were
    // slower in every measured case
    return bytes > (512L << 20) ? 1 : 0;
}

static std::mutex g_xr_mu;
static bool g_xr_override = false;
static XmTile g_xr_tile;

void xr_set_override(const XmTile* t)
{
    std::lock_guard<std::mutex> lk(g_xr_mu);
    g_xr_override = (t != nullptr);
    if (t) g_xr_tile = *t;
}

XmTile xr_default_tile(int R, int dtype, const int n[3], long nin)
{
    {
        std::lock_guard<std::mutex> lk(g_xr_mu);
        if (g_xr_override) return g_xr_tile;
    }
    XmTile t;
    t.tz = n[2];
    t.nseg = 0;
    t.py = 1;
    t.xchunk = n[0];
    const int PZ = xr_pz(dtype);
    const int TZT = std::max(1, n[2] / PZ);
    // rows per thread: 2R+2 queue slots of PY x 16 bytes must stay well
    // inside 256 VGPRs; one row per thread for 16-byte complex points
    const int pymax = R >= 3 ? 2 : 4;
    const int pypref = xr_es(dtype) >= 16 ? 1 : 2;
    // a segment count whose tile height divides n1, with 192-256 threads
    // per block.  Measured (full-row sweeps): 256 x 96^3 f64 nn=3 best at
    // 4 x 48 threads, 2 rows each, 24-32-plane x chunks (0.753 ms against
    // 0.885 for 8 x 48 x 1); 128 x 128^3 c128 H apply best at 2 x 128
    // threads, 1 row, 32-plane chunks (2.77 ms against 3.85 for 4 x 128).
    // Small batches (a single Poisson level, below 32M points) keep up to
    // 384 threads and one row per thread: with few tiles, occupancy wins
    // over the halo rows saved (measured on the 192^3 V-cycles).
    const bool small = (double)n[0] * n[1] * n[2] * (double)nin < 32.0e6;
    int best_score = -1 << 30;
    for (int py = pymax; py >= 1; py >>= 1)
        for (int s = (2 * R + XR_HROWS - 1) / XR_HROWS; s * TZT <= XR_MAXT;
             s++) {
            if (n[1] % (s * py) != 0) continue;
            const int nt = s * TZT;
            const int score =
                small ? 2 * std::min(nt, 384) - (nt > 384 ? 64 : 0) +
                            (s >= 2 * R ? 30 : 0) - 8 * (py - 1)
                      : -std::abs(nt - 224) - (py == pypref ? 0 : 40) +
                            (nt >= 128 ? 0 : -64);
            if (score > best_score) {
                best_score = score;
                t.nseg = s;
                t.py = py;
            }
        }
    if (t.nseg == 0) {
        t.nseg = 1;  // nothing fits: xr_tile_ok rejects it
        return t;
    }
    // x chunks of 16-32 planes, more for small batches (~16 blocks per CU
    // of 256), never below max(4R, 12) planes
    const long tiles = (long)(n[1] / (t.nseg * t.py)) * nin;
    // (measured: 16-plane chunks best for 8-byte data, 32 for c128)
    const int xtarget = xr_es(dtype) >= 16 ? 32 : 16;
    long chunks = std::max<long>(1, (n[0] + xtarget / 2) / xtarget);
    if (tiles * chunks < 4096)
        chunks = std::max<long>(chunks, (4096 + tiles - 1) / tiles);
    chunks = std::max<long>(
        1, std::min<long>(chunks, n[0] / std::max(4 * R, 12)));
    t.xchunk = (int)((n[0] + chunks - 1) / chunks);
    return t;
}

bool row_stencil_fits(const AxisStencil& st, int dtype, const int n[3],
                      long nin)
{
    if (st.R < 1 || st.R > MAXR_FAST) return false;
    return xr_tile_ok(xr_default_tile(st.R, dtype, n, nin), st.R, dtype, n);
}

long row_stencil_max_blocks(const AxisStencil& st, int dtype, const int n[3],
                            long nin)
{
    const XmTile t = xr_default_tile(st.R, dtype, n, nin);
    const int TY = t.nseg * t.py;
    return (long)((n[1] + TY - 1) / TY) * ((n[0] + t.xchunk - 1) / t.xchunk) *
           nin;
}

int launch_row_stencil(const AxisStencil& st, int dtype, const int n[3],
                       const BC& bc, const void* in, void* out, long nin,
                       int epi, const EpiArgs& ea, hipStream_t stream,
                       int* nblocks_out)
{
    const XmTile tl = xr_default_tile(st.R, dtype, n, nin);
    if (!xr_tile_ok(tl, st.R, dtype, n)) {
        set_error("x-row tile does not fit this grid");
        return ENOTSUP_;
    }
    switch (dtype) {
    case 0: return row_launch_f64(st, tl, n, bc, in, out, nin, epi, ea, stream, nblocks_out);
    case 1: return row_launch_c128(st, tl, n, bc, in, out, nin, epi, ea, stream, nblocks_out);
    case 2: return row_launch_f32(st, tl, n, bc, in, out, nin, epi, ea, stream, nblocks_out);
    case 3: return row_launch_c64(st, tl, n, bc, in, out, nin, epi, ea, stream, nblocks_out);
    }
    set_error("x-row: unsupported dtype %d", dtype);
    return ENOTSUP_;
}

}  // namespace gfd
