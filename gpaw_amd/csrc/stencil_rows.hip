// Tile choice and dispatch of the full-row x-marching stencil kernel
// (k_xrow, rows_kernel.h): a tile spans whole z rows and TY rows in y and
// marches along x.
#include "rows.h"

#include <algorithm>
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
    // inside 256 VGPRs
    const int pymax = R >= 3 ? 2 : 4;
    // a segment count whose tile height divides n1: as many threads as
    // possible up to 384 (measured best on 96^3 x 256 grids: 8 x 48
    // threads, one row each, 0.77 ms against 0.81 for 6 x 48 and 0.92 for
    // 6 x 48 with two rows), then one halo row per thread (s >= 2R), then
    // fewer rows per thread
    int best_score = -1;
    for (int py = pymax; py >= 1; py >>= 1)
        for (int s = (2 * R + XR_HROWS - 1) / XR_HROWS; s * TZT <= XR_MAXT;
             s++) {
            if (n[1] % (s * py) != 0) continue;
            const int nt = s * TZT;
            const int score = 2 * std::min(nt, 384) - (nt > 384 ? 64 : 0) +
                              (s >= 2 * R ? 30 : 0) - 8 * (py - 1);
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
    const long tiles = (long)(n[1] / (t.nseg * t.py)) * nin;
    int chunks = 1;
    // x chunks until ~16 blocks per CU (256 CUs)
    if (tiles < 4096)
        chunks = (int)std::max<long>(
            1, std::min<long>((4096 + tiles - 1) / tiles, n[0] / (4 * R)));
    t.xchunk = (n[0] + chunks - 1) / chunks;
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
