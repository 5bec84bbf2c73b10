// Tile choice and dispatch of the full-row x-marching stencil kernel
// (k_xrow, rows_kernel.h): a tile spans whole z rows and TY rows in y and
// marches along x.
#include "rows.h"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

namespace gfd {

static int xr_es(int dtype) { return dtype == 1 ? 16 : (dtype == 2 ? 4 : 8); }  // 3: c64
static int xr_pz(int dtype) { return 16 / xr_es(dtype); }

static bool xr_tile_ok(const XmTile& t, int R, int dtype, const int n[3])
{
    if (t.wide == 2 && (dtype != 1 || t.py > 2)) return false;
    const int PZ = xr_pz(dtype) * (t.wide == 2 ? 2 : 1);
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

// Smallest x chunk any tile (heuristic or tuned) uses: the block count
// bound below must hold for every tile launch_row_stencil may pick.
constexpr int XR_XCMIN = 4;

long row_stencil_max_blocks(const AxisStencil& st, int dtype, const int n[3],
                            long nin)
{
    (void)st;
    (void)dtype;
    return (long)n[1] * ((n[0] + XR_XCMIN - 1) / XR_XCMIN) * nin;
}

// ---------------------------------------------------------------- tuner
// The first launch of a (radius, type, shape, batch) times candidate tiles
// with a PLAIN pass from the call's input into its output (overwritten by
// the real pass right after) and keeps the fastest.  Not while the stream
// is being captured, nor when the output aliases an input.

struct XrKey {
    int R, dtype, n0, n1, n2;
    long nin;
    int axes;  // axes the stencil has (bit per axis): a one-axis stencil
               // runs without LDS in k_xrow and tunes differently
    bool operator<(const XrKey& o) const
    {
        if (axes != o.axes) return axes < o.axes;
        if (R != o.R) return R < o.R;
        if (dtype != o.dtype) return dtype < o.dtype;
        if (n0 != o.n0) return n0 < o.n0;
        if (n1 != o.n1) return n1 < o.n1;
        if (n2 != o.n2) return n2 < o.n2;
        return nin < o.nin;
    }
};

struct XrTuned {
    XmTile row;           // fastest k_xrow tile (every epilogue)
    XmTile dma;           // fastest k_xdma tile (plain passes), dma 0: none
};

static std::mutex g_tune_mu;
static std::map<XrKey, XrTuned> g_tuned;
static bool g_autotune = true;
static bool g_xdma = true;
// calls that must not allocate (events included) take a tile the tuner
// found earlier, or the heuristic one, on this thread
static thread_local bool t_no_tune = false;

void xr_no_tune_here(bool on) { t_no_tune = on; }

void xr_set_dma(bool on)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_xdma = on;
    g_tuned.clear();
}

void xr_set_autotune(bool on)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_autotune = on;
    if (!on) g_tuned.clear();
}

static int row_launch(const AxisStencil& st, const XmTile& tl, int dtype,
                      const int n[3], const BC& bc, const void* in, void* out,
                      long nin, int epi, const EpiArgs& ea, hipStream_t stream,
                      int* nblocks_out)
{
    if (tl.dma > 0) {
        if (epi != EPI_PLAIN) {
            set_error("x-dma: plain passes only");
            return ENOTSUP_;
        }
        return row_dma_launch(st, tl, dtype, n, bc, in, out, nin, stream,
                              nblocks_out);
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

// Every tile the tuner times for a radius, type and shape (after the
// heuristic tile, which it also times): k_xrow tiles, then (dma) k_xdma
// tiles.
std::vector<XmTile> xr_candidates(int R, int dtype, const int n[3], bool dma)
{
    const int PZ = xr_pz(dtype);
    const int TZT = std::max(1, n[2] / PZ);
    std::vector<XmTile> cand;
    const int pymax = R >= 3 ? 2 : 4;
    static const int xcs[] = {4, 8, 12, 16, 24, 32, 48, 64};
    for (int py = 1; py <= pymax; py <<= 1)
        for (int sg = (2 * R + XR_HROWS - 1) / XR_HROWS; sg * TZT <= XR_MAXT;
             sg++) {
            if (n[1] % (sg * py) != 0) continue;
            if (sg * TZT < 64 && sg * TZT < n[2] / PZ * (n[1] / py)) {
                // fewer than a wave of threads while more rows exist
                continue;
            }
            for (int xc : xcs) {
                if (xc < std::max(2 * R, XR_XCMIN) || xc > n[0]) continue;
                XmTile t;
                t.tz = n[2];
                t.nseg = sg;
                t.py = py;
                t.xchunk = xc;
                if (xr_tile_ok(t, R, dtype, n)) cand.push_back(t);
            }
            XmTile t;
            t.tz = n[2];
            t.nseg = sg;
            t.py = py;
            t.xchunk = n[0];
            if (xr_tile_ok(t, R, dtype, n)) cand.push_back(t);
        }
    if (dtype == 1 && n[2] % 2 == 0) {
        // two complex points per thread: half the threads per row, so
        // taller tiles (fewer halo rows per output row)
        const int TZW = n[2] / 2;
        for (int py = 1; py <= 2; py++)
            for (int sg = (2 * R + XR_HROWS - 1) / XR_HROWS; sg * TZW <= XR_MAXT;
                 sg++) {
                if (n[1] % (sg * py) != 0 || sg * TZW < 64) continue;
                for (int xc : {16, 32, n[0]}) {
                    if (xc < std::max(2 * R, XR_XCMIN) || xc > n[0]) continue;
                    XmTile t;
                    t.tz = n[2];
                    t.nseg = sg;
                    t.py = py;
                    t.xchunk = xc;
                    t.wide = 2;
                    if (xr_tile_ok(t, R, dtype, n)) cand.push_back(t);
                }
            }
    }
    if (dma && (dtype == 0 || dtype == 2)) {
        // LDS-DMA ring: 1-3 planes in flight, x chunks of 16-64 planes or
        // the whole grid (measured on 256 x 96^3 f64 nn=3: tall tiles and
        // whole-grid chunks win, 8 x 48 threads x 4 rows, 2 planes ahead;
        // 3 ahead keeps two blocks per CU only for slots up to 20 KB)
        static const int dxcs[] = {16, 32, 1 << 30};
        for (int py = 1; py <= 4; py++)
            for (int sg = 1; sg * TZT <= XR_MAXT; sg++) {
                if (n[1] % (sg * py) != 0 || sg * TZT < 64) continue;
                for (int A = 1; A <= 3; A++)
                    for (int xc : dxcs)
                        for (int order = 0; order <= 2; order += 2) {
                            if (xc >= n[0] && xc != dxcs[2]) continue;
                            XmTile t;
                            t.tz = n[2];
                            t.nseg = sg;
                            t.py = py;
                            t.xchunk = std::min(xc, n[0]);
                            t.dma = A;
                            t.order = order;
                            if (row_dma_tile_ok(t, R, dtype, n))
                                cand.push_back(t);
                        }
            }
    }
    return cand;
}

static XrTuned xr_tune(const AxisStencil& st, int dtype, const int n[3],
                       const BC& bc, const void* in, void* out, long nin,
                       hipStream_t stream, const XmTile& heur, bool dma)
{
    const int R = st.R;
    // time on at most ~32M points of the batch (the ranking of tiles
    // does not change beyond a few waves per CU)
    const long per = (long)n[0] * n[1] * n[2];
    const long nt_in = std::max(1L, std::min(nin, (32L << 20) / per));
    std::vector<XmTile> cand;
    cand.push_back(heur);
    for (const XmTile& t : xr_candidates(R, dtype, n, dma)) cand.push_back(t);
    XrTuned res;
    res.row = heur;
    res.dma = heur;
    res.dma.dma = 0;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return res;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return res;
    }
    EpiArgs ea{};
    ea.c0 = st.c0;
    float best_ms = 1e30f, best_dma_ms = 1e30f;
    // every candidate on the same grids, enough for at least 1024 blocks
    // (4 per CU) of the candidate with the fewest blocks per grid where the
    // batch has them: such tiles would otherwise be timed on a partly idle
    // device
    long min_per_grid = 1L << 40;
    for (const XmTile& t : cand)
        min_per_grid = std::min(min_per_grid,
                                (long)(n[1] / (t.nseg * t.py)) *
                                    ((n[0] + t.xchunk - 1) / t.xchunk));
    const long nt_c = std::min(
        nin, std::max(nt_in, (1024 + min_per_grid - 1) /
                                 std::max(1L, min_per_grid)));
    for (const XmTile& t : cand) {
        // one untimed pass, then two timed
        if (row_launch(st, t, dtype, n, bc, in, out, nt_c, EPI_PLAIN, ea,
                       stream, nullptr) != OK)
            continue;
        (void)hipEventRecord(e0, stream);
        int st2 = OK;
        for (int r = 0; r < 2 && st2 == OK; r++)
            st2 = row_launch(st, t, dtype, n, bc, in, out, nt_c, EPI_PLAIN,
                             ea, stream, nullptr);
        (void)hipEventRecord(e1, stream);
        if (hipEventSynchronize(e1) != hipSuccess || st2 != OK) {
            (void)hipGetLastError();
            continue;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (t.dma == 0 && ms < best_ms) {
            best_ms = ms;
            res.row = t;
        }
        if (t.dma > 0 && ms < best_dma_ms) {
            best_dma_ms = ms;
            res.dma = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_error("");
    // the DMA tile serves plain passes only where it is faster
    if (res.dma.dma > 0 && !(best_dma_ms < best_ms)) res.dma.dma = 0;
    return res;
}

static XmTile xr_pick_tile(const AxisStencil& st, int dtype, const int n[3],
                           const BC& bc, const void* in, void* out, long nin,
                           int epi, const EpiArgs& ea, hipStream_t stream)
{
    const bool plain = epi == EPI_PLAIN;
    auto choose = [&](const XrTuned& t) {
        return plain && t.dma.dma > 0 ? t.dma : t.row;
    };
    {
        std::lock_guard<std::mutex> lk(g_xr_mu);
        if (g_xr_override) {
            XmTile t = g_xr_tile;
            if (!plain) t.dma = 0;
            return t;
        }
    }
    const XmTile heur = xr_default_tile(st.R, dtype, n, nin);
    const XrKey key{st.R, dtype, n[0], n[1], n[2], nin,
                    st.has[0] | st.has[1] << 1 | st.has[2] << 2};
    bool dma;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        if (!g_autotune) return heur;
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) return choose(it->second);
        if (t_no_tune) return heur;
        dma = g_xdma;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return heur;
    }
    // candidates are timed as PLAIN passes of the call's dtype into `out`:
    // not where out holds another element size (RESN writes FP32 from FP64
    // input, UPDW FP64 from FP32 compute) or aliases an input
    if (cs != hipStreamCaptureStatusNone || in == out ||
        (ea.aux && ea.aux == out) || (ea.pot && (const void*)ea.pot == out) ||
        epi == EPI_RESN || epi == EPI_UPDW)
        return heur;
    const XrTuned t =
        xr_tune(st, dtype, n, bc, in, out, nin, stream, heur, dma);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = t;
    return choose(t);
}

int launch_row_stencil(const AxisStencil& st, int dtype, const int n[3],
                       const BC& bc, const void* in, void* out, long nin,
                       int epi, const EpiArgs& ea, hipStream_t stream,
                       int* nblocks_out)
{
    const XmTile tl =
        xr_pick_tile(st, dtype, n, bc, in, out, nin, epi, ea, stream);
    if (tl.dma > 0 ? !row_dma_tile_ok(tl, st.R, dtype, n)
                   : !xr_tile_ok(tl, st.R, dtype, n)) {
        set_error("x-row tile does not fit this grid");
        return ENOTSUP_;
    }
    return row_launch(st, tl, dtype, n, bc, in, out, nin, epi, ea, stream,
                      nblocks_out);
}

}  // namespace gfd
