// Full-row x-marching stencil kernels: constants shared by the tile
// logic (stencil_rows.hip) and the per-data-type launchers (rows_*.hip,
// one translation unit each so they compile in parallel).
#pragma once

#include "stencil.h"

namespace gfd {

constexpr int XR_MAXT = 512;
constexpr int XR_HROWS = 3;  // halo rows loaded per thread per plane
constexpr int XR_LOOK = 1;   // own planes loaded ahead of the stencil reach

// Non-temporal hints of the full-row kernel for arrays of `bytes`: bit 0
// output stores, bit 1 own-row loads (xr_set_nt in stencil.h).
int xr_nt_flags(long bytes);

// Launch k_xrow for an already validated tile; ENOTSUP_ for epilogues the
// data type does not instantiate.
int row_launch_f64(const AxisStencil& st, const XmTile& tl, const int n[3],
                   const BC& bc, const void* in, void* out, long nin, int epi,
                   const EpiArgs& ea, hipStream_t stream, int* nblocks_out);
int row_launch_c128(const AxisStencil& st, const XmTile& tl, const int n[3],
                   const BC& bc, const void* in, void* out, long nin, int epi,
                   const EpiArgs& ea, hipStream_t stream, int* nblocks_out);
int row_launch_f32(const AxisStencil& st, const XmTile& tl, const int n[3],
                   const BC& bc, const void* in, void* out, long nin, int epi,
                   const EpiArgs& ea, hipStream_t stream, int* nblocks_out);
int row_launch_c64(const AxisStencil& st, const XmTile& tl, const int n[3],
                   const BC& bc, const void* in, void* out, long nin, int epi,
                   const EpiArgs& ea, hipStream_t stream, int* nblocks_out);

// k_xdma (rows_dma.h): real data, EPI_PLAIN only; tile.dma planes ahead.
int row_dma_launch(const AxisStencil& st, const XmTile& tl, int dtype,
                   const int n[3], const BC& bc, const void* in, void* out,
                   long nin, hipStream_t stream, int* nblocks_out);
bool row_dma_tile_ok(const XmTile& t, int R, int dtype, const int n[3]);

}  // namespace gfd
