// k_xdma instantiations for double data (float: rows_dma32.hip).
#include "rows_dma.h"

namespace gfd {

int row_dma_launch_f32(const AxisStencil& st, const XmTile& tl,
                       const int n[3], const BC& bc, const void* in,
                       void* out, long nin, hipStream_t stream,
                       int* nblocks_out);

int row_dma_launch(const AxisStencil& st, const XmTile& tl, int dtype,
                   const int n[3], const BC& bc, const void* in, void* out,
                   long nin, hipStream_t stream, int* nblocks_out)
{
    if (dtype == 0)
        return launch_xd<double>(st, tl, n, bc, in, out, nin, stream,
                                 nblocks_out);
    if (dtype == 2)
        return row_dma_launch_f32(st, tl, n, bc, in, out, nin, stream,
                                  nblocks_out);
    set_error("x-dma: real data only");
    return ENOTSUP_;
}

bool row_dma_tile_ok(const XmTile& t, int R, int dtype, const int n[3])
{
    if (dtype != 0 && dtype != 2) return false;
    const int es = dtype == 0 ? 8 : 4;
    const int pz = 16 / es;
    return xd_lds_bytes(t, R, es, n) > 0 && n[1] % (t.nseg * t.py) == 0 &&
           (t.py >= 1 && t.py <= 4) && R <= n[0] && R <= n[1] &&
           t.xchunk >= 1 && n[2] >= ((R + pz - 1) / pz) * pz;
}

}  // namespace gfd
