// k_xdma instantiations for float data.
#include "rows_dma.h"

namespace gfd {

int row_dma_launch_f32(const AxisStencil& st, const XmTile& tl,
                       const int n[3], const BC& bc, const void* in,
                       void* out, long nin, hipStream_t stream,
                       int* nblocks_out)
{
    return launch_xd<float>(st, tl, n, bc, in, out, nin, stream,
                            nblocks_out);
}

}  // namespace gfd
