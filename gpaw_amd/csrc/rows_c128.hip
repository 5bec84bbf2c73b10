// k_xrow instantiations for double2 data.
#include "rows_kernel.h"

namespace gfd {

int row_launch_c128(const AxisStencil& st, const XmTile& tl, const int n[3],
                   const BC& bc, const void* in, void* out, long nin, int epi,
                   const EpiArgs& ea, hipStream_t stream, int* nblocks_out)
{
    switch (epi) {
    case EPI_PLAIN:
        return launch_xr_r<double2, EPI_PLAIN>(st, tl, n, bc, in, out, nin, ea,
                                         stream, nblocks_out);
    case EPI_RESID:
        return launch_xr_r<double2, EPI_RESID>(st, tl, n, bc, in, out, nin, ea,
                                         stream, nblocks_out);
    case EPI_UPDATE:
        return launch_xr_r<double2, EPI_UPDATE>(st, tl, n, bc, in, out, nin, ea,
                                         stream, nblocks_out);
    case EPI_POT:
        return launch_xr_r<double2, EPI_POT>(st, tl, n, bc, in, out, nin, ea,
                                         stream, nblocks_out);
    }
    set_error("x-row: epilogue %d not built for double2", epi);
    return ENOTSUP_;
}

}  // namespace gfd
