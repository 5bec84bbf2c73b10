// k_xrow instantiations for double data.
#include "rows_kernel.h"

namespace gfd {

int row_launch_f64(const AxisStencil& st, const XmTile& tl, const int n[3],
                   const BC& bc, const void* in, void* out, long nin, int epi,
                   const EpiArgs& ea, hipStream_t stream, int* nblocks_out)
{
    switch (epi) {
    case EPI_PLAIN:
        return launch_xr_r<double, EPI_PLAIN>(st, tl, n, bc, in, out, nin, ea,
                                         stream, nblocks_out);
    case EPI_RELAX:
        return launch_xr_r<double, EPI_RELAX>(st, tl, n, bc, in, out, nin, ea,
                                         stream, nblocks_out);
    case EPI_RESN:
        return launch_xr_r<double, EPI_RESN>(st, tl, n, bc, in, out, nin, ea,
                                             stream, nblocks_out);
    case EPI_RESID:
        return launch_xr_r<double, EPI_RESID>(st, tl, n, bc, in, out, nin, ea,
                                         stream, nblocks_out);
    case EPI_UPDATE:
        return launch_xr_r<double, EPI_UPDATE>(st, tl, n, bc, in, out, nin, ea,
                                         stream, nblocks_out);
    case EPI_POT:
        return launch_xr_r<double, EPI_POT>(st, tl, n, bc, in, out, nin, ea,
                                         stream, nblocks_out);
    }
    set_error("x-row: epilogue %d not built for double", epi);
    return ENOTSUP_;
}

}  // namespace gfd
