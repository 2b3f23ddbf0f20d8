// sl_raster_plain_dbg.hip -- k_raster instances: filled triangles with the parity attachments.
#include "sl_raster_impl.h"

namespace slg {

template void launch_raster_cfg<0, 1>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<1, 1>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<2, 1>(const RasterParams&, int, int, hipStream_t);

}  // namespace slg
