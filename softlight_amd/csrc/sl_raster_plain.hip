// sl_raster_plain.hip -- k_raster instances: filled triangles, no parity attachments (the timed variants).
#include "sl_raster_impl.h"

namespace slg {

template void launch_raster_cfg<0, 0>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<1, 0>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<2, 0>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_depth_cfg<0>(const RasterParams&, int, int, hipStream_t);

}  // namespace slg
