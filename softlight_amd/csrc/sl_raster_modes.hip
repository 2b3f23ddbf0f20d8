// sl_raster_modes.hip -- k_raster instances: batches with wireframe, line or point records.
#include "sl_raster_impl.h"

namespace slg {

template void launch_raster_cfg<0, 4>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<1, 4>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<2, 4>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_depth_cfg<1>(const RasterParams&, int, int, hipStream_t);

}  // namespace slg
