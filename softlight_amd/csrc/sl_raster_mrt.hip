// sl_raster_mrt.hip -- k_raster instances: batches with several fragment outputs.
#include "sl_raster_impl.h"

namespace slg {

template void launch_raster_cfg<1, 12>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<2, 12>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<1, 13>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<2, 13>(const RasterParams&, int, int, hipStream_t);

}  // namespace slg
