// sl_raster_vis.hip -- k_raster instances: the visibility attachment of the differentiable draws.
#include "sl_raster_impl.h"

namespace slg {

template void launch_raster_cfg<2, 2>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<2, 3>(const RasterParams&, int, int, hipStream_t);

}  // namespace slg
