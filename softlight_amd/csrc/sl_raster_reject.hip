// sl_raster_reject.hip -- k_raster instances: filled triangles with per-wave depth rejection (AUX bit 4).
#include "sl_raster_impl.h"

namespace slg {

template void launch_raster_cfg<0, 16>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<1, 16>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<2, 16>(const RasterParams&, int, int, hipStream_t);
// with the debug buffers (AUX bit 0): the parity tests' per-pixel order checks
template void launch_raster_cfg<0, 17>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<1, 17>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<2, 17>(const RasterParams&, int, int, hipStream_t);

}  // namespace slg
