// sl_raster_modes_dbg.hip -- k_raster instances: batches with wireframe, line or point records, with the parity attachments.
#include "sl_raster_impl.h"

namespace slg {

template void launch_raster_cfg<0, 5>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<1, 5>(const RasterParams&, int, int, hipStream_t);
template void launch_raster_cfg<2, 5>(const RasterParams&, int, int, hipStream_t);

}  // namespace slg
