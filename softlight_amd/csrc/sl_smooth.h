// sl_smooth.h -- host-side launchers of the kernels in sl_smooth.hip.
#pragma once

#include "sl_device.h"

namespace slg {

// Each launcher runs the boundary pass first (p.bounds: W * H float2).
hipError_t launch_smooth(const SmoothParams& p, hipStream_t s);
hipError_t launch_smooth_backward(const SmoothParams& p, hipStream_t s);

}  // namespace slg
