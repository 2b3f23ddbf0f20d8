// sl_backward.h -- host-side launchers of the kernels in sl_backward.hip.
#pragma once

#include "sl_device.h"

namespace slg {

hipError_t launch_backward(const BackwardParams& p, hipStream_t s);
hipError_t launch_loss_l2_grad(const float4* color, const float4* target, float4* grad, uint64_t n, float scale, hipStream_t s);

}  // namespace slg
