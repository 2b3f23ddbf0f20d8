// sl_kernels.h -- host-side launchers of the kernels in sl_kernels.hip.
#pragma once

#include "sl_device.h"

namespace slg {

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n, hipStream_t s);
hipError_t launch_setup(const SetupParams& p, uint64_t chunks, hipStream_t s);
hipError_t launch_clip(const SetupParams& p, uint32_t jobs, hipStream_t s);
hipError_t launch_ranges(const uint32_t* keys, uint32_t n, uint2* range, hipStream_t s);
hipError_t sort_pairs(void* temp, size_t& tempBytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                      uint32_t* vout, uint32_t n, int endBit, hipStream_t s);
hipError_t launch_raster(const RasterParams& p, int numTiles, int cfg, bool debug, hipStream_t s);

}  // namespace slg
