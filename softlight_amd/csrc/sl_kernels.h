// sl_kernels.h -- host-side launchers of the kernels in sl_kernels.hip.
#pragma once

#include "sl_device.h"

namespace slg {

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n, hipStream_t s);
// Clears the control block before a frame, keeping (and adding the last
// frame's overflow / error flags to) FrameCtrl::sticky.
hipError_t launch_vertex(const SetupParams& p, hipStream_t s);
hipError_t launch_ctrl_reset(FrameCtrl* ctrl, hipStream_t s);
hipError_t launch_setup(const SetupParams& p, uint64_t chunks, hipStream_t s);
hipError_t launch_chunk_scan(const SetupParams& p, hipStream_t s);
// tris / prims: the batch holds triangle draws / point or line draws.
hipError_t launch_emit(const SetupParams& p, bool tris, bool prims, hipStream_t s);
hipError_t launch_clip(const SetupParams& p, uint32_t jobs, hipStream_t s);
// Binning splits the entries into at most kMaxBinBlocks blocks of `per`
// entries (a multiple of 64, at least kMinBinBlock).
constexpr uint32_t kMaxBinBlocks = 512;
constexpr uint32_t kMinBinBlock = 1024;
struct BinPlan { uint32_t blocks; uint32_t per; };
BinPlan bin_plan(uint32_t entries);
// With items non-null the tile scan also writes the raster's work items of
// the heavy-tile split (see RasterItem): tiles of more than splitAbove
// records cut into parts of about partRecs records, at most maxParts parts.
hipError_t launch_binning(const uint32_t* keys, const uint32_t* vals, const FrameCtrl* ctrl, uint32_t cap,
                          BinPlan plan, uint32_t* counts, uint32_t* totals, uint2* range, uint32_t* out,
                          uint32_t* order, int numTiles, hipStream_t s, RasterItem* items = nullptr,
                          uint32_t* itemCount = nullptr, uint32_t* partDone = nullptr, uint32_t partRecs = 1,
                          uint32_t splitAbove = 0, uint32_t maxParts = 0);
// reject: the filled-triangle variants (plain and debug) with depth rejection (k_raster AUX bit 4)
hipError_t launch_raster(const RasterParams& p, int numTiles, int cfg, bool debug, hipStream_t s, bool reject = false);
// k_raster of one colour format (CF: 0 none, 1 8-bit/packed, 2 float) and
// variant (AUX bits, see k_raster) on tile configuration `cfg`; defined in
// sl_raster_impl.h and instantiated in the sl_raster_*.hip units.
template <int CF, int AUX>
void launch_raster_cfg(const RasterParams& p, int grid, int cfg, hipStream_t s);
// The heavy-tile split's first pass (k_raster MODE 1) on tile configuration
// `cfg`, over the parts of split tiles; EXTRA: the batch holds wireframe,
// line or point records.
template <int EXTRA>
void launch_raster_depth_cfg(const RasterParams& p, int grid, int cfg, hipStream_t s);
hipError_t launch_raster_depth(const RasterParams& p, int grid, int cfg, hipStream_t s);

hipError_t launch_blit(const BlitParams& p, int srcType, int dstType, hipStream_t s);

}  // namespace slg
