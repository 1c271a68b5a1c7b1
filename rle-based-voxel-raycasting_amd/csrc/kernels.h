// Host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace rlevox {

// block: 1 = wavefront per line, 2 = two wavefronts per line, else lanes per
// block of the lane-per-line kernel.  run_cache: LDS run-cache words of the
// wavefront kernels (24, 40 or 56; 0 = 40 with two wavefronts per line, 56
// with one; the counting variant always uses 40).
hipError_t launch_raycast_plain(const SceneDev& sc, const rlevox_raymap& rm, uint32_t* raybuf,
                                int32_t* clip, int block, int run_cache, hipStream_t s,
                                unsigned long long* counts = nullptr);
hipError_t launch_compose(const rlevox_raymap& rm, const uint32_t* raybuf, const int32_t* clip,
                          uint16_t* col, int32_t* dep, hipStream_t s);
// Smallest world row of a drawn run's top over all columns of one level,
// atomically min-ed into *out (which must hold INT32_MAX or a larger row).
hipError_t launch_column_tops(const MipDev& m, int level, int32_t* out, hipStream_t s);
hipError_t launch_gapfill(int W, int H, const uint16_t* cin, const int32_t* dep, uint16_t* cout,
                          hipStream_t s);

}  // namespace rlevox
