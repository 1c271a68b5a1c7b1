// Host-side launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace rlevox {

hipError_t launch_raycast_plain(const SceneDev& sc, const rlevox_raymap& rm, uint32_t* raybuf,
                                int32_t* clip, int block, hipStream_t s,
                                unsigned long long* counts = nullptr);
hipError_t launch_compose(const rlevox_raymap& rm, const uint32_t* raybuf, const int32_t* clip,
                          uint16_t* col, int32_t* dep, hipStream_t s);
hipError_t launch_gapfill(int W, int H, const uint16_t* cin, const int32_t* dep, uint16_t* cout,
                          hipStream_t s);

}  // namespace rlevox
