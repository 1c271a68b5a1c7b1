// Internal declarations shared by the host and device halves of librlevox.
#pragma once

#include <cstdint>
#include <string>
#include <thread>
#include <vector>
#include <functional>

#include "../../include/rlevox.h"

namespace rlevox {

// Records a message for rlevox_last_error() and returns `st`.
rlevox_status fail(rlevox_status st, const char* fmt, ...);
void clear_error();

// Runs body(i) for i in [0, n) on `threads` host threads (0 = hardware
// concurrency), in chunks handed out dynamically.
void parallel_for(int64_t n, int threads, int64_t chunk, const std::function<void(int64_t)>& body);

rlevox_status build_ptrmap(const uint16_t* slabs, uint64_t nslabs, int sx, int sz, int shift,
                           uint32_t* out);

// The mip-switch loop (Cuda_Render.h:343-365) doubles the int mapswitch
// once per halving of the camera row above 512, then while the distance
// walked (at most the far distance, <= 2^22) exceeds it.  Poses whose
// doublings would pass 2^30 are rejected, so the int never overflows (an
// overflowed mapswitch would keep the loop going forever).
inline bool mapswitch_bounded(int mapswitch, float camera_row) {
  int n = 0;
  float y = camera_row;
  while (y > 512.0f && n < 40) {
    y = y * 0.5f;
    n++;
  }
  return mapswitch >= 1 && n < 40 && ((int64_t)mapswitch << n) <= (1ll << 30);
}

struct HostLevel {
  int32_t sx = 0, sy = 0, sz = 0;
  int32_t shift = 0;  // offset_shift of the level
  std::vector<uint32_t> ptrmap;
  std::vector<uint16_t> slabs;
};

// Saturating float->int with NaN->0: the semantics of gfx950's v_cvt_i32_f32
// and of CUDA's cvt.rzi.s32.f32, written out so host code agrees with it.
inline int f2i(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f < -2147483648.0f) return INT32_MIN;
  return (int)f;
}

rlevox_status raymap_compute(const rlevox_pose& pose, int W, int H, int mip_distance,
                             int far_distance, rlevox_raymap& out);

}  // namespace rlevox

struct rlevox_hscene {
  int32_t nmips = 0;
  rlevox::HostLevel lv[RLEVOX_MAX_MIPS];
};
