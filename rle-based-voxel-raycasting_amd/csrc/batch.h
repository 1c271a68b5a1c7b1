// Batched rendering: many views (any scene, any resolution) per launch.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"
#include "tile_queue.h"

namespace rlevox {

// One view of a batch, in device memory.
struct ViewDesc {
  rlevox_raymap rm;
  const SceneDev* scene;  // device copy of the view's scene descriptor
  uint32_t* raybuf;       // map_line_count * rs texels
  int32_t* clip;          // 2 per line
  uint16_t* ctmp;         // W*H composed colour before gap-fill
  uint16_t* colour;       // W*H output colour
  int32_t* depth;         // W*H output depth
  int64_t line_begin;     // first global line index of this view
  int64_t tile_begin;     // first global 64x4 pixel tile of this view
  int32_t post;           // RLEVOX_POST_* bits
  int32_t pad;
};

// Batch header, read by the kernels so a captured graph fits any batch.
struct BatchHeader {
  int64_t total_lines;
  int64_t total_tiles;
  int32_t nviews;
  int32_t max_words;  // bitmap words of the largest view
  uint32_t queue;     // next line to hand out (reset by k_batch_reset)
  uint32_t pad;
  TileQueueWords compose_q;  // tile queues (reset by k_batch_reset)
  TileQueueWords gapfill_q;
};

// par_elems: the raycast replays a column's elements with the lane-parallel
// preparation (the plain kernels' replay) instead of one by one; run_cache:
// LDS run-cache words per candidate column (24, 40 or 56).
hipError_t launch_batch(BatchHeader* d_hdr, const ViewDesc* d_views, int raycast_grid, int post_grid,
                        bool par_elems, int run_cache, hipStream_t s);

}  // namespace rlevox
