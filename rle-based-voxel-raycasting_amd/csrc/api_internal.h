// Context and scene objects behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <vector>

#include "device.h"
#include "internal.h"

namespace rlevox {

// A device buffer that only grows.
struct Scratch {
  void* ptr = nullptr;
  size_t size = 0;
  // When set, a buffer replaced by a larger one goes to this list instead of
  // hipFree (which waits for the whole device): its owner frees the list
  // when nothing is in flight.
  std::vector<void*>* retire = nullptr;
  uint64_t grown = 0;  // replacements by a larger buffer
  rlevox_status ensure(size_t bytes);
  void release();
};

rlevox_status hip_fail(hipError_t e, const char* what);
hipError_t dev_malloc(void** p, size_t bytes);

inline bool tuning_valid(const rlevox_tuning& t) {
  return (t.lines_per_wave == 1 || t.lines_per_wave == 16 || t.lines_per_wave == 32 ||
          t.lines_per_wave == 64) &&
         t.waves_per_block >= (t.lines_per_wave == 1 ? 0 : 1) && t.waves_per_block <= 4 &&
         (t.persistent_waves == 0 || (t.persistent_waves >= 4 && t.persistent_waves <= 32)) &&
         (t.order == 0 || t.order == 1) &&
         (t.run_cache == 0 || t.run_cache == 24 || t.run_cache == 40 || t.run_cache == 56) &&
         (t.batch_run_cache == 0 || t.batch_run_cache == 24 || t.batch_run_cache == 40 ||
          t.batch_run_cache == 56) &&
         (t.batch_par >= 0 && t.batch_par <= 2);
}

inline int post_flags(const rlevox_render_params* p) {
  return (p && p->post) ? p->post : (RLEVOX_POST_COMPOSE | RLEVOX_POST_GAPFILL);
}

inline size_t raybuf_bytes(const rlevox_raymap& rm) {
  return (size_t)(rm.map_line_count > 0 ? rm.map_line_count : 1) * (size_t)rm.rs * 4;
}
inline size_t clip_bytes(const rlevox_raymap& rm) {
  return (size_t)(rm.map_line_count > 0 ? rm.map_line_count : 1) * 8;
}

// A caller-supplied parameter block must keep every kernel loop bounded and
// every access inside the buffers sized from it.
inline rlevox_status check_raymap(const rlevox_raymap& rm) {
  if (rm.W < 16 || rm.H < 16 || rm.H > rm.W || rm.W > RLEVOX_MAX_RS || rm.rs != rm.W ||
      rm.maxres != rm.rs)
    return fail(RLEVOX_ERR_INVALID_ARG, "raymap: bad size fields");
  int64_t sum = 0;
  for (int j = 0; j < 4; j++) {
    if (rm.res[j] < 0 || rm.res[j] > 3 * rm.rs)
      return fail(RLEVOX_ERR_INVALID_ARG, "raymap: res[%d] out of range", j);
    sum += rm.res[j];
  }
  if (sum != rm.map_line_count) return fail(RLEVOX_ERR_INVALID_ARG, "raymap: line count mismatch");
  if (rm.mapswitch < 1 || rm.mapswitch > (1 << 24) || rm.far_distance < 1 ||
      rm.far_distance > (1 << 22) || !mapswitch_bounded(rm.mapswitch, rm.position[1]))
    return fail(RLEVOX_ERR_INVALID_ARG, "raymap: mip/far distance out of range");
  return RLEVOX_OK;
}

rlevox_status ctx_bind(rlevox_ctx* ctx);

// Device state of one in-flight batch: ray buffers, descriptors, pinned
// staging and the captured graph (replayed for any batch: the kernels read
// the batch shape from the descriptor block, whose address never changes).
struct BatchSlot {
  static constexpr int kMaxViews = 4096;
  static constexpr int kWavesPerCU = 16;
  rlevox_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  Scratch scratch, descs;
  void* staging = nullptr;
  size_t staging_size = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nviews = 0;
  int waves = kWavesPerCU;        // persistent wavefronts per CU of this batch
  bool par_elems = false;         // raycast with the lane-parallel element replay
  int run_cache = 40;             // LDS run-cache words of the raycast
  int graph_waves = 0;            // the values the captured graph was made with
  bool graph_par = false;
  int graph_rc = 0;
  uint64_t rays = 0;
  rlevox_status init(rlevox_ctx* c);
  rlevox_status prepare(const rlevox_view* views, int n, const rlevox_render_params* params,
                        const rlevox_tuning& tuning);
  rlevox_status run(bool use_graph, hipStream_t s);
  void release();
};

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t count, hipStream_t s);
hipError_t launch_copy_u128(void* dst, const void* src, uint64_t n16, hipStream_t s);

}  // namespace rlevox

struct rlevox_ctx {
  int device = 0;
  int num_cus = 0;
  size_t max_lds = 65536;  // LDS bytes one workgroup may take
  char arch[64] = {0};
  std::mutex mu;  // guards the scratch buffers and tuning
  rlevox::Scratch raybuf, clip, ctmp, counts;
  rlevox::BatchSlot batch;
  rlevox_tuning tuning{1, 0, 0, 1, 0, 0, 0};
  // Autotuned settings per (scene, W, H); others use `tuning`.
  struct TuneKey {
    const rlevox_scene* scene;
    int W, H;
    bool operator<(const TuneKey& o) const {
      if (scene != o.scene) return scene < o.scene;
      if (W != o.W) return W < o.W;
      return H < o.H;
    }
  };
  std::map<TuneKey, rlevox_tuning> tuned;
  // The tuning of a single render (caller holds mu).
  rlevox_tuning tuning_for(const rlevox_scene* s, int W, int H) const {
    auto it = tuned.find(TuneKey{s, W, H});
    return it == tuned.end() ? tuning : it->second;
  }
  // Widest ray-buffer line (screen width) up to which the automatic choice
  // walks each line with two wavefronts, one on the DDA and one drawing.
  // Measured over the 100-pose flythrough of the 4096^2 terrain (mean
  // raycast ms, two wavefronts / one with a 56-word run cache): 640x480
  // 0.091 / 0.122, 1280x720 0.100 / 0.128, 1920x1080 0.118 / 0.141,
  // 2048x1152 0.124 / 0.147, 2560x1440 0.142 / 0.160 (two better at every
  // line count, up to 4096 and more lines), 3200x1800 0.210 / 0.186,
  // 3840x2160 0.251 / 0.229.
  static constexpr int kTwoWaveMaxWidth = 2560;
  // 1 selects the wavefront-per-line kernel, 2 its two-wavefront form;
  // otherwise lanes per block (16 or more), halved until the block's
  // per-lane bitmaps (ceil(rs/32) words each) fit the workgroup's LDS.
  int tuning_block(const rlevox_tuning& t, int lines, int rs) const {
    if (t.lines_per_wave == 1) {
      if (t.waves_per_block == 0) return rs <= kTwoWaveMaxWidth ? 2 : 1;
      return t.waves_per_block >= 2 ? 2 : 1;
    }
    int block = t.lines_per_wave * t.waves_per_block;
    const size_t lane_bytes = (size_t)rlevox::bitmap_words(rs) * sizeof(uint32_t);
    while (block > 16 && (size_t)block * lane_bytes > max_lds) block >>= 1;
    return block;
  }
  void release_all() {
    raybuf.release();
    clip.release();
    ctmp.release();
    counts.release();
    batch.release();
  }
};

struct rlevox_scene {
  rlevox_ctx* ctx = nullptr;
  int nmips = 0;
  int sy[RLEVOX_MAX_MIPS] = {0};
  rlevox::SceneDev dev;
  rlevox::SceneDev* d_dev = nullptr;  // device copy, read by batched kernels
  std::vector<void*> allocs;
  uint64_t bytes = 0;
  float mean_runs = 0;  // runs per level-0 column, from a sample taken at upload
};
