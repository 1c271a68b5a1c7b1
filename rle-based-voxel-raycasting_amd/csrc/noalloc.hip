// rlevox_render_noalloc: a whole-frame render that allocates nothing, for
// callers that cannot call hipMalloc (code under HIP graph capture, serving
// loops that forbid allocation).  It has no counterpart in the reference.
//
// Its scratch (ray buffer, line clip ranges, gap-fill temporary) is the
// file-scope device array `scratch` below, of RLEVOX_NOALLOC_SCRATCH bytes,
// which every GPU the library runs on holds until the process ends.  In host
// code the array's name is only a stand-in: the device address comes from
// hipGetSymbolAddress, looked up once per device with that device current
// and kept per device.  Every call on a GPU shares that GPU's array, which
// keeps the last call's contents (nothing here needs it cleared: each line's
// clip range is written before compose reads the line, and compose writes
// every pixel gap-fill reads); calls that overlap on one GPU race.  The call
// is stream-ordered: it enqueues the raycast, compose and gap-fill on
// `stream` and returns without waiting, so it can be captured into a graph.
#include <hip/hip_runtime.h>

#include <mutex>

#include "api_internal.h"
#include "kernels.h"

namespace rlevox {
namespace {

__device__ __align__(256) unsigned char scratch[RLEVOX_NOALLOC_SCRATCH];

std::mutex g_mu;
void* g_scratch[64] = {nullptr};  // per device id: the array's device address

inline size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// The device address of `scratch` on the current device `dev`.
rlevox_status scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= 64) return fail(RLEVOX_ERR_INVALID_ARG, "render_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "render_noalloc: hipGetSymbolAddress");
    if (!p) return fail(RLEVOX_ERR_DEVICE, "render_noalloc: no device address for the scratch");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return RLEVOX_OK;
}

}  // namespace
}  // namespace rlevox

using namespace rlevox;

extern "C" rlevox_status rlevox_render_noalloc(rlevox_ctx* ctx, const rlevox_scene* scene,
                                               const rlevox_pose* pose, int32_t W, int32_t H,
                                               const rlevox_render_params* params,
                                               uint16_t* d_colour, int32_t* d_depth,
                                               void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!scene || !pose || !d_colour || !d_depth)
    return fail(RLEVOX_ERR_INVALID_ARG, "render_noalloc: null argument");
  if (scene->ctx != ctx)
    return fail(RLEVOX_ERR_INVALID_ARG, "render_noalloc: scene belongs to another context");
  rlevox_raymap rm;
  st = raymap_compute(*pose, W, H, params ? params->mip_distance : 0,
                      params ? params->far_distance : 0, rm);
  if (st != RLEVOX_OK) return st;
  if ((st = check_raymap(rm)) != RLEVOX_OK) return st;
  const int post = post_flags(params);
  const size_t rb = round256(raybuf_bytes(rm)), cb = round256(clip_bytes(rm));
  const size_t tb = (post & RLEVOX_POST_GAPFILL) ? round256((size_t)W * H * 2) : 0;
  if (rb + cb + tb > (size_t)RLEVOX_NOALLOC_SCRATCH)
    return fail(RLEVOX_ERR_OVER_BUDGET,
                "render_noalloc: a %dx%d view needs %zu scratch bytes, the fixed scratch holds %zu",
                W, H, rb + cb + tb, (size_t)RLEVOX_NOALLOC_SCRATCH);
  void* base = nullptr;
  if ((st = scratch_address(ctx->device, &base)) != RLEVOX_OK) return st;
  uint32_t* raybuf = (uint32_t*)base;
  int32_t* clip = (int32_t*)((char*)base + rb);
  uint16_t* cmp = tb ? (uint16_t*)((char*)base + rb + cb) : d_colour;
  rlevox_tuning t;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    t = ctx->tuning_for(scene, W, H);
  }
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = launch_raycast_plain(scene->dev, rm, raybuf, clip,
                                      ctx->tuning_block(t, rm.map_line_count, rm.rs), t.run_cache, s);
  if (e != hipSuccess) return hip_fail(e, "render_noalloc: raycast launch");
  e = launch_compose(rm, raybuf, clip, cmp, d_depth, s);
  if (e != hipSuccess) return hip_fail(e, "render_noalloc: compose launch");
  if (tb) {
    e = launch_gapfill(W, H, cmp, d_depth, d_colour, s);
    if (e != hipSuccess) return hip_fail(e, "render_noalloc: gapfill launch");
  }
  return RLEVOX_OK;
}
