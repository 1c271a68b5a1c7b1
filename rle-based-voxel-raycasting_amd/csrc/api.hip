// C ABI of the device half: contexts, resident scenes, per-frame render and
// the device plumbing used by hosts that do not link the HIP runtime.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <vector>

#include "api_internal.h"
#include "kernels.h"

using namespace rlevox;

namespace rlevox {

// Device allocations of the library go through here so that a test can make
// the n-th one fail as if the device were out of memory.
static std::atomic<int> g_fail_alloc{0};

hipError_t dev_malloc(void** p, size_t bytes) {
  int n = g_fail_alloc.load();
  while (n > 0 && !g_fail_alloc.compare_exchange_weak(n, n - 1)) {
  }
  if (n == 1) return hipErrorOutOfMemory;
  return hipMalloc(p, bytes);
}

rlevox_status hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? RLEVOX_ERR_OUT_OF_MEMORY : RLEVOX_ERR_DEVICE,
              "%s: %s", what, hipGetErrorString(e));
}

rlevox_status ctx_bind(rlevox_ctx* ctx) {
  if (!ctx) return fail(RLEVOX_ERR_INVALID_ARG, "null context");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return RLEVOX_OK;
}

rlevox_status Scratch::ensure(size_t bytes) {
  if (bytes <= size) return RLEVOX_OK;
  if (ptr) {
    if (retire) retire->push_back(ptr);
    else hipFree(ptr);
    ptr = nullptr;
    size = 0;
    grown++;
  }
  size_t want = bytes + bytes / 4;
  hipError_t e = dev_malloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    return hip_fail(e, "hipMalloc(scratch)");
  }
  size = want;
  return RLEVOX_OK;
}

void Scratch::release() {
  if (ptr) hipFree(ptr);
  ptr = nullptr;
  size = 0;
}

}  // namespace rlevox

extern "C" {

rlevox_status rlevox_ctx_create(int32_t device, rlevox_ctx** out) {
  clear_error();
  if (!out) return fail(RLEVOX_ERR_INVALID_ARG, "ctx_create: null out");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
    return fail(RLEVOX_ERR_NO_DEVICE, "ctx_create: no HIP device (%s)",
                e == hipSuccess ? "count 0" : hipGetErrorString(e));
  if (device < 0 || device >= n)
    return fail(RLEVOX_ERR_INVALID_ARG, "ctx_create: device %d of %d", device, n);
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  rlevox_ctx* c = new rlevox_ctx();
  c->device = device;
  c->num_cus = prop.multiProcessorCount;
  c->max_lds = prop.sharedMemPerBlock;
  std::snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
  *out = c;
  return RLEVOX_OK;
}

void rlevox_ctx_destroy(rlevox_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  hipDeviceSynchronize();
  ctx->release_all();
  delete ctx;
}

rlevox_status rlevox_scene_upload(rlevox_ctx* ctx, const rlevox_mip* mips, int32_t nmips,
                                  rlevox_scene** out) {
  clear_error();
  if (!out) return fail(RLEVOX_ERR_INVALID_ARG, "scene_upload: null out");
  *out = nullptr;
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  // The pointer map may be omitted: rebuild it as RLE4::load does.
  std::vector<std::vector<uint32_t>> rebuilt(nmips > 0 ? nmips : 0);
  std::vector<rlevox_mip> m(mips, mips + (nmips > 0 && mips ? nmips : 0));
  for (int i = 0; i < (int)m.size(); i++) {
    if (!m[i].ptrmap && m[i].slabs && m[i].sx > 0 && m[i].sz > 0) {
      rebuilt[i].resize((size_t)2 * m[i].sx * m[i].sz);
      st = rlevox_build_ptrmap(m[i].slabs, m[i].nslabs, m[i].sx, m[i].sz, m[i].offset_shift,
                               rebuilt[i].data());
      if (st != RLEVOX_OK) return st;
      m[i].ptrmap = rebuilt[i].data();
    }
  }
  st = rlevox_scene_validate(m.data(), nmips);
  if (st != RLEVOX_OK) return st;
  rlevox_scene* s = new rlevox_scene();
  s->ctx = ctx;
  s->nmips = nmips;
  std::memset(&s->dev, 0, sizeof(s->dev));
  s->dev.nmips = nmips;
  for (int i = 0; i < nmips; i++) {
    const size_t map_bytes = (size_t)8 * m[i].sx * m[i].sz;
    // padded so whole-cache-line reads that start inside the array stay in it
    const size_t slab_bytes = (((size_t)m[i].nslabs * 2 + 255) & ~(size_t)255) + 512;
    void* dm = nullptr;
    void* ds = nullptr;
    hipError_t e = dev_malloc(&dm, map_bytes);
    if (e == hipSuccess) e = dev_malloc(&ds, slab_bytes);
    if (e == hipSuccess) e = hipMemcpy(dm, m[i].ptrmap, map_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ds, m[i].slabs, (size_t)m[i].nslabs * 2, hipMemcpyHostToDevice);
    if (dm) s->allocs.push_back(dm);
    if (ds) s->allocs.push_back(ds);
    if (e != hipSuccess) {
      rlevox_scene_free(s);
      return hip_fail(e, "scene_upload");
    }
    s->dev.mip[i].map = (const uint2*)dm;
    s->dev.mip[i].slabs = (const uint16_t*)ds;
    s->dev.mip[i].sx = m[i].sx;
    s->dev.mip[i].sz = m[i].sz;
    s->dev.mip[i].shift = m[i].offset_shift;
    s->sy[i] = m[i].sy;
    s->bytes += map_bytes + slab_bytes;
  }
  // Device copy of the descriptor; behind it, scratch for the highest run
  // top of each level (the early end of a line's walk needs the top of the
  // levels still to come).
  void* dd = nullptr;
  hipError_t e = dev_malloc(&dd, sizeof(SceneDev) + sizeof(int32_t) * RLEVOX_MAX_MIPS);
  if (e == hipSuccess) {
    s->allocs.push_back(dd);
    int32_t* d_top = (int32_t*)((char*)dd + sizeof(SceneDev));
    e = launch_fill_u32((uint32_t*)d_top, (uint32_t)INT32_MAX, RLEVOX_MAX_MIPS, nullptr);
    for (int i = 0; i < nmips && e == hipSuccess; i++)
      e = launch_column_tops(s->dev.mip[i], i, d_top + i, nullptr);
    int32_t top[RLEVOX_MAX_MIPS];
    if (e == hipSuccess) e = hipMemcpy(top, d_top, sizeof(top), hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
      int32_t t = m[0].sy;  // a level without drawn runs: nothing above the column bottom
      for (int i = nmips - 1; i >= 0; i--) {
        t = std::min(t, top[i]);
        s->dev.mip[i].top = t;
      }
      s->dev.sy = m[0].sy;
      e = hipMemcpy(dd, &s->dev, sizeof(SceneDev), hipMemcpyHostToDevice);
    }
  }
  if (e != hipSuccess) {
    rlevox_scene_free(s);
    return hip_fail(e, "scene_upload");
  }
  s->d_dev = (SceneDev*)dd;
  {  // runs per level-0 column, from up to 65536 evenly spaced columns
    const rlevox_mip& m0 = m[0];
    const uint64_t cols = (uint64_t)m0.sx * (uint64_t)m0.sz;
    const uint64_t step = std::max<uint64_t>(1, cols / 65536);
    uint64_t runs = 0, k = 0;
    for (uint64_t c = 0; c < cols; c += step, k++)
      runs += m0.slabs[(uint64_t)m0.ptrmap[2 * c] << m0.offset_shift];
    s->mean_runs = k ? (float)((double)runs / (double)k) : 0.0f;
  }
  *out = s;
  return RLEVOX_OK;
}

void rlevox_scene_free(rlevox_scene* s) {
  if (!s) return;
  if (s->ctx) {
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    for (auto it = s->ctx->tuned.begin(); it != s->ctx->tuned.end();)
      it = it->first.scene == s ? s->ctx->tuned.erase(it) : std::next(it);
  }
  if (s->ctx) hipSetDevice(s->ctx->device);
  hipDeviceSynchronize();
  for (void* p : s->allocs) hipFree(p);
  delete s;
}

uint64_t rlevox_scene_device_bytes(const rlevox_scene* s) { return s ? s->bytes : 0; }

rlevox_status rlevox_render_raybuf(rlevox_ctx* ctx, const rlevox_scene* scene,
                                   const rlevox_raymap* rm, uint32_t* d_raybuf,
                                   int32_t* d_line_clip, void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!scene || !rm || !d_raybuf || !d_line_clip)
    return fail(RLEVOX_ERR_INVALID_ARG, "render_raybuf: null argument");
  st = check_raymap(*rm);
  if (st != RLEVOX_OK) return st;
  rlevox_tuning t;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    t = ctx->tuning_for(scene, rm->W, rm->H);
  }
  hipError_t e = launch_raycast_plain(scene->dev, *rm, d_raybuf, d_line_clip,
                                      ctx->tuning_block(t, rm->map_line_count, rm->rs), t.run_cache,
                                      (hipStream_t)stream);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "raycast launch");
}

rlevox_status rlevox_raycast_counters(rlevox_ctx* ctx, const rlevox_scene* scene,
                                      const rlevox_raymap* rm, uint32_t* d_raybuf,
                                      int32_t* d_line_clip, rlevox_counters* out) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!scene || !rm || !d_raybuf || !d_line_clip || !out)
    return fail(RLEVOX_ERR_INVALID_ARG, "raycast_counters: null argument");
  st = check_raymap(*rm);
  if (st != RLEVOX_OK) return st;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if ((st = ctx->counts.ensure(sizeof(rlevox_counters))) != RLEVOX_OK) return st;
  unsigned long long* d = (unsigned long long*)ctx->counts.ptr;
  hipError_t e = launch_fill_u32((uint32_t*)d, 0, sizeof(rlevox_counters) / 4, nullptr);
  if (e == hipSuccess)
    e = launch_raycast_plain(scene->dev, *rm, d_raybuf, d_line_clip,
                             ctx->tuning_block(ctx->tuning_for(scene, rm->W, rm->H), rm->map_line_count,
                                              rm->rs),
                             0, nullptr, d);
  if (e == hipSuccess) e = hipMemcpy(out, d, sizeof(rlevox_counters), hipMemcpyDeviceToHost);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "raycast_counters");
}

rlevox_status rlevox_compose(rlevox_ctx* ctx, const rlevox_raymap* rm, const uint32_t* d_raybuf,
                             const int32_t* d_line_clip, uint16_t* d_colour, int32_t* d_depth,
                             void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!rm || !d_raybuf || !d_line_clip || !d_colour || !d_depth)
    return fail(RLEVOX_ERR_INVALID_ARG, "compose: null argument");
  st = check_raymap(*rm);
  if (st != RLEVOX_OK) return st;
  hipError_t e = launch_compose(*rm, d_raybuf, d_line_clip, d_colour, d_depth, (hipStream_t)stream);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "compose launch");
}

rlevox_status rlevox_gapfill(rlevox_ctx* ctx, int32_t W, int32_t H, const uint16_t* d_colour_in,
                             const int32_t* d_depth, uint16_t* d_colour_out, void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!d_colour_in || !d_depth || !d_colour_out || W < 1 || H < 1 || W > RLEVOX_MAX_RS ||
      H > RLEVOX_MAX_RS)
    return fail(RLEVOX_ERR_INVALID_ARG, "gapfill: bad argument");
  if (d_colour_in == d_colour_out)
    return fail(RLEVOX_ERR_INVALID_ARG, "gapfill: input and output colour buffers must differ");
  hipError_t e = launch_gapfill(W, H, d_colour_in, d_depth, d_colour_out, (hipStream_t)stream);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "gapfill launch");
}

rlevox_status rlevox_render(rlevox_ctx* ctx, const rlevox_scene* scene, const rlevox_pose* pose,
                            int32_t W, int32_t H, const rlevox_render_params* params,
                            uint16_t* d_colour, int32_t* d_depth, void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!scene || !pose || !d_colour || !d_depth)
    return fail(RLEVOX_ERR_INVALID_ARG, "render: null argument");
  rlevox_raymap rm;
  st = raymap_compute(*pose, W, H, params ? params->mip_distance : 0,
                      params ? params->far_distance : 0, rm);
  if (st != RLEVOX_OK) return st;
  st = check_raymap(rm);
  if (st != RLEVOX_OK) return st;
  const int post = post_flags(params);
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t rb = raybuf_bytes(rm), cb = clip_bytes(rm);
  const size_t tb = (post & RLEVOX_POST_GAPFILL) ? (size_t)W * H * 2 : 0;
  if ((st = ctx->raybuf.ensure(rb)) != RLEVOX_OK) return st;
  if ((st = ctx->clip.ensure(cb)) != RLEVOX_OK) return st;
  if (tb && (st = ctx->ctmp.ensure(tb)) != RLEVOX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  uint32_t* raybuf = (uint32_t*)ctx->raybuf.ptr;
  int32_t* clip = (int32_t*)ctx->clip.ptr;
  const rlevox_tuning t = ctx->tuning_for(scene, W, H);
  hipError_t e = launch_raycast_plain(scene->dev, rm, raybuf, clip,
                                      ctx->tuning_block(t, rm.map_line_count, rm.rs), t.run_cache, s);
  if (e != hipSuccess) return hip_fail(e, "raycast launch");
  uint16_t* cmp = tb ? (uint16_t*)ctx->ctmp.ptr : d_colour;
  e = launch_compose(rm, raybuf, clip, cmp, d_depth, s);
  if (e != hipSuccess) return hip_fail(e, "compose launch");
  if (tb) {
    e = launch_gapfill(W, H, cmp, d_depth, d_colour, s);
    if (e != hipSuccess) return hip_fail(e, "gapfill launch");
  }
  // The scratch buffers are reused by the next call on this context.
  e = hipStreamSynchronize(s);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "render");
}

rlevox_status rlevox_set_tuning(rlevox_ctx* ctx, const rlevox_tuning* t) {
  clear_error();
  if (!ctx || !t) return fail(RLEVOX_ERR_INVALID_ARG, "set_tuning: null argument");
  if (!tuning_valid(*t)) return fail(RLEVOX_ERR_INVALID_ARG, "set_tuning: value out of range");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning = *t;
  ctx->tuned.clear();
  return RLEVOX_OK;
}

rlevox_status rlevox_get_tuning(rlevox_ctx* ctx, rlevox_tuning* t) {
  clear_error();
  if (!ctx || !t) return fail(RLEVOX_ERR_INVALID_ARG, "get_tuning: null argument");
  std::lock_guard<std::mutex> lock(ctx->mu);
  *t = ctx->tuning;
  return RLEVOX_OK;
}

rlevox_status rlevox_get_tuning_for(rlevox_ctx* ctx, const rlevox_scene* scene, int32_t W,
                                    int32_t H, rlevox_tuning* t) {
  clear_error();
  if (!ctx || !scene || !t) return fail(RLEVOX_ERR_INVALID_ARG, "get_tuning_for: null argument");
  std::lock_guard<std::mutex> lock(ctx->mu);
  *t = ctx->tuning_for(scene, W, H);
  return RLEVOX_OK;
}

// ---------------------------------------------------------------- plumbing

rlevox_status rlevox_dev_alloc(rlevox_ctx* ctx, uint64_t bytes, void** out) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!out) return fail(RLEVOX_ERR_INVALID_ARG, "dev_alloc: null out");
  hipError_t e = dev_malloc(out, bytes ? bytes : 1);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "hipMalloc");
}

void rlevox_debug_fail_alloc(int32_t n) { g_fail_alloc.store(n > 0 ? n : 0); }

rlevox_status rlevox_dev_free(rlevox_ctx* ctx, void* p) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipError_t e = hipFree(p);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "hipFree");
}

rlevox_status rlevox_dev_fill_u32(rlevox_ctx* ctx, void* p, uint32_t value, uint64_t count,
                                  void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!p && count) return fail(RLEVOX_ERR_INVALID_ARG, "dev_fill: null pointer");
  hipError_t e = launch_fill_u32((uint32_t*)p, value, count, (hipStream_t)stream);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "fill");
}

rlevox_status rlevox_dev_copy(rlevox_ctx* ctx, void* dst, const void* src, uint64_t bytes,
                              void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if ((bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15))
    return fail(RLEVOX_ERR_INVALID_ARG, "dev_copy: pointers and size must be multiples of 16");
  if (bytes && (!dst || !src)) return fail(RLEVOX_ERR_INVALID_ARG, "dev_copy: null pointer");
  hipError_t e = launch_copy_u128(dst, src, bytes / 16, (hipStream_t)stream);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "copy");
}

rlevox_status rlevox_memcpy_htod(rlevox_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "memcpy htod");
}

rlevox_status rlevox_memcpy_dtoh(rlevox_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "memcpy dtoh");
}

rlevox_status rlevox_stream_create(rlevox_ctx* ctx, void** out) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipStream_t s;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  *out = (void*)s;
  return RLEVOX_OK;
}

rlevox_status rlevox_stream_destroy(rlevox_ctx* ctx, void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipError_t e = hipStreamDestroy((hipStream_t)stream);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "hipStreamDestroy");
}

rlevox_status rlevox_stream_sync(rlevox_ctx* ctx, void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "hipStreamSynchronize");
}

rlevox_status rlevox_event_create(rlevox_ctx* ctx, void** out) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipEvent_t ev;
  hipError_t e = hipEventCreate(&ev);
  if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
  *out = (void*)ev;
  return RLEVOX_OK;
}

rlevox_status rlevox_event_destroy(rlevox_ctx* ctx, void* ev) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipError_t e = hipEventDestroy((hipEvent_t)ev);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "hipEventDestroy");
}

rlevox_status rlevox_event_record(rlevox_ctx* ctx, void* ev, void* stream) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipError_t e = hipEventRecord((hipEvent_t)ev, (hipStream_t)stream);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "hipEventRecord");
}

rlevox_status rlevox_event_elapsed_ms(rlevox_ctx* ctx, void* start, void* stop, float* ms) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  hipError_t e = hipEventSynchronize((hipEvent_t)stop);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "hipEventElapsedTime");
}

}  // extern "C"
