// Batched rendering of many views in three launches: a persistent raycast
// whose wavefronts pull (view, line) items from a device work queue, then
// compose and gap-fill over the pixel tiles of every view.  Each view's
// output is bit-identical to rlevox_render's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "api_internal.h"
#include "batch.h"
#include "kernels.h"
#include "wave_line.h"

namespace rlevox {

__global__ void k_batch_reset(BatchHeader* h) {
  if (threadIdx.x == 0) h->queue = 0;
  if (threadIdx.x < kTileQueues) {
    h->compose_q.w[threadIdx.x].v = 0;
    h->gapfill_q.w[threadIdx.x].v = 0;
  }
}

// Largest i with begin(i) <= g (views are ordered by begin).
template <bool TILES>
__device__ inline int find_view(const ViewDesc* __restrict__ v, int n, int64_t g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t b = TILES ? v[mid].tile_begin : v[mid].line_begin;
    if (b <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Persistent: each wavefront takes the next line of the batch until the
// queue is drained; every wavefront reaches the exit test each iteration.
// Register use is capped for four wavefronts per SIMD (16 per CU, the
// default launch): two cold values spill, and the batch runs faster than
// at the three wavefronts per SIMD its natural registers allow.  PAR: the replay prepares a column's elements across the lanes and
// visits only those that can act, as the plain kernels do; it spills more
// and pays only where columns hold many runs (served mix with three 16384^2
// scenes resident: 4165 -> 4769 Mrays/s; 64 views over 4096^2 terrains:
// 10.36 -> 10.86 ms), so a batch takes it by its scenes' runs per column
// unless the tuning says otherwise.  RC: LDS run-cache words per column.
template <bool PAR, int RC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_raycast_batch(BatchHeader* __restrict__ h,
                                                      const ViewDesc* __restrict__ views) {
  static_assert(kDdaLds * 4 <= 64 * RC * 2, "DDA scratch must fit in the run cache");
  __shared__ uint32_t bits[(RLEVOX_MAX_RS + 31) / 32];
  __shared__ uint4 rcache[64 * RC / 8];
  __shared__ int sweep[64 + kWalkEndWords];
  float* dda = reinterpret_cast<float*>(rcache);  // see kDdaLds
  const int lane = threadIdx.x;
  const int64_t total = h->total_lines;
  const int nv = h->nviews;
  int v = 0;  // a wavefront's lines only grow, so its view index only advances
  for (;;) {
    uint32_t g = 0;
    if (lane == 0) g = atomicAdd(&h->queue, 1u);
    g = (uint32_t)__builtin_amdgcn_readlane((int)g, 0);
    if ((int64_t)g >= total) break;
    while (v + 1 < nv && views[v + 1].line_begin <= (int64_t)g) v++;
    const ViewDesc& vd = views[v];
    // lines of a view from the last to the first, as the single-view kernels
    const int x = vd.rm.map_line_count - 1 - (int)((int64_t)g - vd.line_begin);
    render_line_wave<false, false, PAR, RC>(x, *vd.scene, vd.rm,
                                          vd.raybuf + (size_t)x * (size_t)vd.rm.rs,
                            vd.clip + 2 * x, bits, dda, reinterpret_cast<uint16_t*>(rcache),
                            sweep, nullptr);
  }
}

struct DeviceAdd {
  __device__ uint64_t operator()(uint64_t* p, uint64_t v) const {
    return atomicAdd((unsigned long long*)p, (unsigned long long)v);
  }
};

// The view holding `tile`, from the view v of the workgroup's previous
// tile: mostly v or the next view (the home queue's tiles only grow), a
// search otherwise (the first tile, tiles taken from other queues).
__device__ inline int view_of_tile(const ViewDesc* __restrict__ views, int nv, int v, int64_t tile) {
  if (v + 1 < nv && views[v + 1].tile_begin <= tile) {
    v++;
    if (v + 1 < nv && views[v + 1].tile_begin <= tile) v = find_view<true>(views, nv, tile);
  } else if (tile < views[v].tile_begin) {
    v = find_view<true>(views, nv, tile);
  }
  return v;
}

// The descriptor of the view a block is working on, copied to LDS when the
// block's tiles move to another view (its tiles only grow, so views change
// rarely): compose's per-tile reads of the ray map's fields then come from
// LDS (C4 batch compose 1.50 -> 1.20 ms).
__device__ inline void stage_view(const ViewDesc* __restrict__ views, int v, ViewDesc& vs) {
  static_assert(sizeof(ViewDesc) % 4 == 0 && sizeof(ViewDesc) / 4 <= 256, "one word per thread");
  __syncthreads();  // every thread is done with the previous view
  if (threadIdx.x < sizeof(ViewDesc) / 4)
    reinterpret_cast<uint32_t*>(&vs)[threadIdx.x] =
        reinterpret_cast<const uint32_t*>(views + v)[threadIdx.x];
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_compose_batch(BatchHeader* __restrict__ h,
                                                       const ViewDesc* __restrict__ views) {
  __shared__ ViewDesc vd;
  __shared__ ComposeTile ct;
  __shared__ int64_t next[2];
  const int nv = h->nviews;
  TileWalk<DeviceAdd> walk;
  if (threadIdx.x == 0) walk.start(&h->compose_q, h->total_tiles, blockIdx.x & (kTileQueues - 1));
  int v = 0, staged = -1;
  for (int it = 0;; it ^= 1) {
    // two slots: a slot is written again only after the barrier that
    // follows every thread's read of it
    if (threadIdx.x == 0) next[it] = walk.take();
    __syncthreads();
    const int64_t tile = next[it];
    if (tile < 0) break;
    v = view_of_tile(views, nv, v, tile);
    if (v != staged) {
      stage_view(views, v, vd);
      staged = v;
    }
    const int W = vd.rm.W, H = vd.rm.H;
    const int local = (int)(tile - vd.tile_begin);  // < 2^18 tiles per view
    const int tiles_x = (W + 63) / 64;
    const int ty = local / tiles_x;
    const int x0 = (local - ty * tiles_x) * 64, y0 = ty * 4;
    // the previous tile's reads of ct ended before this iteration's first
    // barrier
    compose_tile_stage(vd.rm, x0, y0, ct);
    __syncthreads();
    const int px = x0 + (threadIdx.x & 63);
    const int row = y0 + (threadIdx.x >> 6);
    if (px < W && row < H) {
      uint16_t c;
      int32_t d;
      compose_pixel(vd.rm, vd.raybuf, vd.clip, px, row, ct.col[threadIdx.x & 63],
                    ct.row[threadIdx.x >> 6], ct.border, c, d);
      uint16_t* out = (vd.post & RLEVOX_POST_GAPFILL) ? vd.ctmp : vd.colour;
      out[(size_t)row * W + px] = c;
      vd.depth[(size_t)row * W + px] = d;
    }
  }
}

__global__ __launch_bounds__(256) void k_gapfill_batch(BatchHeader* __restrict__ h,
                                                       const ViewDesc* __restrict__ views) {
  __shared__ GapTile gt;
  __shared__ int64_t next[2];
  const int nv = h->nviews;
  TileWalk<DeviceAdd> walk;
  if (threadIdx.x == 0) walk.start(&h->gapfill_q, h->total_tiles, blockIdx.x & (kTileQueues - 1));
  int v = 0;
  for (int it = 0;; it ^= 1) {
    if (threadIdx.x == 0) next[it] = walk.take();
    __syncthreads();
    const int64_t tile = next[it];
    if (tile < 0) break;
    v = view_of_tile(views, nv, v, tile);
    const ViewDesc& vd = views[v];  // a few fields per tile: read in place
    if (!(vd.post & RLEVOX_POST_GAPFILL)) continue;  // uniform over the block
    const int W = vd.rm.W;
    const int local = (int)(tile - vd.tile_begin);  // < 2^18 tiles per view
    const int tiles_x = (W + 63) / 64;
    const int ty = local / tiles_x;
    gapfill_tile(W, vd.rm.H, vd.ctmp, vd.depth, vd.colour, (local - ty * tiles_x) * 64, ty * 4,
                 gt);
  }
}

template <bool PAR>
static void launch_raycast_batch(BatchHeader* d_hdr, const ViewDesc* d_views, int grid, int rc,
                                 hipStream_t s) {
  if (rc == 24)
    hipLaunchKernelGGL((k_raycast_batch<PAR, 24>), dim3(grid), dim3(64), 0, s, d_hdr, d_views);
  else if (rc == 56)
    hipLaunchKernelGGL((k_raycast_batch<PAR, 56>), dim3(grid), dim3(64), 0, s, d_hdr, d_views);
  else
    hipLaunchKernelGGL((k_raycast_batch<PAR, 40>), dim3(grid), dim3(64), 0, s, d_hdr, d_views);
}

hipError_t launch_batch(BatchHeader* d_hdr, const ViewDesc* d_views, int raycast_grid, int post_grid,
                        bool par_elems, int run_cache, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, d_hdr);
  if (par_elems)
    launch_raycast_batch<true>(d_hdr, d_views, raycast_grid, run_cache, s);
  else
    launch_raycast_batch<false>(d_hdr, d_views, raycast_grid, run_cache, s);
  hipLaunchKernelGGL(k_compose_batch, dim3(post_grid), dim3(256), 0, s, d_hdr, d_views);
  hipLaunchKernelGGL(k_gapfill_batch, dim3(post_grid), dim3(256), 0, s, d_hdr, d_views);
  return hipGetLastError();
}

static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

rlevox_status BatchSlot::init(rlevox_ctx* c) {
  ctx = c;
  if (!stream) {
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "batch stream");
  }
  return RLEVOX_OK;
}

void BatchSlot::release() {
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  scratch.release();
  descs.release();
  if (staging) hipHostFree(staging);
  staging = nullptr;
  staging_size = 0;
  if (stream) hipStreamDestroy(stream);
  stream = nullptr;
}

rlevox_status BatchSlot::prepare(const rlevox_view* views, int n, const rlevox_render_params* params,
                                 const rlevox_tuning& tuning) {
  if (!views || n < 1 || n > kMaxViews)
    return fail(RLEVOX_ERR_INVALID_ARG, "render_batch: need 1..%d views", kMaxViews);
  const int post = post_flags(params);
  std::vector<rlevox_raymap> rms(n);
  std::vector<double> cost(n);
  for (int i = 0; i < n; i++) {
    const rlevox_view& v = views[i];
    if (!v.scene || !v.d_colour || !v.d_depth)
      return fail(RLEVOX_ERR_INVALID_ARG, "render_batch: view %d has a null pointer", i);
    if (v.scene->ctx != ctx)
      return fail(RLEVOX_ERR_INVALID_ARG, "render_batch: view %d scene belongs to another context", i);
    rlevox_status st = raymap_compute(v.pose, v.W, v.H, params ? params->mip_distance : 0,
                                      params ? params->far_distance : 0, rms[i]);
    if (st == RLEVOX_OK) st = check_raymap(rms[i]);
    if (st != RLEVOX_OK) return st;
    cost[i] = (double)rms[i].map_line_count * rms[i].rs;
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  waves = tuning.persistent_waves ? tuning.persistent_waves : kWavesPerCU;
  if (tuning.order == 1)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  // Lane-parallel element replay when some scene's columns hold more than
  // kParRuns runs on average (the generated 4096^2 terrains: 5.6-5.9; the
  // 16384^2 terrain's settings give 7.6 already at 4096^2).
  constexpr float kParRuns = 7.0f;
  par_elems = false;
  for (int i = 0; i < n; i++) par_elems = par_elems || views[i].scene->mean_runs > kParRuns;
  if (tuning.batch_par) par_elems = tuning.batch_par == 2;
  run_cache = tuning.batch_run_cache ? tuning.batch_run_cache : 40;
  // scratch layout: per view ray buffer, line clip ranges, composed colour
  size_t bytes = 0;
  std::vector<size_t> off_rb(n), off_clip(n), off_ct(n);
  for (int k = 0; k < n; k++) {
    const int i = order[k];
    off_rb[k] = bytes;
    bytes += align256(raybuf_bytes(rms[i]));
    off_clip[k] = bytes;
    bytes += align256(clip_bytes(rms[i]));
    off_ct[k] = bytes;
    bytes += (post & RLEVOX_POST_GAPFILL) ? align256((size_t)rms[i].W * rms[i].H * 2) : 0;
  }
  rlevox_status st = scratch.ensure(bytes ? bytes : 256);
  if (st != RLEVOX_OK) return st;
  const size_t dbytes = align256(sizeof(BatchHeader)) + sizeof(ViewDesc) * (size_t)kMaxViews;
  if (descs.size < dbytes) {
    // the descriptor block never moves after this, so a captured graph stays valid
    if ((st = descs.ensure(dbytes)) != RLEVOX_OK) return st;
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  if (staging_size < dbytes) {
    if (staging) hipHostFree(staging);
    staging = nullptr;
    hipError_t e = hipHostMalloc(&staging, dbytes, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(batch staging)");
    staging_size = dbytes;
  }
  BatchHeader* hh = (BatchHeader*)staging;
  ViewDesc* hv = (ViewDesc*)((char*)staging + align256(sizeof(BatchHeader)));
  char* base = (char*)scratch.ptr;
  int64_t lines = 0, tiles = 0;
  int max_rs = 16;
  rays = 0;
  for (int k = 0; k < n; k++) {
    const int i = order[k];
    ViewDesc& d = hv[k];
    std::memset(&d, 0, sizeof(d));
    d.rm = rms[i];
    d.scene = views[i].scene->d_dev;
    d.raybuf = (uint32_t*)(base + off_rb[k]);
    d.clip = (int32_t*)(base + off_clip[k]);
    d.ctmp = (post & RLEVOX_POST_GAPFILL) ? (uint16_t*)(base + off_ct[k]) : nullptr;
    d.colour = views[i].d_colour;
    d.depth = views[i].d_depth;
    d.line_begin = lines;
    d.tile_begin = tiles;
    d.post = post;
    lines += rms[i].map_line_count;
    tiles += (int64_t)((rms[i].W + 63) / 64) * ((rms[i].H + 3) / 4);
    max_rs = std::max(max_rs, rms[i].rs);
    rays += (uint64_t)rms[i].W * rms[i].H;
  }
  if (lines > 0xffffffffll) return fail(RLEVOX_ERR_INVALID_ARG, "render_batch: too many lines");
  std::memset(hh, 0, sizeof(*hh));
  hh->total_lines = lines;
  hh->total_tiles = tiles;
  hh->nviews = n;
  hh->max_words = bitmap_words(max_rs);
  nviews = n;
  return RLEVOX_OK;
}

rlevox_status BatchSlot::run(bool use_graph, hipStream_t s) {
  BatchHeader* d_hdr = (BatchHeader*)descs.ptr;
  ViewDesc* d_views = (ViewDesc*)((char*)descs.ptr + align256(sizeof(BatchHeader)));
  const size_t copy = align256(sizeof(BatchHeader)) + sizeof(ViewDesc) * (size_t)nviews;
  hipError_t e = hipMemcpyAsync(descs.ptr, staging, copy, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "batch descriptors");
  const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
  const int ray_grid = cus * waves;
  const int post_grid = cus * 8;
  if (!use_graph) {
    e = launch_batch(d_hdr, d_views, ray_grid, post_grid, par_elems, run_cache, s);
    return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "batch launch");
  }
  if (exec && (graph_waves != waves || graph_par != par_elems || graph_rc != run_cache)) {
    // the kernels or their grid changed: capture again
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  if (!exec) {
    graph_waves = waves;
    graph_par = par_elems;
    graph_rc = run_cache;
    hipStream_t cap = nullptr;
    e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t le = launch_batch(d_hdr, d_views, ray_grid, post_grid, par_elems, run_cache, cap);
      e = hipStreamEndCapture(cap, &graph);
      if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (cap) hipStreamDestroy(cap);
    if (e != hipSuccess) {
      if (graph) hipGraphDestroy(graph);
      graph = nullptr;
      exec = nullptr;
      return hip_fail(e, "batch graph capture");
    }
  }
  e = hipGraphLaunch(exec, s);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "batch graph launch");
}

static rlevox_status render_batch_impl(rlevox_ctx* ctx, const rlevox_view* views, int32_t nviews,
                                       const rlevox_render_params* params, void* stream,
                                       bool graph) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if ((st = ctx->batch.init(ctx)) != RLEVOX_OK) return st;
  if ((st = ctx->batch.prepare(views, nviews, params, ctx->tuning)) != RLEVOX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  if ((st = ctx->batch.run(graph, s)) != RLEVOX_OK) return st;
  // the slot's scratch and staging are reused by the next call
  hipError_t e = hipStreamSynchronize(s);
  return e == hipSuccess ? RLEVOX_OK : hip_fail(e, "render_batch");
}

}  // namespace rlevox

using namespace rlevox;

extern "C" rlevox_status rlevox_render_batch(rlevox_ctx* ctx, const rlevox_view* views,
                                             int32_t nviews, const rlevox_render_params* params,
                                             void* stream) {
  return render_batch_impl(ctx, views, nviews, params, stream, false);
}

extern "C" rlevox_status rlevox_render_batch_graph(rlevox_ctx* ctx, const rlevox_view* views,
                                                   int32_t nviews, const rlevox_render_params* params,
                                                   void* stream) {
  return render_batch_impl(ctx, views, nviews, params, stream, true);
}
