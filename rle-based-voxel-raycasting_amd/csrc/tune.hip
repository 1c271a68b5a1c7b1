// Autotuner: times candidate kernel configurations on sample views and keeps
// the fastest.  The reference fixes its launch shape at compile time
// (Cuda_Main.cu:163-178: 64-thread blocks, one thread per ray line); here
// the shape is a runtime choice per context, and every candidate produces
// bit-identical framebuffers, so the choice only affects speed.
#include <hip/hip_runtime.h>

#include <vector>

#include "api_internal.h"

using namespace rlevox;

namespace {

// Candidates: one or two wavefronts per line (lanes take consecutive DDA
// steps) with an LDS run cache of 24, 40 or 56 words per column, and one
// lane per line in blocks of 16..256 lanes.
const rlevox_tuning kCandidates[] = {
    {1, 1, 0, 1, 40}, {1, 2, 0, 1, 40}, {1, 1, 0, 1, 24}, {1, 2, 0, 1, 24}, {1, 1, 0, 1, 56},
    {1, 2, 0, 1, 56}, {64, 1, 0, 1},    {64, 2, 0, 1},    {64, 4, 0, 1},    {32, 1, 0, 1},
    {32, 2, 0, 1},    {16, 1, 0, 1},    {16, 4, 0, 1},
};

// The plain autotuner leaves the batched path's settings as they are.
void keep_batch_fields(rlevox_tuning& t, const rlevox_tuning& saved) {
  t.persistent_waves = saved.persistent_waves;
  t.order = saved.order;
  t.batch_run_cache = saved.batch_run_cache;
  t.batch_par = saved.batch_par;
}

rlevox_status time_views(rlevox_ctx* ctx, const rlevox_view* views, int n,
                         const rlevox_render_params* params, hipStream_t s, hipEvent_t a,
                         hipEvent_t b, int reps, double* ms) {
  hipError_t e = hipEventRecord(a, s);
  if (e != hipSuccess) return hip_fail(e, "autotune event");
  for (int r = 0; r < reps; r++)
    for (int i = 0; i < n; i++) {
      const rlevox_view& v = views[i];
      rlevox_status st = rlevox_render(ctx, v.scene, &v.pose, v.W, v.H, params, v.d_colour,
                                       v.d_depth, s);
      if (st != RLEVOX_OK) return st;
    }
  e = hipEventRecord(b, s);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
  if (e != hipSuccess) return hip_fail(e, "autotune timing");
  *ms = (double)t / reps;
  return RLEVOX_OK;
}

}  // namespace

extern "C" rlevox_status rlevox_autotune(rlevox_ctx* ctx, const rlevox_view* views, int32_t nviews,
                                         const rlevox_render_params* params, rlevox_tuning* best,
                                         double* best_ms) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!views || nviews < 1 || !best)
    return fail(RLEVOX_ERR_INVALID_ARG, "autotune: need views and an output");
  for (int i = 0; i < nviews; i++)
    if (!views[i].scene || !views[i].d_colour || !views[i].d_depth)
      return fail(RLEVOX_ERR_INVALID_ARG, "autotune: view %d has a null pointer", i);
  rlevox_tuning saved;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    saved = ctx->tuning;
  }
  // the views grouped by (scene, W, H), in order of first appearance
  std::vector<rlevox_ctx::TuneKey> keys;
  std::vector<std::vector<rlevox_view>> groups;
  for (int i = 0; i < nviews; i++) {
    const rlevox_ctx::TuneKey k{views[i].scene, views[i].W, views[i].H};
    size_t g = 0;
    while (g < keys.size() && (keys[g] < k || k < keys[g])) g++;
    if (g == keys.size()) {
      keys.push_back(k);
      groups.emplace_back();
    }
    groups[g].push_back(views[i]);
  }
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e != hipSuccess) st = hip_fail(e, "autotune setup");
  const int ncand = (int)(sizeof(kCandidates) / sizeof(kCandidates[0]));
  for (size_t g = 0; st == RLEVOX_OK && g < keys.size(); g++) {
    double best_t = 0;
    int best_i = -1;
    for (int c = 0; c < ncand; c++) {
      rlevox_tuning t = kCandidates[c];
      keep_batch_fields(t, saved);
      {
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->tuned[keys[g]] = t;
      }
      double ms = 0;
      const rlevox_view* gv = groups[g].data();
      const int gn = (int)groups[g].size();
      // one untimed pass (code-object load, scratch growth), then the timed ones
      if ((st = time_views(ctx, gv, gn, params, s, a, b, 1, &ms)) != RLEVOX_OK) break;
      if ((st = time_views(ctx, gv, gn, params, s, a, b, 2, &ms)) != RLEVOX_OK) break;
      if (best_i < 0 || ms < best_t) {
        best_t = ms;
        best_i = c;
      }
    }
    if (st != RLEVOX_OK) break;
    rlevox_tuning t = kCandidates[best_i];
    keep_batch_fields(t, saved);
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tuned[keys[g]] = t;
    if (g == 0) {
      *best = t;
      if (best_ms) *best_ms = best_t;
    }
  }
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (st == RLEVOX_OK) {
      ctx->tuning = *best;  // also the default for keys not tuned here
    } else {
      for (const auto& k : keys) ctx->tuned.erase(k);
      ctx->tuning = saved;
    }
  }
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  if (s) hipStreamDestroy(s);
  return st;
}

extern "C" rlevox_status rlevox_autotune_batch(rlevox_ctx* ctx, const rlevox_view* views,
                                               int32_t nviews, const rlevox_render_params* params,
                                               rlevox_tuning* best, double* best_ms) {
  clear_error();
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  if (!views || nviews < 1 || nviews > BatchSlot::kMaxViews || !best)
    return fail(RLEVOX_ERR_INVALID_ARG, "autotune_batch: need 1..%d views and an output",
                BatchSlot::kMaxViews);
  rlevox_tuning saved;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    saved = ctx->tuning;
  }
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e != hipSuccess) st = hip_fail(e, "autotune_batch setup");
  // the batch shape: persistent wavefronts per CU, view order, run-cache
  // depth and element replay, all combinations
  const int kWaves[] = {8, 12, 16, 24}, kOrder[] = {0, 1}, kRc[] = {24, 40, 56}, kPar[] = {1, 2};
  rlevox_tuning win = saved;
  double win_ms = 0;
  bool have = false;
  for (int w : kWaves)
    for (int o : kOrder)
      for (int rc : kRc)
        for (int par : kPar) {
          if (st != RLEVOX_OK) break;
          rlevox_tuning t = saved;
          t.persistent_waves = w;
          t.order = o;
          t.batch_run_cache = rc;
          t.batch_par = par;
          {
            std::lock_guard<std::mutex> lock(ctx->mu);
            ctx->tuning = t;
          }
          // one untimed batch (code-object load, scratch growth), then two timed
          double ms = 0;
          for (int rep = 0; rep < 3 && st == RLEVOX_OK; rep++) {
            if (rep == 1) e = hipEventRecord(a, s);
            if (e == hipSuccess) st = rlevox_render_batch(ctx, views, nviews, params, s);
            else st = hip_fail(e, "autotune_batch event");
          }
          if (st != RLEVOX_OK) break;
          e = hipEventRecord(b, s);
          if (e == hipSuccess) e = hipEventSynchronize(b);
          float tm = 0;
          if (e == hipSuccess) e = hipEventElapsedTime(&tm, a, b);
          if (e != hipSuccess) {
            st = hip_fail(e, "autotune_batch timing");
            break;
          }
          ms = (double)tm / 2;
          if (!have || ms < win_ms) {
            win = t;
            win_ms = ms;
            have = true;
          }
        }
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tuning = st == RLEVOX_OK ? win : saved;
    if (st == RLEVOX_OK) *best = win;
  }
  if (st == RLEVOX_OK && best_ms) *best_ms = win_ms;
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  if (s) hipStreamDestroy(s);
  return st;
}
