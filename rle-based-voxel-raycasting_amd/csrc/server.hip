// Served request loop: continuous batching of render requests over resident
// scenes.
//
// The reference renders one view per frame, synchronously, from its display
// callback (Cuda_Main.cu:183-271, main.cpp:769-873).  Here many requests of
// any size, scene and pose arrive asynchronously; a worker thread packs the
// queued ones into batches and keeps up to `streams` batches in flight at
// once, each on its own lane:
//
//   lane = a BatchSlot (own stream, ray-buffer scratch, descriptor block and
//          captured hipGraph of reset + raycast + compose + gap-fill),
//          a device framebuffer arena and a pinned host staging area.
//
// A batch completes as a whole, so a small request packed with a large one
// waits for it.  With large_lanes > 0 the first lanes take only large
// requests (W*H >= large_pixels), with a persistent raycast of large_waves
// wavefronts per CU so the GPU keeps room for the others, and the remaining
// lanes take only the smaller requests; batch_pixels caps the pixels of one
// batch, which bounds how long it holds its lane.
//
// Admission is by device memory: a request holds its ray buffer, line clip
// ranges, gap-fill temporary and framebuffers from launch to completion, and
// a batch only takes requests while the bytes in flight stay within
// memory_budget.  Requests that arrive while batches run join the next batch
// on the next free lane, so their work overlaps the running ones on the
// device.  A request either names its own device framebuffer (device
// delivery: rendered in place) or is rendered into the lane's arena and
// copied to host memory (host delivery).  Completion is signalled by a host
// function enqueued behind each batch's kernels and copies; latency runs
// from submit to that point.
//
// Every framebuffer is the one rlevox_render produces for the same request:
// the batch kernels run the same per-line and per-pixel code.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "api_internal.h"

namespace rlevox {
namespace {

using Clock = std::chrono::steady_clock;

inline double ms_since(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

enum ReqState { QUEUED, RUNNING, DONE, FAILED, CANCELLED };

struct Req {
  uint64_t ticket = 0;
  const rlevox_scene* scene = nullptr;
  rlevox_pose pose{};
  int W = 0, H = 0;
  bool large = false;            // served by the large lanes (large_lanes > 0)
  uint16_t* d_colour = nullptr;  // caller's device framebuffer (device delivery)
  int32_t* d_depth = nullptr;
  size_t bytes = 0;  // device bytes held while in flight
  Clock::time_point t_arrival, t_done;
  ReqState state = QUEUED;
  bool cancel = false;
  rlevox_status status = RLEVOX_OK;
  std::string error;
  std::vector<uint16_t> colour;
  std::vector<int32_t> depth;
};

// Device bytes a request holds from launch to completion: its ray buffer,
// line clip ranges, gap-fill temporary and, for host delivery, the
// framebuffer in the lane's arena.
inline size_t request_bytes(const rlevox_raymap& rm, int post, bool host_delivery) {
  const size_t px = (size_t)rm.W * rm.H;
  size_t b = align256(raybuf_bytes(rm)) + align256(clip_bytes(rm));
  if (post & RLEVOX_POST_GAPFILL) b += align256(px * 2);
  if (host_delivery) b += align256(px * 2) + align256(px * 4);
  return b;
}

struct Lane;

struct Completion {
  rlevox_server* srv;
  Lane* lane;
};

struct Lane {
  BatchSlot slot;
  Scratch fb;  // colour + depth of the batch's host-delivered requests
  void* pinned = nullptr;
  size_t pinned_size = 0;
  // Buffers replaced by larger ones while other lanes may have work in
  // flight: freed when no batch is in flight (hipFree and hipHostFree wait
  // for the whole device, which would stall every lane).
  std::vector<void*> retired_dev, retired_host;
  uint64_t pinned_grown = 0;
  std::vector<Req*> reqs;
  std::vector<size_t> fb_off;  // per host-delivered request: colour at off, depth after it
  size_t bytes = 0;
  bool large = false;     // takes only large requests (else only the others)
  bool busy = false;
  bool finished = false;  // set by the completion host function
  Completion done{nullptr, nullptr};
};

}  // namespace
}  // namespace rlevox

using namespace rlevox;

struct rlevox_server {
  rlevox_ctx* ctx = nullptr;
  rlevox_server_config cfg{};
  int post = 0;
  std::mutex mu;
  std::condition_variable cv_work;  // worker: queue, completions, stop
  std::condition_variable cv_done;  // waiters: a request finished
  std::deque<Req*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  uint64_t next_ticket = 1;
  std::vector<std::unique_ptr<Lane>> lanes;
  size_t inflight_bytes = 0;
  int inflight_batches = 0;
  Clock::time_point busy_since;
  rlevox_server_stats stats{};
  bool stop = false;
  std::string fatal;  // a device error that stopped the loop
  std::thread worker;
};

namespace rlevox {
namespace {

// Host function behind a batch: no HIP calls.  It notifies while holding the
// lock, so once the worker has seen `finished` this function no longer
// touches the server (which the worker may then tear down).
void on_batch_done(void* p) {
  Completion* c = (Completion*)p;
  std::lock_guard<std::mutex> lock(c->srv->mu);
  c->lane->finished = true;
  c->srv->cv_work.notify_one();
}

// Packs and launches one batch on `lane`.  Called by the worker without the
// server lock; the lane and its requests are owned by the worker meanwhile.
rlevox_status launch_lane(rlevox_server* srv, Lane& lane, const rlevox_tuning& tuning) {
  const int n = (int)lane.reqs.size();
  size_t fb_bytes = 0;
  lane.fb_off.resize(n);
  for (int i = 0; i < n; i++) {
    const Req* r = lane.reqs[i];
    const size_t px = (size_t)r->W * r->H;
    lane.fb_off[i] = fb_bytes;
    if (!r->d_colour) fb_bytes += align256(px * 2) + align256(px * 4);
  }
  rlevox_status st = fb_bytes ? lane.fb.ensure(fb_bytes) : RLEVOX_OK;
  if (st != RLEVOX_OK) return st;
  if (fb_bytes && lane.pinned_size < fb_bytes) {
    if (lane.pinned) {
      lane.retired_host.push_back(lane.pinned);
      lane.pinned_grown++;
    }
    lane.pinned = nullptr;
    lane.pinned_size = 0;
    const size_t want = fb_bytes + fb_bytes / 4;
    hipError_t e = hipHostMalloc(&lane.pinned, want, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(server staging)");
    lane.pinned_size = want;
  }
  std::vector<rlevox_view> views(n);
  char* fb = (char*)lane.fb.ptr;
  for (int i = 0; i < n; i++) {
    const Req* r = lane.reqs[i];
    rlevox_view& v = views[i];
    v.scene = r->scene;
    v.pose = r->pose;
    v.W = r->W;
    v.H = r->H;
    if (r->d_colour) {
      v.d_colour = r->d_colour;
      v.d_depth = r->d_depth;
    } else {
      v.d_colour = (uint16_t*)(fb + lane.fb_off[i]);
      v.d_depth = (int32_t*)(fb + lane.fb_off[i] + align256((size_t)r->W * r->H * 2));
    }
  }
  if ((st = lane.slot.init(srv->ctx)) != RLEVOX_OK) return st;
  if ((st = lane.slot.prepare(views.data(), n, &srv->cfg.params, tuning)) != RLEVOX_OK) return st;
  hipStream_t s = lane.slot.stream;
  if ((st = lane.slot.run(true, s)) != RLEVOX_OK) return st;
  hipError_t e = hipSuccess;
  if (fb_bytes) e = hipMemcpyAsync(lane.pinned, lane.fb.ptr, fb_bytes, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return hip_fail(e, "server framebuffer copy");
  lane.done = Completion{srv, &lane};
  e = hipLaunchHostFunc(s, on_batch_done, &lane.done);
  if (e != hipSuccess) return hip_fail(e, "server completion callback");
  return RLEVOX_OK;
}

// Moves a finished lane's framebuffers to its requests.  Server lock held.
void finish_lane(rlevox_server* srv, Lane& lane) {
  const auto now = Clock::now();
  const char* base = (const char*)lane.pinned;
  for (size_t i = 0; i < lane.reqs.size(); i++) {
    Req* r = lane.reqs[i];
    const size_t px = (size_t)r->W * r->H;
    r->t_done = now;
    if (r->cancel) {
      r->state = CANCELLED;
      continue;
    }
    if (!r->d_colour) {
      r->colour.resize(px);
      r->depth.resize(px);
      std::memcpy(r->colour.data(), base + lane.fb_off[i], px * 2);
      std::memcpy(r->depth.data(), base + lane.fb_off[i] + align256(px * 2), px * 4);
    }
    r->state = DONE;
    srv->stats.completed++;
    srv->stats.rays += px;
  }
  srv->inflight_bytes -= lane.bytes;
  lane.bytes = 0;
  lane.reqs.clear();
  lane.busy = false;
  lane.finished = false;
  if (--srv->inflight_batches == 0)
    srv->stats.busy_seconds += ms_since(srv->busy_since, now) * 1e-3;
}

// Frees the buffers lanes replaced by larger ones.  Called by the worker
// with no batch in flight (or at destroy).
void free_retired(rlevox_server* srv) {
  for (auto& lp : srv->lanes) {
    for (void* p : lp->retired_dev) hipFree(p);
    for (void* p : lp->retired_host) hipHostFree(p);
    lp->retired_dev.clear();
    lp->retired_host.clear();
  }
}

void fail_requests(rlevox_server* srv, const std::vector<Req*>& rs, rlevox_status st,
                   const std::string& msg) {
  for (Req* r : rs) {
    r->state = FAILED;
    r->status = st;
    r->error = msg;
    r->t_done = Clock::now();
  }
  srv->cv_done.notify_all();
}

void worker_main(rlevox_server* srv) {
  if (ctx_bind(srv->ctx) != RLEVOX_OK) {
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->fatal = "server: cannot bind the device: " + std::string(rlevox_last_error());
    std::vector<Req*> rest(srv->queue.begin(), srv->queue.end());
    srv->queue.clear();
    fail_requests(srv, rest, RLEVOX_ERR_DEVICE, srv->fatal);
    srv->cv_done.notify_all();
    return;
  }
  std::unique_lock<std::mutex> lock(srv->mu);
  for (;;) {
    bool progress = false;
    for (auto& lp : srv->lanes) {
      if (lp->busy && lp->finished) {
        finish_lane(srv, *lp);
        progress = true;
      }
    }
    if (progress) srv->cv_done.notify_all();
    if (srv->inflight_batches == 0) free_retired(srv);  // nothing in flight: no stall
    // drop cancelled requests at the head of the queue
    while (!srv->queue.empty() && srv->queue.front()->cancel) {
      srv->queue.front()->state = CANCELLED;
      srv->queue.pop_front();
      progress = true;
    }
    if (progress) srv->cv_done.notify_all();
    if (srv->stop && srv->inflight_batches == 0) break;
    // pack queued requests of a free lane's class in arrival order while
    // memory, the batch's view count and its pixel cap allow
    Lane* free_lane = nullptr;
    std::vector<Req*> take;
    size_t bytes = 0;
    if (!srv->stop && !srv->queue.empty() && srv->fatal.empty()) {
      const bool split = srv->cfg.large_lanes > 0;
      for (auto& lp : srv->lanes) {
        if (lp->busy) continue;
        uint64_t pixels = 0;
        for (auto it = srv->queue.begin(); it != srv->queue.end() &&
                                           (int)take.size() < srv->cfg.max_batch_views;) {
          Req* r = *it;
          if (r->cancel) {
            r->state = CANCELLED;
            it = srv->queue.erase(it);
            continue;
          }
          if (split && r->large != lp->large) {
            ++it;
            continue;
          }
          const uint64_t px = (uint64_t)r->W * (uint64_t)r->H;
          if (!take.empty() && srv->cfg.batch_pixels && pixels + px > srv->cfg.batch_pixels) break;
          if (srv->inflight_bytes + bytes + r->bytes > srv->cfg.memory_budget) break;
          bytes += r->bytes;
          pixels += px;
          take.push_back(r);
          it = srv->queue.erase(it);
        }
        if (!take.empty()) {
          free_lane = lp.get();
          break;
        }
      }
    }
    if (!take.empty()) {
      for (Req* r : take) r->state = RUNNING;
      free_lane->reqs = take;
      free_lane->bytes = bytes;
      free_lane->busy = true;
      free_lane->finished = false;
      srv->inflight_bytes += bytes;
      if (srv->inflight_batches++ == 0) srv->busy_since = Clock::now();
      srv->stats.batches++;
      rlevox_tuning tuning;
      {
        std::lock_guard<std::mutex> tl(srv->ctx->mu);
        tuning = srv->ctx->tuning;
      }
      if (free_lane->large && srv->cfg.large_waves > 0) tuning.persistent_waves = srv->cfg.large_waves;
      lock.unlock();
      rlevox_status st = launch_lane(srv, *free_lane, tuning);
      std::string msg = st == RLEVOX_OK ? std::string() : std::string(rlevox_last_error());
      lock.lock();
      if (st != RLEVOX_OK) {
        // nothing of this batch is pending on the device unless the launch
        // itself went through; wait for the lane's stream before reuse
        lock.unlock();
        hipStreamSynchronize(free_lane->slot.stream);
        lock.lock();
        fail_requests(srv, free_lane->reqs, st, msg);
        srv->inflight_bytes -= free_lane->bytes;
        free_lane->bytes = 0;
        free_lane->reqs.clear();
        free_lane->busy = false;
        if (--srv->inflight_batches == 0)
          srv->stats.busy_seconds += ms_since(srv->busy_since, Clock::now()) * 1e-3;
        if (st == RLEVOX_ERR_DEVICE) srv->fatal = msg;
        srv->cv_done.notify_all();
      }
      continue;
    }
    if (!srv->fatal.empty()) {  // fail whatever is queued
      std::vector<Req*> rest(srv->queue.begin(), srv->queue.end());
      srv->queue.clear();
      fail_requests(srv, rest, RLEVOX_ERR_DEVICE, srv->fatal);
      srv->cv_done.notify_all();
    }
    srv->cv_work.wait(lock);
  }
}

}  // namespace
}  // namespace rlevox

extern "C" {

rlevox_status rlevox_server_create(rlevox_ctx* ctx, const rlevox_server_config* cfg,
                                   rlevox_server** out) {
  clear_error();
  if (!ctx || !cfg || !out) return fail(RLEVOX_ERR_INVALID_ARG, "server_create: null argument");
  *out = nullptr;
  if (cfg->memory_budget < (1u << 20))
    return fail(RLEVOX_ERR_INVALID_ARG, "server_create: memory budget below 1 MiB");
  if (cfg->max_batch_views < 1 || cfg->max_batch_views > BatchSlot::kMaxViews)
    return fail(RLEVOX_ERR_INVALID_ARG, "server_create: max_batch_views must be 1..%d",
                BatchSlot::kMaxViews);
  if (cfg->max_queue < 1) return fail(RLEVOX_ERR_INVALID_ARG, "server_create: max_queue < 1");
  if (cfg->streams < 1 || cfg->streams > 8)
    return fail(RLEVOX_ERR_INVALID_ARG, "server_create: streams must be 1..8");
  if (cfg->large_lanes < 0 || (cfg->large_lanes > 0 && cfg->large_lanes >= cfg->streams))
    return fail(RLEVOX_ERR_INVALID_ARG, "server_create: large_lanes must be 0..streams-1");
  if (cfg->large_lanes > 0 && cfg->large_pixels < 1)
    return fail(RLEVOX_ERR_INVALID_ARG, "server_create: large_lanes needs large_pixels >= 1");
  if (cfg->large_waves != 0 && (cfg->large_waves < 4 || cfg->large_waves > 32))
    return fail(RLEVOX_ERR_INVALID_ARG, "server_create: large_waves must be 0 or 4..32");
  const int post = post_flags(&cfg->params);
  if (post & ~(RLEVOX_POST_COMPOSE | RLEVOX_POST_GAPFILL) || !(post & RLEVOX_POST_COMPOSE))
    return fail(RLEVOX_ERR_INVALID_ARG, "server_create: post flags must include compose");
  rlevox_status st = ctx_bind(ctx);
  if (st != RLEVOX_OK) return st;
  auto* srv = new rlevox_server();
  srv->ctx = ctx;
  srv->cfg = *cfg;
  srv->post = post;
  for (int i = 0; i < cfg->streams; i++) {
    srv->lanes.emplace_back(new Lane());
    Lane& ln = *srv->lanes.back();
    ln.large = i < cfg->large_lanes;
    ln.fb.retire = &ln.retired_dev;
    ln.slot.scratch.retire = &ln.retired_dev;
    if ((st = srv->lanes.back()->slot.init(ctx)) != RLEVOX_OK) {
      for (auto& l : srv->lanes) l->slot.release();
      delete srv;
      return st;
    }
  }
  srv->worker = std::thread(worker_main, srv);
  *out = srv;
  return RLEVOX_OK;
}

rlevox_status rlevox_server_submit(rlevox_server* srv, const rlevox_scene* scene,
                                   const rlevox_pose* pose, int32_t W, int32_t H,
                                   uint16_t* d_colour, int32_t* d_depth, uint64_t* ticket) {
  clear_error();
  if (!srv || !scene || !pose || !ticket)
    return fail(RLEVOX_ERR_INVALID_ARG, "server_submit: null argument");
  *ticket = 0;
  if ((d_colour == nullptr) != (d_depth == nullptr))
    return fail(RLEVOX_ERR_INVALID_ARG, "server_submit: give both device framebuffers or neither");
  auto reject = [&](rlevox_status st) {
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->stats.rejected++;
    return st;
  };
  if (scene->ctx != srv->ctx)
    return reject(fail(RLEVOX_ERR_INVALID_ARG, "server_submit: scene belongs to another context"));
  rlevox_raymap rm;
  rlevox_status st = raymap_compute(*pose, W, H, srv->cfg.params.mip_distance,
                                    srv->cfg.params.far_distance, rm);
  if (st == RLEVOX_OK) st = check_raymap(rm);
  if (st != RLEVOX_OK) return reject(st);
  const size_t bytes = request_bytes(rm, srv->post, d_colour == nullptr);
  if (bytes > srv->cfg.memory_budget)
    return reject(fail(RLEVOX_ERR_OVER_BUDGET,
                       "server_submit: a %dx%d request needs %zu device bytes, budget %llu", W, H,
                       bytes, (unsigned long long)srv->cfg.memory_budget));
  std::unique_ptr<Req> r(new Req());
  r->scene = scene;
  r->pose = *pose;
  r->W = W;
  r->H = H;
  r->large = srv->cfg.large_lanes > 0 && (int64_t)W * H >= srv->cfg.large_pixels;
  r->d_colour = d_colour;
  r->d_depth = d_depth;
  r->bytes = bytes;
  r->t_arrival = Clock::now();
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    if (!srv->fatal.empty()) {
      srv->stats.rejected++;
      return fail(RLEVOX_ERR_DEVICE, "%s", srv->fatal.c_str());
    }
    if ((int)srv->queue.size() >= srv->cfg.max_queue) {
      srv->stats.rejected++;
      return fail(RLEVOX_ERR_QUEUE_FULL, "server_submit: %d requests pending", srv->cfg.max_queue);
    }
    r->ticket = srv->next_ticket++;
    *ticket = r->ticket;
    srv->queue.push_back(r.get());
    srv->reqs.emplace(r->ticket, std::move(r));
    srv->stats.submitted++;
  }
  srv->cv_work.notify_one();
  return RLEVOX_OK;
}

rlevox_status rlevox_server_wait(rlevox_server* srv, uint64_t ticket, int32_t timeout_ms,
                                 uint16_t* h_colour, int32_t* h_depth, double* latency_ms) {
  clear_error();
  if (!srv) return fail(RLEVOX_ERR_INVALID_ARG, "server_wait: null server");
  std::unique_lock<std::mutex> lock(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end())
    return fail(RLEVOX_ERR_INVALID_ARG, "server_wait: unknown ticket %llu", (unsigned long long)ticket);
  Req* r = it->second.get();
  auto ready = [&] { return r->state == DONE || r->state == FAILED || r->state == CANCELLED; };
  if (timeout_ms < 0) {
    srv->cv_done.wait(lock, ready);
  } else if (!srv->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready)) {
    return fail(RLEVOX_ERR_NOT_READY, "server_wait: request %llu not finished",
                (unsigned long long)ticket);
  }
  rlevox_status st = RLEVOX_OK;
  if (r->state == CANCELLED) {
    st = fail(RLEVOX_ERR_CANCELLED, "server_wait: request %llu was cancelled",
              (unsigned long long)ticket);
  } else if (r->state == FAILED) {
    st = fail(r->status, "%s", r->error.c_str());
  } else {
    const size_t px = (size_t)r->W * r->H;
    if (!r->d_colour) {
      if (h_colour) std::memcpy(h_colour, r->colour.data(), px * 2);
      if (h_depth) std::memcpy(h_depth, r->depth.data(), px * 4);
    }
    if (latency_ms) *latency_ms = ms_since(r->t_arrival, r->t_done);
  }
  srv->reqs.erase(it);  // a ticket is consumed by the wait that sees it finish
  return st;
}

rlevox_status rlevox_server_cancel(rlevox_server* srv, uint64_t ticket) {
  clear_error();
  if (!srv) return fail(RLEVOX_ERR_INVALID_ARG, "server_cancel: null server");
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end())
      return fail(RLEVOX_ERR_INVALID_ARG, "server_cancel: unknown ticket %llu",
                  (unsigned long long)ticket);
    Req* r = it->second.get();
    if (r->state == DONE || r->state == FAILED) return RLEVOX_OK;  // too late: result stands
    r->cancel = true;
    if (r->state == QUEUED) {
      for (auto q = srv->queue.begin(); q != srv->queue.end(); ++q)
        if (*q == r) {
          srv->queue.erase(q);
          break;
        }
      r->state = CANCELLED;
    }
  }
  srv->cv_done.notify_all();
  return RLEVOX_OK;
}

rlevox_status rlevox_server_stats_get(rlevox_server* srv, rlevox_server_stats* out) {
  clear_error();
  if (!srv || !out) return fail(RLEVOX_ERR_INVALID_ARG, "server_stats_get: null argument");
  std::lock_guard<std::mutex> lock(srv->mu);
  *out = srv->stats;
  if (srv->inflight_batches > 0) out->busy_seconds += ms_since(srv->busy_since, Clock::now()) * 1e-3;
  out->buffer_growths = 0;
  for (auto& lp : srv->lanes)
    out->buffer_growths += lp->fb.grown + lp->slot.scratch.grown + lp->pinned_grown;
  return RLEVOX_OK;
}

void rlevox_server_destroy(rlevox_server* srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->stop = true;
    for (Req* r : srv->queue) r->state = CANCELLED;
    srv->queue.clear();
  }
  srv->cv_work.notify_all();
  srv->cv_done.notify_all();
  if (srv->worker.joinable()) srv->worker.join();  // the worker drains in-flight batches first
  ctx_bind(srv->ctx);
  free_retired(srv);
  for (auto& l : srv->lanes) {
    l->slot.release();
    l->fb.release();
    if (l->pinned) hipHostFree(l->pinned);
    l->pinned = nullptr;
  }
  delete srv;
}

}  // extern "C"
