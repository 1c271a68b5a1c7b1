/*
 * rlevox.h - C ABI of the MI355X-native RLE voxel-column raycaster.
 *
 * Every entry point takes plain pointers and sizes; device pointers are HIP
 * device addresses and `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  No entry point exits or hangs on bad input: each returns an
 * rlevox_status and leaves a message for rlevox_last_error() (per thread).
 *
 * Coordinate and data conventions are those of the reference:
 *   - a scene is a stack of mip levels, each a grid of vertical run-length
 *     encoded columns (struct Map4, RLE-Raycaster/src/Rle4.h:7-21);
 *   - per column the slab array holds [count, ncolours, elem_1..elem_count,
 *     colour_1..colour_ncolours] with elem = drawn<<10 | skip
 *     (RLE-Raycaster/src/Rle4.cpp:166-188);
 *   - the pointer map holds two uint32 per column, x fastest:
 *     [slab offset, count | first_elem<<16] (RLE-Raycaster/src/Rle4.cpp:295-314);
 *   - y grows downward inside a column (row 0 is the top of the volume) and a
 *     camera at world row Yc has pose.pos[1] = -Yc (RLE-Raycaster/src/main.cpp:282-347);
 *   - a ray-buffer texel is colour16 | (depth & 0xfffe) << 16, and a texel
 *     cleared but not drawn holds RLEVOX_CLEAR_TEXEL
 *     (RLE-Raycaster/src/Cuda_Render.h:257,708,723).
 */
#ifndef RLEVOX_H
#define RLEVOX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RLEVOX_ABI_VERSION 3
#define RLEVOX_MAX_MIPS 16
/* Cleared ray-buffer texel (Cuda_Render.h:257). Its byte 2 (0xff) marks sky. */
#define RLEVOX_CLEAR_TEXEL 0x00ff8844u
/* Framebuffer value of a pixel that shows no voxel. */
#define RLEVOX_SKY_COLOUR 0x8844u
#define RLEVOX_SKY_DEPTH 0x7fffffff
/* Largest offset_shift of a mip level (2^36 slab words = 128 GiB). */
#define RLEVOX_MAX_OFFSET_SHIFT 4
/* Largest render size (screen width) accepted. */
#define RLEVOX_MAX_RS 8192

typedef enum {
  RLEVOX_OK = 0,
  RLEVOX_ERR_INVALID_ARG = 1,
  RLEVOX_ERR_INVALID_SCENE = 2,
  RLEVOX_ERR_OUT_OF_MEMORY = 3,
  RLEVOX_ERR_DEVICE = 4,
  RLEVOX_ERR_NO_DEVICE = 5,
  RLEVOX_ERR_QUEUE_FULL = 6,
  RLEVOX_ERR_NOT_READY = 7,
  RLEVOX_ERR_CANCELLED = 8,
  RLEVOX_ERR_IO = 9,
  RLEVOX_ERR_OVER_BUDGET = 10
} rlevox_status;

/* ------------------------------------------------------------------ types */

typedef struct rlevox_ctx rlevox_ctx;       /* one per device */
typedef struct rlevox_scene rlevox_scene;   /* a scene resident in HBM */
typedef struct rlevox_hscene rlevox_hscene; /* a scene in host memory */
typedef struct rlevox_server rlevox_server; /* continuous-batching request loop */

/* One mip level (Map4, Rle4.h:7-21).  ptrmap may be NULL on upload: it is
 * then rebuilt from the slabs as RLE4::load does (Rle4.cpp:295-314).
 * offset_shift widens the 32-bit slab offsets: each column starts at a
 * multiple of 2^offset_shift slab words and its pointer-map entry holds the
 * word offset >> offset_shift (0 = the reference's layout, 32-bit offsets in
 * uint16 units; 1..4 let one level hold up to 2^(32+shift) words; padding
 * words between columns are never read). */
typedef struct {
  int32_t sx, sy, sz;
  int32_t offset_shift;
  const uint32_t* ptrmap; /* 2*sx*sz words */
  const uint16_t* slabs;  /* nslabs words */
  uint64_t nslabs;
} rlevox_mip;

/* Camera pose: pos = world position, rot[0] = pitch (positive looks down),
 * rot[1] = yaw, rot[2] = roll (unsupported by the method, must be 0;
 * RayMap.h:120). */
typedef struct {
  float pos[3];
  float rot[3];
} rlevox_pose;

#define RLEVOX_POST_COMPOSE 1 /* ray buffer -> screen (colorize pass) */
#define RLEVOX_POST_GAPFILL 2 /* depth-guided smoothing (soft pass) */

typedef struct {
  int32_t mip_distance; /* MIP_DISTANCE (core.h:9); 0 = 1024 */
  int32_t far_distance; /* RAYS_DISTANCE (core.h:8); 0 = 80000 */
  int32_t post;         /* RLEVOX_POST_* bits; 0 = compose + gapfill */
  int32_t reserved;
} rlevox_render_params;

/* Per-frame parameter block: the fields of RayMap_GPU (RayMap.h:16-54) that
 * render_line reads, plus the per-frame constants the reference recomputes in
 * every thread (Cuda_Render.h:77-78,188-191,335) and the compose uniforms
 * (main.cpp:575-603), all computed once on the host. 4-byte fields only. */
typedef struct {
  int32_t W, H;            /* screen size */
  int32_t rs;              /* render size = ray-buffer line length = W */
  int32_t map_line_count;  /* ray lines = res[0]+res[1]+res[2]+res[3] */
  int32_t res[4];          /* lines per screen segment */
  int32_t maxres;          /* = rs (RAYS_CASTED_RES/4 in the reference) */
  int32_t mapswitch;       /* mip distance scaled by pitch (Cuda_Render.h:335) */
  int32_t far_distance;    /* z_far */
  int32_t reserved0;
  float rotation[3];
  float position[3];
  float border, clip_min, clip_max;
  float vp2d[3];           /* vanishing point on the image plane */
  float p4[3];             /* frustum origin */
  float p_2d[8][3];
  float p_no[8][3];        /* segment end points */
  float to3d[4][4];        /* image plane -> frustum space, row vectors */
  float p_ofs_min[4], p_ofs_max[4];
  float sin_x, cos_x, sin_y, cos_y; /* of pitch and yaw */
  float cos_myaw, sin_myaw;         /* cos(-yaw), sin(-yaw) */
  float vanish[2];                  /* compose uniform `vanish` */
  float ofs_add[4];                 /* compose uniform `ofs_add` */
  float rot_x_greater_zero;         /* compose uniform */
  float ray_ratio;                  /* compose uniform res_x_y_ray_ratio.w */
} rlevox_raymap;

/* Exact per-frame work counts of the raycast stage (the DETAIL_BENCH counters
 * of Cuda_Render.h:14-22, extended). */
typedef struct {
  uint64_t lines;       /* lines that cleared their range */
  uint64_t dda_steps;   /* DDA iterations */
  uint64_t ptr_loads;   /* pointer-map entries read (8 B each) */
  uint64_t rle_loads;   /* slab elements read after the cached first (2 B) */
  uint64_t elems;       /* RLE elements examined */
  uint64_t spans;       /* spans drawn */
  uint64_t col_loads;   /* colours read = texels drawn (2 B) */
  uint64_t cleared;     /* texels cleared (4 B) */
  /* The same for the walk as the kernels make it: a line's walk ends as soon
   * as no later step can draw (the fields above count the reference's walk
   * to the far distance; spans, colours and cleared texels are equal). */
  uint64_t walk_steps;
  uint64_t walk_ptr_loads;
  uint64_t walk_rle_loads;
  uint64_t walk_elems;
} rlevox_counters;

/* Seeded synthetic terrain: value-noise heightfield, layered caves, per-voxel
 * colour, mips by the >=3-of-8 rule. */
typedef struct {
  int32_t size_xz;       /* columns per side, power of two */
  int32_t size_y;        /* voxels per column, <= 4096 */
  int32_t nmips;         /* mip levels to build */
  int32_t octaves;       /* heightfield octaves */
  uint64_t seed;
  float persistence;     /* amplitude ratio between octaves */
  float hmin, hmax;      /* terrain height range in voxels above the bottom */
  float cave_layers;     /* mean cave layers per column */
  int32_t cave_max_layers;
  int32_t threads;       /* 0 = all */
} rlevox_gen_params;

/* A request handed to the batched renderer. */
typedef struct {
  const rlevox_scene* scene;
  rlevox_pose pose;
  int32_t W, H;
  uint16_t* d_colour; /* W*H device */
  int32_t* d_depth;   /* W*H device */
} rlevox_view;

/* Tuning knobs of the raycast kernels (autotuner output).  Every value gives
 * bit-identical framebuffers; only speed changes. */
typedef struct {
  int32_t lines_per_wave;  /* 1 = one wavefront walks each line (lanes take
                              consecutive DDA steps); 16, 32 or 64 = one lane
                              per line */
  int32_t waves_per_block; /* 1..4; with lines_per_wave 1: 1 = one wavefront
                              both walks and draws a line, 2..4 = two
                              wavefronts per line, one walking the DDA ahead
                              of the one drawing, 0 = two when the frame
                              has at most 14 lines per compute unit, else
                              one (single renders: the batched kernels keep
                              one wavefront per line, which uses a full GPU
                              better) */
  int32_t persistent_waves; /* wavefronts per CU of the persistent batch kernel
                               (0 = default 16, else 4..32) */
  int32_t order;           /* 0 = view order, 1 = longest views first */
  int32_t run_cache;       /* LDS run-cache words per candidate column of the
                              wavefront-per-line kernels (24, 40 or 56; 0 =
                              40 with two wavefronts per line, 56 with one):
                              deeper holds more of a column's runs and
                              colours in LDS, shallower loads less per
                              column */
  int32_t batch_run_cache; /* the same for the persistent batch kernel (0 = 40) */
  int32_t batch_par;       /* element replay of the persistent batch kernel:
                              0 = chosen by the scenes' runs per column
                              (lane-parallel above 7), 1 = element by
                              element, 2 = lane-parallel preparation */
} rlevox_tuning;

typedef struct {
  uint64_t memory_budget;  /* device bytes for in-flight framebuffers + ray buffers */
  int32_t max_batch_views; /* views packed into one launch */
  int32_t max_queue;       /* pending requests before submit says QUEUE_FULL */
  int32_t streams;         /* batches in flight at once (1..8) */
  int32_t large_lanes;     /* 0: every lane takes any request.  k (1..streams-1):
                              the first k lanes take only large requests and
                              the others only the rest, so a small request
                              never shares a batch with a large one */
  rlevox_render_params params;
  uint64_t batch_pixels;   /* a batch stops taking requests once their W*H
                              sum reaches this (it always takes one; 0 = no
                              limit): bounds how long a batch holds a lane */
  int64_t large_pixels;    /* W*H from which a request is large (large_lanes) */
  int32_t large_waves;     /* persistent raycast wavefronts per CU in a large
                              lane's batches (0 = the batch default; fewer
                              leave room on the GPU to the other lanes) */
  int32_t reserved;
} rlevox_server_config;

typedef struct {
  uint64_t submitted, completed, rejected, batches;
  uint64_t rays;                 /* W*H summed over completed requests */
  double busy_seconds;           /* wall time with at least one batch in flight */
  uint64_t buffer_growths;       /* lane buffers replaced by larger ones (the
                                    old ones are freed once nothing is in flight) */
} rlevox_server_stats;

/* ------------------------------------------------------------ utilities */

const char* rlevox_last_error(void);
int32_t rlevox_abi_version(void);

/* ------------------------------------------------------ host-side inputs */

/* RayMap::get_ray_map + set_border (RayMap.h:66,98-402; main.cpp:769-779).
 * Needs no device.  W >= H, both >= 16, W <= RLEVOX_MAX_RS. */
rlevox_status rlevox_raymap_compute(const rlevox_pose* pose, int32_t W, int32_t H,
                                    const rlevox_render_params* params,
                                    rlevox_raymap* out);

/* Pointer-map rebuild of RLE4::load (Rle4.cpp:295-314); with offset_shift > 0
 * each column starts at the next multiple of 2^offset_shift words. */
rlevox_status rlevox_build_ptrmap(const uint16_t* slabs, uint64_t nslabs, int32_t sx,
                                  int32_t sz, int32_t offset_shift, uint32_t* ptrmap_out);

/* Re-packs a scene with columns aligned to 2^offset_shift slab words (the
 * decoded voxels are unchanged).  offset_shift >= 3 puts every column header
 * on a 16-byte boundary, which the raycast's run cache loads whole; larger
 * shifts also lift the 2^32-word limit of a level.  No counterpart in the
 * reference (its layout is offset_shift 0). */
rlevox_status rlevox_repack(const rlevox_mip* mips, int32_t nmips, int32_t offset_shift,
                            rlevox_hscene** out);

/* Checks that every column is well formed and stays inside its slab array. */
rlevox_status rlevox_scene_validate(const rlevox_mip* mips, int32_t nmips);

rlevox_status rlevox_gen_scene(const rlevox_gen_params* params, rlevox_hscene** out);

/* RLE4::compress (Rle4.cpp:52-198) + Tree::get_mipmap (tree.h:23-85) for a
 * dense volume: solid/colour are indexed x + y*sx + z*sx*sy. */
rlevox_status rlevox_encode_dense(int32_t sx, int32_t sy, int32_t sz, const uint8_t* solid,
                                  const uint16_t* colour, int32_t nmips,
                                  rlevox_hscene** out);

/* .rle4 files (RLE4::save / RLE4::load, Rle4.cpp:220-384).  The format has
 * no field for offset_shift: only offset_shift 0 levels are saved. */
rlevox_status rlevox_rle4_save(const char* path, const rlevox_mip* mips, int32_t nmips);
rlevox_status rlevox_rle4_load(const char* path, rlevox_hscene** out);

int32_t rlevox_hscene_nmips(const rlevox_hscene* s);
rlevox_status rlevox_hscene_mip(const rlevox_hscene* s, int32_t m, rlevox_mip* out);
void rlevox_hscene_free(rlevox_hscene* s);

/* ----------------------------------------------------- device: context */

rlevox_status rlevox_ctx_create(int32_t device, rlevox_ctx** out);
void rlevox_ctx_destroy(rlevox_ctx* ctx);

/* Scene upload (RLE4::all_to_gpu / copy_to_gpu, Rle4.cpp:432-448).  The
 * library copies the host arrays; the caller may free them afterwards. */
rlevox_status rlevox_scene_upload(rlevox_ctx* ctx, const rlevox_mip* mips, int32_t nmips,
                                  rlevox_scene** out);
void rlevox_scene_free(rlevox_scene* scene);
uint64_t rlevox_scene_device_bytes(const rlevox_scene* scene);

/* ------------------------------------------------------ device: render */

/* Raycast stage only: the cuda_main_render2 counterpart (Cuda_Main.cu:124,
 * 183-271).  d_raybuf holds map_line_count*rs texels; d_line_clip receives
 * the cleared range [lo, hi] of every line (hi < lo: nothing cleared).
 * Texels outside a line's range are left untouched. */
rlevox_status rlevox_render_raybuf(rlevox_ctx* ctx, const rlevox_scene* scene,
                                   const rlevox_raymap* rm, uint32_t* d_raybuf,
                                   int32_t* d_line_clip, void* stream);

/* rlevox_render_raybuf plus the exact work counts of the frame (the
 * DETAIL_BENCH counters, Cuda_Render.h:14-22 and Cuda_Main.cu:244-266);
 * synchronous.  The counting kernel is a separate instantiation, so the
 * plain entry point pays nothing for it. */
rlevox_status rlevox_raycast_counters(rlevox_ctx* ctx, const rlevox_scene* scene,
                                      const rlevox_raymap* rm, uint32_t* d_raybuf,
                                      int32_t* d_line_clip, rlevox_counters* out);

/* Compose (colorize_buddha_soft.frag:12-92): ray buffer -> W*H colour16 +
 * int32 depth, row 0 at the top of the image.  d_raybuf holds
 * map_line_count*rs texels (any texel may be read; only those inside a
 * line's cleared range are used). */
rlevox_status rlevox_compose(rlevox_ctx* ctx, const rlevox_raymap* rm, const uint32_t* d_raybuf,
                             const int32_t* d_line_clip, uint16_t* d_colour,
                             int32_t* d_depth, void* stream);

/* Gap-fill / smoothing (soft.frag:8-74) restated on integer framebuffers. */
rlevox_status rlevox_gapfill(rlevox_ctx* ctx, int32_t W, int32_t H, const uint16_t* d_colour_in,
                             const int32_t* d_depth, uint16_t* d_colour_out, void* stream);

/* Whole frame: pose -> framebuffer (device pointers), scratch owned by ctx.
 * Returns when the framebuffer is complete (the context's scratch is reused
 * by the next call), as the reference's per-frame render does; the stage
 * entry points above are stream-ordered and return at once. */
rlevox_status rlevox_render(rlevox_ctx* ctx, const rlevox_scene* scene, const rlevox_pose* pose,
                            int32_t W, int32_t H, const rlevox_render_params* params,
                            uint16_t* d_colour, int32_t* d_depth, void* stream);

/* rlevox_render without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, serving loops that forbid hipMalloc).  Its scratch (ray buffer,
 * line clip ranges, gap-fill temporary) is a fixed array of
 * RLEVOX_NOALLOC_SCRATCH bytes that each GPU the library runs on holds until
 * the process ends.  A view needs 4*W bytes per ray line, 8 bytes per line
 * and, with gap-fill, 2*W*H bytes (each rounded up to 256); one needing more
 * than the array fails with OVER_BUDGET (the 1920x1080 flythrough views
 * need 20-30 MB).  Stream-ordered: it enqueues the frame
 * on `stream` and returns without waiting.  All calls on one GPU share that
 * GPU's array, so calls that overlap on one GPU race. */
#define RLEVOX_NOALLOC_SCRATCH (64u << 20)
rlevox_status rlevox_render_noalloc(rlevox_ctx* ctx, const rlevox_scene* scene,
                                    const rlevox_pose* pose, int32_t W, int32_t H,
                                    const rlevox_render_params* params, uint16_t* d_colour,
                                    int32_t* d_depth, void* stream);

/* Many views (any scenes, any sizes) in one persistent launch over a device
 * work queue; each framebuffer is identical to rlevox_render's. */
rlevox_status rlevox_render_batch(rlevox_ctx* ctx, const rlevox_view* views, int32_t nviews,
                                  const rlevox_render_params* params, void* stream);

/* Same as rlevox_render_batch, with the three stages captured once as a
 * hipGraph per batch shape and replayed. */
rlevox_status rlevox_render_batch_graph(rlevox_ctx* ctx, const rlevox_view* views,
                                        int32_t nviews, const rlevox_render_params* params,
                                        void* stream);

/* The context's tuning: used by batches and the server, and by single
 * renders of a (scene, W, H) the autotuner has not measured.  Setting it
 * also forgets every autotuned (scene, W, H) entry. */
rlevox_status rlevox_set_tuning(rlevox_ctx* ctx, const rlevox_tuning* t);
rlevox_status rlevox_get_tuning(rlevox_ctx* ctx, rlevox_tuning* t);
/* The tuning a single render of (scene, W, H) uses (no counterpart in the
 * reference, whose launch shape is fixed at compile time). */
rlevox_status rlevox_get_tuning_for(rlevox_ctx* ctx, const rlevox_scene* scene, int32_t W,
                                    int32_t H, rlevox_tuning* t);

/* Autotuner: for each (scene, W, H) among the `nviews` sample views, times
 * the candidate tunings on those views and keeps the fastest for later
 * renders of that (scene, W, H) (dropped when the scene is freed or the
 * tuning is set).  *best receives the first view's winner, which also
 * becomes the context's tuning. */
rlevox_status rlevox_autotune(rlevox_ctx* ctx, const rlevox_view* views, int32_t nviews,
                              const rlevox_render_params* params, rlevox_tuning* best,
                              double* best_ms);
/* Autotuner of the batched path (no counterpart in the reference, which
 * fixes its launch shape at compile time: core.h:10, Cuda_Main.cu:195-202,
 * 226): times rlevox_render_batch on the given views as one batch for every
 * combination of persistent wavefronts per CU (8, 12, 16, 24), view order
 * (0, 1), batch run-cache depth (24, 40, 56) and element replay (1, 2), and
 * makes the fastest the context's batch settings (persistent_waves, order,
 * batch_run_cache, batch_par of its tuning), which rlevox_render_batch,
 * rlevox_render_batch_graph and servers created on the context use.  The
 * views' framebuffers are rendered into.  *best receives the context's
 * tuning after the change, *best_ms the winner's batch time. */
rlevox_status rlevox_autotune_batch(rlevox_ctx* ctx, const rlevox_view* views, int32_t nviews,
                                    const rlevox_render_params* params, rlevox_tuning* best,
                                    double* best_ms);

/* ------------------------------------------------------ device: serving */

rlevox_status rlevox_server_create(rlevox_ctx* ctx, const rlevox_server_config* cfg,
                                   rlevox_server** out);
/* Queues one request; *ticket identifies it.  With d_colour and d_depth
 * (W*H device arrays, both or neither) the framebuffer is rendered straight
 * into them and stays on the device; without them the server renders into
 * its own arena and copies the result to host memory for rlevox_server_wait.
 * QUEUE_FULL when max_queue requests are pending, OVER_BUDGET when the
 * request alone exceeds the memory budget (ray buffer, line clip ranges,
 * gap-fill temporary and, for host delivery, the framebuffer). */
rlevox_status rlevox_server_submit(rlevox_server* srv, const rlevox_scene* scene,
                                   const rlevox_pose* pose, int32_t W, int32_t H,
                                   uint16_t* d_colour, int32_t* d_depth, uint64_t* ticket);
/* Waits up to timeout_ms (-1 = forever) for the request and consumes its
 * ticket; copies a host-delivered framebuffer to the host arrays (may be
 * NULL; ignored for device delivery) and the latency (submit to framebuffer
 * ready: on the device for device delivery, in host memory otherwise) to
 * *latency_ms.  NOT_READY on time-out (the ticket stays valid), CANCELLED
 * for a cancelled request.  No call on a server may be in progress when it
 * is destroyed. */
rlevox_status rlevox_server_wait(rlevox_server* srv, uint64_t ticket, int32_t timeout_ms,
                                 uint16_t* h_colour, int32_t* h_depth, double* latency_ms);
rlevox_status rlevox_server_cancel(rlevox_server* srv, uint64_t ticket);
rlevox_status rlevox_server_stats_get(rlevox_server* srv, rlevox_server_stats* out);
void rlevox_server_destroy(rlevox_server* srv);

/* ------------------------------------------------- device: plumbing */

rlevox_status rlevox_dev_alloc(rlevox_ctx* ctx, uint64_t bytes, void** out);
rlevox_status rlevox_dev_free(rlevox_ctx* ctx, void* p);

/* Testing aid, no counterpart in the reference: the n-th device allocation
 * the library makes from now on (scene upload, scratch growth, dev_alloc,
 * batch and server buffers) fails as out of memory; n <= 0 disarms.
 * Process-wide. */
void rlevox_debug_fail_alloc(int32_t n);
rlevox_status rlevox_dev_fill_u32(rlevox_ctx* ctx, void* p, uint32_t value, uint64_t count, void* stream);
/* Device-to-device copy by a plain streaming kernel (16-byte aligned
 * pointers, size a multiple of 16): exact read+write traffic, used to
 * calibrate hardware byte counters.  No counterpart in the reference. */
rlevox_status rlevox_dev_copy(rlevox_ctx* ctx, void* dst, const void* src, uint64_t bytes, void* stream);
rlevox_status rlevox_memcpy_htod(rlevox_ctx* ctx, void* dst, const void* src, uint64_t bytes);
rlevox_status rlevox_memcpy_dtoh(rlevox_ctx* ctx, void* dst, const void* src, uint64_t bytes);
rlevox_status rlevox_stream_create(rlevox_ctx* ctx, void** out);
rlevox_status rlevox_stream_destroy(rlevox_ctx* ctx, void* stream);
rlevox_status rlevox_stream_sync(rlevox_ctx* ctx, void* stream);
rlevox_status rlevox_event_create(rlevox_ctx* ctx, void** out);
rlevox_status rlevox_event_destroy(rlevox_ctx* ctx, void* ev);
rlevox_status rlevox_event_record(rlevox_ctx* ctx, void* ev, void* stream);
rlevox_status rlevox_event_elapsed_ms(rlevox_ctx* ctx, void* start, void* stop, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* RLEVOX_H */
