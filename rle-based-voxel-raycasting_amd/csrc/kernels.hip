// HIP kernels for gfx950: plain per-view raycast, compose, gap-fill, and the
// batched persistent raycast over a device work queue.
#include <hip/hip_runtime.h>

#include "device.h"
#include "kernels.h"
#include "wave_line.h"

namespace rlevox {

// ------------------------------------------------------------------ plain
// One lane per ray line (cudaRender, Cuda_Main.cu:150-181); each lane keeps
// its line bitmap in LDS, word w of lane l at lds[w*blockDim + l] so lanes
// reading the same word index hit distinct banks.
template <bool COUNT>
__global__ __launch_bounds__(256) void k_raycast_plain(SceneDev sc, rlevox_raymap rm,
                                                       uint32_t* __restrict__ raybuf,
                                                       int32_t* __restrict__ clip,
                                                       unsigned long long* counts) {
  extern __shared__ uint32_t lds[];
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= rm.map_line_count) return;
  const LineBits bits{lds + threadIdx.x, (int)blockDim.x};
  render_line<COUNT>(x, sc, rm, raybuf + (size_t)x * (size_t)rm.rs, clip + 2 * x, bits, counts);
}

// Blocks take the ray lines from the last to the first: the lines of the
// last segments of the ray map start first, which for views looking down
// puts the long lines (those reaching the sky, walked to the far distance)
// at the head of the launch.  Measured 1-6% faster than index order on
// 1080p and 4K flythroughs and random views, equal on grazing views.
//
// One wavefront (= one block) per ray line.  Registers are capped for four
// wavefronts per SIMD (4096 resident lines on 256 CUs); uncapped the
// compiler takes 140 and only three fit.
template <bool COUNT, int RC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_raycast_wave(SceneDev sc, rlevox_raymap rm,
                                                     uint32_t* __restrict__ raybuf,
                                                     int32_t* __restrict__ clip,
                                                     unsigned long long* counts) {
  static_assert(kDdaLds * 4 <= 64 * RC * 2, "DDA scratch must fit in the run cache");
  __shared__ uint4 rcache[64 * RC / 8];
  __shared__ int sweep[64 + kWalkEndWords];
  float* dda = reinterpret_cast<float*>(rcache);  // see kDdaLds
  extern __shared__ uint32_t lds[];
  if ((int)blockIdx.x >= rm.map_line_count) return;
  const int x = rm.map_line_count - 1 - (int)blockIdx.x;  // last line first
  render_line_wave<COUNT, false, true, RC>(x, sc, rm, raybuf + (size_t)x * (size_t)rm.rs,
                                           clip + 2 * x, lds, dda,
                                           reinterpret_cast<uint16_t*>(rcache), sweep, counts);
}

// Two wavefronts (= one block) per ray line: wavefront 0 walks the DDA and
// gathers the pointer map, wavefront 1 projects, caches and draws.
template <int RC>
__global__ __launch_bounds__(128) void k_raycast_wave2(SceneDev sc, rlevox_raymap rm,
                                                       uint32_t* __restrict__ raybuf,
                                                       int32_t* __restrict__ clip) {
  __shared__ uint4 rcache[64 * RC / 8];
  __shared__ int sweep[64 + kWalkEndWords];
  __shared__ float dda[kDdaLds];
  __shared__ ChunkRing ring;
  extern __shared__ uint32_t lds[];
  if ((int)blockIdx.x >= rm.map_line_count) return;
  const int x = rm.map_line_count - 1 - (int)blockIdx.x;  // last line first
  render_line_wave<false, true, true, RC>(x, sc, rm, raybuf + (size_t)x * (size_t)rm.rs,
                                          clip + 2 * x, lds, dda,
                                          reinterpret_cast<uint16_t*>(rcache), sweep, nullptr,
                                          &ring);
}

__global__ __launch_bounds__(256) void k_compose(rlevox_raymap rm, const uint32_t* __restrict__ raybuf,
                                                 const int32_t* __restrict__ clip,
                                                 uint16_t* __restrict__ col,
                                                 int32_t* __restrict__ dep) {
  __shared__ ComposeTile ct;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4;
  compose_tile_stage(rm, x0, y0, ct);
  __syncthreads();
  const int px = x0 + (threadIdx.x & 63);
  const int row = y0 + (threadIdx.x >> 6);
  if (px >= rm.W || row >= rm.H) return;
  uint16_t c;
  int32_t d;
  compose_pixel(rm, raybuf, clip, px, row, ct.col[threadIdx.x & 63], ct.row[threadIdx.x >> 6],
                ct.border, c, d);
  col[(size_t)row * rm.W + px] = c;
  dep[(size_t)row * rm.W + px] = d;
}

// Tile of block b of a grid of n tiles, gx per row: blocks are dispatched
// round robin over the 8 XCDs, and each XCD takes groups of 8 tile rows (32
// pixel rows), the groups dealt round robin, so the border rows that the
// tiles of a group share come through the same L2 (1080p gap-fill HBM fetch
// 99 -> 27 MB at the same speed; contiguous eighths of the image per XCD
// fetched 15 MB but ran 2% slower, the eighths differing in work).
__device__ inline int xcd_tile(int b, int n, int gx) {
  const int G = 8 * gx, full = n / (8 * G) * (8 * G);
  if (b >= full) return b;  // the last partial round in dispatch order
  const int x = b & 7, k = b >> 3;
  return ((k / G) * 8 + x) * G + (k % G);
}

__global__ __launch_bounds__(256) void k_gapfill(int W, int H, const uint16_t* __restrict__ cin,
                                                 const int32_t* __restrict__ dep,
                                                 uint16_t* __restrict__ cout) {
  __shared__ GapTile tile;
  const int gx = (int)gridDim.x;
  const int t = xcd_tile((int)(blockIdx.y * gridDim.x + blockIdx.x), (int)(gridDim.x * gridDim.y), gx);
  const int ty = t / gx;
  gapfill_tile(W, H, cin, dep, cout, (t - ty * gx) * 64, ty * 4, tile);
}

template <int RC>
static void launch_wave(const SceneDev& sc, const rlevox_raymap& rm, uint32_t* raybuf,
                        int32_t* clip, bool two, hipStream_t s) {
  const size_t lds = (size_t)bitmap_words(rm.rs) * sizeof(uint32_t);
  if (two)
    hipLaunchKernelGGL(k_raycast_wave2<RC>, dim3(rm.map_line_count), dim3(128), lds, s, sc, rm,
                       raybuf, clip);
  else
    hipLaunchKernelGGL((k_raycast_wave<false, RC>), dim3(rm.map_line_count), dim3(64), lds, s, sc,
                       rm, raybuf, clip, nullptr);
}

hipError_t launch_raycast_plain(const SceneDev& sc, const rlevox_raymap& rm, uint32_t* raybuf,
                                int32_t* clip, int block, int run_cache, hipStream_t s,
                                unsigned long long* counts) {
  const int lines = rm.map_line_count;
  if (lines <= 0) return hipSuccess;
  if (block == 1 || block == 2) {  // wavefront kernels
    // Default depth: 40 words with two wavefronts per line (1080p
    // flythrough: 0.1504 ms vs 0.1509 / 0.1514 with 24 / 56), 56 with one,
    // where frames have many long lines (4K flythrough 0.267 vs 0.272 ms,
    // 4K grazing views of the 16384^2 terrain 8.03 vs 8.10 ms).
    if (run_cache == 0) run_cache = block == 2 ? kRunCache : 56;
    if (counts) {  // counting walks with one wavefront and the default run cache
      const size_t lds = (size_t)bitmap_words(rm.rs) * sizeof(uint32_t);
      hipLaunchKernelGGL((k_raycast_wave<true, kRunCache>), dim3(lines), dim3(64), lds, s, sc, rm,
                         raybuf, clip, counts);
    } else if (run_cache == 24) {
      launch_wave<24>(sc, rm, raybuf, clip, block == 2, s);
    } else if (run_cache == 56) {
      launch_wave<56>(sc, rm, raybuf, clip, block == 2, s);
    } else {
      launch_wave<kRunCache>(sc, rm, raybuf, clip, block == 2, s);
    }
    return hipGetLastError();
  }
  const int grid = (lines + block - 1) / block;
  const size_t lds = (size_t)bitmap_words(rm.rs) * block * sizeof(uint32_t);
  if (counts)
    hipLaunchKernelGGL(k_raycast_plain<true>, dim3(grid), dim3(block), lds, s, sc, rm, raybuf,
                       clip, counts);
  else
    hipLaunchKernelGGL(k_raycast_plain<false>, dim3(grid), dim3(block), lds, s, sc, rm, raybuf,
                       clip, counts);
  return hipGetLastError();
}

// The world row of each column's first drawn run (the element walk of
// Cuda_Render.h:511-524 up to the first element with drawn voxels), reduced
// to the minimum over the level.  Columns are validated at upload, so the
// walk stays inside the slab array.
__global__ __launch_bounds__(256) void k_column_tops(const uint2* __restrict__ map,
                                                     const uint16_t* __restrict__ slabs, int shift,
                                                     int level, uint64_t ncols, int32_t* out) {
  int best = INT32_MAX;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncols; c += stride) {
    const uint2 e = map[c];
    const uint32_t n = e.y & 0xffffu;
    const uint16_t* runs = slabs + 2 + ((uint64_t)e.x << shift);
    uint32_t w = e.y >> 16;
    int y = 0;
    for (uint32_t k = 0; k < n; k++) {
      if (k > 0) w = runs[k];
      y += (int)(w & 1023u) << level;
      if (w >> 10) {
        best = min(best, y);
        break;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) best = min(best, __shfl_xor(best, o));
  if ((threadIdx.x & 63) == 0 && best != INT32_MAX) atomicMin(out, best);
}

hipError_t launch_column_tops(const MipDev& m, int level, int32_t* out, hipStream_t s) {
  const uint64_t ncols = (uint64_t)m.sx * (uint64_t)m.sz;
  if (ncols == 0) return hipSuccess;
  uint64_t blocks = (ncols + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_column_tops, dim3((unsigned)blocks), dim3(256), 0, s, m.map, m.slabs,
                     m.shift, level, ncols, out);
  return hipGetLastError();
}

hipError_t launch_compose(const rlevox_raymap& rm, const uint32_t* raybuf, const int32_t* clip,
                          uint16_t* col, int32_t* dep, hipStream_t s) {
  dim3 grid((rm.W + 63) / 64, (rm.H + 3) / 4);
  hipLaunchKernelGGL(k_compose, grid, dim3(256), 0, s, rm, raybuf, clip, col, dep);
  return hipGetLastError();
}

hipError_t launch_gapfill(int W, int H, const uint16_t* cin, const int32_t* dep, uint16_t* cout,
                          hipStream_t s) {
  dim3 grid((W + 63) / 64, (H + 3) / 4);
  hipLaunchKernelGGL(k_gapfill, grid, dim3(256), 0, s, W, H, cin, dep, cout);
  return hipGetLastError();
}

}  // namespace rlevox

namespace rlevox {

// Fills count uint32 words; used instead of hipMemset so the same code can
// be captured into a graph.
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Copies n 16-byte words: a plain streaming kernel of known traffic, used to
// calibrate hardware byte counters against an exact read+write volume.
__global__ __launch_bounds__(256) void k_copy_u128(uint4* __restrict__ dst,
                                                  const uint4* __restrict__ src, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

hipError_t launch_copy_u128(void* dst, const void* src, uint64_t n16, hipStream_t s) {
  if (n16 == 0) return hipSuccess;
  uint64_t blocks = (n16 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_copy_u128, dim3((unsigned)blocks), dim3(256), 0, s, (uint4*)dst,
                     (const uint4*)src, n16);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, value, count);
  return hipGetLastError();
}

}  // namespace rlevox
