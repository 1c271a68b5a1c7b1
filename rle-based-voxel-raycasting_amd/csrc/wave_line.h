// Wavefront-per-line raycast: the 64 lanes of one wavefront walk one ray
// line together and produce results bit-identical to render_line().
//
//  1. DDA chunk: up to 64 DDA steps per pass, lane k taking step k.  The
//     crossing distances and positions are repeated fp32 additions
//     (Cuda_Render.h:398-414) of six independent sequences; their terms
//     0..63 are computed side by side with the same additions the walk
//     makes, and the x/z order is the merge order of the two distance
//     sequences (ties go to x: the walk takes z only when strictly closer).
//  2. Gather: each lane loads its cell's pointer-map entry, one round trip
//     for the whole chunk, issued one chunk ahead (software pipeline).
//  3. Stateless projection per lane: clip-top inputs and F, the projected
//     top of the first drawn element in front of the eye.
//  4. Run cache: each step that may act copies the head of its column
//     (header, runs, colours) into LDS; it then precomputes, in parallel,
//     the span parameters of that first element and F2, the projected top
//     of the next element the loop would test.
//  5. Ordered replay: with the current floating horizon ymin, a step whose
//     clip-top test culls it or whose F <= ymin changes no line state (the
//     element loop breaks first, Cuda_Render.h:467,543).  A ballot finds the
//     next step that acts; only those steps are replayed, in order, using
//     the precomputed values, and the loop over further elements runs only
//     when F2 > ymin.
//     Runs of acting steps of the simple kind (a span drawn from the
//     horizon up, the top clip untouched, the next element not drawn) are
//     applied at once by a horizon sweep: the horizon before each step is
//     a prefix maximum across the lanes (DPP lane moves), and each lane
//     draws one row of the swept rows.
//  6. Spans are drawn 64 pixels per pass: lane k applies the per-pixel
//     increments k times from the span start, exactly as the pixel loop
//     does (Cuda_Render.h:687), and the line's LDS bitmap is updated from a
//     ballot of the drawn pixels.
//  7. The walk ends once no later step can draw (walk_end.h), leaving the
//     line as the full walk to the far distance does.
//  8. Two-wavefront form (k_raycast_wave2): one wavefront walks the DDA and
//     gathers (1-2), the other projects, caches and draws (3-7); chunks pass
//     between them through a three-slot LDS ring with one workgroup barrier
//     per chunk.
#pragma once

#include <type_traits>

#include "device.h"

namespace rlevox {

__device__ inline float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Marks a value that is the same in every lane as uniform, so the compiler
// keeps it in a scalar register and branches on it with scalar branches
// instead of exec masks (only for values that really are uniform).
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float uni_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ inline uint64_t uni64(uint64_t v) {
  return ((uint64_t)(uint32_t)uni((int)(v >> 32)) << 32) | (uint32_t)uni((int)(uint32_t)v);
}

__device__ inline bool lds_bit(const uint32_t* bits, int y) { return (bits[y >> 5] >> (y & 31)) & 1u; }

// LDS hand-off between the lanes of one wavefront, on LDS that only this
// wavefront uses (hand-offs between the two wavefronts of a line go through
// the ring and its workgroup barrier).  A wavefront's LDS instructions
// execute in program order, so a write by one lane is seen by a later read
// of another lane once the compiler keeps both on their side of this point.
// Unlike a workgroup barrier, this emits no s_barrier and does not wait for
// the wavefront's outstanding global loads (the prefetches stay in flight).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Slab words cached per step (from the 16-byte boundary at or below the
// column header): covers header, elements and colours of typical columns.
// The default depth; the wavefront-per-line kernels are also built for 24
// and 56 words (a tuning choice).
constexpr int kRunCache = 40;
// LDS floats per wavefront used by the DDA chunk.  The kernels place them in
// the run cache's memory: the DDA of chunk c+1 runs after chunk c's replay
// has finished with the cache and before chunk c+1's cache is stored.
constexpr int kDdaLds = 6 * 68 + 64;
static_assert(kDdaLds * 4 <= 64 * kRunCache * 2, "DDA scratch must fit in the run cache");

// First y in [from, to) whose bit is clear, or `to` (from <= to).
__device__ inline int first_clear(const uint32_t* bits, int from, int to, int lane) {
  for (int base = from; base < to; base += 64) {
    const int y = base + lane;
    const uint64_t m = uni64(__ballot(y < to && !lds_bit(bits, y)));
    if (m) return uni(base + __builtin_ctzll(m));
  }
  return to;
}

// Number of entries of the increasing 64-array `arr` that precede value v:
// (LE ? arr[l] <= v : arr[l] < v).  `guess` is an estimate of the answer.
template <bool LE>
__device__ inline int rank_in(const float* arr, float v, int guess) {
  int j = min(max(guess, 0), 64);
  auto before = [&](int l) { return LE ? !(v < arr[l]) : (arr[l] < v); };
  while (j > 0 && !before(j - 1)) j--;
  while (j < 64 && before(j)) j++;
  return j;
}

// Wavefront scans with DPP lane moves (register reads across lanes, no LDS
// round trip): row_shr:n (0x110 + n) inside each 16-lane row, then
// row_bcast:15 (0x142, into rows 1 and 3) and row_bcast:31 (0x143, into
// rows 2 and 3) across rows, wave_shr:1 (0x138) for the exclusive form.  A
// lane whose source lies outside keeps `old`, the operation's identity.
// Call with every lane of the wavefront active.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline int dpp_move(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xf, false);
}

// Exclusive prefix maximum over the lanes of the wavefront (lane 0 gets
// `init`).
__device__ inline int prefix_max_excl(int v, int init, int lane) {
  v = max(v, dpp_move<0x111>(INT32_MIN, v));
  v = max(v, dpp_move<0x112>(INT32_MIN, v));
  v = max(v, dpp_move<0x114>(INT32_MIN, v));
  v = max(v, dpp_move<0x118>(INT32_MIN, v));
  v = max(v, dpp_move<0x142, 0xa>(INT32_MIN, v));
  v = max(v, dpp_move<0x143, 0xc>(INT32_MIN, v));
  const int prev = dpp_move<0x138>(INT32_MIN, v);
  return lane == 0 ? init : max(init, prev);
}

// Inclusive prefix sum of v over lanes 0..lane.
__device__ inline int prefix_sum_incl(int v, int lane) {
  (void)lane;
  v += dpp_move<0x111>(0, v);
  v += dpp_move<0x112>(0, v);
  v += dpp_move<0x114>(0, v);
  v += dpp_move<0x118>(0, v);
  v += dpp_move<0x142, 0xa>(0, v);
  v += dpp_move<0x143, 0xc>(0, v);
  return v;
}

// ------------------------------------------------------------ span pixels

// Sets the bitmap bits of rows base..base+63 given by mask m.  Lane 0 issues
// LDS atomic ORs whose results nobody waits for.  Only rows below the span
// end (< rs) are ever in m, so the words stay inside the bitmap.
__device__ inline void bits_or(uint32_t* bits, int base, uint64_t m, int lane) {
  if (m && lane == 0) {
    const int w0 = base >> 5, off = base & 31;
    const uint32_t lo = (uint32_t)(m << off);
    if (lo) atomicOr(&bits[w0], lo);
    const uint32_t mid = off ? (uint32_t)(m >> (32 - off)) : (uint32_t)(m >> 32);
    if (mid) atomicOr(&bits[w0 + 1], mid);
    const uint32_t hi = off ? (uint32_t)(m >> (64 - off)) : 0u;
    if (hi) atomicOr(&bits[w0 + 2], hi);
  }
}

// Draws pixels [y0, s2) of one span (uniform arguments; Cuda_Render.h:
// 678-733).  Pixel y0+k has had k per-pixel increments.
__device__ inline int draw_span(int y0, int s2, float uzb, float ozb, float u2dz, float onedz2,
                                float tlo, float thi, const uint16_t* cached, int ccap,
                                const uint16_t* colours, int cofs, uint32_t* __restrict__ row,
                                uint32_t* bits, int lane) {
  int drawn = 0;
  for (int base = y0; base < s2; base += 64) {
    float uz = uzb, oz = ozb;
    const int adds = min(lane, s2 - 1 - base);
    for (int i = 0; i < adds; i++) {
      uz += u2dz;
      oz += onedz2;
    }
    const int y = base + lane;
    const bool draw = y < s2 && !lds_bit(bits, y);
    if (draw) {
      const float fu = fminf(fmaxf(uz / oz, tlo), thi);
      const uint32_t u = (uint32_t)d_f2i(fu);
      const uint32_t depth = (uint32_t)d_f2i(1.0f / oz) & 0xfffeu;
      const int cw = cofs + (int)u;
      const uint32_t col = cw < ccap ? cached[cw] : colours[u];
      row[y] = col + (depth << 16);
    }
    const uint64_t dm = __ballot(draw);
    drawn += (int)__popcll(dm);
    bits_or(bits, base, dm, lane);  // pixel base+i is bit i of dm
    wave_sync();
    if (base + 64 < s2) {
      uzb = readlane_f(uz, 63) + u2dz;
      ozb = readlane_f(oz, 63) + onedz2;
    }
  }
  return drawn;
}

// The clear rows of a span [s1, e2): for spans of at most 64 rows one LDS
// read and one ballot give both the first clear row y0 (which fixes the
// span's start value, Cuda_Render.h:678-680) and the rows to draw.
struct SpanWindow {
  int y0;      // first clear row, or e2
  uint64_t m;  // fused: clear rows, bit j = row s1 + j
  bool fused;
};

__device__ inline SpanWindow span_window(const uint32_t* bits, int s1, int e2, int lane) {
  SpanWindow w;
  if (e2 - s1 <= 64) {
    const bool clear = lane < e2 - s1 && !lds_bit(bits, s1 + lane);
    w.m = uni64(__ballot(clear));
    w.y0 = uni(w.m ? s1 + (int)__builtin_ctzll(w.m) : e2);
    w.fused = true;
  } else {
    w.y0 = first_clear(bits, s1, e2, lane);
    w.m = 0;
    w.fused = false;
  }
  return w;
}

// Draws the clear rows of [w.y0, e2) (Cuda_Render.h:687-733): pixel y0+k has
// had k per-pixel increments.
__device__ inline int span_draw(const SpanWindow& w, int s1, int e2, float uzb, float ozb,
                                float u2dz, float onedz2, float tlo, float thi,
                                const uint16_t* cached, int ccap, const uint16_t* colours,
                                int cofs, uint32_t* __restrict__ row, uint32_t* bits, int lane) {
  if (!w.fused)
    return draw_span(w.y0, e2, uzb, ozb, u2dz, onedz2, tlo, thi, cached, ccap, colours, cofs, row,
                     bits, lane);
  if ((w.m >> lane) & 1) {
    float uz = uzb, oz = ozb;
    for (int i = w.y0 - s1; i < lane; i++) {
      uz += u2dz;
      oz += onedz2;
    }
    const float fu = fminf(fmaxf(uz / oz, tlo), thi);
    const uint32_t u = (uint32_t)d_f2i(fu);
    const uint32_t depth = (uint32_t)d_f2i(1.0f / oz) & 0xfffeu;
    const int cw = cofs + (int)u;
    const uint32_t col = cw < ccap ? cached[cw] : colours[u];
    row[s1 + lane] = col + (depth << 16);
  }
  bits_or(bits, s1, w.m, lane);
  wave_sync();
  return (int)__popcll(w.m);
}

// Uniform DDA walk state of one line (Cuda_Render.h:270-367).
struct DdaWalk {
  float dd0, dd1, i0x, i0y, i1x, i1y;  // next crossing distances / positions
  float g0x, g0y, g1x, g1y, gd0, gd1;  // their increments
  float px, py, dnow;                  // last step taken
  int idx;
  int mip, gridx, gridz, mapswitch, shift;
  int dzlog;  // dz = 2^dzlog
  float dz, z, yswitch;
  const uint2* map;
  const uint16_t* slab;
};

// One chunk of up to 64 DDA steps: lane k holds step k and its gathered
// pointer-map entry.  The mip level is constant within a chunk.
struct DdaChunk {
  int L, mip, shift;
  const uint16_t* slab;
  float db, dn, pbx, pby;  // per lane: distances and position before the step
  int ib;
  uint2 ent;
};

// Advances the walk by one chunk.  Returns false when the line has reached
// the far distance.
__device__ inline bool dda_next_chunk(DdaWalk& w, DdaChunk& c, const SceneDev& sc, int zfar,
                                      float* dl, int lane) {
  // distance mip switch (Cuda_Render.h:343-365), uniform
  while (w.z > (float)w.mapswitch || w.yswitch > 512.0f) {
    w.yswitch = w.yswitch * 0.5f;
    if (w.mip < sc.nmips - 1) {
      w.mip++;
      w.gridx >>= 1;
      w.gridz >>= 1;
      w.map = sc.mip[w.mip].map;
      w.slab = sc.mip[w.mip].slabs;
      w.shift = sc.mip[w.mip].shift;
    }
    w.g0x *= 2;
    w.g0y *= 2;
    w.g1x *= 2;
    w.g1y *= 2;
    w.gd0 *= 2;
    w.gd1 *= 2;
    w.mapswitch *= 2;
    w.dz *= 2;
    w.dzlog++;
  }
  // Iterations left before the next mip switch and before z passes z_far.
  // z, dz = 2^dzlog and mapswitch are exact integers below 2^31, and
  // z <= mapswitch here; a negative distance to z_far ends the line either
  // way, so the shifts decide exactly as the divisions would.
  const int zi = (int)w.z;
  const int n_mip = ((w.mapswitch - zi) >> w.dzlog) + 1;
  const int n_far = (zfar - zi) >> w.dzlog;
  if (n_far <= 0) return false;
  const int L = min(64, min(n_mip, n_far));
  c.L = L;
  c.mip = w.mip;
  c.shift = w.shift;
  c.slab = w.slab;
  c.db = c.dn = c.pbx = c.pby = 0;
  c.ib = 0;
  // The six crossing sequences (distances, x and z positions of the x and z
  // crossings) are independent recurrences s <- s + g.  Lane j < 6 walks
  // sequence j for 63 additions, storing every term to LDS (one add and one
  // store per term for all six sequences at once); each term is made by the
  // same additions the sequential walk makes (the k-th x-crossing term has
  // seen k additions whatever the interleaving).  Lane k then reads term k.
  // The walk's order of x and z steps is the merge of the two distance
  // sequences (ties go to x: the walk takes z only when strictly closer).
  // Layout: six sequences of 65 terms at a stride of 68 floats (the writers,
  // lanes 0..5, hit different LDS banks), then 64 merge slots.
  constexpr int kSeq = 68;
  float* sA = dl;
  float* sB = dl + kSeq;
  float* sX0 = dl + 2 * kSeq;
  float* sY0 = dl + 3 * kSeq;
  float* sX1 = dl + 4 * kSeq;
  float* sY1 = dl + 5 * kSeq;
  int* slot = reinterpret_cast<int*>(dl + 6 * kSeq);
  {
    if (lane < 6) {
      float t = w.dd0, g = w.gd0;
      t = lane == 1 ? w.dd1 : t;
      g = lane == 1 ? w.gd1 : g;
      t = lane == 2 ? w.i0x : t;
      g = lane == 2 ? w.g0x : g;
      t = lane == 3 ? w.i0y : t;
      g = lane == 3 ? w.g0y : g;
      t = lane == 4 ? w.i1x : t;
      g = lane == 4 ? w.g1x : g;
      t = lane == 5 ? w.i1y : t;
      g = lane == 5 ? w.g1y : g;
      float* dst = dl + lane * kSeq;
      dst[0] = t;
#pragma unroll
      for (int k = 1; k <= 64; k++) {  // term 64 is the state after a full chunk
        t = t + g;
        dst[k] = t;
      }
    }
    wave_sync();
    const float ak = sA[lane], bk = sB[lane];
    // Merge ranks: rx = lane + #{j : sB[j] < ak}, rz = lane + #{j : sA[j] <= bk}.
    // The sequences grow by about gd per term, so a guess g is off by at most
    // one in practice; three probes around it settle the rank exactly
    // (the sequences are sorted), else the lane searches.
    const float qx = (ak - w.dd1) * __builtin_amdgcn_rcpf(w.gd1);
    const float qz = (bk - w.dd0) * __builtin_amdgcn_rcpf(w.gd0);
    const int gx = (int)fminf(fmaxf(ceilf(qx), 0.0f), 64.0f);
    const int gz = (int)fminf(fmaxf(ceilf(qz), 0.0f), 64.0f);
    const float bm = sB[max(gx - 1, 0)], b0 = sB[gx], bp = sB[min(gx + 1, 64)];
    const float am = sA[max(gz - 1, 0)], a0 = sA[gz], ap = sA[min(gz + 1, 64)];
    const bool okx = (gx == 0 || bm < ak) && (gx >= 63 || !(bp < ak));
    const bool okz = (gz == 0 || !(bk < am)) && (gz >= 63 || bk < ap);
    int cx = gx + ((gx < 64 && b0 < ak) ? 1 : 0);
    int cz = gz + ((gz < 64 && !(bk < a0)) ? 1 : 0);
    if (!okx) cx = rank_in<false>(sB, ak, gx);
    if (!okz) cz = rank_in<true>(sA, bk, gz);
    const int rx = lane + cx;
    const int rz = lane + cz;
    if (rx < L) slot[rx] = lane;
    if (rz < L) slot[rz] = 64 + lane;
    const int nx = uni((int)__popcll(__ballot(rx < L)));
    const int nz = L - nx;
    // The walk's state after the chunk, read in one round trip: the next
    // terms (index nx of the x sequences, nz of the z sequences) and the last
    // step, which is the later of x-term nx-1 and z-term nz-1 (a z-term
    // follows an equal x-term).
    const float nA = sA[nx], nX0 = sX0[nx], nY0 = sY0[nx];
    const float nB = sB[nz], nX1 = sX1[nz], nY1 = sY1[nz];
    const float lA = sA[max(nx - 1, 0)], lX0 = sX0[max(nx - 1, 0)], lY0 = sY0[max(nx - 1, 0)];
    const float lB = sB[max(nz - 1, 0)], lX1 = sX1[max(nz - 1, 0)], lY1 = sY1[max(nz - 1, 0)];
    wave_sync();
    const int cc = lane < L ? slot[lane] : 0;
    const int p = (lane < L && lane > 0) ? slot[lane - 1] : 0;
    if (lane < L) {
      c.dn = cc < 64 ? sA[cc] : sB[cc - 64];
      if (lane == 0) {
        c.db = w.dnow;
        c.pbx = w.px;
        c.pby = w.py;
        c.ib = w.idx;
      } else {
        c.db = p < 64 ? sA[p] : sB[p - 64];
        c.pbx = p < 64 ? sX0[p] : sX1[p - 64];
        c.pby = p < 64 ? sY0[p] : sY1[p - 64];
        c.ib = p < 64 ? 0 : 1;
      }
    }
    const bool last_z = nz > 0 && (nx == 0 || !(lB < lA));
    w.dnow = uni_f(last_z ? lB : lA);
    w.px = uni_f(last_z ? lX1 : lX0);
    w.py = uni_f(last_z ? lY1 : lY0);
    w.idx = last_z ? 1 : 0;
    w.dd0 = uni_f(nA);
    w.i0x = uni_f(nX0);
    w.i0y = uni_f(nY0);
    w.dd1 = uni_f(nB);
    w.i1x = uni_f(nX1);
    w.i1y = uni_f(nY1);
    wave_sync();
  }
  w.z += (float)L * w.dz;
  return true;
}

// Cell of each step of the chunk just walked (Cuda_Render.h:416-442) and the
// load of its pointer-map entry (left in flight).
__device__ inline void dda_gather(const DdaWalk& w, DdaChunk& c, float vpx, float vpz,
                                  float fix_xf, float fix_yf, int lane) {
  const int fix_x = (int)((float)(1 - c.ib) * fix_xf);
  const int fix_z = (int)((float)c.ib * fix_yf);
  const int voxel_x = d_f2i(vpx + c.pbx) + fix_x;
  const int voxel_z = d_f2i(vpz + c.pby) + fix_z;
  const int vx = (voxel_x >> w.mip) & (w.gridx - 1);
  const int vz = (voxel_z >> w.mip) & (w.gridz - 1);
  c.ent = make_uint2(0u, 0u);
  if (lane < c.L) c.ent = w.map[vx + vz * w.gridx];
}

// Chunks handed from the producer wavefront (DDA + pointer gather) to the
// consumer wavefront (projection, run cache, replay) of a two-wavefront
// line: three slots used round robin.
struct ChunkRing {
  float db[3][64], dn[3][64];
  uint2 ent[3][64];
  const uint16_t* slab[3];
  int L[3], mip[3], shift[3], have[3];
  int stop[3];  // written by the consumer in the round that uses the slot
};

// The fields the consumer reads: distances, level and the gathered entry.
__device__ inline void ring_put_dda(ChunkRing& r, int s, const DdaChunk& c, bool have, int lane) {
  if (have) {
    r.db[s][lane] = c.db;
    r.dn[s][lane] = c.dn;
  }
  if (lane == 0) {
    r.slab[s] = c.slab;
    r.L[s] = have ? c.L : 0;
    r.mip[s] = c.mip;
    r.shift[s] = c.shift;
    r.have[s] = have ? 1 : 0;
  }
}

__device__ inline void ring_put_ent(ChunkRing& r, int s, const DdaChunk& c, int lane) {
  r.ent[s][lane] = c.ent;  // waits for the gather
}

__device__ inline void ring_get(const ChunkRing& r, int s, DdaChunk& c, int lane) {
  c.L = uni(r.L[s]);
  c.mip = uni(r.mip[s]);
  c.shift = uni(r.shift[s]);
  c.slab = (const uint16_t*)uni64((uint64_t)r.slab[s]);
  c.db = r.db[s][lane];
  c.dn = r.dn[s][lane];
  c.ent = r.ent[s][lane];
}

// TWO: the line is walked by a 128-thread block, wavefront 0 producing the
// DDA chunks (and their pointer gathers) two chunks ahead of wavefront 1,
// which consumes them; the two halves of the work overlap.  Otherwise one
// wavefront does both, software-pipelined.  PAR_ELEMS: the ordered replay
// precomputes a column's remaining elements across the lanes (pays where
// steps often show several elements, as in grazing views of caved terrain;
// the persistent batch kernel leaves it off, where its registers spill).
// LDS arguments: bits = the line's bitmap, dl = DDA scratch, rc = run cache,
// sweep_lds = 64 ints for the horizon sweep followed by kWalkEndWords words
// for the early-end test's constants.
template <bool COUNT, bool TWO = false, bool PAR_ELEMS = true, int RC = kRunCache>
__device__ inline void render_line_wave(int x, const SceneDev& sc, const rlevox_raymap& rm,
                                        uint32_t* __restrict__ row, int32_t* __restrict__ clip,
                                        uint32_t* bits, float* dl, uint16_t* rc, int* sweep_lds,
                                        unsigned long long* counts, ChunkRing* ring = nullptr) {
  static_assert(!(COUNT && TWO), "the counting variant walks with one wavefront");
  const int lane = threadIdx.x & 63;
  const bool producer = TWO && threadIdx.x < 64;  // wavefront 0 of a two-wavefront line
  const int rs = rm.rs;
  const float half = (float)(rs / 2);
  if (lane == 0 && !producer) {
    clip[0] = 0;
    clip[1] = -1;
  }

  // ---- line plane (Cuda_Render.h:128-172), uniform
  const int r0 = rm.res[0], r1 = r0 + rm.res[1], r2 = r1 + rm.res[2];
  int q = 0;
  if (x >= r2) q = 3;
  else if (x >= r1) q = 2;
  else if (x >= r0) q = 1;
  float qofs = (float)x;
  if (q == 1) qofs -= (float)r0;
  if (q == 2) qofs -= (float)r1;
  if (q == 3) qofs -= (float)r2;
  const float a = qofs / (float)rm.res[q];
  F3 p1 = f3ld(rm.p_2d[5]);
  const F3 e0 = f3ld(rm.p_no[2 * q]);
  F3 p2 = e0 + (f3ld(rm.p_no[2 * q + 1]) - e0) * a;
  {
    const float s2 = line_scale(p1, p2, rm.clip_max, rm.clip_min);
    const F3 c2 = p1 + (p2 - p1) * s2;
    const float s1 = line_scale(p2, p1, rm.clip_max, rm.clip_min);
    const F3 c1 = p2 + (p1 - p2) * s1;
    p1 = c1;
    p2 = c2;
  }
  const F3 p1_3d = mat_apply(rm.to3d, p1 * 4.0f);
  const F3 p2_3d = mat_apply(rm.to3d, p2 * 4.0f);
  F3 dir = (p1_3d + p2_3d) * 0.5f - f3ld(rm.p4);
  dir.y = 0;
  {
    const float d = dir.x * dir.x + dir.y * dir.y + dir.z * dir.z;
    const float inv = 1.0f / sqrtf(d);
    dir = dir * inv;
  }
  const float ray_x = rm.cos_myaw * dir.x + rm.sin_myaw * dir.z;
  const float ray_z = rm.cos_myaw * dir.z - rm.sin_myaw * dir.x;
  const bool vertical = q < 2;
  const float sin_x = rm.sin_x, cos_x = rm.cos_x;
  bool reverse = false;
  if (vertical && ray_z <= 0) reverse = true;
  if (!vertical && ray_x <= 0 && sin_x > 0) reverse = true;
  if (!vertical && ray_x > 0 && sin_x < 0) reverse = true;
  float hmul = reverse ? -half : half;
  if (vertical) hmul = -hmul;
  const int p_add = reverse ? 1 : -2;
  int p1x = d_f2i((float)rs * p1.x) + p_add;
  int p1y = d_f2i((float)rs * p1.y) + p_add;
  int p2x = d_f2i((float)rs * p2.x) - p_add;
  int p2y = d_f2i((float)rs * p2.y) - p_add;
  p1x = min(max(p1x, 0), rs - 1);
  p1y = min(max(p1y, 0), rs - 1);
  p2x = min(max(p2x, 0), rs - 1);
  p2y = min(max(p2y, 0), rs - 1);
  if (p1y == p2y) return;
  int ymin, ymax;
  if (vertical) {
    ymin = rs - 1 - p1y;
    ymax = rs - 1 - p2y;
  } else {
    ymin = rs - 1 - p1x;
    ymax = rs - 1 - p2x;
  }
  if (reverse) {
    ymin = rs - 1 - ymin;
    ymax = rs - 1 - ymax;
  }
  if (ymin > ymax) {
    const int t = ymin;
    ymin = ymax;
    ymax = t;
  }
  if (ymin >= ymax) return;

  // ---- clear the range and the bitmap (Cuda_Render.h:255-264), coalesced
  if (!producer) {
    for (int n = ymin + lane; n <= ymax; n += 64) row[n] = RLEVOX_CLEAR_TEXEL;
    const int words = bitmap_words(rs);
    for (int w = lane; w < words; w += 64) bits[w] = 0;
    if (lane == 0) {
      clip[0] = ymin;
      clip[1] = ymax;
    }
    wave_sync();
  }
  LineCounts cnt;
  if (COUNT) cnt.cleared = (uint64_t)(ymax - ymin + 1);

  // ---- DDA setup (Cuda_Render.h:270-304), uniform
  const float vpx = rm.position[0], vpy = rm.position[1], vpz = rm.position[2];
  const float ddx = ray_x * rm.cos_y + ray_z * rm.sin_y;
  const float ddy = ray_x * rm.sin_y - ray_z * rm.cos_y;
  float fix_xf = -1, fix_yf = -1, sign_x = -1, sign_y = -1;
  float frac_x = vpx - (float)d_f2i(vpx);
  float frac_y = vpz - (float)d_f2i(vpz);
  if (ddx >= 0) {
    fix_xf = 0;
    sign_x = 1;
    frac_x = 1 - frac_x;
  }
  if (ddy >= 0) {
    fix_yf = 0;
    sign_y = 1;
    frac_y = 1 - frac_y;
  }
  DdaWalk wk;
  wk.g0y = ddy / fabsf(ddx);
  wk.g0x = sign_x;
  wk.g1x = ddx / fabsf(ddy);
  wk.g1y = sign_y;
  wk.i0x = wk.g0x * frac_x;
  wk.i0y = wk.g0y * frac_x;
  wk.i1x = wk.g1x * frac_y;
  wk.i1y = wk.g1y * frac_y;
  wk.gd0 = sqrtf(wk.g0x * wk.g0x + wk.g0y * wk.g0y);
  wk.gd1 = sqrtf(wk.g1x * wk.g1x + wk.g1y * wk.g1y);
  wk.dd0 = sqrtf(wk.i0x * wk.i0x + wk.i0y * wk.i0y);
  wk.dd1 = sqrtf(wk.i1x * wk.i1x + wk.i1y * wk.i1y);
  wk.px = wk.py = wk.dnow = 0;
  wk.idx = 0;
  wk.mip = 0;
  wk.gridx = sc.mip[0].sx;
  wk.gridz = sc.mip[0].sz;
  wk.map = sc.mip[0].map;
  wk.slab = sc.mip[0].slabs;
  wk.shift = sc.mip[0].shift;
  wk.mapswitch = rm.mapswitch;
  wk.dz = 1;
  wk.dzlog = 0;
  wk.z = 0;
  wk.yswitch = vpy;
  const float z_add = sin_x;
  const float y_add = (vertical ? cos_x : 0.0f) * hmul;
  const int zfar = rm.far_distance;
  const float mountain = vpy;
  // the early-end test's constants, in LDS behind the sweep marks
  WalkEnd* wend = reinterpret_cast<WalkEnd*>(sweep_lds + 64);
  float wfloor = 0.0f;  // held in a scalar register across the walk
  if (!producer) {
    if (lane == 0) wend->init(cos_x, sin_x, hmul, ray_x, ray_z, vertical, mountain, half);
    wave_sync();
    wfloor = uni_f(wend->floor_());
  }

  // Software pipeline: while chunk c's run-cache loads are in flight the DDA
  // of chunk c+1 runs, and while chunk c is replayed the pointer-map gather
  // of chunk c+1 is in flight.  The walk does
  // not depend on the replay (only the line's end does), so a chunk
  // computed past the end of the line is simply dropped.
  // The chunk loop is unrolled twice over two register sets (chunk c in one,
  // c+1 in the other), so the gathered entries never have to be copied (a
  // copy would wait for the load).
  auto chunk_step = [&](auto produce, const DdaChunk& cur, DdaChunk& nxt, bool& have_next) -> bool {
    if (ymin >= ymax) return false;
    const int L = cur.L, mip = cur.mip, shift = cur.shift;
    const uint16_t* slab = cur.slab;
    const uint2 ent = cur.ent;
    const uint64_t wofs = (uint64_t)ent.x << shift;  // column's slab word offset

    // ---- per-lane stateless projection (Cuda_Render.h:444-466)
    const bool valid = lane < L;
    const float ddelta = cur.dn - cur.db;
    const float vs_x = ray_x * cur.db;
    const float vs_z = ray_z * cur.db;
    const float corr_x = ray_x * ddelta;
    const float corr_z = ray_z * ddelta;
    const float pz = cos_x * vs_z + sin_x * mountain;
    float py3 = vertical ? cos_x * mountain - sin_x * vs_z : vs_x;
    py3 *= hmul;
    const float czz = cos_x * corr_z;
    float cyy = vertical ? -sin_x * corr_z : corr_x;
    cyy *= hmul;
    const float top = pz * half + py3;
    const uint32_t slen = ent.y & 0xffffu;
    const uint16_t* runs = slab + 2 + wofs;

    // F: projected top of the first drawn element in front of the eye.  A
    // step whose cell top is already below the horizon with pz > 0 stays
    // culled as the horizon rises (the replay's clip-top test, monotone in
    // ymin for pz > 0): it can never act, and its F is not needed.
    const bool culled0 = pz > 0.0f && top <= pz * (float)ymin;
    int F = INT32_MIN, kF = -1, ygF = 0;
    uint32_t nel = 0;
    float texF = 0;
    int textureF = 0;
    float ytopF = 0, ybotF = 0;
    if (!culled0) {
      uint32_t w = ent.y >> 16;
      int yg = 0;
      float tex = 0;
      for (uint32_t k = 0; k < slen; k++) {
        if (k > 0) w = runs[k];
        nel++;
        int dr = (int)(w >> 10);
        yg += (int)(w & 1023u) << mip;
        if (dr == 0) continue;
        const int texture = (int)tex;
        tex += (float)dr;
        dr <<= mip;
        const float ytop = (float)yg;
        yg += dr;
        float zz1 = pz, yy1 = py3;
        if (mountain + ytop >= 0) {
          zz1 += czz;
          yy1 += cyy;
        }
        const float pz1 = zz1 + z_add * ytop;
        if (pz1 <= 0) continue;
        const float py1 = yy1 + y_add * ytop;
        F = d_f2i(half + py1 / pz1);
        kF = (int)k;
        ygF = yg;
        texF = tex;
        textureF = texture;
        ytopF = ytop;
        ybotF = (float)yg;
        break;
      }
    }

    // ---- run cache: the head of each candidate column, fetched now and
    // stored to LDS after the next chunk's DDA (its latency overlaps that
    // work).  The horizon only rises, so F > ymin now is necessary for
    // acting; the clip-top test is monotone in ymin only for pz > 0.
    const bool cand = valid && F > ymin && !culled0;
    const uint64_t rc_base = wofs & ~(uint64_t)7;  // 16-byte aligned start
    const int rc_delta = (int)(wofs - rc_base);
    const uint16_t* cached = rc + lane * RC + rc_delta;
    const int ccap = cand ? RC - rc_delta : 0;
    static_assert(RC % 8 == 0 && RC >= 24, "the run cache is fetched in 16-byte words");
    constexpr int kWords = RC / 8;
    typedef unsigned int rc_word __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) const rc_word g_rc_word;
    rc_word rcw[kWords];  // loaded and stored only for candidates
    if (cand) {
      // Global loads: in the drawing wavefront of a two-wavefront line the
      // slab pointer comes through LDS, and flat loads would make each LDS
      // wait until they return wait for them too.
      const g_rc_word* src = (const g_rc_word*)(slab + rc_base);
#pragma unroll
      for (int i = 0; i < kWords; i++) rcw[i] = src[i];
    }

    // ---- next chunk's DDA (the run cache loads are in flight meanwhile)
    if constexpr (decltype(produce)::value) have_next = dda_next_chunk(wk, nxt, sc, zfar, dl, lane);

    // The stores wait for the loads: a walking wavefront stores after the DDA
    // above, a drawing-only one after the span parameters below.
    auto store_rc = [&]() {
      if (cand) {
        rc_word* dst = reinterpret_cast<rc_word*>(rc + lane * RC);
#pragma unroll
        for (int i = 0; i < kWords; i++) dst[i] = rcw[i];
      }
    };
    if constexpr (decltype(produce)::value) store_rc();
    // ---- next chunk's pointer gather: in flight during this chunk's replay
    if constexpr (decltype(produce)::value) {
      if (have_next) dda_gather(wk, nxt, vpx, vpz, fix_xf, fix_yf, lane);
    }
    // first element's lower end and span parameters (Cuda_Render.h:554-660)
    int s1F = 0;
    bool drawF = false;
    float s1rF = 0, u1zF = 0, u2dzF = 0, onez1F = 0, onedz2F = 0;
    int F2 = INT32_MIN;
    if (cand) {
      float zz2 = pz, yy2 = py3;
      if (mountain + ybotF < 0) {
        zz2 += czz;
        yy2 += cyy;
      }
      const float pz2 = zz2 + z_add * ybotF;
      if (pz2 > 0) {
        drawF = true;
        const float py2 = yy2 + y_add * ybotF;
        s1F = d_f2i(half + py2 / pz2 - 1.0f);
        const float pz1r = pz + z_add * ytopF;
        const float py1r = py3 + y_add * ytopF;
        const float pz2r = pz + z_add * ybotF;
        const float py2r = py3 + y_add * ybotF;
        const float s2r = half + py1r / pz1r;
        s1rF = half + py2r / pz2r;
        u1zF = texF / pz2r;
        u2dzF = (float)textureF / pz1r - u1zF;
        onez1F = 1.0f / pz2r;
        onedz2F = 1.0f / pz1r - onez1F;
        u2dzF /= s2r - s1rF;
        onedz2F /= s2r - s1rF;
      }
    }
    if constexpr (!decltype(produce)::value) store_rc();
    wave_sync();  // run cache filled
    if (cand) {  // F2: projected top of the next element the loop tests
      int yg = ygF;
      for (uint32_t k = (uint32_t)kF + 1; k < slen; k++) {
        const int ci = (int)k + 2;
        const uint32_t w = ci < ccap ? cached[ci] : runs[k];
        int dr = (int)(w >> 10);
        yg += (int)(w & 1023u) << mip;
        if (dr == 0) continue;
        dr <<= mip;
        const float ytop = (float)yg;
        yg += dr;
        float zz1 = pz, yy1 = py3;
        if (mountain + ytop >= 0) {
          zz1 += czz;
          yy1 += cyy;
        }
        const float pz1 = zz1 + z_add * ytop;
        if (pz1 <= 0) continue;
        const float py1 = yy1 + y_add * ytop;
        F2 = d_f2i(half + py1 / pz1);
        break;
      }
    }

    // ---- ordered replay of the steps that act
    int cur_step = 0;
    bool done = false;
    for (;;) {
      ymin = uni(ymin);
      ymax = uni(ymax);
      if (ymin >= ymax) {
        done = true;
        break;
      }
      const bool culled = top <= pz * (float)ymin;
      const bool act = valid && lane >= cur_step && !culled && F > ymin;
      const uint64_t am = uni64(__ballot(act));
      if (!COUNT && (am & (am - 1))) {
        // ---- horizon sweep: a run of acting steps applied at once.  The
        // typical step draws its first element's span from the horizon up
        // (s1 <= ymin), leaves the top clip alone (F < ymax) and breaks at
        // its next element (F2 <= F): its only effects are ymin = F and the
        // rows [ymin, F) drawn.  The horizon before step t is then the prefix
        // maximum M_t of the earlier actors' F.  Each lane checks that, under
        // M_t, it acts exactly when the speculation says it does and that
        // its step is of that simple kind; the steps before the first lane
        // failing a check, whose spans together stay within 64 rows that are
        // all clear in the bitmap (row F included, where the horizon stops),
        // are applied here exactly as the ordered replay would apply them.
        const int Fs = act ? F : INT32_MIN;
        const int M = prefix_max_excl(Fs, ymin, lane);
        const bool spec_actor = act && F > M;
        const bool true_actor = valid && lane >= cur_step && !(top <= pz * (float)M) && F > M;
        const bool simple = cand && drawF && s1F <= M && F < ymax && F2 <= F && F - ymin <= 63;
        const bool ok = spec_actor == true_actor && (!spec_actor || simple);
        const uint64_t bad = uni64(__ballot(lane >= cur_step && !ok));
        const int limit = uni(bad ? (int)__builtin_ctzll(bad) : 64);
        const uint64_t acc = uni64(__ballot(spec_actor && lane < limit));
        if (acc) {
          const int last = 63 - (int)__builtin_clzll(acc);
          const int fend = uni(readlane_i(F, last));
          // rows ymin..fend (fend < ymin + 64) must all be clear
          const bool rclear = lane > fend - ymin || !lds_bit(bits, ymin + lane);
          if (uni64(__ballot(rclear)) == ~0ull) {
            // row r = ymin + j belongs to the accepted step with the highest
            // start row <= j (starts are marked in LDS)
            sweep_lds[lane] = -1;
            wave_sync();
            if ((acc >> lane) & 1) sweep_lds[M - ymin] = lane;
            wave_sync();
            const int mark = sweep_lds[lane];
            const uint64_t smask = uni64(__ballot(mark >= 0));
            const int nrows = fend - ymin;
            const bool mine = lane < nrows;
            const uint64_t upto = lane >= 63 ? ~0ull : ((2ull << lane) - 1ull);
            const int srow = 63 - (int)__builtin_clzll((smask & upto) | 1ull);
            const int a0 = __shfl(mark, srow);
            const int a = mine ? a0 : 0;
            const float s1r = __shfl(s1rF, a), u1z = __shfl(u1zF, a), u2dz = __shfl(u2dzF, a);
            const float onez1 = __shfl(onez1F, a), onedz2 = __shfl(onedz2F, a);
            const float thi = __shfl(texF, a) - 1.0f;
            const float tlo = (float)__shfl(textureF, a);
            const uint32_t aofs = (uint32_t)__shfl((int)ent.x, a);
            const uint32_t alen = (uint32_t)__shfl((int)ent.y, a) & 0xffffu;
            if (mine) {
              const int y0 = ymin + srow;
              const float mult = (float)(y0 + 1) - s1r;
              float uz = u1z + u2dz * mult, oz = onez1 + onedz2 * mult;
              for (int i = srow; i < lane; i++) {
                uz += u2dz;
                oz += onedz2;
              }
              const float fu = fminf(fmaxf(uz / oz, tlo), thi);
              const uint32_t u = (uint32_t)d_f2i(fu);
              const uint32_t depth = (uint32_t)d_f2i(1.0f / oz) & 0xfffeu;
              const uint64_t wofs = (uint64_t)aofs << shift;
              const int tdelta = (int)(wofs & 7u);
              const int cw = 2 + (int)alen + tdelta + (int)u;
              const uint32_t col = cw < RC ? rc[a * RC + cw]
                                                  : slab[2 + wofs + alen + u];
              row[ymin + lane] = col + (depth << 16);
            }
            bits_or(bits, ymin, nrows >= 64 ? ~0ull : ((1ull << nrows) - 1ull), lane);
            wave_sync();
            ymin = fend;
            cur_step = limit;
            continue;
          }
        }
      }
      const int t = uni(am ? (int)__builtin_ctzll(am) : L);
      if (COUNT) {
        // no-op steps [cur_step, t): pointer loads and elements walked to the break
        const bool noop_loaded = valid && lane >= cur_step && lane < t && !culled;
        cnt.steps += (uint64_t)(t - cur_step);
        cnt.ptr += (uint64_t)__popcll(__ballot(noop_loaded));
        unsigned long long e = noop_loaded ? nel : 0, r = (noop_loaded && nel) ? nel - 1 : 0;
        for (int o = 32; o > 0; o >>= 1) {
          e += __shfl_xor(e, o);
          r += __shfl_xor(r, o);
        }
        cnt.elems += e;
        cnt.rle += r;
      }
      if (!am) break;
      if (COUNT) {
        cnt.steps++;
        cnt.ptr++;
      }

      const uint64_t tofs = (uint64_t)(uint32_t)readlane_i((int)ent.x, t) << shift;
      const uint32_t tlf = (uint32_t)readlane_i((int)ent.y, t);
      const uint32_t tslen = tlf & 0xffffu;
      const uint16_t* truns = slab + 2 + tofs;
      const uint16_t* tcol = truns + tslen;
      const bool tcached = readlane_i((int)cand, t) != 0;
      const int tdelta = (int)(tofs & 7u);
      const uint16_t* tcache = rc + t * RC + tdelta;
      const int tcap = tcached ? RC - tdelta : 0;
      const float tpz = readlane_f(pz, t), tpy = readlane_f(py3, t);
      const float tczz = readlane_f(czz, t), tcyy = readlane_f(cyy, t);

      // element loop state after the first drawn element in front
      uint32_t k0;
      int ygen;
      float tex;
      const bool fast = !COUNT && tcached;
      if (fast) {
        // ---- first drawn element: precomputed (Cuda_Render.h:540-733)
        const int s2 = readlane_i(F, t);  // > ymin by the ballot
        const bool tdraw = readlane_i((int)drawF, t) != 0;
        int s1 = readlane_i(s1F, t);
        if (tdraw && s1 < ymax) {
          int e2 = s2;
          if (e2 >= ymax) {
            e2 = ymax;
            ymax = s1;
          }
          if (s1 <= ymin) {
            s1 = ymin;
            ymin = e2;
            if (ymax > ymin) ymin = first_clear(bits, ymin, ymax, lane);
          }
          const SpanWindow sw = span_window(bits, s1, e2, lane);
          if (sw.y0 < e2) {
            const float s1r = readlane_f(s1rF, t);
            const float u1z = readlane_f(u1zF, t), u2dz = readlane_f(u2dzF, t);
            const float onez1 = readlane_f(onez1F, t), onedz2 = readlane_f(onedz2F, t);
            const float mult = (float)(sw.y0 + 1) - s1r;
            const float uzb = u1z + u2dz * mult;
            const float ozb = onez1 + onedz2 * mult;
            const float thi = readlane_f(texF, t) - 1.0f;
            const float tlo = (float)readlane_i(textureF, t);
            span_draw(sw, s1, e2, uzb, ozb, u2dz, onedz2, tlo, thi, tcache, tcap, tcol,
                      2 + (int)tslen, row, bits, lane);
          }
        }
        // the next element tested breaks unless it reaches above ymin
        if (readlane_i(F2, t) <= ymin) {
          cur_step = t + 1;
          continue;
        }
        k0 = (uint32_t)readlane_i(kF, t) + 1;
        ygen = readlane_i(ygF, t);
        tex = readlane_f(texF, t);
      } else {
        k0 = 0;
        ygen = 0;
        tex = 0;
      }

      // ---- remaining elements (Cuda_Render.h:511-734)
      if constexpr (!COUNT && PAR_ELEMS) {
        // Lane j precomputes element k0 + base + j of the column: its run
        // word, its rows and texture range from prefix sums over the lanes
        // (integers, so equal to the sequential sums; tex holds integers
        // below 2^24, so its float sums are exact too), and every projection
        // and span parameter the sequential loop would compute for it, with
        // the same operations in the same order.  The uniform loop below
        // then applies the horizon clip, the bitmap and the drawing to one
        // element after another, reading those values by lane.
        bool stop = false;
        for (uint32_t base = k0; base < tslen && !stop; base += 64) {
          const uint32_t k = base + (uint32_t)lane;
          const bool in = k < tslen;
          uint32_t w = 0;
          if (in) w = k == 0 ? tlf >> 16 : ((int)k + 2 < tcap ? tcache[k + 2] : truns[k]);
          const int edrawn = (int)(w >> 10);
          const int eskip = (int)(w & 1023u) << mip;
          const int yinc = prefix_sum_incl(eskip + (edrawn << mip), lane);
          const int tinc = prefix_sum_incl(edrawn, lane);
          const int ytop_i = ygen + yinc - (edrawn << mip);  // after this element's skip
          const int texture = (int)(tex + (float)(tinc - edrawn));
          const float etex = tex + (float)tinc;  // tex after this element
          const float ytop = (float)ytop_i;
          const float ybot = (float)(ytop_i + (edrawn << mip));
          float zz1 = tpz, yy1 = tpy;
          if (mountain + ytop >= 0) {
            zz1 += tczz;
            yy1 += tcyy;
          }
          const float pz1 = zz1 + z_add * ytop;
          const float py1 = yy1 + y_add * ytop;
          const int es2 = d_f2i(half + py1 / pz1);
          float zz2 = tpz, yy2 = tpy;
          if (mountain + ybot < 0) {
            zz2 += tczz;
            yy2 += tcyy;
          }
          const float pz2 = zz2 + z_add * ybot;
          const float py2 = yy2 + y_add * ybot;
          const int es1 = d_f2i(half + py2 / pz2 - 1.0f);
          const float pz1r = tpz + z_add * ytop;
          const float py1r = tpy + y_add * ytop;
          const float pz2r = tpz + z_add * ybot;
          const float py2r = tpy + y_add * ybot;
          const float s2r = half + py1r / pz1r;
          const float es1r = half + py2r / pz2r;
          const float eu1z = etex / pz2r;
          float eu2dz = (float)texture / pz1r - eu1z;
          const float eonez1 = 1.0f / pz2r;
          float eonedz2 = 1.0f / pz1r - eonez1;
          eu2dz /= s2r - es1r;
          eonedz2 /= s2r - es1r;
          // flags: drawn, in front at the top, in front at the bottom
          const int eflags = (in && edrawn != 0 ? 1 : 0) | (pz1 > 0 ? 2 : 0) | (pz2 > 0 ? 4 : 0);
          const int nb = (int)min(64u, tslen - base);
          // Only elements that can act are visited.  Within the loop ymax
          // only falls and ymin never rises above the ymax of this point,
          // so an element that can neither draw (needs bottom row s1 <
          // ymax) nor stop the column (needs top row s2 <= ymin) from here
          // on stays inert: the loop below would reach one of its
          // `continue`s for it.
          const bool can_act = lane < nb && (eflags & 3) == 3 &&
                               (es2 <= ymax || ((eflags & 4) && es1 < ymax));
          for (uint64_t act = __ballot(can_act); act != 0; act &= act - 1) {
            const int j = (int)__builtin_ctzll(act);
            const int f = readlane_i(eflags, j);
            if (!(f & 1) || !(f & 2)) continue;
            int s2 = readlane_i(es2, j);
            if (s2 <= ymin) {
              stop = true;
              break;
            }
            if (!(f & 4)) continue;
            int s1 = readlane_i(es1, j);
            if (s1 >= ymax) continue;
            if (s2 >= ymax) {
              s2 = ymax;
              ymax = s1;
            }
            if (s1 <= ymin) {
              s1 = ymin;
              ymin = s2;
              if (ymax > ymin) ymin = first_clear(bits, ymin, ymax, lane);
            }
            const SpanWindow sw = span_window(bits, s1, s2, lane);
            if (sw.y0 >= s2) continue;
            const float s1r = readlane_f(es1r, j);
            const float u2dz = readlane_f(eu2dz, j), onedz2 = readlane_f(eonedz2, j);
            const float mult = (float)(sw.y0 + 1) - s1r;
            const float uzb = readlane_f(eu1z, j) + u2dz * mult;
            const float ozb = readlane_f(eonez1, j) + onedz2 * mult;
            span_draw(sw, s1, s2, uzb, ozb, u2dz, onedz2, (float)readlane_i(texture, j),
                      readlane_f(etex, j) - 1.0f, tcache, tcap, tcol, 2 + (int)tslen, row, bits,
                      lane);
          }
          ygen += readlane_i(yinc, 63);
          tex += (float)readlane_i(tinc, 63);
        }
        cur_step = t + 1;
        continue;
      }
      // ---- remaining elements (Cuda_Render.h:511-734), uniform
      for (uint32_t k = k0; k < tslen; k++) {
        uint32_t w;
        if (k == 0) w = tlf >> 16;
        else w = ((int)k + 2 < tcap) ? tcache[k + 2] : truns[k];
        if (COUNT) {
          cnt.elems++;
          if (k > 0) cnt.rle++;
        }
        int drawn = (int)(w >> 10);
        ygen += (int)(w & 1023u) << mip;
        if (drawn == 0) continue;
        const int texture = (int)tex;
        tex += (float)drawn;
        drawn <<= mip;
        const float ytop = (float)ygen;
        ygen += drawn;
        const float ybot = (float)ygen;
        float zz1 = tpz, yy1 = tpy;
        if (mountain + ytop >= 0) {
          zz1 += tczz;
          yy1 += tcyy;
        }
        const float pz1 = zz1 + z_add * ytop;
        if (pz1 <= 0) continue;
        const float py1 = yy1 + y_add * ytop;
        int s2 = d_f2i(half + py1 / pz1);
        if (s2 <= ymin) break;
        float zz2 = tpz, yy2 = tpy;
        if (mountain + ybot < 0) {
          zz2 += tczz;
          yy2 += tcyy;
        }
        const float pz2 = zz2 + z_add * ybot;
        if (pz2 <= 0) continue;
        const float py2 = yy2 + y_add * ybot;
        int s1 = d_f2i(half + py2 / pz2 - 1.0f);
        if (s1 >= ymax) continue;
        if (s2 >= ymax) {
          s2 = ymax;
          ymax = s1;
        }
        if (s1 <= ymin) {
          s1 = ymin;
          ymin = s2;
          if (ymax > ymin) ymin = first_clear(bits, ymin, ymax, lane);
        }
        const SpanWindow sw = span_window(bits, s1, s2, lane);
        if (sw.y0 >= s2) continue;
        const int y0 = sw.y0;
        const float pz1r = tpz + z_add * ytop;
        const float py1r = tpy + y_add * ytop;
        const float pz2r = tpz + z_add * ybot;
        const float py2r = tpy + y_add * ybot;
        const float s2r = half + py1r / pz1r;
        const float s1r = half + py2r / pz2r;
        const float u1z = tex / pz2r;
        float u2dz = (float)texture / pz1r - u1z;
        const float onez1 = 1.0f / pz2r;
        float onedz2 = 1.0f / pz1r - onez1;
        u2dz /= s2r - s1r;
        onedz2 /= s2r - s1r;
        const float mult = (float)(y0 + 1) - s1r;
        const float uzb = u1z + u2dz * mult;
        const float ozb = onez1 + onedz2 * mult;
        if (COUNT) cnt.spans++;
        const int d = span_draw(sw, s1, s2, uzb, ozb, u2dz, onedz2, (float)texture, tex - 1.0f,
                                tcache, tcap, tcol, 2 + (int)tslen, row, bits, lane);
        if (COUNT) cnt.col += (uint64_t)d;
      }
      cur_step = t + 1;
    }
    // ---- early end of the walk (WalkEnd): later steps lie beyond this
    // chunk's last crossing distance
    if (!done && !cnt.ended && wfloor < (float)ymin + (1.0f - WalkEnd::kAbs) &&
        wend->done(readlane_f(cur.dn, L - 1), ymin, sc.mip[mip].top, sc.sy)) {
      if (COUNT) cnt.end_walk();  // the counting variant walks on to count the reference's work
      else done = true;
    }
    return !done;
  };
  if constexpr (!TWO) {
    DdaChunk ca, cb;
    bool ha = dda_next_chunk(wk, ca, sc, zfar, dl, lane), hb = false;
    if (ha) dda_gather(wk, ca, vpx, vpz, fix_xf, fix_yf, lane);
    const std::true_type yes;
    for (;;) {
      if (!ha || !chunk_step(yes, ca, cb, hb)) break;
      if (!hb || !chunk_step(yes, cb, ca, ha)) break;
    }
  } else {
    // Round i: slot i%3 holds chunk i complete, slot (i+1)%3 chunk i+1's DDA
    // (its gather in flight in the producer).  The producer walks chunk i+2
    // and stores chunk i+1's entries while the consumer processes chunk i;
    // a block barrier ends the round.  Both wavefronts read the same flags
    // after each barrier (have[i%3] last written two rounds before, stop[i%3]
    // in this round), so they run the same rounds and the same barriers.
    ChunkRing& r = *ring;
    if (producer) {
      DdaChunk p0, p1;
      const bool h0 = dda_next_chunk(wk, p0, sc, zfar, dl, lane);
      ring_put_dda(r, 0, p0, h0, lane);
      if (h0) dda_gather(wk, p0, vpx, vpz, fix_xf, fix_yf, lane);
      const bool h1 = h0 && dda_next_chunk(wk, p1, sc, zfar, dl, lane);
      ring_put_dda(r, 1, p1, h1, lane);
      if (h0) ring_put_ent(r, 0, p0, lane);
      if (h1) dda_gather(wk, p1, vpx, vpz, fix_xf, fix_yf, lane);
      __syncthreads();
      // two register sets: `pend` holds the chunk whose gather is in flight
      auto produce_round = [&](int s, DdaChunk& pend, bool& hpend, DdaChunk& nxt) {
        const int s1 = s == 2 ? 0 : s + 1, s2 = s1 == 2 ? 0 : s1 + 1;
        const bool h2 = hpend && dda_next_chunk(wk, nxt, sc, zfar, dl, lane);
        ring_put_dda(r, s2, nxt, h2, lane);
        if (hpend) ring_put_ent(r, s1, pend, lane);
        if (h2) dda_gather(wk, nxt, vpx, vpz, fix_xf, fix_yf, lane);
        hpend = h2;
      };
      bool hp = h1;
      int s = 0;
      for (;;) {
        if (!uni(r.have[s])) break;
        produce_round(s, p1, hp, p0);
        __syncthreads();
        if (uni(r.stop[s])) break;
        s = s == 2 ? 0 : s + 1;
        if (!uni(r.have[s])) break;
        produce_round(s, p0, hp, p1);
        __syncthreads();
        if (uni(r.stop[s])) break;
        s = s == 2 ? 0 : s + 1;
      }
    } else {
      __syncthreads();
      const std::false_type no;
      DdaChunk cur, unused;
      bool hu = false;
      int s = 0;
      for (;;) {
        if (!uni(r.have[s])) break;
        ring_get(r, s, cur, lane);
        const bool more = chunk_step(no, cur, unused, hu);
        // A slot's flag is read after this round's barrier and rewritten
        // three rounds later, so no wavefront can see a later round's value.
        if (lane == 0) r.stop[s] = more ? 0 : 1;
        __syncthreads();
        if (uni(r.stop[s])) break;
        s = s == 2 ? 0 : s + 1;
      }
    }
  }
  if (COUNT && lane == 0) cnt.flush(counts);
}

}  // namespace rlevox
