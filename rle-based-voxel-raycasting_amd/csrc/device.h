// Device-side data layout and the per-line raycast routine shared by the
// plain and batched kernels.
//
// Arithmetic contract (kept identical on every path so framebuffers are
// bit-exact): fp32 in the reference's operation order, no FMA contraction
// (-ffp-contract=off), IEEE division and square root, saturating float->int
// conversion (v_cvt_i32_f32), NaN-ignoring fminf/fmaxf, and no
// transcendental function on the device (the ray map carries them).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rlevox.h"

#define RLEVOX_HD __host__ __device__
#include "walk_end.h"
#include "recip_div.h"

namespace rlevox {

// One mip level in HBM: the pointer map as 8-byte entries, the slab array.
struct MipDev {
  const uint2* map;
  const uint16_t* slabs;
  int32_t sx, sz;
  int32_t shift;  // column word offset = map[c].x << shift
  int32_t top;    // highest run top (smallest world row) of this level and all coarser ones
};

struct SceneDev {
  MipDev mip[RLEVOX_MAX_MIPS];
  int32_t nmips;
  int32_t sy;  // column height in world rows (level 0)
};

// Ray-buffer line bitmap words for render size rs.
__host__ __device__ inline int bitmap_words(int rs) { return (rs + 31) >> 5; }

// Saturating float -> int32 with NaN -> 0, branch-free: NaN is replaced
// first, the median-of-three clamps to [-2^31, 2^31 - 128] (the largest float
// below 2^31), where the conversion is exact truncation, and values from 2^31
// up select INT32_MAX.
__device__ inline int d_f2i(float f) {
  const float g = f != f ? 0.0f : f;
  const int r = (int)__builtin_amdgcn_fmed3f(g, -2147483648.0f, 2147483520.0f);
  return g >= 2147483648.0f ? INT32_MAX : r;
}

// LDS bitmap of one line, word w of this lane at base[w * stride].
struct LineBits {
  uint32_t* base;
  int stride;
  __device__ bool test(int y) const { return (base[(y >> 5) * stride] >> (y & 31)) & 1u; }
  __device__ void set(int y) const { base[(y >> 5) * stride] |= 1u << (y & 31); }
  __device__ void clear(int words) const {
    for (int w = 0; w < words; w++) base[w * stride] = 0;
  }
};

struct F3 {
  float x, y, z;
};
__device__ inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ inline F3 f3ld(const float* p) { return F3{p[0], p[1], p[2]}; }
__device__ inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }

// LineScale / ClipLine (Cuda_Render.h:39-65)
__device__ inline float line_scale(F3 in, F3 ctr, float cmax, float cmin) {
  float sx = 1, sy = 1;
  if (ctr.x > 1) sx = (1 - in.x) / (ctr.x - in.x);
  if (ctr.x < 0) sx = in.x / (in.x - ctr.x);
  if (ctr.y > cmax) sy = (cmax - in.y) / (ctr.y - in.y);
  if (ctr.y < cmin) sy = (-cmin + in.y) / (in.y - ctr.y);
  return (sx < sy) ? sx : sy;
}

// MatMul (Cuda_Render.h:67-73), row-vector convention.
__device__ inline F3 mat_apply(const float (*m)[4], F3 v) {
  return f3(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z + m[3][0],
            m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z + m[3][1],
            m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z + m[3][2]);
}

// Exact work counts of one line (the DETAIL_BENCH counters of
// Cuda_Render.h:14-22, extended); only the counting kernel variant keeps them.
struct LineCounts {
  uint64_t steps = 0, ptr = 0, rle = 0, elems = 0, spans = 0, col = 0, cleared = 0;
  // the same at the point where the kernels end the walk
  bool ended = false;
  uint64_t wsteps = 0, wptr = 0, wrle = 0, welems = 0;
  __device__ void end_walk() {
    if (ended) return;
    ended = true;
    wsteps = steps;
    wptr = ptr;
    wrle = rle;
    welems = elems;
  }
  __device__ void flush(unsigned long long* counts) {
    end_walk();
    atomicAdd(counts + 0, 1ull);
    atomicAdd(counts + 1, (unsigned long long)steps);
    atomicAdd(counts + 2, (unsigned long long)ptr);
    atomicAdd(counts + 3, (unsigned long long)rle);
    atomicAdd(counts + 4, (unsigned long long)elems);
    atomicAdd(counts + 5, (unsigned long long)spans);
    atomicAdd(counts + 6, (unsigned long long)col);
    atomicAdd(counts + 7, (unsigned long long)cleared);
    atomicAdd(counts + 8, (unsigned long long)wsteps);
    atomicAdd(counts + 9, (unsigned long long)wptr);
    atomicAdd(counts + 10, (unsigned long long)wrle);
    atomicAdd(counts + 11, (unsigned long long)welems);
  }
};

// Raycasts ray line x of one view into its ray-buffer row and records the
// cleared range in clip[0..1] (hi < lo when nothing was cleared).
// Restates Render::render_line (Cuda_Render.h:96-736) with the active
// configuration of core.h:24-32 (floating horizon, extended floating
// horizon, LDS bitmap).  With COUNT the work counts are added to `counts`
// (rlevox_counters layout) at the end of the line.
template <bool COUNT>
__device__ inline void render_line(int x, const SceneDev& sc, const rlevox_raymap& rm,
                                   uint32_t* __restrict__ row, int32_t* __restrict__ clip,
                                   const LineBits bits, unsigned long long* counts = nullptr) {
  LineCounts cnt;
  const int rs = rm.rs;
  const float half = (float)(rs / 2);
  clip[0] = 0;
  clip[1] = -1;

  // ---- line plane (Cuda_Render.h:128-172)
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
  // ---- reverse texturing (Cuda_Render.h:203-210)
  bool reverse = false;
  if (vertical && ray_z <= 0) reverse = true;
  if (!vertical && ray_x <= 0 && sin_x > 0) reverse = true;
  if (!vertical && ray_x > 0 && sin_x < 0) reverse = true;
  float hmul = reverse ? -half : half;
  if (vertical) hmul = -hmul;

  // ---- screen-space clip range (Cuda_Render.h:215-250)
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

  // ---- clear (Cuda_Render.h:255-264)
  for (int n = ymin; n <= ymax; n++) row[n] = RLEVOX_CLEAR_TEXEL;
  bits.clear(bitmap_words(rs));
  clip[0] = ymin;
  clip[1] = ymax;
  if (COUNT) cnt.cleared = (uint64_t)(ymax - ymin + 1);

  // ---- DDA setup (Cuda_Render.h:270-304)
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
  float g0y = ddy / fabsf(ddx), g0x = sign_x;
  float g1x = ddx / fabsf(ddy), g1y = sign_y;
  float i0x = g0x * frac_x, i0y = g0y * frac_x;
  float i1x = g1x * frac_y, i1y = g1y * frac_y;
  float gd0 = sqrtf(g0x * g0x + g0y * g0y);
  float gd1 = sqrtf(g1x * g1x + g1y * g1y);
  float dd0 = sqrtf(i0x * i0x + i0y * i0y);
  float dd1 = sqrtf(i1x * i1x + i1y * i1y);
  float pbx = 0, pby = 0, dbefore = 0;
  float px = 0, py = 0, dnow = 0;
  int idx = 0, idx_before = 0;

  int mip = 0;
  int gridx = sc.mip[0].sx, gridz = sc.mip[0].sz;
  const uint2* map = sc.mip[0].map;
  const uint16_t* slab = sc.mip[0].slabs;
  int shift = sc.mip[0].shift;
  const float z_add = sin_x;
  const float y_add = (vertical ? cos_x : 0.0f) * hmul;
  const float zfar = (float)rm.far_distance;
  int mapswitch = rm.mapswitch;
  float dz = 1, z = 0;
  float yswitch = vpy;
  const float mountain = vpy;
  WalkEnd wend;
  wend.init(cos_x, sin_x, hmul, ray_x, ray_z, vertical, mountain, half);
  int since_end_test = 0;

  for (;;) {
    // ---- distance mip switch (Cuda_Render.h:343-367)
    while (z > (float)mapswitch || yswitch > 512.0f) {
      yswitch = yswitch * 0.5f;
      if (mip < sc.nmips - 1) {
        mip++;
        gridx >>= 1;
        gridz >>= 1;
        map = sc.mip[mip].map;
        slab = sc.mip[mip].slabs;
        shift = sc.mip[mip].shift;
      }
      g0x *= 2;
      g0y *= 2;
      g1x *= 2;
      g1y *= 2;
      gd0 *= 2;
      gd1 *= 2;
      mapswitch *= 2;
      dz *= 2;
    }
    // ---- early end of the walk (WalkEnd), tested every 64 steps
    if (++since_end_test == 64) {
      since_end_test = 0;
      if (ymin < ymax && !cnt.ended && wend.done(dnow, ymin, sc.mip[mip].top, sc.sy)) {
        if (!COUNT) break;
        cnt.end_walk();  // the counting variant walks on to count the reference's work
      }
    }
    z += dz;
    if (z > zfar) break;
    if (ymin >= ymax) break;
    if (COUNT) cnt.steps++;

    // ---- DDA step (Cuda_Render.h:376-414)
    dbefore = dnow;
    pbx = px;
    pby = py;
    idx_before = idx;
    if (dd1 < dd0) {
      dnow = dd1;
      idx = 1;
      dd1 += gd1;
      px = i1x;
      py = i1y;
      i1x += g1x;
      i1y += g1y;
    } else {
      dnow = dd0;
      idx = 0;
      px = i0x;
      py = i0y;
      dd0 += gd0;
      i0x += g0x;
      i0y += g0y;
    }
    const int fix_x = (int)((float)(1 - idx_before) * fix_xf);
    const int fix_z = (int)((float)idx_before * fix_yf);
    const float ddelta = dnow - dbefore;
    const float vs_x = ray_x * dbefore;
    const float vs_z = ray_z * dbefore;
    const int voxel_x = d_f2i(vpx + pbx) + fix_x;
    const int voxel_z = d_f2i(vpz + pby) + fix_z;
    const int vx = (voxel_x >> mip) & (gridx - 1);
    const int vz = (voxel_z >> mip) & (gridz - 1);
    const float corr_x = ray_x * ddelta;
    const float corr_z = ray_z * ddelta;
    const float pz = cos_x * vs_z + sin_x * mountain;
    float py3 = vertical ? cos_x * mountain - sin_x * vs_z : vs_x;
    py3 *= hmul;
    // ---- clip top (Cuda_Render.h:467)
    if (pz * half + py3 <= pz * (float)ymin) continue;

    // ---- pointer-map gather (Cuda_Render.h:474-478)
    const uint2 ent = map[vx + vz * gridx];
    if (COUNT) cnt.ptr++;
    const uint32_t slen = ent.y & 0xffffu;
    const float czz = cos_x * corr_z;
    float cyy = vertical ? -sin_x * corr_z : corr_x;
    cyy *= hmul;
    const uint16_t* runs = slab + 2 + ((uint64_t)ent.x << shift);
    const uint16_t* colours = runs + slen;
    uint32_t w = ent.y >> 16;
    int ygen = 0;
    float tex = 0;

    // ---- RLE elements (Cuda_Render.h:511-734)
    for (uint32_t k = 0; k < slen; k++) {
      if (k > 0) w = runs[k];
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

      float zz1 = pz, yy1 = py3;
      if (mountain + ytop >= 0) {
        zz1 += czz;
        yy1 += cyy;
      }
      const float pz1 = zz1 + z_add * ytop;
      if (pz1 <= 0) continue;
      const float py1 = yy1 + y_add * ytop;
      int s2 = d_f2i(half + py1 / pz1);
      if (s2 <= ymin) break;

      float zz2 = pz, yy2 = py3;
      if (mountain + ybot < 0) {
        zz2 += czz;
        yy2 += cyy;
      }
      const float pz2 = zz2 + z_add * ybot;
      if (pz2 <= 0) continue;
      const float py2 = yy2 + y_add * ybot;
      int s1 = d_f2i(half + py2 / pz2 - 1.0f);
      if (s1 >= ymax) continue;

      // ---- floating horizon (Cuda_Render.h:563-580)
      if (s2 >= ymax) {
        s2 = ymax;
        ymax = s1;
      }
      if (s1 <= ymin) {
        s1 = ymin;
        ymin = s2;
        while (ymax > ymin && bits.test(ymin)) ++ymin;
      }
      int y = s1;
      while (y < s2 && bits.test(y)) ++y;
      if (y >= s2) continue;

      // ---- perspective-correct span (Cuda_Render.h:642-733)
      const float pz1r = pz + z_add * ytop;
      const float py1r = py3 + y_add * ytop;
      const float pz2r = pz + z_add * ybot;
      const float py2r = py3 + y_add * ybot;
      const float s2r = half + py1r / pz1r;
      const float s1r = half + py2r / pz2r;
      const float u1z = tex / pz2r;
      float u2dz = (float)texture / pz1r - u1z;
      const float onez1 = 1.0f / pz2r;
      float onedz2 = 1.0f / pz1r - onez1;
      u2dz /= s2r - s1r;
      onedz2 /= s2r - s1r;
      const float mult = (float)(y + 1) - s1r;
      float uz = u1z + u2dz * mult;
      float onez = onez1 + onedz2 * mult;
      if (COUNT) cnt.spans++;
      const float tlo = (float)texture, thi = tex - 1.0f;
      for (; y < s2; ++y, uz += u2dz, onez += onedz2) {
        if (bits.test(y)) continue;
        const float fu = fminf(fmaxf(uz / onez, tlo), thi);
        const uint32_t u = (uint32_t)d_f2i(fu);
        const uint32_t depth = (uint32_t)d_f2i(1.0f / onez) & 0xfffeu;
        row[y] = (uint32_t)colours[u] + (depth << 16);
        bits.set(y);
        if (COUNT) cnt.col++;
      }
    }
  }
  if (COUNT) cnt.flush(counts);
}

// The compose pass (colorize_buddha_soft.frag:12-92) per screen pixel;
// weights of exactly 0/1 are applied as selections.  Row 0 = image top.
// The shader's quotients that depend on the column alone, on the row alone
// or on neither are computed apart (the same operations on the same
// operands), so that a tile can share them between its pixels.
struct ComposeCol {
  float scx;   // (px + 0.5) / RESX
  float cterm; // |scx - vanish.x| * RESX / RESY
};
struct ComposeRow {
  float scy;   // (H - 1 - prow + 0.5) / RESY
  float ox;    // (H - 1 - prow + 0.5) / RESX
};

__device__ inline ComposeCol compose_col(const rlevox_raymap& rm, int px) {
  const float RESX = (float)rm.W, RESY = (float)rm.H;
  const float scx = ((float)px + 0.5f) / RESX;
  return ComposeCol{scx, fabsf(scx - rm.vanish[0]) * RESX / RESY};
}

__device__ inline ComposeRow compose_row(const rlevox_raymap& rm, int prow) {
  const float RESX = (float)rm.W, RESY = (float)rm.H;
  const float fcy = (float)(rm.H - 1 - prow) + 0.5f;
  return ComposeRow{fcy / RESY, fcy / RESX};
}

__device__ inline float compose_border(const rlevox_raymap& rm) {
  const float RESX = (float)rm.W, RESY = (float)rm.H;
  return (RESX - RESY) / (RESX * 2.0f);
}

// The shared quotients of one 64x4 tile (columns x0..x0+63, rows
// y0..y0+3): the columns by the first wavefront, the rows and the border by
// the next two.  Call from all 256 threads; a barrier must follow before
// the pixels read it.
struct ComposeTile {
  ComposeCol col[64];
  ComposeRow row[4];
  float border;
};

__device__ inline void compose_tile_stage(const rlevox_raymap& rm, int x0, int y0,
                                          ComposeTile& t) {
  const int i = threadIdx.x;
  if (i < 64)
    t.col[i] = compose_col(rm, x0 + i);
  else if (i < 68)
    t.row[i - 64] = compose_row(rm, y0 + i - 64);
  else if (i == 128)
    t.border = compose_border(rm);
}

__device__ inline void compose_pixel(const rlevox_raymap& rm, const uint32_t* __restrict__ raybuf,
                                     const int32_t* __restrict__ line_clip, int px, int prow,
                                     ComposeCol col, ComposeRow rw, float border,
                                     uint16_t& out_c, int32_t& out_d) {
  const float RESX = (float)rm.W, RESY = (float)rm.H;
  const int rs = rm.rs;
  const float vx = rm.vanish[0], vy = rm.vanish[1];
  const float scx = col.scx, scy = rw.scy;
  const float scx1 = scx - vx, scy1 = scy - vy;
  const bool upper = !(0.0f < scy1);
  const bool left = !(0.0f < scx1);
  const float e = fabsf(scy1) - col.cterm;
  const bool ostep = !(0.0f < e);
  const float o2 = ostep ? scx : rw.ox;  // (ostep ? fcx : fcy) / RESX
  const bool rxg = rm.rot_x_greater_zero != 0.0f;
  float ty, tx;
  if (!ostep) {
    const float s = upper ? 1.0f : 0.0f;
    float ang2 = scx1 * fabsf(1.0f - s - vy) / scy1;
    ang2 = upper ? ang2 + (1.0f - vx) : ang2 + vx;
    ty = upper ? (rm.ofs_add[1] + ang2) : (rm.ofs_add[0] + 1.0f - ang2);
    const float ob = o2 + border;
    const bool first = upper ? !rxg : rxg;
    tx = first ? ob : 1.0f - ob;
  } else {
    const float t = left ? 1.0f : 0.0f;
    float ang3 = scy1 * fabsf(1.0f - t - vx) / scx1;
    ang3 = left ? ang3 + (1.0f - vy) : ang3 + vy;
    ang3 = ang3 * RESY / RESX + border;
    ty = left ? (rm.ofs_add[3] + ang3) : (rm.ofs_add[2] + 1.0f - ang3);
    const bool first = left ? !rxg : rxg;
    tx = first ? o2 : 1.0f - o2;
  }
  ty = ty * rm.ray_ratio * 0.25f;
  uint16_t c = RLEVOX_SKY_COLOUR;
  int32_t d = RLEVOX_SKY_DEPTH;
  if (isfinite(tx) && isfinite(ty)) {
    int tc = d_f2i(floorf(tx * (float)rs));
    int tr = d_f2i(floorf(ty * (float)(4 * rs)));
    tc = min(max(tc, 0), rs - 1);
    tr = min(max(tr, 0), 4 * rs - 1);
    if (tr < rm.map_line_count) {
      // Both loads go out together: the texel is inside the ray buffer
      // (map_line_count x rs) for any tc, and is used only inside the
      // line's cleared range.
      // Read as global memory: the batch kernels take these pointers from
      // an LDS copy of the view, where the compiler would emit flat loads
      // and wait on both counters between them.
      typedef __attribute__((address_space(1))) const uint64_t g_u64;
      typedef __attribute__((address_space(1))) const uint32_t g_u32;
      const uint32_t t = *((g_u32*)raybuf + (size_t)tr * (size_t)rs + (size_t)tc);
      const uint64_t lcw = *(g_u64*)(line_clip + 2 * tr);  // {lo, hi}, 8-byte aligned
      const int2 lc = make_int2((int)(uint32_t)lcw, (int)(uint32_t)(lcw >> 32));
      // non-short-circuit test: no branch for the compiler to sink a load into
      const bool use = (tc >= lc.x) & (tc <= lc.y) & (((t >> 16) & 0xffu) != 0xffu);
      c = use ? (uint16_t)(t & 0xffffu) : c;
      d = use ? (int32_t)(t >> 16) : d;
    }
  }
  out_c = c;
  out_d = d;
}

// Gap-fill / smoothing of one pixel (soft.frag:8-74 on integer buffers):
// ring of 7 taps for the nearest depth, radius 78720/d in 1/256 px, 5x5 taps
// within 1/0.7 of that depth averaged into the shade byte when more than 5
// qualify.
__constant__ const int kRingDx[7] = {0, 9, 9, 2, -8, -10, -3};
__constant__ const int kRingDy[7] = {10, 6, -4, -10, -7, 3, 10};

// Global-memory views of the gap-fill's buffers: the batch kernels read the
// pointers from a view descriptor, where the compiler would emit flat loads.
typedef __attribute__((address_space(1))) const int32_t g_i32;
typedef __attribute__((address_space(1))) const uint16_t g_u16;
typedef int v4i32 __attribute__((ext_vector_type(4)));
typedef uint32_t v2u32 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const v4i32 g_v4i32;
typedef __attribute__((address_space(1))) const v2u32 g_v2u32;
typedef __attribute__((address_space(1))) uint16_t g_u16_out;

__device__ inline int floor_div_pos(int a, int b) {  // b > 0
  const int q = a / b, r = a % b;
  return (r != 0 && r < 0) ? q - 1 : q;
}

__device__ inline uint16_t gapfill_pixel(int W, int H, g_u16* __restrict__ cin,
                                         g_i32* __restrict__ dep, int x, int prow) {
  const uint32_t i = (uint32_t)prow * (uint32_t)W + (uint32_t)x;
  const uint16_t c0 = cin[i];
  int dmin = dep[i];
  for (int k = 0; k < 7; k++) {
    const int xx = min(max(x + kRingDx[k], 0), W - 1);
    const int yy = min(max(prow - kRingDy[k], 0), H - 1);
    dmin = min(dmin, dep[(uint32_t)yy * (uint32_t)W + (uint32_t)xx]);
  }
  if (dmin >= RLEVOX_SKY_DEPTH) return c0;
  const int r_q8 = 78720 / max(dmin, 1);
  if (r_q8 < 42) return c0;
  int n = 1, sum = c0 & 0xff;
  for (int a = 0; a < 5; a++) {
    const int ox = floor_div_pos((2 * a - 5) * r_q8 + 640, 1280);
    const int xx = min(max(x + ox, 0), W - 1);
    for (int b = 0; b < 5; b++) {
      const int oy = floor_div_pos((2 * b - 5) * r_q8 + 640, 1280);
      const int yy = min(max(prow - oy, 0), H - 1);
      const uint32_t t = (uint32_t)yy * (uint32_t)W + (uint32_t)xx;
      const int d = dep[t];
      if (d < RLEVOX_SKY_DEPTH && (int64_t)10 * dmin >= (int64_t)7 * d) {
        n++;
        sum += cin[t] & 0xff;
      }
    }
  }
  if (n > 6) return (uint16_t)((c0 & 0xff00) | ((sum + n / 2) / n));
  return c0;
}

// The same gap-fill for one 64x4 tile of a 256-thread block, with the tile
// and a kGapHalo-pixel border staged in LDS (coordinates clamped to the
// image as each tap clamps them).  The ring (offsets up to 10 px) and every
// 5x5 tap set of radius up to kGapHalo px read LDS; pixels nearer than that
// (larger radii) read their taps from memory.  Call from all 256 threads.
constexpr int kGapHalo = 12;
constexpr int kGapTW = 64 + 2 * kGapHalo, kGapTH = 4 + 2 * kGapHalo;
// One staged pixel: its depth, and its tap word (shade byte | 1 << 16 |
// colour high byte << 24).  Summing the tap words of the qualifying taps
// counts them in bits 16-23 and adds their shades in bits 0-15 (at most 26
// pixels of 255), whatever the high bytes add above.
struct GapTile {
  uint2 p[kGapTH][kGapTW];
};

__device__ inline uint32_t gap_word(uint32_t c) {
  return (c & 0xffu) | (1u << 16) | ((c >> 8) << 24);
}

// The hardware reciprocal (v_rcp_f32, at most 1 ulp off) for recip_div.h.
struct HwRcp {
  __device__ float operator()(float x) const { return __builtin_amdgcn_rcpf(x); }
};

__device__ inline void gapfill_tile(int W, int H, const uint16_t* __restrict__ cin_,
                                    const int32_t* __restrict__ dep_, uint16_t* __restrict__ cout_,
                                    int x0, int y0, GapTile& t) {
  // W, H <= 8192: pixel indices and byte offsets fit in 32 bits
  g_u16* cin = (g_u16*)cin_;
  g_i32* dep = (g_i32*)dep_;
  // uniform: a tile whose border lies inside the image needs no clamping
  const bool inner = x0 >= kGapHalo && y0 >= kGapHalo && x0 - kGapHalo + kGapTW <= W &&
                     y0 - kGapHalo + kGapTH <= H;
  if (inner && (W & 3) == 0) {
    // four pixels per load (x0 - kGapHalo and W are multiples of 4); every
    // load of the thread is issued before the first store
    constexpr int kQ = kGapTW / 4, kQuads = kQ * kGapTH, kStage = (kQuads + 255) / 256;
    static_assert(kGapTW % 4 == 0 && kGapHalo % 4 == 0, "quads must align");
    const uint32_t g0 = (uint32_t)(y0 - kGapHalo) * (uint32_t)W + (uint32_t)(x0 - kGapHalo);
    v4i32 dv[kStage];
    v2u32 cv[kStage];
#pragma unroll
    for (int k = 0; k < kStage; k++) {
      const int i = threadIdx.x + k * 256;
      if (i < kQuads) {
        const int ty = i / kQ, tq = i - ty * kQ;
        const uint32_t o = g0 + (uint32_t)ty * (uint32_t)W + 4u * (uint32_t)tq;
        dv[k] = *(g_v4i32*)(dep + o);
        cv[k] = *(g_v2u32*)(cin + o);
      }
    }
#pragma unroll
    for (int k = 0; k < kStage; k++) {
      const int i = threadIdx.x + k * 256;
      if (i < kQuads) {
        const int ty = i / kQ, tq = i - ty * kQ;
        uint4* dst = reinterpret_cast<uint4*>(&t.p[ty][4 * tq]);
        dst[0] = make_uint4((uint32_t)dv[k].x, gap_word(cv[k].x & 0xffffu), (uint32_t)dv[k].y,
                            gap_word(cv[k].x >> 16));
        dst[1] = make_uint4((uint32_t)dv[k].z, gap_word(cv[k].y & 0xffffu), (uint32_t)dv[k].w,
                            gap_word(cv[k].y >> 16));
      }
    }
  } else {
    constexpr int kStage = (kGapTW * kGapTH + 255) / 256;
    int32_t dv[kStage];
    uint16_t cv[kStage];
#pragma unroll
    for (int k = 0; k < kStage; k++) {
      const int i = threadIdx.x + k * 256;
      if (i < kGapTW * kGapTH) {
        const int ty = i / kGapTW, tx = i - ty * kGapTW;
        const int gx = min(max(x0 - kGapHalo + tx, 0), W - 1);
        const int gy = min(max(y0 - kGapHalo + ty, 0), H - 1);
        const uint32_t g = (uint32_t)gy * (uint32_t)W + (uint32_t)gx;
        dv[k] = dep[g];
        cv[k] = cin[g];
      }
    }
#pragma unroll
    for (int k = 0; k < kStage; k++) {
      const int i = threadIdx.x + k * 256;
      if (i < kGapTW * kGapTH) {
        const int ty = i / kGapTW, tx = i - ty * kGapTW;
        t.p[ty][tx] = make_uint2((uint32_t)dv[k], gap_word(cv[k]));
      }
    }
  }
  __syncthreads();
  const int lx = (threadIdx.x & 63) + kGapHalo, ly = (threadIdx.x >> 6) + kGapHalo;
  const int x = x0 + (threadIdx.x & 63), prow = y0 + (threadIdx.x >> 6);
  if (x < W && prow < H) {
    const uint2 ctr = t.p[ly][lx];
    int dmin = (int)ctr.x;
    for (int k = 0; k < 7; k++) dmin = min(dmin, (int)t.p[ly - kRingDy[k]][lx + kRingDx[k]].x);
    uint16_t out = (uint16_t)((ctr.y & 0xffu) | ((ctr.y >> 24) << 8));  // the pixel's colour
    if (dmin < RLEVOX_SKY_DEPTH) {
      const int dm = max(dmin, 1);
      const int r_q8 = div_78720(dm, HwRcp());  // floor(78720 / dm)
      if (r_q8 >= 42) {
        // tap offsets floor(((2a-5) r + 640) / 1280); the numerators stay
        // above -655360, so a bias of 512*1280 keeps the division unsigned
        int o[5];
#pragma unroll
        for (int a = 0; a < 5; a++)
          o[a] = (int)((uint32_t)((2 * a - 5) * r_q8 + 640 + 655360) / 1280u) - 512;
        if (o[0] >= -kGapHalo) {
          // r_q8 >= 42 means dmin <= 1874: a tap qualifies when
          // 10*dmin >= 7*d, i.e. d <= floor(10*dmin/7) (below the sky depth).
          // Branch-free: every tap is read (so the 25 LDS reads can be in
          // flight together) and a failing tap adds 0.  The centre pixel
          // counts once with its own shade.
          const int dlim = (10 * dmin) / 7;
          uint32_t acc = ctr.y;
          int rowo[5];
#pragma unroll
          for (int b = 0; b < 5; b++) rowo[b] = (ly - o[b]) * kGapTW + lx;
          const uint2* tp = &t.p[0][0];
#pragma unroll
          for (int a = 0; a < 5; a++) {
#pragma unroll
            for (int b = 0; b < 5; b++) {
              const uint2 q = tp[rowo[b] + o[a]];
              acc += (int)q.x <= dlim ? q.y : 0u;
            }
          }
          const int n = (int)((acc >> 16) & 0xffu), sum = (int)(acc & 0xffffu);
          // n <= 26, sum + n/2 <= 26 * 255 + 13
          if (n > 6) out = (uint16_t)((out & 0xff00) | div_small(sum + n / 2, n, HwRcp()));
        } else {
          out = gapfill_pixel(W, H, cin, dep, x, prow);
        }
      }
    }
    *((g_u16_out*)cout_ + ((uint32_t)prow * (uint32_t)W + (uint32_t)x)) = out;
  }
  __syncthreads();  // the block's next tile reuses the LDS
}

}  // namespace rlevox
