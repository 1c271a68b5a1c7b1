// ORACLE - test infrastructure only.
//
// A CPU restatement of the reference's raycast path, used solely as the
// checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
// Nothing in the product library links, loads or calls this code.
//
// Reference files (sp4cerat/RLE-based-Voxel-Raycasting, RLE-Raycaster/...):
//   src/RayMap.h:64-402          get_ray_map                 -> orc_raymap
//   inc/mathlib/_matrix44.h      rotate_x/rotate_y/translate/invert_simpler/set/operator*
//   src/VecMath.h:60-81,150-166  vec3f dot/cross/angle/length
//   src/Cuda_Render.h:39-94      LineScale/ClipLine/MatMul/vec3f_rot_y
//   src/Cuda_Render.h:96-736     render_line                 -> orc_render_line
//   src/Cuda_Main.cu:150-181     line index mapping          -> orc_raycast
//   inc/cutil_math.h:58-61,1184-1188  normalize = v * (1/sqrtf(dot(v,v)))
//   bin/shader/colorize_buddha_soft.frag:12-92                -> orc_compose
//   bin/shader/soft.frag:8-74                                 -> orc_gapfill
//   src/Rle4.cpp:52-198,295-314  RLE4::compress / load       -> orc_encode_dense
//   src/tree.h:23-85             Tree::get_mipmap            -> orc_mip_dense
//
// The reference itself cannot be compiled here (it needs the missing
// src.BestFitMem/bmalloc.h, CUDA's vector types and Windows headers), so this
// restatement is pinned by independent checks: an independent per-pixel 3D
// raycaster (brute.cpp) compared image-wide, closed-form scenes, codec round
// trips and planted mistakes (ORC_PLANT) that those checks must catch.
//
// Arithmetic contract shared with the HIP kernels (the parity hazards):
//   * fp32 throughout in the source's operation order, no FMA contraction
//     (built with -ffp-contract=off), IEEE division and sqrt;
//   * float->int conversion saturates and maps NaN to 0 (f2i), which is what
//     CUDA's cvt.rzi and gfx950's v_cvt_i32_f32 do;
//   * min/max on floats are fminf/fmaxf (NaN-ignoring), as CUDA's min/max;
//   * every transcendental of render_line depends only on the pose and is
//     computed once per frame, in double, here in orc_raymap;
//   * normalize() is 1/sqrtf, not a hardware rsqrt;
//   * the per-line bitmap holds ceil(rs/32) private words (the reference's
//     31-word slice aliases the next thread above row 991);
//   * the dead store ofs_cache_start[0]=0 (Cuda_Render.h:112) is dropped.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/rlevox.h"

#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

namespace {

struct V3 {
  float x, y, z;
};
inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
// vec3f::dot (VecMath.h:60-61): a.x*x + a.y*y + a.z*z
inline float dot(V3 self, V3 a) { return a.x * self.x + a.y * self.y + a.z * self.z; }
inline V3 ld(const float* p) { return mk(p[0], p[1], p[2]); }
inline void st(float* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// Saturating float->int with NaN->0 (parity hazard H4).
inline int f2i(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f < -2147483648.0f) return INT32_MIN;
  return (int)f;
}

struct M44 {
  float m[4][4];
};
inline void m_ident(M44& a) {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) a.m[i][j] = (i == j) ? 1.0f : 0.0f;
}
// n_sin/n_cos (nmath.h:33-34) evaluated in double, rounded to float.
inline float fsin(float a) { return (float)std::sin((double)a); }
inline float fcos(float a) { return (float)std::cos((double)a); }
// _matrix44::rotate_x (_matrix44.h:529-541)
inline void m_rotate_x(M44& a, float ang) {
  float c = fcos(ang), s = fsin(ang);
  for (int i = 0; i < 4; i++) {
    float mi1 = a.m[i][1], mi2 = a.m[i][2];
    a.m[i][1] = mi1 * c + mi2 * -s;
    a.m[i][2] = mi1 * s + mi2 * c;
  }
}
// _matrix44::rotate_y (_matrix44.h:547-559)
inline void m_rotate_y(M44& a, float ang) {
  float c = fcos(ang), s = fsin(ang);
  for (int i = 0; i < 4; i++) {
    float mi0 = a.m[i][0], mi2 = a.m[i][2];
    a.m[i][0] = mi0 * c + mi2 * s;
    a.m[i][2] = mi0 * -s + mi2 * c;
  }
}
// operator*(matrix44, vector3) (_matrix44.h:863-870)
inline V3 m_mul(const M44& a, V3 v) {
  return mk(a.m[0][0] * v.x + a.m[1][0] * v.y + a.m[2][0] * v.z + a.m[3][0],
            a.m[0][1] * v.x + a.m[1][1] * v.y + a.m[2][1] * v.z + a.m[3][1],
            a.m[0][2] * v.x + a.m[1][2] * v.y + a.m[2][2] * v.z + a.m[3][2]);
}
// _matrix44::invert_simpler (_matrix44.h:431-443)
inline void m_invert_simpler(M44& a) {
  std::swap(a.m[0][1], a.m[1][0]);
  std::swap(a.m[0][2], a.m[2][0]);
  std::swap(a.m[1][2], a.m[2][1]);
  float m30 = -(a.m[0][0] * a.m[3][0] + a.m[1][0] * a.m[3][1] + a.m[2][0] * a.m[3][2]);
  float m31 = -(a.m[0][1] * a.m[3][0] + a.m[1][1] * a.m[3][1] + a.m[2][1] * a.m[3][2]);
  a.m[3][2] = -(a.m[0][2] * a.m[3][0] + a.m[1][2] * a.m[3][1] + a.m[2][2] * a.m[3][2]);
  a.m[3][1] = m31;
  a.m[3][0] = m30;
}
// vec3f::length (VecMath.h:150-153)
inline float vlength(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
// vec3f::angle (VecMath.h:71-81); `self` is the object, v the argument.
inline float vangle(V3 self, V3 v) {
  float d = v.x * self.x + v.y * self.y + v.z * self.z;
  float len = vlength(v) * vlength(self);
  if (len == 0) len = 0.00001f;
  float input = d / len;
  if (input < -1) input = -1;
  if (input > 1) input = 1;
  return (float)std::acos((double)input);
}

// RayMap::LineScale (RayMap.h:70-83) == Render::LineScale (Cuda_Render.h:39-52)
inline float line_scale(V3 input, V3 center, float clip_max, float clip_min) {
  float scale_x = 1, scale_y = 1;
  if (center.x > 1) scale_x = (1 - input.x) / (center.x - input.x);
  if (center.x < 0) scale_x = input.x / (input.x - center.x);
  if (center.y > clip_max) scale_y = (clip_max - input.y) / (center.y - input.y);
  if (center.y < clip_min) scale_y = (-clip_min + input.y) / (input.y - center.y);
  return (scale_x < scale_y) ? scale_x : scale_y;
}
// ClipLine (Cuda_Render.h:54-65)
inline void clip_line(V3& p1, V3& p2, float clip_max, float clip_min) {
  float scale = line_scale(p1, p2, clip_max, clip_min);
  V3 c2 = p1 + (p2 - p1) * scale;
  scale = line_scale(p2, p1, clip_max, clip_min);
  V3 c1 = p2 + (p1 - p2) * scale;
  p1 = c1;
  p2 = c2;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- ray map
// RayMap::get_ray_map (RayMap.h:98-402) with set_border((W-H)/(2W))
// (the formula left at main.cpp:776; 0.125 at 1024x768 as main.cpp:775),
// RAYS_CASTED_RES/4 generalised to rs = W, plus the per-frame constants of
// render_line (Cuda_Render.h:77-78,188-191,335) and the compose uniforms
// (main.cpp:575-603).
int orc_raymap(const rlevox_pose* pose, int W, int H, int mip_distance, int far_distance,
               rlevox_raymap* out) {
  if (!pose || !out || W < 16 || H < 16 || H > W || W > RLEVOX_MAX_RS) return 1;
  std::memset(out, 0, sizeof(*out));
  rlevox_raymap& r = *out;
  r.W = W;
  r.H = H;
  r.rs = W;
  r.maxres = W;
  r.far_distance = far_distance > 0 ? far_distance : 80000;
  const int mipd = mip_distance > 0 ? mip_distance : 1024;
  const float border = ((float)W - (float)H) / ((float)W * 2.0f);
  r.border = border;
  r.clip_min = border;
  r.clip_max = 1 - border;
  V3 rot = ld(pose->rot), pos = ld(pose->pos);
  st(r.rotation, rot);
  st(r.position, pos);

  V3 p[6] = {mk(1, 1, 1), mk(-1, 1, 1), mk(-1, -1, 1), mk(1, -1, 1), mk(0, 0, 0), mk(0, 0, 0)};
  M44 m;
  m_ident(m);
  m_rotate_x(m, rot.x);
  m_rotate_y(m, rot.y);
  m.m[3][0] += 3;  // translate(vector3(3,2,0)) (_matrix44.h:583-588)
  m.m[3][1] += 2;
  m.m[3][2] += 0;
  for (int i = 0; i < 5; i++) p[i] = m_mul(m, p[i]);

  V3 down = mk(0, -1, 0);
  V3 view = (p[0] + p[2]) / 2.0f - p[4];
  float alpha = float(M_PI) / 2.0f - vangle(down, view);
  float scale = 1.0f / fsin(alpha);
  p[5] = p[4] + down * scale;

  V3 a = p[1] - p[0], b = p[3] - p[0];
  V3 nrm = mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
  V3 d1 = p[1] - p[0], d2 = p[3] - p[0], d3 = nrm, d4 = p[0];
  // _vector3::dot (_vector3.h:387-390) then operator*= (_vector3.h:190-195)
  auto dot3 = [](V3 u) { return u.x * u.x + u.y * u.y + u.z * u.z; };
  d1 = d1 * (1.0f / dot3(d1));
  d2 = d2 * (1.0f / dot3(d2));
  d3 = d3 * (1.0f / dot3(d3));
  M44 to2d;
  // set(vector4,...) from vector3 rows (w = 1) (_matrix44.h:255-261, _vector4.h:120-125)
  V3 rows[4] = {d1, d2, d3, d4};
  for (int i = 0; i < 4; i++) {
    to2d.m[i][0] = rows[i].x;
    to2d.m[i][1] = rows[i].y;
    to2d.m[i][2] = rows[i].z;
    to2d.m[i][3] = 1.0f;
  }
  std::memcpy(r.to3d, to2d.m, sizeof(r.to3d));
  m_invert_simpler(to2d);
  V3 p2d[8];
  for (int i = 0; i < 8; i++) p2d[i] = mk(0, 0, 0);
  for (int i = 0; i < 6; i++) p2d[i] = m_mul(to2d, p[i]);
  st(r.vp2d, p2d[5]);

  const int maxres = W;
  const int safety = 2;
  const float ys_min = border, ys_max = 1 - border;
  const V3 plist[4] = {mk(-1, -1, 0), mk(1, -1, 0), mk(1, 1, 0), mk(-1, 1, 0)};
  const V3 plist2[4] = {mk(0, ys_min, 0), mk(1, ys_min, 0), mk(1, ys_max, 0), mk(0, ys_max, 0)};
  V3 pno[8];
  for (int i = 0; i < 8; i++) pno[i] = mk(0, 0, 0);
  int res[4] = {0, 0, 0, 0};
  float ofs_min[4] = {0, 0, 0, 0}, ofs_max[4] = {0, 0, 0, 0};
  p2d[5].z = 0;
  const V3 c = p2d[5];

  // Upper part (RayMap.h:209-240)
  if (c.y > border) {
    pno[0] = c + plist[0] * std::fabs(c.y - border);
    pno[1] = c + plist[1] * std::fabs(c.y - border);
    V3 p_no_in = pno[1];
    if (c.x > 1) {
      if (pno[1].x > 1) pno[1].x = 1;
    } else if (dot(pno[0] - c, plist2[2] - c) > 0) {
      pno[1] = c + (plist2[2] - c) * std::fabs((plist2[0].y - c.y) / (plist2[2].y - c.y));
    }
    if (c.x < 0) {
      if (pno[0].x < 0) pno[0].x = 0;
    } else if (dot(p_no_in - c, plist2[3] - c) > 0) {
      pno[0] = c + (plist2[3] - c) * std::fabs((plist2[1].y - c.y) / (plist2[3].y - c.y));
    }
    pno[0].y = pno[1].y = plist2[0].y;
    pno[0].x = float(f2i((float)maxres * pno[0].x) - safety) / (float)maxres;
    pno[1].x = float(f2i((float)maxres * pno[1].x) + safety) / (float)maxres;
    res[0] = f2i((float)maxres * std::fabs(pno[0].x - pno[1].x));
    if (pno[0].x - pno[1].x > 0) res[0] = 0;
    if (res[0] > maxres * 3) res[0] = maxres * 3;
    ofs_min[0] = pno[0].x;
    ofs_max[0] = pno[1].x;
  }
  // Lower part (RayMap.h:243-277)
  if (c.y < 1 - border) {
    pno[2] = c + plist[3] * std::fabs(c.y - 1 + border);
    pno[3] = c + plist[2] * std::fabs(c.y - 1 + border);
    V3 p_no_in = pno[2];
    if (c.x < 0) {
      if (pno[2].x < 0) pno[2].x = 0;
    } else if (dot(pno[3] - c, plist2[0] - c) > 0) {
      pno[2] = c + (plist2[0] - c) * std::fabs((plist2[2].y - c.y) / (plist2[0].y - c.y));
    }
    if (c.x > 1) {
      if (pno[3].x > 1) pno[3].x = 1;
    } else {
      V3 delta = plist2[1] - c;
      if (dot(p_no_in - c, delta) > 0) {
        pno[3] = c + delta * std::fabs((plist2[3].y - c.y) / delta.y);
      }
    }
    pno[2].y = pno[3].y = plist2[2].y;
    pno[2].x = float(f2i((float)maxres * pno[2].x) - safety) / (float)maxres;
    pno[3].x = float(f2i((float)maxres * pno[3].x) + safety) / (float)maxres;
    res[1] = f2i((float)maxres * std::fabs(pno[2].x - pno[3].x));
    if (pno[2].x - pno[3].x > 0) res[1] = 0;
    if (res[1] > maxres * 3) res[1] = maxres * 3;
    ofs_min[1] = pno[2].x;
    ofs_max[1] = pno[3].x;
  }
  // Left part (RayMap.h:280-313)
  if (c.x > 0) {
    pno[4] = c + plist[0] * std::fabs(c.x);
    pno[5] = c + plist[3] * std::fabs(c.x);
    V3 p_no_in = pno[5];
    if (c.y > 1 - border) {
      if (pno[5].y > 1 - border) pno[5].y = 1 - border;
    } else {
      const int id = 2;
      if (dot(pno[4] - c, plist2[id] - c) > 0)
        pno[5] = c + (plist2[id] - c) *
                         std::fabs((plist2[id ^ 1].x - c.x) / (plist2[id].x - c.x));
    }
    if (c.y < border) {
      if (pno[4].y < border) pno[4].y = border;
    } else {
      const int id = 1;
      if (dot(p_no_in - c, plist2[id] - c) > 0)
        pno[4] = c + (plist2[id] - c) *
                         std::fabs((plist2[id ^ 1].x - c.x) / (plist2[id].x - c.x));
    }
    pno[4].x = pno[5].x = 0;
    pno[4].y = float(f2i((float)maxres * pno[4].y) - safety) / (float)maxres;
    pno[5].y = float(f2i((float)maxres * pno[5].y) + safety) / (float)maxres;
    res[2] = f2i((float)maxres * std::fabs(pno[4].y - pno[5].y));
    if (pno[4].y - pno[5].y > 0) res[2] = 0;
    if (res[2] > maxres * 3) res[2] = maxres * 3;
    ofs_min[2] = pno[4].y;
    ofs_max[2] = pno[5].y;
  }
  // Right part (RayMap.h:316-352)
  if (c.x < 1) {
    pno[6] = c + plist[1] * std::fabs(1 - c.x);
    pno[7] = c + plist[2] * std::fabs(1 - c.x);
    V3 p_no_in = pno[7];
    if (c.y > 1 - border) {
      if (pno[7].y > 1 - border) pno[7].y = 1 - border;
    } else {
      const int id = 3;
      if (dot(pno[6] - c, plist2[id] - c) > 0)
        pno[7] = c + (plist2[id] - c) *
                         std::fabs((plist2[id ^ 1].x - c.x) / (plist2[id].x - c.x));
    }
    if (c.y < border) {
      if (pno[6].y < border) pno[6].y = border;
    } else {
      const int id = 0;
      if (dot(p_no_in - c, plist2[id] - c) > 0)
        pno[6] = c + (plist2[id] - c) *
                         std::fabs((plist2[id ^ 1].x - c.x) / (plist2[id].x - c.x));
    }
    pno[6].x = pno[7].x = 1;
    pno[6].y = float(f2i((float)maxres * pno[6].y) - safety) / (float)maxres;
    pno[7].y = float(f2i((float)maxres * pno[7].y) + safety) / (float)maxres;
    res[3] = f2i((float)maxres * std::fabs(pno[6].y - pno[7].y));
    if (pno[6].y - pno[7].y > 0) res[3] = 0;
    if (res[3] > maxres * 3) res[3] = maxres * 3;
    ofs_min[3] = pno[6].y;
    ofs_max[3] = pno[7].y;
  }
  st(r.p4, p[4]);
  for (int i = 0; i < 8; i++) {
    st(r.p_2d[i], p2d[i]);
    st(r.p_no[i], pno[i]);
  }
  for (int j = 0; j < 4; j++) {
    r.res[j] = res[j];
    r.p_ofs_min[j] = ofs_min[j];
    r.p_ofs_max[j] = ofs_max[j];
  }
  r.map_line_count = res[0] + res[1] + res[2] + res[3];

  // Per-frame constants of render_line.
  r.sin_x = fsin(rot.x);
  r.cos_x = fcos(rot.x);
  r.sin_y = fsin(rot.y);
  r.cos_y = fcos(rot.y);
  r.cos_myaw = fcos(-rot.y);
  r.sin_myaw = fsin(-rot.y);
  // mapswitch = MIP_DISTANCE * (0.25*(4-abs(viewrot.x))) (Cuda_Render.h:182,335)
  r.mapswitch = (int)((double)mipd * (0.25 * (double)(4.0f - std::fabs(rot.x))));
  if (r.mapswitch < 1) return 1;  // the mip-switch loop could not terminate
  {  // and its int doublings must not overflow (see the camera-row halvings)
    int n = 0;
    float y = r.position[1];
    while (y > 512.0f && n < 40) {
      y = y * 0.5f;
      n++;
    }
    if (n >= 40 || ((int64_t)r.mapswitch << n) > (1ll << 30)) return 1;
  }

  // Compose uniforms (main.cpp:575-603), RAYS_CASTED = RAYS_CASTED_RES = 4*rs.
  r.vanish[0] = 1 - r.vp2d[0];
  r.vanish[1] = ((1 - r.vp2d[1]) - border) * (float)W / (float)H;
  const float rays_res = (float)(4 * W);
  float ofs1 = 4 * (float)res[0] / rays_res;
  float ofs2 = 4 * (float)res[1] / rays_res + ofs1;
  float ofs3 = 4 * (float)res[2] / rays_res + ofs2;
  r.ofs_add[0] = -ofs_min[0];
  r.ofs_add[1] = -ofs_min[1] + ofs1;
  r.ofs_add[2] = -ofs_min[2] + ofs2;
  r.ofs_add[3] = -ofs_min[3] + ofs3;
  r.rot_x_greater_zero = (rot.x > 0) ? 1.0f : 0.0f;
  r.ray_ratio = rays_res / rays_res;
  return 0;
}

// Hooks a host check can define before including this file to observe the
// walk of each line (tests/native/walk_end_check.cpp); empty otherwise.
#ifndef ORC_HOOK_LINE_BEGIN
#define ORC_HOOK_LINE_BEGIN
#endif
#ifndef ORC_HOOK_STEP
#define ORC_HOOK_STEP
#endif
#ifndef ORC_HOOK_ACT
#define ORC_HOOK_ACT
#endif
#ifndef ORC_HOOK_LINE_END
#define ORC_HOOK_LINE_END
#endif

// ------------------------------------------------------------ render line
struct OrcScene {
  const rlevox_mip* mips;
  int nmips;
};

// The plane of ray line x (Cuda_Render.h:107-172): its horizontal ray
// direction in the camera's yaw frame, whether it belongs to an upper/lower
// segment (`vertical`), and its clipped 2D end points.
struct LinePlane {
  float ray_x, ray_z;
  bool vertical;
  V3 start2d, end2d;
};

static LinePlane line_plane(int x, const rlevox_raymap& rm) {
  int rays[4];
  rays[0] = rm.res[0];
  rays[1] = rm.res[1] + rays[0];
  rays[2] = rm.res[2] + rays[1];
  rays[3] = rm.res[3] + rays[2];
  int quadrant = 0;
  if (x >= rays[2]) quadrant = 3;
  else if (x >= rays[1]) quadrant = 2;
  else if (x >= rays[0]) quadrant = 1;
  float quadrant_ofs = (float)x;
  if (quadrant >= 1) quadrant_ofs -= (float)rays[quadrant - 1];
  float quadrant_num = (float)rm.res[quadrant];
  float a = quadrant_ofs / quadrant_num;
  const int j = quadrant;
  V3 p1 = ld(rm.p_2d[5]);
  V3 p2 = ld(rm.p_no[j * 2]) + (ld(rm.p_no[j * 2 + 1]) - ld(rm.p_no[j * 2])) * a;
  clip_line(p1, p2, rm.clip_max, rm.clip_min);
  M44 to3d;
  std::memcpy(to3d.m, rm.to3d, sizeof(to3d.m));
  V3 p1_3d = m_mul(to3d, p1 * 4.0f);
  V3 p2_3d = m_mul(to3d, p2 * 4.0f);
  V3 delta = (p1_3d + p2_3d) * 0.5f - ld(rm.p4);
  delta.y = 0;
  {  // normalize (cutil_math.h:1184-1188) with rsqrtf = 1/sqrtf (cutil_math.h:58-61)
    float d = delta.x * delta.x + delta.y * delta.y + delta.z * delta.z;
    float inv_len = 1.0f / std::sqrt(d);
    delta = delta * inv_len;
  }
  // vec3f_rot_y(viewrot.y) (Cuda_Render.h:75-80)
  LinePlane lp;
  lp.ray_x = rm.cos_myaw * delta.x + rm.sin_myaw * delta.z;
  lp.ray_z = rm.cos_myaw * delta.z - rm.sin_myaw * delta.x;
  lp.start2d = p1;
  lp.end2d = p2;
  lp.vertical = (1 - (j >> 1)) != 0;
  return lp;
}

static void orc_render_line(int x, const OrcScene& sc, const rlevox_raymap& rm, uint32_t* raybuf,
                            int32_t* line_clip, uint32_t* y_cache, rlevox_counters* ctr) {
  const int res_x = rm.rs, res_y = rm.rs;
  const float res_x2 = (float)(res_x / 2);
  const float res_y2 = (float)(res_y / 2);
  uint32_t* out = raybuf + (size_t)x * (size_t)res_y;
  line_clip[2 * x] = 0;
  line_clip[2 * x + 1] = -1;

  const LinePlane lp = line_plane(x, rm);
  const float ray_x = lp.ray_x, ray_z = lp.ray_z;
  const V3 start2d = lp.start2d, end2d = lp.end2d;
  const bool vertical = lp.vertical;

  // Render vars (Cuda_Render.h:177-191)
  int mip_lvl = 0;
  int y_clip_min = 0;
  int y_clip_max = res_y - 1;
  const int z_far = rm.far_distance;
  float dz = 1;
  int mapswitch = rm.mapswitch;
  const float sin_x = rm.sin_x, cos_x = rm.cos_x, sin_y = rm.sin_y, cos_y = rm.cos_y;

  // Reverse texturing (Cuda_Render.h:203-210)
  bool reverse = false;
  if (vertical && ray_z <= 0) reverse = true;
  if (!vertical && ray_x <= 0 && sin_x > 0) reverse = true;
  if (!vertical && ray_x > 0 && sin_x < 0) reverse = true;
  float res_x2_mul_reverse = reverse ? -res_x2 : res_x2;
  if (vertical) res_x2_mul_reverse = -res_x2_mul_reverse;

  // Screen-space clipping (Cuda_Render.h:215-250)
#if ORC_PLANT == 8
  int p_add = reverse ? -2 : 1;
#else
  int p_add = reverse ? 1 : -2;
#endif
  int p1x = f2i((float)res_x * start2d.x) + p_add;
  int p1y = f2i((float)res_y * start2d.y) + p_add;
  int p2x = f2i((float)res_x * end2d.x) - p_add;
  int p2y = f2i((float)res_y * end2d.y) - p_add;
  if (p1x < 0) p1x = 0;
  if (p1x >= res_x) p1x = res_x - 1;
  if (p1y < 0) p1y = 0;
  if (p1y >= res_y) p1y = res_y - 1;
  if (p2x < 0) p2x = 0;
  if (p2x >= res_x) p2x = res_x - 1;
  if (p2y < 0) p2y = 0;
  if (p2y >= res_y) p2y = res_y - 1;
  if (p1y == p2y) return;
  y_clip_min = res_x - 1 - p1x;
  y_clip_max = res_x - 1 - p2x;
  if (vertical) {
    y_clip_min = res_y - 1 - p1y;
    y_clip_max = res_y - 1 - p2y;
  }
  if (reverse) {
    y_clip_min = res_y - 1 - y_clip_min;
    y_clip_max = res_y - 1 - y_clip_max;
  }
  if (y_clip_min > y_clip_max) std::swap(y_clip_min, y_clip_max);
  if (y_clip_min >= y_clip_max) return;

  // Clear the line and its bitmap (Cuda_Render.h:255-264)
  for (int n = y_clip_min; n <= y_clip_max; n++) out[n] = RLEVOX_CLEAR_TEXEL;
  const int words = (res_y + 31) >> 5;
  for (int n = 0; n < words; n++) y_cache[n] = 0;
  line_clip[2 * x] = y_clip_min;
  line_clip[2 * x + 1] = y_clip_max;
  uint64_t c_steps = 0, c_ptr = 0, c_rle = 0, c_elem = 0, c_span = 0, c_col = 0;
  const uint64_t c_clear = (uint64_t)(y_clip_max - y_clip_min + 1);

  // DDA setup (Cuda_Render.h:270-304)
  const float vpx = rm.position[0], vpy = rm.position[1], vpz = rm.position[2];
  float dx = ray_x * cos_y + ray_z * sin_y;
  float dy = ray_x * sin_y - ray_z * cos_y;
  float fix_x_f = -1, fix_y_f = -1;
  float frac_x = vpx - (float)f2i(vpx);
  float frac_y = vpz - (float)f2i(vpz);
  float sign_x = -1, sign_y = -1;
  if (dx >= 0) {
    fix_x_f = 0;
    sign_x = 1;
    frac_x = 1 - frac_x;
  }
  if (dy >= 0) {
    fix_y_f = 0;
    sign_y = 1;
    frac_y = 1 - frac_y;
  }
  float grad0_y = dy / std::fabs(dx), grad0_x = sign_x;
  float grad1_x = dx / std::fabs(dy), grad1_y = sign_y;
  float isect0_x = grad0_x * frac_x, isect0_y = grad0_y * frac_x;
  float isect1_x = grad1_x * frac_y, isect1_y = grad1_y * frac_y;
  float grad_dist0 = std::sqrt(grad0_x * grad0_x + grad0_y * grad0_y);
  float grad_dist1 = std::sqrt(grad1_x * grad1_x + grad1_y * grad1_y);
  float dds_dist0 = std::sqrt(isect0_x * isect0_x + isect0_y * isect0_y);
  float dds_dist1 = std::sqrt(isect1_x * isect1_x + isect1_y * isect1_y);
  float pos_before_x = 0, pos_before_y = 0, dds_dist_before = 0;
  float pos_x = 0, pos_y = 0, dds_dist_now = 0;
  int index = 0, index_before = 0;

  int gridx = sc.mips[0].sx, gridz = sc.mips[0].sz;
  const float pos3d_z_add = sin_x;
  float pos3d_y_add = vertical ? cos_x : 0;
  pos3d_y_add *= res_x2_mul_reverse;
  const uint32_t* map_ptr = sc.mips[0].ptrmap;
  const uint16_t* slab_ptr = sc.mips[0].slabs;
  int offset_shift = sc.mips[0].offset_shift;  // widened offsets (0 = reference layout)
  float z = 0;
  float y_map_switch = vpy;

  auto bit = [&](int y) -> bool { return (y_cache[y >> 5] >> (y & 31)) & 1u; };

  ORC_HOOK_LINE_BEGIN
  for (;;) {
    // Mip switch (Cuda_Render.h:343-367)
    while (z > (float)mapswitch || y_map_switch > 512.0f) {
      y_map_switch = y_map_switch * 0.5f;
      if (mip_lvl < sc.nmips - 1) {
        mip_lvl++;
        gridx >>= 1;
        gridz >>= 1;
        map_ptr = sc.mips[mip_lvl].ptrmap;
        slab_ptr = sc.mips[mip_lvl].slabs;
        offset_shift = sc.mips[mip_lvl].offset_shift;
      }
      grad0_x *= 2;
      grad0_y *= 2;
      grad1_x *= 2;
      grad1_y *= 2;
      grad_dist0 *= 2;
      grad_dist1 *= 2;
      mapswitch *= 2;
      dz *= 2;
    }
    ORC_HOOK_STEP
    z += dz;
    if (z > (float)z_far) break;
    if (y_clip_min >= y_clip_max) break;
    c_steps++;

    // DDA step (Cuda_Render.h:376-414)
    dds_dist_before = dds_dist_now;
    pos_before_x = pos_x;
    pos_before_y = pos_y;
    index_before = index;
    if (dds_dist1 < dds_dist0) {
      dds_dist_now = dds_dist1;
      index = 1;
      dds_dist1 += grad_dist1;
      pos_x = isect1_x;
      pos_y = isect1_y;
      isect1_x += grad1_x;
      isect1_y += grad1_y;
    } else {
      dds_dist_now = dds_dist0;
      index = 0;
      pos_x = isect0_x;
      pos_y = isect0_y;
      dds_dist0 += grad_dist0;
      isect0_x += grad0_x;
      isect0_y += grad0_y;
    }
    // skipme is never set (Cuda_Render.h:331,416,543)
    int fix_x = (int)((float)(1 - index_before) * fix_x_f);
    int fix_z = (int)((float)index_before * fix_y_f);
#if ORC_PLANT == 3
    std::swap(fix_x, fix_z);
#endif
    float dds_dist_delta = dds_dist_now - dds_dist_before;
    float vs_x = ray_x * dds_dist_before;
    float vs_z = ray_z * dds_dist_before;
    int voxel_x = f2i(vpx + pos_before_x) + fix_x;
    int voxel_z = f2i(vpz + pos_before_y) + fix_z;
    int vx = (voxel_x >> mip_lvl) & (gridx - 1);
    int vz = (voxel_z >> mip_lvl) & (gridz - 1);
    const float mountain = vpy;
    float correct_x = ray_x * dds_dist_delta;
    float correct_z = ray_z * dds_dist_delta;
    float pos3d_z = cos_x * vs_z + sin_x * mountain;
    float pos3d_y = vertical ? cos_x * mountain - sin_x * vs_z : vs_x;
    pos3d_y *= res_x2_mul_reverse;
    // Clip top (Cuda_Render.h:467)
    if (pos3d_z * res_y2 + pos3d_y <= pos3d_z * (float)y_clip_min) continue;

    // Pointer-map gather (Cuda_Render.h:474-478)
    const uint32_t* ent = map_ptr + 2 * ((size_t)vx + (size_t)vz * (size_t)gridx);
    c_ptr++;
    const uint64_t slab_offset = (uint64_t)ent[0] << offset_shift;
    const uint32_t len_first = ent[1];
    const uint16_t slen = (uint16_t)len_first;
    float corr_zz = cos_x * correct_z;
    float corr_yy = vertical ? -sin_x * correct_z : correct_x;
#if ORC_PLANT == 2
    corr_zz = -corr_zz;
#endif
    corr_yy *= res_x2_mul_reverse;
    int sti_general = 0;
    int sti_skip = 0;
    uint16_t sti_ = (uint16_t)(len_first >> 16);
    const uint16_t* slabs = slab_ptr + 2 + slab_offset;
    const uint16_t* send = slabs + slen;
    float tex = 0;

    // RLE element loop (Cuda_Render.h:511-734)
    for (int k = 0; k < (int)slen; k++) {
      if (k > 0) {
        sti_ = slabs[k];
        c_rle++;
      }
      c_elem++;
      sti_skip = sti_ >> 10;
      sti_general += (sti_ & 1023) << mip_lvl;
      if (sti_skip == 0) continue;
      int texture = (int)tex;
      tex += (float)sti_skip;
      sti_skip <<= mip_lvl;
      float sti_general_sti_skip = (float)sti_general;
      sti_general += sti_skip;
      const float sg = (float)sti_general;

      float correct_zz1 = pos3d_z, correct_yy1 = pos3d_y;
      if (mountain + sti_general_sti_skip >= 0) {
        correct_zz1 += corr_zz;
        correct_yy1 += corr_yy;
      }
      float pos3d_z1 = correct_zz1 + pos3d_z_add * sti_general_sti_skip;
      if (pos3d_z1 <= 0) continue;
      float pos3d_y1 = correct_yy1 + pos3d_y_add * sti_general_sti_skip;
      int scr_y2 = f2i(res_y2 + pos3d_y1 / pos3d_z1);
      if (scr_y2 <= y_clip_min) break;

      float correct_zz2 = pos3d_z, correct_yy2 = pos3d_y;
      if (mountain + sg < 0) {
        correct_zz2 += corr_zz;
        correct_yy2 += corr_yy;
      }
      float pos3d_z2 = correct_zz2 + pos3d_z_add * sg;
      if (pos3d_z2 <= 0) continue;
      float pos3d_y2 = correct_yy2 + pos3d_y_add * sg;
#if ORC_PLANT == 6
      int scr_y1 = f2i(res_y2 + pos3d_y2 / pos3d_z2);
#else
      int scr_y1 = f2i(res_y2 + pos3d_y2 / pos3d_z2 - 1.0f);
#endif
      if (scr_y1 >= y_clip_max) continue;

      ORC_HOOK_ACT
      // Floating horizon (Cuda_Render.h:563-580)
      if (scr_y2 >= y_clip_max) {
        scr_y2 = y_clip_max;
        y_clip_max = scr_y1;
      }
      if (scr_y1 <= y_clip_min) {
        scr_y1 = y_clip_min;
        y_clip_min = scr_y2;
        while (y_clip_max > y_clip_min && bit(y_clip_min)) ++y_clip_min;
      }
      int y = scr_y1;
      // Bitmap skip (Cuda_Render.h:639-640)
      while (y < scr_y2 && bit(y)) ++y;
      if (y >= scr_y2) continue;

      // Perspective-correct span (Cuda_Render.h:642-681)
      float pos3d_z1r = pos3d_z + pos3d_z_add * sti_general_sti_skip;
      float pos3d_y1r = pos3d_y + pos3d_y_add * sti_general_sti_skip;
      float pos3d_z2r = pos3d_z + pos3d_z_add * sg;
      float pos3d_y2r = pos3d_y + pos3d_y_add * sg;
      float scr_y2r = res_y2 + pos3d_y1r / pos3d_z1r;
      float scr_y1r = res_y2 + pos3d_y2r / pos3d_z2r;
      float u1z = tex / pos3d_z2r;
      float u2dz = (float)texture / pos3d_z1r - u1z;
      float onez1 = 1.0f / pos3d_z2r;
      float onedz2 = 1.0f / pos3d_z1r - onez1;
      u2dz /= scr_y2r - scr_y1r;
      onedz2 /= scr_y2r - scr_y1r;
      float mult = (float)(y + 1) - scr_y1r;
      float uz = u1z + u2dz * mult;
      float onez = onez1 + onedz2 * mult;
      c_span++;
      // Pixel loop (Cuda_Render.h:687-733)
      for (; y < scr_y2; ++y, uz += u2dz, onez += onedz2) {
        if (bit(y)) continue;
        float fu = std::fmin(std::fmax(uz / onez, (float)texture), tex - 1.0f);
#if ORC_PLANT == 1
        fu = (float)texture;
#elif ORC_PLANT == 7
        fu = std::fmin(std::fmax(uz / onez, (float)texture), tex);
#endif
        uint32_t u = (uint32_t)f2i(fu);
#if ORC_PLANT == 9
        uint32_t real_z = (uint32_t)f2i(1.0f / onez + 0.5f) & 0xfffeu;
#else
        uint32_t real_z = (uint32_t)f2i(1.0f / onez) & 0xfffeu;
#endif
        uint32_t color16 = send[u];
        out[y] = color16 + (real_z << 16);
        c_col++;
        y_cache[y >> 5] |= 1u << (y & 31);
      }
    }
  }
  ORC_HOOK_LINE_END
  if (ctr) {
    ctr->lines += 1;
    ctr->dda_steps += c_steps;
    ctr->ptr_loads += c_ptr;
    ctr->rle_loads += c_rle;
    ctr->elems += c_elem;
    ctr->spans += c_span;
    ctr->col_loads += c_col;
    ctr->cleared += c_clear;
  }
}

// All lines of one frame (cudaRender, Cuda_Main.cu:150-181: thread x = line x).
// raybuf: map_line_count*rs texels; line_clip: 2 per line.
void orc_raycast(const rlevox_mip* mips, int nmips, const rlevox_raymap* rm, uint32_t* raybuf,
                 int32_t* line_clip, rlevox_counters* ctr, int threads) {
  OrcScene sc{mips, nmips};
  const int lines = rm->map_line_count;
  const int words = (rm->rs + 31) >> 5;
  rlevox_counters total;
  std::memset(&total, 0, sizeof(total));
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
  {
    std::vector<uint32_t> cache(words);
    rlevox_counters local;
    std::memset(&local, 0, sizeof(local));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int x = 0; x < lines; x++) orc_render_line(x, sc, *rm, raybuf, line_clip, cache.data(), &local);
#ifdef _OPENMP
#pragma omp critical
#endif
    {
      total.lines += local.lines;
      total.dda_steps += local.dda_steps;
      total.ptr_loads += local.ptr_loads;
      total.rle_loads += local.rle_loads;
      total.elems += local.elems;
      total.spans += local.spans;
      total.col_loads += local.col_loads;
      total.cleared += local.cleared;
    }
  }
  if (ctr) *ctr = total;
}

// ---------------------------------------------------------------- compose
// colorize_buddha_soft.frag:12-92 per screen pixel, fp32 in the shader's
// operation order.  Terms weighted by exactly 0 or 1 (step() results) are
// applied as selections, which is bit-identical whenever they are finite.
// Texel fetch = GL_NEAREST with GL_CLAMP_TO_EDGE (GL_Main.cpp:154-157) on a
// texture rs wide and 4*rs lines high (main.cpp:459,474-482).  Row 0 of the
// output is the top of the image (gl_FragCoord.y = H - row - 0.5).
static inline void compose_pixel(const rlevox_raymap& rm, const uint32_t* raybuf,
                                 const int32_t* line_clip, int px, int row, uint16_t* col,
                                 int32_t* dep) {
  const float RESX = (float)rm.W, RESY = (float)rm.H;
  const int rs = rm.rs;
  const float border = (RESX - RESY) / (RESX * 2.0f);
  const float vx = rm.vanish[0], vy = rm.vanish[1];
  const float fcx = (float)px + 0.5f;
  const float fcy = (float)(rm.H - 1 - row) + 0.5f;
  const float scx = fcx / RESX, scy = fcy / RESY;
  const float scx1 = scx - vx, scy1 = scy - vy;
  const bool upper = !(0.0f < scy1);  // step(scy1, 0.0)
  const bool left = !(0.0f < scx1);   // step(scx1, 0.0)
  const float e = std::fabs(scy1) - std::fabs(scx1) * RESX / RESY;
  const bool ostep = !(0.0f < e);
  const float o2 = (ostep ? fcx : fcy) / RESX;
  const bool rxg = rm.rot_x_greater_zero != 0.0f;
  float ty, tx;
  if (!ostep) {
    const float s = upper ? 1.0f : 0.0f;
    float ang2 = scx1 * std::fabs(1.0f - s - vy) / scy1;
    ang2 = upper ? ang2 + (1.0f - vx) : ang2 + vx;
    const float x_pre = ang2;
    ty = upper ? (rm.ofs_add[1] + x_pre) : (rm.ofs_add[0] + 1.0f - x_pre);
    const float ob = o2 + border;
    // seg_up_x / seg_dn_x (colorize_buddha_soft.frag:68-77)
    const bool first = upper ? !rxg : rxg;
    tx = first ? ob : 1.0f - ob;
#if ORC_PLANT == 4
    tx = 1.0f - tx;
#endif
  } else {
    const float t = left ? 1.0f : 0.0f;
    float ang3 = scy1 * std::fabs(1.0f - t - vx) / scx1;
    ang3 = left ? ang3 + (1.0f - vy) : ang3 + vy;
    ang3 = ang3 * RESY / RESX + border;
    const float x_pre = ang3;
    ty = left ? (rm.ofs_add[3] + x_pre) : (rm.ofs_add[2] + 1.0f - x_pre);
    const bool first = left ? !rxg : rxg;
    tx = first ? o2 : 1.0f - o2;
  }
  ty = ty * rm.ray_ratio * 0.25f;
  uint16_t c = RLEVOX_SKY_COLOUR;
  int32_t d = RLEVOX_SKY_DEPTH;
  if (std::isfinite(tx) && std::isfinite(ty)) {
    int tc = f2i(std::floor(tx * (float)rs));
#if ORC_PLANT == 10
    int tr = f2i(std::floor(ty * (float)(4 * rs) + 0.5f));
#else
    int tr = f2i(std::floor(ty * (float)(4 * rs)));
#endif
    tc = tc < 0 ? 0 : (tc > rs - 1 ? rs - 1 : tc);
    tr = tr < 0 ? 0 : (tr > 4 * rs - 1 ? 4 * rs - 1 : tr);
    if (tr < rm.map_line_count && tc >= line_clip[2 * tr] && tc <= line_clip[2 * tr + 1]) {
      const uint32_t t = raybuf[(size_t)tr * (size_t)rs + (size_t)tc];
      if (((t >> 16) & 0xffu) != 0xffu) {  // c.z != 1.0 (frag:89)
        c = (uint16_t)(t & 0xffffu);
        d = (int32_t)(t >> 16);
      }
    }
  }
  col[(size_t)row * rm.W + px] = c;
  dep[(size_t)row * rm.W + px] = d;
}

void orc_compose(const rlevox_raymap* rm, const uint32_t* raybuf, const int32_t* line_clip,
                 uint16_t* col, int32_t* dep) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int row = 0; row < rm->H; row++)
    for (int px = 0; px < rm->W; px++) compose_pixel(*rm, raybuf, line_clip, px, row, col, dep);
}

// ---------------------------------------------------------------- gap-fill
// soft.frag:8-74 restated on integer framebuffers.  w = 1/depth, so the
// shader's max-w ring search becomes a min-depth search and its
// `w >= 0.7*radmax` test becomes 10*dmin >= 7*d.  Radii are in pixels of the
// 2048-wide target: 0.0023 * 2048 * 0.001 * 65280 / d = 307.5/d, kept in
// 1/256 px (78720/d); the shader filters when that exceeds 0.00008*2048 px
// (42/256 px).  Taps round to the nearest pixel and clamp to the image.  The
// low byte of colour16 (shade) is averaged with rounding; the high byte and
// the depth stay the centre pixel's.
static const int kRingDx[7] = {0, 9, 9, 2, -8, -10, -3};
static const int kRingDy[7] = {10, 6, -4, -10, -7, 3, 10};

static inline int floor_div(int a, int b) {  // b > 0
  int q = a / b, r = a % b;
  return (r != 0 && r < 0) ? q - 1 : q;
}

void orc_gapfill(int W, int H, const uint16_t* cin, const int32_t* dep, uint16_t* cout) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int row = 0; row < H; row++) {
    for (int x = 0; x < W; x++) {
      const size_t i = (size_t)row * W + x;
      const uint16_t c0 = cin[i];
      int dmin = dep[i];
      for (int k = 0; k < 7; k++) {
        int xx = std::min(std::max(x + kRingDx[k], 0), W - 1);
        int yy = std::min(std::max(row - kRingDy[k], 0), H - 1);
        int d = dep[(size_t)yy * W + xx];
        if (d < dmin) dmin = d;
      }
      if (dmin >= RLEVOX_SKY_DEPTH) {
        cout[i] = c0;
        continue;
      }
      const int dm = dmin < 1 ? 1 : dmin;
      const int r_q8 = 78720 / dm;
      if (r_q8 < 42) {
        cout[i] = c0;
        continue;
      }
      int n = 1;
      int sum = c0 & 0xff;
      for (int a = 0; a < 5; a++) {
        const int ox = floor_div((2 * a - 5) * r_q8 + 640, 1280);
        for (int b = 0; b < 5; b++) {
          const int oy = floor_div((2 * b - 5) * r_q8 + 640, 1280);
          int xx = std::min(std::max(x + ox, 0), W - 1);
#if ORC_PLANT == 5
          int yy = std::min(std::max(row + oy, 0), H - 1);
#else
          int yy = std::min(std::max(row - oy, 0), H - 1);
#endif
          const size_t t = (size_t)yy * W + xx;
          const int d = dep[t];
          if (d < RLEVOX_SKY_DEPTH && (int64_t)10 * dmin >= (int64_t)7 * d) {
            n++;
            sum += cin[t] & 0xff;
          }
        }
      }
      if (n > 6)
        cout[i] = (uint16_t)((c0 & 0xff00) | ((sum + n / 2) / n));
      else
        cout[i] = c0;
    }
  }
}

// ------------------------------------------------------------- RLE codec
// Surface test of RLE4::compress (Rle4.cpp:100-132): a solid voxel is stored
// when fewer than 27 voxels of its 3x3x3 block (inside the volume) are solid.
void orc_surface_dense(int sx, int sy, int sz, const uint8_t* solid, uint8_t* surf) {
  const size_t sxy = (size_t)sx * sy;
  for (int k = 0; k < sz; k++)
    for (int j = 0; j < sy; j++)
      for (int i = 0; i < sx; i++) {
        const size_t idx = i + j * (size_t)sx + k * sxy;
        if (!solid[idx]) {
          surf[idx] = 0;
          continue;
        }
        int cnt = 0;
        for (int a = -1; a < 2; a++)
          for (int b = -1; b < 2; b++)
            for (int c = -1; c < 2; c++) {
              int x = i + a, y = j + b, z = k + c;
              if (x < 0 || y < 0 || z < 0 || x >= sx || y >= sy || z >= sz) continue;
              if (solid[x + y * (size_t)sx + z * sxy]) cnt++;
            }
        surf[idx] = cnt < 27 ? 1 : 0;
      }
}

// RLE4::compress slab layout and run splitting (Rle4.cpp:78-196) followed by
// the pointer-map rebuild of RLE4::load (Rle4.cpp:295-314).  Colours come
// from `colour` instead of encoded normals.  Returns the slab count, or -1
// when `cap` is too small.
int64_t orc_encode_dense(int sx, int sy, int sz, const uint8_t* solid, const uint16_t* colour,
                         uint16_t* slabs_out, uint64_t cap, uint32_t* ptrmap_out) {
  std::vector<uint8_t> surf((size_t)sx * sy * sz);
  orc_surface_dense(sx, sy, sz, solid, surf.data());
  std::vector<uint16_t> slab;
  std::vector<uint16_t> texture;
  std::vector<uint32_t> hdr((size_t)sx * sz);
  const size_t sxy = (size_t)sx * sy;
  for (int k = 0; k < sz; k++)
    for (int i = 0; i < sx; i++) {
      hdr[i + (size_t)k * sx] = (uint32_t)slab.size();
      size_t slab_len = slab.size();
      slab.push_back(0);
      size_t tex_len = slab.size();
      slab.push_back(0);
      int skip = 0, solidn = 0;
      texture.clear();
      for (int j = 0; j < sy + 1; j++) {
        bool store = false;
        if (j < sy) {
          const size_t idx = i + j * (size_t)sx + k * sxy;
          if (surf[idx]) {
            texture.push_back(colour[idx]);
            store = true;
          }
        }
        if (solidn > 0 && !store) {
          while (skip > 1023) {
            slab.push_back(1023);
            skip -= 1023;
          }
          while (solidn > 63) {
            slab.push_back((uint16_t)(63 * 1024 + (skip & 1023)));
            solidn -= 63;
            skip = 0;
          }
          slab.push_back((uint16_t)((solidn & 63) * 1024 + (skip & 1023)));
          solidn = 0;
          skip = 0;
        }
        if (store) solidn++;
        else skip++;
      }
      slab[slab_len] = (uint16_t)(slab.size() - (slab_len + 2));
      slab[tex_len] = (uint16_t)texture.size();
      for (size_t t = 0; t < texture.size(); t++) slab.push_back(texture[t]);
    }
  if (slab.size() > cap) return -1;
  std::memcpy(slabs_out, slab.data(), slab.size() * 2);
  // RLE4::load pointer map (Rle4.cpp:295-314)
  size_t ofs = 0;
  for (int zz = 0; zz < sz; zz++)
    for (int xx = 0; xx < sx; xx++) {
      const size_t c = xx + (size_t)zz * sx;
      ptrmap_out[2 * c] = (uint32_t)ofs;
      uint32_t count = slab[ofs];
      uint32_t firstrle = 0;
      if (ofs + 2 < slab.size()) firstrle = slab[ofs + 2];
      ptrmap_out[2 * c + 1] = count + (firstrle << 16);
      ofs += (size_t)slab[ofs] + slab[ofs + 1] + 2;
    }
  return (int64_t)slab.size();
}

// Tree::get_mipmap (tree.h:23-85): a coarse voxel is set when at least 3 of
// its 8 children are; it takes the colour of that third child, visited in
// (subx, suby, subz) order as the material bits are (tree.h:54-75).
void orc_mip_dense(int sx, int sy, int sz, const uint8_t* solid, const uint16_t* colour,
                   uint8_t* solid_out, uint16_t* colour_out) {
  const int hx = sx / 2, hy = sy / 2, hz = sz / 2;
  for (int k = 0; k < hz; k++)
    for (int j = 0; j < hy; j++)
      for (int i = 0; i < hx; i++) {
        const size_t o = i + j * (size_t)hx + k * (size_t)hx * hy;
        int mincnt = 3;
        bool bit0 = false;
        uint16_t c = 0;
        for (int subx = 0; subx < 2 && !bit0; subx++)
          for (int suby = 0; suby < 2 && !bit0; suby++)
            for (int subz = 0; subz < 2 && !bit0; subz++) {
              const size_t idx = (i * 2 + subx) + (j * 2 + suby) * (size_t)sx +
                                 (k * 2 + subz) * (size_t)sx * sy;
              if (solid[idx]) {
                mincnt--;
                if (mincnt == 0) {
                  bit0 = true;
                  c = colour[idx];
                }
              }
            }
        solid_out[o] = bit0 ? 1 : 0;
        colour_out[o] = bit0 ? c : 0;
      }
}

// Decodes one mip back to a dense volume: surf[idx] = 1 for every stored
// voxel, col[idx] = its colour (the read loop of RLE4::load's debug dump,
// Rle4.cpp:339-377).  Returns 0, or 1 if a column is malformed.
int orc_decode(const rlevox_mip* mip, uint8_t* surf, uint16_t* col) {
  const int sx = mip->sx, sy = mip->sy, sz = mip->sz;
  const size_t sxy = (size_t)sx * sy;
  std::memset(surf, 0, sxy * sz);
  std::memset(col, 0, sxy * sz * 2);
  for (int k = 0; k < sz; k++)
    for (int i = 0; i < sx; i++) {
      const size_t c = i + (size_t)k * sx;
      const uint64_t ofs = (uint64_t)mip->ptrmap[2 * c] << mip->offset_shift;
      if (ofs + 2 > mip->nslabs) return 1;
      const uint16_t* p = mip->slabs + ofs;
      const int count = p[0], ncol = p[1];
      if (ofs + 2 + count + ncol > mip->nslabs) return 1;
      int y = 0, t = 0;
      for (int e = 0; e < count; e++) {
        const uint16_t w = p[2 + e];
        y += w & 1023;
        const int drawn = w >> 10;
        for (int d = 0; d < drawn; d++, y++, t++) {
          if (y >= sy || t >= ncol) return 1;
          surf[i + y * (size_t)sx + k * sxy] = 1;
          col[i + y * (size_t)sx + k * sxy] = p[2 + count + t];
        }
      }
      if (t != ncol) return 1;
    }
  return 0;
}

int orc_plant(void) { return ORC_PLANT; }

// Line-plane parameters of every ray line, for checks of the per-line
// arithmetic downstream of them: out[7*x ..] = ray_x, ray_z, vertical (0/1),
// start2d.x, start2d.y, end2d.x, end2d.y.
void orc_line_planes(const rlevox_raymap* rm, float* out) {
  for (int x = 0; x < rm->map_line_count; x++) {
    const LinePlane lp = line_plane(x, *rm);
    float* o = out + 7 * (size_t)x;
    o[0] = lp.ray_x;
    o[1] = lp.ray_z;
    o[2] = lp.vertical ? 1.0f : 0.0f;
    o[3] = lp.start2d.x;
    o[4] = lp.start2d.y;
    o[5] = lp.end2d.x;
    o[6] = lp.end2d.y;
  }
}

}  // extern "C"
