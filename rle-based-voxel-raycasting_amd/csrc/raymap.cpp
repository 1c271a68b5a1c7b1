// Per-frame view setup: the screen is split into up to four fans of ray
// lines radiating from the vanishing point of the down vector, one fan per
// screen side.  Restates RayMap::get_ray_map (RLE-Raycaster/src/RayMap.h:
// 98-402) with set_border((W-H)/(2W)) (main.cpp:776) and the line budget
// RAYS_CASTED_RES/4 generalised to the render size rs = W (core.h:5-7).
//
// All fp32 operations keep the reference's order; transcendental functions
// are evaluated in double and rounded once.  The block also carries the
// per-frame constants that render_line would otherwise recompute in every
// thread (Cuda_Render.h:77-78,188-191,335) and the compose uniforms
// (main.cpp:575-603), so the device code contains no transcendental call.
#include <cmath>
#include <cstring>

#include "internal.h"

namespace rlevox {
namespace {

struct Vec {
  float x, y, z;
};
inline Vec vec(float x, float y, float z) { return Vec{x, y, z}; }
inline Vec add(Vec a, Vec b) { return vec(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec sub(Vec a, Vec b) { return vec(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec scale(Vec a, float s) { return vec(a.x * s, a.y * s, a.z * s); }
// vec3f::dot (VecMath.h:60-61) with `me` as the object: a.x*x + a.y*y + a.z*z
inline float vdot(Vec me, Vec a) { return a.x * me.x + a.y * me.y + a.z * me.z; }
inline float rsin(float a) { return (float)std::sin((double)a); }
inline float rcos(float a) { return (float)std::cos((double)a); }

// Row-vector 4x4 matrix of the RadonLabs mathlib (_matrix44.h).
struct Mat {
  float m[4][4];
  void identity() {
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) m[i][j] = i == j ? 1.0f : 0.0f;
  }
  void rotate_x(float a) {  // _matrix44.h:529-541
    const float c = rcos(a), s = rsin(a);
    for (int i = 0; i < 4; i++) {
      const float m1 = m[i][1], m2 = m[i][2];
      m[i][1] = m1 * c + m2 * -s;
      m[i][2] = m1 * s + m2 * c;
    }
  }
  void rotate_y(float a) {  // _matrix44.h:547-559
    const float c = rcos(a), s = rsin(a);
    for (int i = 0; i < 4; i++) {
      const float m0 = m[i][0], m2 = m[i][2];
      m[i][0] = m0 * c + m2 * s;
      m[i][2] = m0 * -s + m2 * c;
    }
  }
  Vec apply(Vec v) const {  // operator*(matrix44, vector3), _matrix44.h:863-870
    return vec(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z + m[3][0],
               m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z + m[3][1],
               m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z + m[3][2]);
  }
  void invert_rigid() {  // invert_simpler, _matrix44.h:431-443
    float t;
    t = m[0][1]; m[0][1] = m[1][0]; m[1][0] = t;
    t = m[0][2]; m[0][2] = m[2][0]; m[2][0] = t;
    t = m[1][2]; m[1][2] = m[2][1]; m[2][1] = t;
    const float r0 = -(m[0][0] * m[3][0] + m[1][0] * m[3][1] + m[2][0] * m[3][2]);
    const float r1 = -(m[0][1] * m[3][0] + m[1][1] * m[3][1] + m[2][1] * m[3][2]);
    m[3][2] = -(m[0][2] * m[3][0] + m[1][2] * m[3][1] + m[2][2] * m[3][2]);
    m[3][1] = r1;
    m[3][0] = r0;
  }
};

inline float length(Vec v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

// vec3f::angle (VecMath.h:71-81)
float angle_between(Vec me, Vec v) {
  const float d = v.x * me.x + v.y * me.y + v.z * me.z;
  float len = length(v) * length(me);
  if (len == 0) len = 0.00001f;
  float c = d / len;
  if (c < -1) c = -1;
  if (c > 1) c = 1;
  return (float)std::acos((double)c);
}

// Snaps a segment end to the line grid (RayMap.h:231-232 and siblings).
inline float snap(int maxres, float v, int delta) {
  return float(f2i((float)maxres * v) + delta) / (float)maxres;
}

// Line count of one segment (RayMap.h:233-236 and siblings).
inline int segment_lines(int maxres, float lo, float hi) {
  int n = f2i((float)maxres * std::fabs(lo - hi));
  if (lo - hi > 0) n = 0;
  if (n > maxres * 3) n = maxres * 3;
  return n;
}

}  // namespace

rlevox_status raymap_compute(const rlevox_pose& pose, int W, int H, int mip_distance,
                             int far_distance, rlevox_raymap& rm) {
  if (W < 16 || H < 16 || H > W || W > RLEVOX_MAX_RS)
    return fail(RLEVOX_ERR_INVALID_ARG, "raymap: need 16 <= H <= W <= %d, got %dx%d",
                RLEVOX_MAX_RS, W, H);
  for (int i = 0; i < 3; i++)
    if (!std::isfinite(pose.pos[i]) || !std::isfinite(pose.rot[i]))
      return fail(RLEVOX_ERR_INVALID_ARG, "raymap: pose has a non-finite component");
  if (pose.rot[2] != 0.0f)
    return fail(RLEVOX_ERR_INVALID_ARG, "raymap: camera roll is not supported by the method");
  // At zero pitch the down vector has no vanishing point (RayMap.h:138-141);
  // the application keeps |pitch| >= 0.01 rad (main.cpp:398-405).
  if (!(std::fabs(pose.rot[0]) >= 0.01f) || !(std::fabs(pose.rot[0]) < 3.13f))
    return fail(RLEVOX_ERR_INVALID_ARG, "raymap: need 0.01 <= |pitch| < 3.13, got %g",
                (double)pose.rot[0]);
  if (std::fabs(pose.pos[0]) > 1.0e6f || std::fabs(pose.pos[1]) > 1.0e6f ||
      std::fabs(pose.pos[2]) > 1.0e6f)
    return fail(RLEVOX_ERR_INVALID_ARG, "raymap: |position| must stay below 1e6");
  const int mipd = mip_distance > 0 ? mip_distance : 1024;
  const int zfar = far_distance > 0 ? far_distance : 80000;
  if (mipd > (1 << 20) || zfar > (1 << 22))
    return fail(RLEVOX_ERR_INVALID_ARG, "raymap: mip/far distance out of range");

  std::memset(&rm, 0, sizeof(rm));
  rm.W = W;
  rm.H = H;
  rm.rs = W;
  rm.maxres = W;
  rm.far_distance = zfar;
  const float border = ((float)W - (float)H) / ((float)W * 2.0f);
  rm.border = border;
  rm.clip_min = border;
  rm.clip_max = 1 - border;
  for (int i = 0; i < 3; i++) {
    rm.rotation[i] = pose.rot[i];
    rm.position[i] = pose.pos[i];
  }

  // Frustum corners, origin and (later) the vanishing point (RayMap.h:107-114)
  Vec fr[6] = {vec(1, 1, 1), vec(-1, 1, 1), vec(-1, -1, 1), vec(1, -1, 1), vec(0, 0, 0), vec(0, 0, 0)};
  Mat cam;
  cam.identity();
  cam.rotate_x(pose.rot[0]);
  cam.rotate_y(pose.rot[1]);
  cam.m[3][0] += 3;  // translate(vector3(3,2,0))
  cam.m[3][1] += 2;
  cam.m[3][2] += 0;
  for (int i = 0; i < 5; i++) fr[i] = cam.apply(fr[i]);

  const Vec down = vec(0, -1, 0);
  const Vec sum = add(fr[0], fr[2]);
  const Vec view_dir = sub(vec(sum.x / 2.0f, sum.y / 2.0f, sum.z / 2.0f), fr[4]);
  const float alpha = float(M_PI) / 2.0f - angle_between(down, view_dir);
  const float dist = 1.0f / rsin(alpha);
  fr[5] = add(fr[4], scale(down, dist));

  // Image-plane basis (RayMap.h:145-161)
  const Vec e1 = sub(fr[1], fr[0]), e2 = sub(fr[3], fr[0]);
  const Vec n = vec(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
  Vec basis[4] = {e1, e2, n, fr[0]};
  for (int i = 0; i < 3; i++) {
    const Vec b = basis[i];
    basis[i] = scale(b, 1.0f / (b.x * b.x + b.y * b.y + b.z * b.z));
  }
  Mat plane;
  for (int i = 0; i < 4; i++) {
    plane.m[i][0] = basis[i].x;
    plane.m[i][1] = basis[i].y;
    plane.m[i][2] = basis[i].z;
    plane.m[i][3] = 1.0f;
  }
  std::memcpy(rm.to3d, plane.m, sizeof(rm.to3d));
  plane.invert_rigid();
  Vec p2d[8];
  for (int i = 0; i < 8; i++) p2d[i] = vec(0, 0, 0);
  for (int i = 0; i < 6; i++) p2d[i] = plane.apply(fr[i]);
  rm.vp2d[0] = p2d[5].x;
  rm.vp2d[1] = p2d[5].y;
  rm.vp2d[2] = p2d[5].z;
  p2d[5].z = 0;
  const Vec vp = p2d[5];

  // Fans of lines per screen side (RayMap.h:176-352)
  const int maxres = W;
  const float ys_min = border, ys_max = 1 - border;
  const Vec dir[4] = {vec(-1, -1, 0), vec(1, -1, 0), vec(1, 1, 0), vec(-1, 1, 0)};
  const Vec corner[4] = {vec(0, ys_min, 0), vec(1, ys_min, 0), vec(1, ys_max, 0), vec(0, ys_max, 0)};
  Vec pn[8];
  for (int i = 0; i < 8; i++) pn[i] = vec(0, 0, 0);
  int res[4] = {0, 0, 0, 0};
  float omin[4] = {0, 0, 0, 0}, omax[4] = {0, 0, 0, 0};
  auto toward = [&](int id, float ratio_num, float ratio_den) {
    return add(vp, scale(sub(corner[id], vp), std::fabs(ratio_num / ratio_den)));
  };

  if (vp.y > border) {  // upper fan (RayMap.h:209-240)
    pn[0] = add(vp, scale(dir[0], std::fabs(vp.y - border)));
    pn[1] = add(vp, scale(dir[1], std::fabs(vp.y - border)));
    const Vec in = pn[1];
    if (vp.x > 1) {
      if (pn[1].x > 1) pn[1].x = 1;
    } else if (vdot(sub(pn[0], vp), sub(corner[2], vp)) > 0) {
      pn[1] = toward(2, corner[0].y - vp.y, corner[2].y - vp.y);
    }
    if (vp.x < 0) {
      if (pn[0].x < 0) pn[0].x = 0;
    } else if (vdot(sub(in, vp), sub(corner[3], vp)) > 0) {
      pn[0] = toward(3, corner[1].y - vp.y, corner[3].y - vp.y);
    }
    pn[0].y = pn[1].y = corner[0].y;
    pn[0].x = snap(maxres, pn[0].x, -2);
    pn[1].x = snap(maxres, pn[1].x, +2);
    res[0] = segment_lines(maxres, pn[0].x, pn[1].x);
    omin[0] = pn[0].x;
    omax[0] = pn[1].x;
  }
  if (vp.y < 1 - border) {  // lower fan (RayMap.h:243-277)
    pn[2] = add(vp, scale(dir[3], std::fabs(vp.y - 1 + border)));
    pn[3] = add(vp, scale(dir[2], std::fabs(vp.y - 1 + border)));
    const Vec in = pn[2];
    if (vp.x < 0) {
      if (pn[2].x < 0) pn[2].x = 0;
    } else if (vdot(sub(pn[3], vp), sub(corner[0], vp)) > 0) {
      pn[2] = toward(0, corner[2].y - vp.y, corner[0].y - vp.y);
    }
    if (vp.x > 1) {
      if (pn[3].x > 1) pn[3].x = 1;
    } else {
      const Vec d = sub(corner[1], vp);
      if (vdot(sub(in, vp), d) > 0) pn[3] = add(vp, scale(d, std::fabs((corner[3].y - vp.y) / d.y)));
    }
    pn[2].y = pn[3].y = corner[2].y;
    pn[2].x = snap(maxres, pn[2].x, -2);
    pn[3].x = snap(maxres, pn[3].x, +2);
    res[1] = segment_lines(maxres, pn[2].x, pn[3].x);
    omin[1] = pn[2].x;
    omax[1] = pn[3].x;
  }
  if (vp.x > 0) {  // left fan (RayMap.h:280-313)
    pn[4] = add(vp, scale(dir[0], std::fabs(vp.x)));
    pn[5] = add(vp, scale(dir[3], std::fabs(vp.x)));
    const Vec in = pn[5];
    if (vp.y > 1 - border) {
      if (pn[5].y > 1 - border) pn[5].y = 1 - border;
    } else if (vdot(sub(pn[4], vp), sub(corner[2], vp)) > 0) {
      pn[5] = toward(2, corner[3].x - vp.x, corner[2].x - vp.x);
    }
    if (vp.y < border) {
      if (pn[4].y < border) pn[4].y = border;
    } else if (vdot(sub(in, vp), sub(corner[1], vp)) > 0) {
      pn[4] = toward(1, corner[0].x - vp.x, corner[1].x - vp.x);
    }
    pn[4].x = pn[5].x = 0;
    pn[4].y = snap(maxres, pn[4].y, -2);
    pn[5].y = snap(maxres, pn[5].y, +2);
    res[2] = segment_lines(maxres, pn[4].y, pn[5].y);
    omin[2] = pn[4].y;
    omax[2] = pn[5].y;
  }
  if (vp.x < 1) {  // right fan (RayMap.h:316-352)
    pn[6] = add(vp, scale(dir[1], std::fabs(1 - vp.x)));
    pn[7] = add(vp, scale(dir[2], std::fabs(1 - vp.x)));
    const Vec in = pn[7];
    if (vp.y > 1 - border) {
      if (pn[7].y > 1 - border) pn[7].y = 1 - border;
    } else if (vdot(sub(pn[6], vp), sub(corner[3], vp)) > 0) {
      pn[7] = toward(3, corner[2].x - vp.x, corner[3].x - vp.x);
    }
    if (vp.y < border) {
      if (pn[6].y < border) pn[6].y = border;
    } else if (vdot(sub(in, vp), sub(corner[0], vp)) > 0) {
      pn[6] = toward(0, corner[1].x - vp.x, corner[0].x - vp.x);
    }
    pn[6].x = pn[7].x = 1;
    pn[6].y = snap(maxres, pn[6].y, -2);
    pn[7].y = snap(maxres, pn[7].y, +2);
    res[3] = segment_lines(maxres, pn[6].y, pn[7].y);
    omin[3] = pn[6].y;
    omax[3] = pn[7].y;
  }

  rm.p4[0] = fr[4].x;
  rm.p4[1] = fr[4].y;
  rm.p4[2] = fr[4].z;
  for (int i = 0; i < 8; i++) {
    rm.p_2d[i][0] = p2d[i].x;
    rm.p_2d[i][1] = p2d[i].y;
    rm.p_2d[i][2] = p2d[i].z;
    rm.p_no[i][0] = pn[i].x;
    rm.p_no[i][1] = pn[i].y;
    rm.p_no[i][2] = pn[i].z;
  }
  for (int j = 0; j < 4; j++) {
    rm.res[j] = res[j];
    rm.p_ofs_min[j] = omin[j];
    rm.p_ofs_max[j] = omax[j];
  }
  rm.map_line_count = res[0] + res[1] + res[2] + res[3];

  rm.sin_x = rsin(pose.rot[0]);
  rm.cos_x = rcos(pose.rot[0]);
  rm.sin_y = rsin(pose.rot[1]);
  rm.cos_y = rcos(pose.rot[1]);
  rm.cos_myaw = rcos(-pose.rot[1]);
  rm.sin_myaw = rsin(-pose.rot[1]);
  rm.mapswitch = (int)((double)mipd * (0.25 * (double)(4.0f - std::fabs(pose.rot[0]))));
  // The mip-switch loop doubles mapswitch until it exceeds the distance
  // walked (Cuda_Render.h:343-365); zero would never grow.
  if (rm.mapswitch < 1)
    return fail(RLEVOX_ERR_INVALID_ARG, "raymap: mip distance %d is too small at pitch %g", mipd,
                (double)pose.rot[0]);
  if (!mapswitch_bounded(rm.mapswitch, rm.position[1]))
    return fail(RLEVOX_ERR_INVALID_ARG, "raymap: mip distance %d too large for camera row %g", mipd,
                (double)rm.position[1]);

  const float lines_tex = (float)(4 * W);  // RAYS_CASTED_RES == RAYS_CASTED == 4*rs
  rm.vanish[0] = 1 - rm.vp2d[0];
  rm.vanish[1] = ((1 - rm.vp2d[1]) - border) * (float)W / (float)H;
  const float o1 = 4 * (float)res[0] / lines_tex;
  const float o2 = 4 * (float)res[1] / lines_tex + o1;
  const float o3 = 4 * (float)res[2] / lines_tex + o2;
  rm.ofs_add[0] = -omin[0];
  rm.ofs_add[1] = -omin[1] + o1;
  rm.ofs_add[2] = -omin[2] + o2;
  rm.ofs_add[3] = -omin[3] + o3;
  rm.rot_x_greater_zero = pose.rot[0] > 0 ? 1.0f : 0.0f;
  rm.ray_ratio = lines_tex / lines_tex;
  return RLEVOX_OK;
}

}  // namespace rlevox

extern "C" rlevox_status rlevox_raymap_compute(const rlevox_pose* pose, int32_t W, int32_t H,
                                               const rlevox_render_params* params,
                                               rlevox_raymap* out) {
  rlevox::clear_error();
  if (!pose || !out) return rlevox::fail(RLEVOX_ERR_INVALID_ARG, "raymap: null argument");
  const int mipd = params ? params->mip_distance : 0;
  const int zfar = params ? params->far_distance : 0;
  return rlevox::raymap_compute(*pose, W, H, mipd, zfar, *out);
}
