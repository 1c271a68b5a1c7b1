// Early end of a line's walk (shared by the raycast kernels; plain C++ so a
// host check can run the same code: the includer defines RLEVOX_HD as the
// function qualifiers it needs).
#pragma once

#include <math.h>

namespace rlevox {

// Early end of a line's walk.  The reference walks every line until its
// window closes or the DDA passes the far distance (Cuda_Render.h:366-372),
// and for lines that keep some sky that is the far distance.  The walk can
// stop as soon as no later step can draw, which leaves the line exactly as
// the full walk does:
//  * every later step lies at a crossing distance d >= D0 (the distances
//    only grow; the far-edge correction adds to them), and every drawn run
//    of the levels still to come has its top at a world row >= top;
//  * a step draws or moves the window only through an element whose top
//    projects above the floating horizon (scr_y2 > ymin, Cuda_Render.h:
//    540-543); otherwise its first drawn element in front of the eye breaks
//    the loop, and the clip-top cull (:467) changes nothing either;
//  * that projection, half + py/pz with pz = cos_x*ray_z*d + sin_x*y and
//    py = hmul*(cos_x*y - sin_x*ray_z*d) (upper and lower segments) or
//    hmul*ray_x*d (left and right), y = mountain + row, is a ratio of affine
//    functions of d for fixed y and of y for fixed d; with pz > 0 on the
//    whole range [D0, inf) x [mountain + top, mountain + sy] it has no pole
//    there, so its supremum is at a corner (d -> inf being the horizon limit).
// The test adds a margin of 1e-4 of the magnitudes involved plus 1/8 pixel,
// far above the fp32 rounding of the reference's own projection and of this
// evaluation (a few units in the last place each).  When even that bound
// truncates to <= ymin the walk is over.
struct WalkEnd {
  static constexpr float kRel = 1e-4f, kAbs = 0.125f;
  float kz, sx, cx, R, rx, rz, half, m, lim;
  int on, vertical;

  // The line's constants; cos_x, sin_x: of the pitch; hmul: the signed half
  // render size multiplying py (Cuda_Render.h:208-210); mountain: eye row
  // offset (viewpos.y).
  RLEVOX_HD void init(float cos_x, float sin_x, float hmul, float ray_x, float ray_z, bool vert,
                       float mountain, float half_) {
    cx = cos_x;
    sx = sin_x;
    R = hmul;
    rx = ray_x;
    rz = ray_z;
    half = half_;
    m = mountain;
    vertical = vert ? 1 : 0;
    kz = cos_x * ray_z;
    on = kz > 0.0f ? 1 : 0;  // pz must grow with d (else far cells fall behind the eye)
    lim = 0.0f;
    if (on) {  // the horizon limit d -> inf, with its margin
      const float pyk = (vert ? -hmul * sin_x * ray_z : hmul * ray_x) / kz;
      const float pya = fabsf(hmul) * (fabsf(sin_x * ray_z) + fabsf(ray_x)) / kz;
      lim = half + pyk + kRel * (fabsf(half) + fabsf(pyk) + pya);
    }
  }

  // Bound (with margin) at distance d and height y; ok = false when pz is
  // not safely positive there.
  RLEVOX_HD float corner(float d, float y, bool& ok) const {
    const float pz = kz * d + sx * y;
    const float pza = kz * d + fabsf(sx * y);
    if (!(pz > 1e-3f * pza) || !(pz > 1e-3f)) {
      ok = false;
      return 0.0f;
    }
    const float py = vertical ? R * (cx * y - sx * rz * d) : R * rx * d;
    const float pya = fabsf(R) * (fabsf(cx * y) + fabsf(sx * rz * d) + fabsf(rx * d));
    // one division: the products' rounding (a few ulps) is far inside the margin
    const float inv = 1.0f / pz;
    const float q = py * inv;
    return half + q + kRel * (fabsf(half) + fabsf(q) * (pza * inv) + pya * inv);
  }

  // A value no bound falls below: while ymin + (1 - kAbs) <= floor_(), done()
  // is false whatever the distance (the bound is at least the horizon
  // limit), so a line looking at its horizon skips the corners.
  RLEVOX_HD float floor_() const { return on ? lim : __builtin_inff(); }

  // True when no step at a distance >= D0 can draw while the floating
  // horizon is at ymin; top: highest run top (world row) of the levels still
  // to come, sy: column height.
  RLEVOX_HD bool done(float D0, int ymin, int top, int sy) const {
    if (!on) return false;
    bool ok = true;
    const float a = corner(D0, m + (float)top, ok);
    const float b = corner(D0, m + (float)sy, ok);
    if (!ok) return false;
    return fmaxf(fmaxf(a, b), lim) < (float)ymin + (1.0f - kAbs);
  }
};
// 32-bit words of a WalkEnd kept in LDS by the wavefront kernels (read once
// per chunk, so its values hold no registers across the walk).
constexpr int kWalkEndWords = (int)(sizeof(WalkEnd) / 4);

}  // namespace rlevox
