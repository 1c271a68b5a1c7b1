// ORACLE - test infrastructure only (checker of the checker).
//
// An independent per-pixel voxel raycaster in double precision.  It shares
// no code and no intermediate quantity with oracle.cpp's line renderer: it
// casts one ray per screen pixel through a dense decoded volume with a 3D
// Amanatides-Woo DDA and reports the first stored voxel.  Tests compare the
// line renderer's composed image with it pixel by pixel (same colour on the
// large majority of pixels, depth within a tolerance), which pins the ray
// map, line setup, DDA, RLE decode, projection, colour indexing and compose
// mapping at once.
//
// Camera model (derived from RayMap.h:107-161 and Cuda_Render.h:159-172,
// 271-272,459-462): pixel (px,row) looks along camera-space
// d = (2*sx-1, (2*sy-1)*H/W, 1) with sx = (px+0.5)/W, sy = (H-row-0.5)/H;
// frustum space f = d * Rx(pitch) * Ry(yaw) (row vectors, _matrix44.h
// rotate_x/rotate_y); voxel space is (x, row_down, z) = (f.x, -f.y, -f.z)
// from the eye at (pos.x, -pos.y, pos.z), and the depth of a point is its
// camera-space d.z multiple.  x and z wrap with the map size (Cuda_Render.h:
// 441-442).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../include/rlevox.h"

extern "C" int orc_brute_render(int sx, int sy, int sz, const uint8_t* surf, const uint16_t* col,
                                const rlevox_pose* pose, int W, int H, double max_hdist,
                                uint16_t* out_col, int32_t* out_depth, float* out_hdist) {
  if ((sx & (sx - 1)) || (sz & (sz - 1))) return 1;
  const double pitch = pose->rot[0], yaw = pose->rot[1];
  const double c = std::cos(pitch), s = std::sin(pitch);
  const double cb = std::cos(yaw), sb = std::sin(yaw);
  const double ox = pose->pos[0], oy = -(double)pose->pos[1], oz = pose->pos[2];
  const size_t sxy = (size_t)sx * sy;
#pragma omp parallel for schedule(dynamic, 4)
  for (int row = 0; row < H; row++) {
    for (int px = 0; px < W; px++) {
      const size_t o = (size_t)row * W + px;
      const double X = 2.0 * ((px + 0.5) / W) - 1.0;
      const double Y = (2.0 * ((H - row - 0.5) / H) - 1.0) * (double)H / (double)W;
      const double Z = 1.0;
      // d * Rx
      const double v0 = X, v1 = c * Y - s * Z, v2 = s * Y + c * Z;
      // * Ry
      const double fx = v0 * cb + v2 * sb, fy = v1, fz = -v0 * sb + v2 * cb;
      const double D[3] = {fx, -fy, -fz};
      const double O[3] = {ox, oy, oz};
      const double hlen = std::sqrt(D[0] * D[0] + D[2] * D[2]);
      out_col[o] = RLEVOX_SKY_COLOUR;
      out_depth[o] = -1;
      out_hdist[o] = -1.0f;
      long cell[3];
      int step[3];
      double tmax[3], tdelta[3];
      for (int a = 0; a < 3; a++) {
        cell[a] = (long)std::floor(O[a]);
        if (D[a] > 0) {
          step[a] = 1;
          tdelta[a] = 1.0 / D[a];
          tmax[a] = ((double)cell[a] + 1.0 - O[a]) / D[a];
        } else if (D[a] < 0) {
          step[a] = -1;
          tdelta[a] = -1.0 / D[a];
          tmax[a] = (O[a] - (double)cell[a]) / -D[a];
        } else {
          step[a] = 0;
          tdelta[a] = INFINITY;
          tmax[a] = INFINITY;
        }
      }
      double t = 0.0;
      const double tlimit = hlen > 0 ? max_hdist / hlen : max_hdist;
      for (int it = 0; it < 1 << 22 && t <= tlimit; it++) {
        const long cy = cell[1];
        if (cy >= 0 && cy < sy) {
          const long cx = cell[0] & (long)(sx - 1);
          const long cz = cell[2] & (long)(sz - 1);
          const size_t idx = (size_t)cx + (size_t)cy * sx + (size_t)cz * sxy;
          if (surf[idx]) {
            out_col[o] = col[idx];
            out_depth[o] = (int32_t)std::floor(t);
            out_hdist[o] = (float)(t * hlen);
            break;
          }
        } else if ((cy < 0 && step[1] <= 0) || (cy >= sy && step[1] >= 0)) {
          break;  // outside the slab of the volume, moving away
        }
        int a = 0;
        if (tmax[1] < tmax[a]) a = 1;
        if (tmax[2] < tmax[a]) a = 2;
        t = tmax[a];
        tmax[a] += tdelta[a];
        cell[a] += step[a];
      }
    }
  }
  return 0;
}
