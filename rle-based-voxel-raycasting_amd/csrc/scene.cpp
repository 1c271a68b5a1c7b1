// Scene construction on the host: occupancy as per-column bitsets, mips by
// the >=3-of-8 rule (Tree::get_mipmap, tree.h:23-85), surface extraction and
// the RLE4 slab layout with its run splitting (RLE4::compress,
// Rle4.cpp:52-198), pointer maps as RLE4::load builds them (Rle4.cpp:295-314).
//
// Everything is bit-parallel over 64 voxels of a column: a column of sy
// voxels is ceil(sy/64) words, bit (y & 63) of word (y >> 6) = voxel row y
// (row 0 at the top).  Levels are encoded in two passes over the columns
// (sizes, then data) so the slab array is written once, in parallel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "internal.h"

namespace rlevox {
namespace {

inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

struct Level {
  int sx = 0, sy = 0, sz = 0, nw = 0;
  std::vector<uint64_t> bits;  // sx*sz columns of nw words, x fastest
  const uint64_t* col(int x, int z) const { return bits.data() + ((size_t)z * sx + x) * nw; }
  uint64_t* col(int x, int z) { return bits.data() + ((size_t)z * sx + x) * nw; }
  void init(int x, int y, int z) {
    sx = x;
    sy = y;
    sz = z;
    nw = (y + 63) >> 6;
    bits.assign((size_t)sx * sz * nw, 0);
  }
};

inline void set_rows(uint64_t* c, int a, int b) {  // rows [a, b)
  for (int y = a; y < b;) {
    const int w = y >> 6, o = y & 63;
    const int n = std::min(64 - o, b - y);
    const uint64_t m = (n == 64) ? ~0ull : (((1ull << n) - 1) << o);
    c[w] |= m;
    y += n;
  }
}
inline void clear_rows(uint64_t* c, int a, int b) {
  for (int y = a; y < b;) {
    const int w = y >> 6, o = y & 63;
    const int n = std::min(64 - o, b - y);
    const uint64_t m = (n == 64) ? ~0ull : (((1ull << n) - 1) << o);
    c[w] &= ~m;
    y += n;
  }
}
inline bool test(const uint64_t* c, int y) { return (c[y >> 6] >> (y & 63)) & 1; }

// Even bits of w packed into the low 32 bits.
inline uint64_t even_bits(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return x;
}

// Coarse level: a voxel is set when at least 3 of its 8 children are.
// Bit-sliced saturating counter over the 8 child planes.
void reduce_level(const Level& f, Level& c, int threads) {
  c.init(f.sx / 2, f.sy / 2, f.sz / 2);
  parallel_for(c.sz, threads, 4, [&](int64_t k) {
    std::vector<uint64_t> plane(8);
    for (int i = 0; i < c.sx; i++) {
      uint64_t* out = c.col(i, (int)k);
      const uint64_t* ch[4] = {f.col(2 * i, 2 * (int)k), f.col(2 * i, 2 * (int)k + 1),
                               f.col(2 * i + 1, 2 * (int)k), f.col(2 * i + 1, 2 * (int)k + 1)};
      for (int w = 0; w < c.nw; w++) {
        for (int q = 0; q < 4; q++) {
          const uint64_t lo = (2 * w < f.nw) ? ch[q][2 * w] : 0;
          const uint64_t hi = (2 * w + 1 < f.nw) ? ch[q][2 * w + 1] : 0;
          plane[2 * q] = even_bits(lo) | (even_bits(hi) << 32);
          plane[2 * q + 1] = even_bits(lo >> 1) | (even_bits(hi >> 1) << 32);
        }
        uint64_t s0 = 0, s1 = 0, s2 = 0;
        for (int p = 0; p < 8; p++) {
          const uint64_t c0 = s0 & plane[p];
          s0 ^= plane[p];
          const uint64_t c1 = s1 & c0;
          s1 ^= c0;
          s2 |= c1;
        }
        out[w] = s2 | (s1 & s0);
      }
      if (c.sy & 63) out[c.nw - 1] &= (1ull << (c.sy & 63)) - 1;
    }
  });
}

// Rows y with y-1, y, y+1 all set (rows outside the volume count as empty).
inline void vertical_core(const uint64_t* a, uint64_t* out, int nw) {
  for (int w = 0; w < nw; w++) {
    const uint64_t up = (a[w] << 1) | (w > 0 ? a[w - 1] >> 63 : 0);          // bit y <- y-1
    const uint64_t dn = (a[w] >> 1) | (w + 1 < nw ? a[w + 1] << 63 : 0);     // bit y <- y+1
    out[w] = a[w] & up & dn;
  }
}

// Surface of column (x,z): solid voxels whose 3x3x3 block (clipped to the
// volume, no wrap) is not entirely solid (Rle4.cpp:100-132).
void surface_column(const Level& L, int x, int z, uint64_t* surf, uint64_t* tmp) {
  const int nw = L.nw;
  const uint64_t* self = L.col(x, z);
  bool edge = x == 0 || z == 0 || x == L.sx - 1 || z == L.sz - 1;
  if (edge) {  // a neighbour outside the map counts as empty
    for (int w = 0; w < nw; w++) surf[w] = self[w];
    return;
  }
  uint64_t* e = surf;
  for (int w = 0; w < nw; w++) e[w] = ~0ull;
  for (int dz = -1; dz <= 1; dz++)
    for (int dx = -1; dx <= 1; dx++) {
      vertical_core(L.col(x + dx, z + dz), tmp, nw);
      for (int w = 0; w < nw; w++) e[w] &= tmp[w];
    }
  for (int w = 0; w < nw; w++) surf[w] = self[w] & ~e[w];
}

inline int next_set(const uint64_t* c, int nw, int y, int sy) {
  if (y >= sy) return sy;
  int w = y >> 6;
  uint64_t v = c[w] & (~0ull << (y & 63));
  while (!v) {
    if (++w >= nw) return sy;
    v = c[w];
  }
  return std::min(sy, (w << 6) + __builtin_ctzll(v));
}
inline int next_clear(const uint64_t* c, int nw, int y, int sy) {
  if (y >= sy) return sy;
  int w = y >> 6;
  uint64_t v = ~c[w] & (~0ull << (y & 63));
  while (!v) {
    if (++w >= nw) return sy;
    v = ~c[w];
  }
  return std::min(sy, (w << 6) + __builtin_ctzll(v));
}

// Elements of one run (RLE4::compress flush, Rle4.cpp:166-182).  `emit` gets
// each element word.
template <class F>
inline void run_elements(int skip, int solid, F&& emit) {
  while (skip > 1023) {
    emit((uint16_t)1023);
    skip -= 1023;
  }
  while (solid > 63) {
    emit((uint16_t)(63 * 1024 + (skip & 1023)));
    solid -= 63;
    skip = 0;
  }
  emit((uint16_t)((solid & 63) * 1024 + (skip & 1023)));
}

// Encodes one level.  colour(x, y, z) gives the colour of a stored voxel in
// this level's coordinates.
template <class ColourFn>
rlevox_status encode_level(const Level& L, ColourFn&& colour, HostLevel& out, int threads) {
  const int64_t ncols = (int64_t)L.sx * L.sz;
  std::vector<uint32_t> size(ncols);
  parallel_for(L.sz, threads, 4, [&](int64_t z) {
    std::vector<uint64_t> surf(L.nw), tmp(L.nw);
    for (int x = 0; x < L.sx; x++) {
      surface_column(L, x, (int)z, surf.data(), tmp.data());
      uint32_t nel = 0, ncol = 0;
      int prev = 0;
      for (int a = next_set(surf.data(), L.nw, 0, L.sy); a < L.sy;) {
        const int b = next_clear(surf.data(), L.nw, a, L.sy);
        run_elements(a - prev, b - a, [&](uint16_t) { nel++; });
        ncol += b - a;
        prev = b;
        a = next_set(surf.data(), L.nw, b, L.sy);
      }
      size[z * L.sx + x] = 2 + nel + ncol;
      if (nel > 65535 || ncol > 65535) size[z * L.sx + x] = 0xffffffffu;
    }
  });
  for (int64_t c = 0; c < ncols; c++)
    if (size[c] == 0xffffffffu)
      return fail(RLEVOX_ERR_INVALID_SCENE, "encode: column %lld has more than 65535 runs or colours",
                  (long long)c);
  // Column starts: contiguous (the reference's layout) while every offset
  // fits 32 bits, else aligned to the smallest 2^shift that makes them fit.
  std::vector<uint64_t> start(ncols + 1);
  int shift = 0;
  for (;; shift++) {
    if (shift > RLEVOX_MAX_OFFSET_SHIFT)
      return fail(RLEVOX_ERR_INVALID_SCENE, "encode: level too large for 32+%d-bit offsets",
                  RLEVOX_MAX_OFFSET_SHIFT);
    const uint64_t align = (1ull << shift) - 1;
    start[0] = 0;
    for (int64_t c = 0; c < ncols; c++) {
      start[c + 1] = start[c] + size[c];
      if (c + 1 < ncols) start[c + 1] = (start[c + 1] + align) & ~align;
    }
    if ((start[ncols - 1] >> shift) <= 0xffffffffull) break;
  }
  out.sx = L.sx;
  out.sy = L.sy;
  out.sz = L.sz;
  out.shift = shift;
  out.slabs.assign(start[ncols], 0);
  out.ptrmap.assign((size_t)2 * ncols, 0);
  parallel_for(L.sz, threads, 4, [&](int64_t z) {
    std::vector<uint64_t> surf(L.nw), tmp(L.nw);
    for (int x = 0; x < L.sx; x++) {
      const int64_t c = z * L.sx + x;
      surface_column(L, x, (int)z, surf.data(), tmp.data());
      uint16_t* p = out.slabs.data() + start[c];
      uint32_t nel = 0;
      int prev = 0;
      for (int a = next_set(surf.data(), L.nw, 0, L.sy); a < L.sy;) {
        const int b = next_clear(surf.data(), L.nw, a, L.sy);
        run_elements(a - prev, b - a, [&](uint16_t w) { p[2 + nel++] = w; });
        prev = b;
        a = next_set(surf.data(), L.nw, b, L.sy);
      }
      uint32_t ncol = 0;
      for (int a = next_set(surf.data(), L.nw, 0, L.sy); a < L.sy;) {
        const int b = next_clear(surf.data(), L.nw, a, L.sy);
        for (int y = a; y < b; y++) p[2 + nel + ncol++] = colour(x, y, (int)z);
        a = next_set(surf.data(), L.nw, b, L.sy);
      }
      p[0] = (uint16_t)nel;
      p[1] = (uint16_t)ncol;
    }
  });
  // Pointer map (Rle4.cpp:295-314).
  const uint64_t total = start[ncols];
  parallel_for(ncols, threads, 65536, [&](int64_t c) {
    const uint64_t ofs = start[c];
    const uint32_t first = (ofs + 2 < total) ? out.slabs[ofs + 2] : 0;
    out.ptrmap[2 * c] = (uint32_t)(ofs >> shift);
    out.ptrmap[2 * c + 1] = (uint32_t)out.slabs[ofs] + (first << 16);
  });
  return RLEVOX_OK;
}

// ------------------------------------------------------------ generator

// Periodic value noise: hashed lattice values, smoothstep interpolation,
// sampled at column centres; the lattice wraps with the map so the tiled
// world (Cuda_Render.h:441-442) has no seams.
struct Noise {
  uint32_t seed;
  int size;
  float lattice(uint32_t salt, int ix, int iz) const {
    const uint32_t h = mix32((uint32_t)ix * 0x9E3779B1u ^ mix32((uint32_t)iz * 0x85EBCA77u ^ salt ^ seed));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
  }
  float value(uint32_t salt, int x, int z, int cell) const {
    const int period = std::max(1, size / cell);
    const float fx = ((float)(x % cell) + 0.5f) / (float)cell;
    const float fz = ((float)(z % cell) + 0.5f) / (float)cell;
    const int ix = (x / cell) % period, iz = (z / cell) % period;
    const int jx = (ix + 1) % period, jz = (iz + 1) % period;
    const float sx = fx * fx * (3 - 2 * fx), sz = fz * fz * (3 - 2 * fz);
    const float a = lattice(salt, ix, iz), b = lattice(salt, jx, iz);
    const float c = lattice(salt, ix, jz), d = lattice(salt, jx, jz);
    const float ab = a + (b - a) * sx, cd = c + (d - c) * sx;
    return ab + (cd - ab) * sz;
  }
};

struct Terrain {
  rlevox_gen_params p;
  Noise n;
  int base_cell;

  explicit Terrain(const rlevox_gen_params& prm) : p(prm) {
    n.seed = mix32((uint32_t)p.seed ^ mix32((uint32_t)(p.seed >> 32) + 0x1234567u));
    n.size = p.size_xz;
    base_cell = std::max(1, p.size_xz / 4);
  }
  // Height in voxels above the bottom of the volume.
  float height(int x, int z) const {
    float sum = 0, norm = 0, amp = 1;
    int cell = base_cell;
    for (int o = 0; o < p.octaves && cell >= 1; o++, cell /= 2) {
      sum += amp * n.value(0x100u + o, x, z, cell);
      norm += amp;
      amp *= p.persistence;
    }
    float v = norm > 0 ? sum / norm : 0.5f;
    v = 0.5f + (v - 0.5f) * 1.8f;  // spread the bell-shaped sum over the range
    v = std::min(1.0f, std::max(0.0f, v));
    return p.hmin + (p.hmax - p.hmin) * v;
  }
  // Cave layers: count ~ geometric with the requested mean (smooth in space),
  // each a horizontal slot whose thickness follows its own noise field.
  int layers(int x, int z) const {
    if (p.cave_layers <= 0) return 0;
    float u = 0.65f * n.value(0x200u, x, z, 128) + 0.35f * n.value(0x201u, x, z, 32);
    u = std::min(0.9999f, std::max(0.0f, (u - 0.15f) / 0.7f));
    const double q = p.cave_layers / (1.0 + p.cave_layers);
    int L = (int)std::floor(std::log(1.0 - u) / std::log(q));
    return std::min(std::max(L, 0), p.cave_max_layers);
  }
  void column(int x, int z, uint64_t* c, int sy) const {
    const float h = height(x, z);
    int top = sy - (int)h;
    top = std::min(std::max(top, 0), sy);
    set_rows(c, top, sy);
    const int L = layers(x, z);
    const int depth = sy - top;
    if (L == 0 || depth < 8) return;
    const float spacing = (float)depth / (float)(L + 1);
    for (int i = 0; i < L; i++) {
      const float t = n.value(0x300u + i, x, z, 48);
      const float half = spacing * 0.42f * std::max(0.0f, (t - 0.2f) / 0.8f);
      if (half < 0.5f) continue;
      const float jitter = (n.value(0x380u + i, x, z, 96) - 0.5f) * spacing * 0.3f;
      const float centre = (float)top + spacing * (float)(i + 1) + jitter;
      int a = (int)std::ceil(centre - half), b = (int)std::floor(centre + half);
      a = std::max(a, top + 2);
      b = std::min(b, sy - 2);
      if (a < b) clear_rows(c, a, b);
    }
  }
  // Colour of a stored voxel at mip-0 coordinates: bits 8-9 a material band
  // by depth, low byte a shade that varies per voxel (the layout of
  // ny | bit1 | bit2, Rle4.cpp:139-161).
  uint16_t colour(int X, int Y, int Z) const {
    const uint32_t h = mix32((uint32_t)X * 73856093u ^ (uint32_t)Z * 19349663u ^
                             (uint32_t)Y * 83492791u ^ n.seed);
    const int sy = p.size_y;
    const uint32_t band = (uint32_t)std::min(3, (Y * 4) / std::max(1, sy));
    const uint32_t shade = 64u + (h & 63u) + (uint32_t)((Y * 127) / std::max(1, sy));
    return (uint16_t)((band << 8) | std::min(255u, shade));
  }
};

}  // namespace
}  // namespace rlevox

using namespace rlevox;

extern "C" rlevox_status rlevox_gen_scene(const rlevox_gen_params* params, rlevox_hscene** out) {
  clear_error();
  if (!params || !out) return fail(RLEVOX_ERR_INVALID_ARG, "gen_scene: null argument");
  *out = nullptr;
  const rlevox_gen_params& p = *params;
  if (p.size_xz < 4 || (p.size_xz & (p.size_xz - 1)) || p.size_xz > 32768 || p.size_y < 4 ||
      p.size_y > 4096 || p.nmips < 1 || p.nmips > RLEVOX_MAX_MIPS || (p.size_xz >> (p.nmips - 1)) < 1 ||
      (p.size_y >> (p.nmips - 1)) < 1 || p.octaves < 1 || !(p.hmax >= p.hmin) || p.hmin < 0 ||
      p.hmax > p.size_y || p.cave_max_layers < 0 || p.cave_max_layers > 1000)
    return fail(RLEVOX_ERR_INVALID_ARG, "gen_scene: parameters out of range");
  Terrain T(p);
  std::unique_ptr<rlevox_hscene> s(new rlevox_hscene());
  Level cur;
  cur.init(p.size_xz, p.size_y, p.size_xz);
  parallel_for(cur.sz, p.threads, 8, [&](int64_t z) {
    for (int x = 0; x < cur.sx; x++) T.column(x, (int)z, cur.col(x, (int)z), cur.sy);
  });
  for (int m = 0; m < p.nmips; m++) {
    if (m > 0) {
      Level next;
      reduce_level(cur, next, p.threads);
      cur = std::move(next);
    }
    const int sh = m;
    rlevox_status st = encode_level(
        cur, [&](int x, int y, int z) { return T.colour(x << sh, y << sh, z << sh); }, s->lv[m],
        p.threads);
    if (st != RLEVOX_OK) return st;
    s->nmips = m + 1;
  }
  *out = s.release();
  return RLEVOX_OK;
}

// Dense volume -> RLE4 mips.  Coarse colours come from the third set child
// in (x, y, z) sub-order, as the material bits do in Tree::get_mipmap.
extern "C" rlevox_status rlevox_encode_dense(int32_t sx, int32_t sy, int32_t sz, const uint8_t* solid,
                                             const uint16_t* colour, int32_t nmips,
                                             rlevox_hscene** out) {
  clear_error();
  if (!solid || !colour || !out || sx < 1 || sy < 1 || sz < 1 || nmips < 1 ||
      nmips > RLEVOX_MAX_MIPS || (sx & (sx - 1)) || (sz & (sz - 1)) || (sx >> (nmips - 1)) < 1 ||
      (sz >> (nmips - 1)) < 1 || (sy >> (nmips - 1)) < 1 || (int64_t)sx * sy * sz > (1ll << 31))
    return fail(RLEVOX_ERR_INVALID_ARG, "encode_dense: bad argument");
  *out = nullptr;
  Level cur;
  cur.init(sx, sy, sz);
  const size_t sxy = (size_t)sx * sy;
  for (int z = 0; z < sz; z++)
    for (int x = 0; x < sx; x++) {
      uint64_t* c = cur.col(x, z);
      for (int y = 0; y < sy; y++)
        if (solid[x + y * (size_t)sx + z * sxy]) c[y >> 6] |= 1ull << (y & 63);
    }
  std::vector<uint16_t> col(colour, colour + sxy * sz);
  std::unique_ptr<rlevox_hscene> s(new rlevox_hscene());
  for (int m = 0; m < nmips; m++) {
    if (m > 0) {
      Level next;
      reduce_level(cur, next, 0);
      std::vector<uint16_t> ncol((size_t)next.sx * next.sy * next.sz, 0);
      for (int k = 0; k < next.sz; k++)
        for (int i = 0; i < next.sx; i++) {
          const uint64_t* pc = next.col(i, k);
          for (int j = 0; j < next.sy; j++) {
            if (!test(pc, j)) continue;
            int seen = 0;
            for (int a = 0; a < 2 && seen < 3; a++)
              for (int b = 0; b < 2 && seen < 3; b++)
                for (int c = 0; c < 2 && seen < 3; c++)
                  if (test(cur.col(2 * i + a, 2 * k + c), 2 * j + b) && ++seen == 3)
                    ncol[i + j * (size_t)next.sx + k * (size_t)next.sx * next.sy] =
                        col[(2 * i + a) + (2 * j + b) * (size_t)cur.sx +
                            (2 * k + c) * (size_t)cur.sx * cur.sy];
          }
        }
      cur = std::move(next);
      col.swap(ncol);
    }
    const size_t lsx = cur.sx, lsxy = (size_t)cur.sx * cur.sy;
    rlevox_status st = encode_level(
        cur, [&](int x, int y, int z) { return col[x + y * lsx + z * lsxy]; }, s->lv[m], 0);
    if (st != RLEVOX_OK) return st;
    s->nmips = m + 1;
  }
  *out = s.release();
  return RLEVOX_OK;
}

// Re-packs levels with columns aligned to 2^offset_shift words.  Column
// contents are copied unchanged, so the decoded scene is the same.
extern "C" rlevox_status rlevox_repack(const rlevox_mip* mips, int32_t nmips, int32_t offset_shift,
                                       rlevox_hscene** out) {
  clear_error();
  if (!mips || !out || offset_shift < 0 || offset_shift > RLEVOX_MAX_OFFSET_SHIFT)
    return fail(RLEVOX_ERR_INVALID_ARG, "repack: bad argument");
  *out = nullptr;
  rlevox_status st = rlevox_scene_validate(mips, nmips);
  if (st != RLEVOX_OK) return st;
  std::unique_ptr<rlevox_hscene> s(new rlevox_hscene());
  for (int m = 0; m < nmips; m++) {
    const rlevox_mip& in = mips[m];
    const int64_t ncols = (int64_t)in.sx * in.sz;
    const uint64_t align = (1ull << offset_shift) - 1;
    std::vector<uint64_t> start(ncols + 1);
    start[0] = 0;
    for (int64_t c = 0; c < ncols; c++) {
      const uint16_t* p = in.slabs + ((uint64_t)in.ptrmap[2 * c] << in.offset_shift);
      start[c + 1] = start[c] + 2 + p[0] + p[1];
      if (c + 1 < ncols) start[c + 1] = (start[c + 1] + align) & ~align;
    }
    if (ncols > 0 && (start[ncols - 1] >> offset_shift) > 0xffffffffull)
      return fail(RLEVOX_ERR_INVALID_ARG, "repack: mip %d needs a larger offset_shift", m);
    HostLevel& l = s->lv[m];
    l.sx = in.sx;
    l.sy = in.sy;
    l.sz = in.sz;
    l.shift = offset_shift;
    l.slabs.assign(start[ncols], 0);
    l.ptrmap.assign((size_t)2 * ncols, 0);
    parallel_for(ncols, 0, 65536, [&](int64_t c) {
      const uint16_t* p = in.slabs + ((uint64_t)in.ptrmap[2 * c] << in.offset_shift);
      std::memcpy(l.slabs.data() + start[c], p, (size_t)(2 + p[0] + p[1]) * 2);
      l.ptrmap[2 * c] = (uint32_t)(start[c] >> offset_shift);
      l.ptrmap[2 * c + 1] = in.ptrmap[2 * c + 1];
    });
    s->nmips = m + 1;
  }
  *out = s.release();
  return RLEVOX_OK;
}
