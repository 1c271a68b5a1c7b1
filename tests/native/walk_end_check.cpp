// Checks the early end of a line's walk (csrc/walk_end.h) against the
// oracle's full walk (Cuda_Render.h:96-736 restated in oracle/oracle.cpp):
// the oracle is included with hooks that evaluate WalkEnd before every DDA
// step, exactly as the kernels would at a chunk boundary, and record the
// first step at which it says the walk may end and the last step that drew
// or moved the line's window.  A step after the claimed end that still acts
// is a violation.  Seeded terrains of two sizes, random poses with eye
// heights from 1 to 2000 voxels above the surface, pitches from -1.3 to
// 1.45 rad and resolutions from 160x120 to 1280x720.  Built and run by
// tests/test_native_checks.py; WALK_END_PLANT=n lowers every run top by n
// rows in the bound, a mistake the check must catch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#define RLEVOX_HD
#include "../../rle-based-voxel-raycasting_amd/csrc/walk_end.h"

#ifndef WALK_END_PLANT
#define WALK_END_PLANT 0
#endif

struct Trace {
  rlevox::WalkEnd we;
  long end_at, last_act, steps;
};
static Trace* g_tr = nullptr;
static int g_top[16];
static int g_sy = 0;

#define ORC_HOOK_LINE_BEGIN                                                                  \
  Trace& tr_ = g_tr[x];                                                                      \
  tr_.we.init(cos_x, sin_x, res_x2_mul_reverse, ray_x, ray_z, vertical, vpy, res_y2);
#define ORC_HOOK_STEP                                                                        \
  if (tr_.end_at < 0 && y_clip_min < y_clip_max &&                                           \
      tr_.we.done(dds_dist_now, y_clip_min, g_top[mip_lvl] + WALK_END_PLANT, g_sy))          \
    tr_.end_at = (long)c_steps;
#define ORC_HOOK_ACT tr_.last_act = (long)c_steps;
#define ORC_HOOK_LINE_END tr_.steps = (long)c_steps;
#include "../../oracle/oracle.cpp"

// Highest run top (world row) of each level and all coarser ones.
static void level_tops(const std::vector<rlevox_mip>& mips) {
  int t = mips[0].sy;
  for (int m = (int)mips.size() - 1; m >= 0; m--) {
    const rlevox_mip& mp = mips[m];
    for (int64_t c = 0; c < (int64_t)mp.sx * mp.sz; c++) {
      const uint32_t ofs = mp.ptrmap[2 * c], lf = mp.ptrmap[2 * c + 1];
      const uint16_t* runs = mp.slabs + ((uint64_t)ofs << mp.offset_shift) + 2;
      uint32_t w = lf >> 16;
      int y = 0;
      for (uint32_t k = 0; k < (lf & 0xffffu); k++) {
        if (k > 0) w = runs[k];
        y += (int)(w & 1023u) << m;
        if (w >> 10) {
          if (y < t) t = y;
          break;
        }
      }
    }
    g_top[m] = t;
  }
  g_sy = mips[0].sy;
}

static int surface_row(const rlevox_mip& m0, int x, int z) {
  const int64_t c = (int64_t)(z & (m0.sz - 1)) * m0.sx + (x & (m0.sx - 1));
  const uint32_t ofs = m0.ptrmap[2 * c];
  const uint16_t* col = m0.slabs + ((uint64_t)ofs << m0.offset_shift);
  int y = 0;
  for (int e = 0; e < col[0]; e++) {
    const uint16_t w = col[2 + e];
    y += w & 1023;
    if (w >> 10) return y;
  }
  return m0.sy;
}

int main(int argc, char** argv) {
  const int views = argc > 1 ? std::atoi(argv[1]) : 40;
  long violations = 0, full = 0, ended = 0, lines_ended = 0;
  const int sizes[][2] = {{160, 120}, {320, 240}, {640, 480}, {512, 512}, {1280, 720}};
  for (int scene = 0; scene < 2; scene++) {
    rlevox_gen_params gp{};
    gp.size_xz = scene == 0 ? 512 : 64;
    gp.size_y = scene == 0 ? 256 : 64;
    gp.nmips = scene == 0 ? 4 : 3;
    gp.octaves = scene == 0 ? 6 : 4;
    gp.seed = 3 + scene;
    gp.persistence = 0.5f;
    gp.hmin = scene == 0 ? 32 : 8;
    gp.hmax = scene == 0 ? 224 : 48;
    gp.cave_layers = 1.0f;
    gp.cave_max_layers = scene == 0 ? 31 : 8;
    rlevox_hscene* hs = nullptr;
    if (rlevox_gen_scene(&gp, &hs) != RLEVOX_OK) {
      std::printf("generator failed: %s\n", rlevox_last_error());
      return 1;
    }
    std::vector<rlevox_mip> mips(rlevox_hscene_nmips(hs));
    for (int i = 0; i < (int)mips.size(); i++) rlevox_hscene_mip(hs, i, &mips[i]);
    level_tops(mips);
    std::mt19937 rng(11 + scene);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    for (int v = 0; v < views; v++) {
      const int W = sizes[v % 5][0], H = sizes[v % 5][1];
      const float x = u(rng) * gp.size_xz, z = u(rng) * gp.size_xz;
      const int kind = v % 4;
      const float alt = kind == 0 ? 1 + 19 * u(rng) : (kind == 1 ? 20 + 280 * u(rng) : 300 + 1700 * u(rng));
      float pitch = (v / 4) % 4 == 0 ? -1.3f + 1.29f * u(rng)
                    : (v / 4) % 4 == 1 ? 0.01f + 0.59f * u(rng)
                    : (v / 4) % 4 == 2 ? 0.6f + 0.85f * u(rng) : -0.1f + 0.2f * u(rng);
      if (std::fabs(pitch) < 0.01f) pitch = 0.01f;
      rlevox_pose p{};
      p.pos[0] = x + gp.size_xz;
      p.pos[1] = alt - (float)surface_row(mips[0], (int)x, (int)z);
      p.pos[2] = z + gp.size_xz;
      p.rot[0] = pitch;
      p.rot[1] = 6.2831853f * u(rng);
      rlevox_raymap rm;
      if (orc_raymap(&p, W, H, 0, 0, &rm) != 0) continue;
      const int lines = rm.map_line_count;
      if (lines <= 0) continue;
      std::vector<Trace> tr(lines);
      for (Trace& t : tr) t.end_at = t.last_act = t.steps = -1;
      g_tr = tr.data();
      std::vector<uint32_t> raybuf((size_t)lines * rm.rs);
      std::vector<int32_t> clip((size_t)lines * 2);
      rlevox_counters ctr;
      orc_raycast(mips.data(), (int)mips.size(), &rm, raybuf.data(), clip.data(), &ctr, 0);
      for (const Trace& t : tr) {
        if (t.end_at >= 0 && t.last_act > t.end_at) violations++;
        if (t.end_at >= 0) lines_ended++;
      }
      full += (long)ctr.dda_steps;
      for (const Trace& t : tr)  // steps of the walk that ends early
        if (t.steps >= 0) ended += t.end_at >= 0 && t.end_at < t.steps ? t.end_at : t.steps;
    }
    rlevox_hscene_free(hs);
  }
  std::printf("lines ended early %ld, steps walked %ld of %ld, violations %ld\n", lines_ended,
              ended, full, violations);
  return 0;
}
