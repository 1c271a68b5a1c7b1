// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).
//
// Built with g++ -fsanitize=address,undefined together with the library's
// host sources (scene generator and encoder, validation, pointer-map
// rebuild, re-packing, .rle4 files, ray map) and the oracle, then driven
// through their C entry points on a small scene, on malformed scenes and on
// out-of-range poses.  Any sanitizer report aborts the run; the last line
// printed is "host sanitize ok" only when every check passed.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/rlevox.h"

extern "C" {
int orc_raymap(const rlevox_pose* pose, int W, int H, int mip_distance, int far_distance,
               rlevox_raymap* out);
void orc_raycast(const rlevox_mip* mips, int nmips, const rlevox_raymap* rm, uint32_t* raybuf,
                 int32_t* line_clip, rlevox_counters* ctr, int threads);
void orc_compose(const rlevox_raymap* rm, const uint32_t* raybuf, const int32_t* line_clip,
                 uint16_t* col, int32_t* dep);
void orc_gapfill(int W, int H, const uint16_t* cin, const int32_t* dep, uint16_t* cout);
}

static int failures = 0;
#define CHECK(cond)                                                     \
  do {                                                                  \
    if (!(cond)) {                                                      \
      std::printf("FAILED %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond, \
                  rlevox_last_error());                                 \
      failures++;                                                       \
    }                                                                   \
  } while (0)

static std::vector<rlevox_mip> mips_of(const rlevox_hscene* s) {
  std::vector<rlevox_mip> m(rlevox_hscene_nmips(s));
  for (int i = 0; i < (int)m.size(); i++) rlevox_hscene_mip(s, i, &m[i]);
  return m;
}

static void render_some(const std::vector<rlevox_mip>& mips, std::mt19937& rng, int frames) {
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  const int sizes[][2] = {{64, 48}, {160, 120}, {320, 240}};
  for (int f = 0; f < frames; f++) {
    rlevox_pose p;
    p.pos[0] = u(rng) * mips[0].sx;
    p.pos[1] = -5.0f - 40.0f * u(rng);
    p.pos[2] = u(rng) * mips[0].sz;
    p.rot[0] = (u(rng) < 0.5f ? -1.0f : 1.0f) * (0.05f + 0.6f * u(rng));
    p.rot[1] = 6.2831853f * u(rng);
    p.rot[2] = 0.0f;
    const int W = sizes[f % 3][0], H = sizes[f % 3][1];
    rlevox_raymap rm, orm;
    CHECK(rlevox_raymap_compute(&p, W, H, nullptr, &rm) == RLEVOX_OK);
    CHECK(orc_raymap(&p, W, H, 0, 0, &orm) == 0);
    CHECK(std::memcmp(&rm, &orm, sizeof(rm)) == 0);
    const int lines = rm.map_line_count > 0 ? rm.map_line_count : 1;
    std::vector<uint32_t> raybuf((size_t)lines * rm.rs, 0xdeadbeefu);
    std::vector<int32_t> clip((size_t)lines * 2);
    std::vector<uint16_t> col((size_t)W * H), fin((size_t)W * H);
    std::vector<int32_t> dep((size_t)W * H);
    rlevox_counters ctr;
    orc_raycast(mips.data(), (int)mips.size(), &rm, raybuf.data(), clip.data(), &ctr, 1);
    orc_compose(&rm, raybuf.data(), clip.data(), col.data(), dep.data());
    orc_gapfill(W, H, col.data(), dep.data(), fin.data());
  }
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  std::mt19937 rng(7);

  // generated terrain: validate, rebuild pointer maps, re-pack, file round trip
  rlevox_gen_params gp{};
  gp.size_xz = 64;
  gp.size_y = 128;
  gp.nmips = 3;
  gp.octaves = 5;
  gp.seed = 3;
  gp.persistence = 0.5f;
  gp.hmin = 20.0f;
  gp.hmax = 100.0f;
  gp.cave_layers = 2.0f;
  gp.cave_max_layers = 6;
  gp.threads = 1;
  rlevox_hscene* hs = nullptr;
  CHECK(rlevox_gen_scene(&gp, &hs) == RLEVOX_OK);
  if (!hs) return 1;
  std::vector<rlevox_mip> mips = mips_of(hs);
  CHECK(rlevox_scene_validate(mips.data(), (int)mips.size()) == RLEVOX_OK);
  for (const rlevox_mip& m : mips) {
    std::vector<uint32_t> pm((size_t)2 * m.sx * m.sz);
    CHECK(rlevox_build_ptrmap(m.slabs, m.nslabs, m.sx, m.sz, m.offset_shift, pm.data()) ==
          RLEVOX_OK);
    CHECK(std::memcmp(pm.data(), m.ptrmap, pm.size() * 4) == 0);
  }
  render_some(mips, rng, 12);

  rlevox_hscene* rp = nullptr;
  CHECK(rlevox_repack(mips.data(), (int)mips.size(), 3, &rp) == RLEVOX_OK);
  if (rp) {
    std::vector<rlevox_mip> rmips = mips_of(rp);
    CHECK(rlevox_scene_validate(rmips.data(), (int)rmips.size()) == RLEVOX_OK);
    render_some(rmips, rng, 6);
    rlevox_hscene_free(rp);
  }

  const std::string path = dir + "/sanitize_scene.rle4";
  CHECK(rlevox_rle4_save(path.c_str(), mips.data(), (int)mips.size()) == RLEVOX_OK);
  rlevox_hscene* ld = nullptr;
  CHECK(rlevox_rle4_load(path.c_str(), &ld) == RLEVOX_OK);
  if (ld) {
    std::vector<rlevox_mip> lm = mips_of(ld);
    CHECK(lm.size() == mips.size());
    for (size_t i = 0; i < lm.size() && i < mips.size(); i++) {
      CHECK(lm[i].nslabs == mips[i].nslabs);
      CHECK(std::memcmp(lm[i].slabs, mips[i].slabs, mips[i].nslabs * 2) == 0);
    }
    rlevox_hscene_free(ld);
  }
  // truncated and garbage files are rejected
  {
    FILE* f = std::fopen(path.c_str(), "r+b");
    if (f) {
      const int32_t bad[4] = {3, 1 << 30, -5, 7};
      std::fseek(f, 4, SEEK_SET);
      std::fwrite(bad, sizeof(bad), 1, f);
      std::fclose(f);
    }
    rlevox_hscene* x = nullptr;
    CHECK(rlevox_rle4_load(path.c_str(), &x) != RLEVOX_OK);
    if (x) rlevox_hscene_free(x);
    const std::string missing = dir + "/does_not_exist.rle4";
    CHECK(rlevox_rle4_load(missing.c_str(), &x) != RLEVOX_OK);
  }

  // malformed scenes: corrupted counts, offsets and sizes are rejected
  {
    const rlevox_mip& m0 = mips[0];
    std::vector<uint16_t> slabs(m0.slabs, m0.slabs + m0.nslabs);
    std::vector<uint32_t> ptr(m0.ptrmap, m0.ptrmap + (size_t)2 * m0.sx * m0.sz);
    std::uniform_int_distribution<int> col((int)0, m0.sx * m0.sz - 1);
    for (int trial = 0; trial < 200; trial++) {
      std::vector<uint16_t> s2 = slabs;
      std::vector<uint32_t> p2 = ptr;
      const int c = col(rng);
      switch (trial % 5) {
        case 0: p2[2 * c] = (uint32_t)m0.nslabs + 5; break;           // offset past the end
        case 1: s2[p2[2 * c]] = 0xffff; break;                        // element count
        case 2: s2[p2[2 * c] + 1] = 0xffff; break;                    // colour count
        case 3: p2[2 * c + 1] ^= 0x00010000u; break;                  // cached first element
        case 4: p2[2 * c + 1] = (p2[2 * c + 1] & 0xffff0000u) | 0xfff0u; break;  // cached count
      }
      rlevox_mip bad = m0;
      bad.slabs = s2.data();
      bad.ptrmap = p2.data();
      std::vector<rlevox_mip> bm = mips;
      bm[0] = bad;
      CHECK(rlevox_scene_validate(bm.data(), (int)bm.size()) != RLEVOX_OK);
    }
    rlevox_mip odd = m0;
    odd.sx = 48;  // not a power of two
    CHECK(rlevox_scene_validate(&odd, 1) != RLEVOX_OK);
    odd = m0;
    odd.nslabs = 3;
    CHECK(rlevox_scene_validate(&odd, 1) != RLEVOX_OK);
    CHECK(rlevox_scene_validate(mips.data(), 0) != RLEVOX_OK);
  }

  // dense encoder on a random volume
  {
    const int sx = 16, sy = 32, sz = 16;
    std::vector<uint8_t> solid((size_t)sx * sy * sz);
    std::vector<uint16_t> colour(solid.size());
    std::uniform_int_distribution<int> bit(0, 3);
    for (size_t i = 0; i < solid.size(); i++) {
      solid[i] = bit(rng) != 0;
      colour[i] = (uint16_t)(i * 2654435761u >> 16);
    }
    rlevox_hscene* es = nullptr;
    CHECK(rlevox_encode_dense(sx, sy, sz, solid.data(), colour.data(), 2, &es) == RLEVOX_OK);
    if (es) {
      std::vector<rlevox_mip> em = mips_of(es);
      CHECK(rlevox_scene_validate(em.data(), (int)em.size()) == RLEVOX_OK);
      render_some(em, rng, 3);
      rlevox_hscene_free(es);
    }
    CHECK(rlevox_encode_dense(15, sy, sz, solid.data(), colour.data(), 2, &es) != RLEVOX_OK);
  }

  // ray maps: out-of-range poses and sizes are rejected, never crash
  {
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    int rejected = 0;
    for (int i = 0; i < 2000; i++) {
      rlevox_pose p;
      for (int k = 0; k < 3; k++) p.pos[k] = 1e6f * u(rng) * u(rng) * u(rng);
      p.rot[0] = 3.2f * u(rng);
      p.rot[1] = 20.0f * u(rng);
      p.rot[2] = (i % 7 == 0) ? 0.1f : 0.0f;
      const int W = 16 + (i * 37) % 3000, H = 16 + (i * 53) % 2000;
      rlevox_raymap rm;
      if (rlevox_raymap_compute(&p, W, H, nullptr, &rm) != RLEVOX_OK) rejected++;
    }
    CHECK(rejected > 0);
    rlevox_pose p{{10, -10, 10}, {0.3f, 0.1f, 0}};
    rlevox_raymap rm;
    CHECK(rlevox_raymap_compute(&p, 15, 15, nullptr, &rm) != RLEVOX_OK);
    CHECK(rlevox_raymap_compute(&p, 100, 200, nullptr, &rm) != RLEVOX_OK);
    CHECK(rlevox_raymap_compute(nullptr, 100, 50, nullptr, &rm) != RLEVOX_OK);
  }

  rlevox_hscene_free(hs);
  if (failures) {
    std::printf("%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("host sanitize ok\n");
  return 0;
}
