// Host-side utilities of librlevox: error reporting, a small thread pool
// helper, pointer-map rebuild, scene validation, host scenes and the .rle4
// file format.
#include <sched.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "internal.h"

namespace rlevox {

static thread_local std::string t_error;

rlevox_status fail(rlevox_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_error = buf;
  return st;
}

void clear_error() { t_error.clear(); }

// Host threads used when a call asks for 0: OMP_NUM_THREADS when set (the
// process's CPU share on a shared machine), else the CPUs this process may
// run on.
int default_threads() {
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int n = std::atoi(e);
    if (n > 0) return n;
  }
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) {
    const int n = CPU_COUNT(&set);
    if (n > 0) return n;
  }
  const int n = (int)std::thread::hardware_concurrency();
  return n > 0 ? n : 1;
}

void parallel_for(int64_t n, int threads, int64_t chunk, const std::function<void(int64_t)>& body) {
  if (n <= 0) return;
  if (threads <= 0) threads = default_threads();
  if (chunk <= 0) chunk = 1;
  const int64_t nchunks = (n + chunk - 1) / chunk;
  if (threads > nchunks) threads = (int)nchunks;
  if (threads <= 1) {
    for (int64_t i = 0; i < n; i++) body(i);
    return;
  }
  std::atomic<int64_t> next(0);
  auto worker = [&]() {
    for (;;) {
      const int64_t c = next.fetch_add(1);
      if (c >= nchunks) break;
      const int64_t lo = c * chunk, hi = std::min(n, lo + chunk);
      for (int64_t i = lo; i < hi; i++) body(i);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; t++) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
}

// Pointer-map rebuild of RLE4::load (Rle4.cpp:295-314): columns are walked x
// fastest; each entry is [offset >> shift, count | first<<16] where `first`
// is the word after the column header (the cached first element).  With a
// shift, each column starts at the next multiple of 2^shift words.
rlevox_status build_ptrmap(const uint16_t* slabs, uint64_t nslabs, int sx, int sz, int shift,
                           uint32_t* out) {
  uint64_t ofs = 0;
  const uint64_t ncols = (uint64_t)sx * (uint64_t)sz;
  const uint64_t align = (1ull << shift) - 1;
  for (uint64_t c = 0; c < ncols; c++) {
    if (ofs + 2 > nslabs)
      return fail(RLEVOX_ERR_INVALID_SCENE, "ptrmap: slabs end inside the header of column %llu",
                  (unsigned long long)c);
    if ((ofs >> shift) > 0xffffffffull)
      return fail(RLEVOX_ERR_INVALID_SCENE, "ptrmap: offset of column %llu exceeds 32+%d bits",
                  (unsigned long long)c, shift);
    const uint32_t count = slabs[ofs];
    const uint32_t first = (ofs + 2 < nslabs) ? slabs[ofs + 2] : 0;
    out[2 * c] = (uint32_t)(ofs >> shift);
    out[2 * c + 1] = count + (first << 16);
    ofs += (uint64_t)slabs[ofs] + slabs[ofs + 1] + 2;
    if (c + 1 < ncols) ofs = (ofs + align) & ~align;
  }
  if (ofs > nslabs)
    return fail(RLEVOX_ERR_INVALID_SCENE, "ptrmap: last column runs past the slab array");
  return RLEVOX_OK;
}

// Checks everything the raycast kernel relies on to stay inside its arrays:
// power-of-two grids halving per mip (heights halve too), every column inside
// the slab array, the cached count/first element equal to the slab's, as many
// colours as drawn voxels and no runs past the column height.
static rlevox_status validate(const rlevox_mip* mips, int nmips) {
  if (!mips || nmips < 1 || nmips > RLEVOX_MAX_MIPS)
    return fail(RLEVOX_ERR_INVALID_ARG, "scene: need 1..%d mips, got %d", RLEVOX_MAX_MIPS, nmips);
  const int sx0 = mips[0].sx, sz0 = mips[0].sz;
  if (sx0 < 1 || sz0 < 1 || (sx0 & (sx0 - 1)) || (sz0 & (sz0 - 1)) || sx0 > 32768 || sz0 > 32768)
    return fail(RLEVOX_ERR_INVALID_SCENE, "scene: mip 0 grid %dx%d must be powers of two <= 32768",
                sx0, sz0);
  for (int m = 0; m < nmips; m++) {
    const rlevox_mip& mp = mips[m];
    if (mp.sx != (sx0 >> m) || mp.sz != (sz0 >> m) || mp.sx < 1 || mp.sz < 1)
      return fail(RLEVOX_ERR_INVALID_SCENE, "scene: mip %d grid %dx%d, expected %dx%d", m, mp.sx,
                  mp.sz, sx0 >> m, sz0 >> m);
    if (mp.sy < 1 || (m > 0 && mp.sy != std::max(1, mips[0].sy >> m)))
      return fail(RLEVOX_ERR_INVALID_SCENE, "scene: mip %d height %d, expected %d", m, mp.sy,
                  std::max(1, mips[0].sy >> m));
    if (!mp.slabs || !mp.ptrmap)
      return fail(RLEVOX_ERR_INVALID_ARG, "scene: mip %d has a null array", m);
    if (mp.offset_shift < 0 || mp.offset_shift > RLEVOX_MAX_OFFSET_SHIFT)
      return fail(RLEVOX_ERR_INVALID_SCENE, "scene: mip %d offset_shift %d outside 0..%d", m,
                  mp.offset_shift, RLEVOX_MAX_OFFSET_SHIFT);
    const int64_t ncols = (int64_t)mp.sx * mp.sz;
    std::atomic<int64_t> bad(-1);
    std::atomic<int> why(0);
    parallel_for(mp.sz, 0, 8, [&](int64_t z) {
      if (bad.load(std::memory_order_relaxed) >= 0) return;
      for (int64_t x = 0; x < mp.sx; x++) {
        const int64_t c = z * mp.sx + x;
        const uint64_t ofs = (uint64_t)mp.ptrmap[2 * c] << mp.offset_shift;
        const uint32_t lf = mp.ptrmap[2 * c + 1];
        int reason = 0;
        if (ofs + 2 > mp.nslabs) reason = 1;
        else {
          const uint32_t count = mp.slabs[ofs], ncol = mp.slabs[ofs + 1];
          if ((lf & 0xffffu) != count) reason = 2;
          else if (ofs + 2 + count + ncol > mp.nslabs) reason = 3;
          else if (count > 0 && (lf >> 16) != mp.slabs[ofs + 2]) reason = 4;
          else {
            uint64_t drawn = 0, height = 0;
            for (uint32_t e = 0; e < count; e++) {
              const uint16_t w = e == 0 ? (uint16_t)(lf >> 16) : mp.slabs[ofs + 2 + e];
              drawn += w >> 10;
              height += (w >> 10) + (w & 1023u);
            }
            if (drawn != ncol) reason = 5;
            else if (height > (uint64_t)mp.sy) reason = 6;
          }
        }
        if (reason) {
          int64_t expect = -1;
          if (bad.compare_exchange_strong(expect, c)) why.store(reason);
          return;
        }
      }
    });
    if (bad.load() >= 0) {
      static const char* msg[] = {"", "header outside the slab array", "count differs from the pointer map",
                                  "runs or colours outside the slab array",
                                  "first element differs from the pointer map",
                                  "colour count differs from the drawn voxels",
                                  "runs extend past the column height"};
      return fail(RLEVOX_ERR_INVALID_SCENE, "scene: mip %d column %lld: %s", m,
                  (long long)bad.load(), msg[why.load()]);
    }
    (void)ncols;
  }
  return RLEVOX_OK;
}

}  // namespace rlevox

using namespace rlevox;

extern "C" {

const char* rlevox_last_error(void) { return t_error.c_str(); }
int32_t rlevox_abi_version(void) { return RLEVOX_ABI_VERSION; }

rlevox_status rlevox_build_ptrmap(const uint16_t* slabs, uint64_t nslabs, int32_t sx, int32_t sz,
                                  int32_t offset_shift, uint32_t* ptrmap_out) {
  clear_error();
  if (!slabs || !ptrmap_out || sx < 1 || sz < 1 || offset_shift < 0 ||
      offset_shift > RLEVOX_MAX_OFFSET_SHIFT)
    return fail(RLEVOX_ERR_INVALID_ARG, "build_ptrmap: bad argument");
  return build_ptrmap(slabs, nslabs, sx, sz, offset_shift, ptrmap_out);
}

rlevox_status rlevox_scene_validate(const rlevox_mip* mips, int32_t nmips) {
  clear_error();
  return validate(mips, nmips);
}

int32_t rlevox_hscene_nmips(const rlevox_hscene* s) { return s ? s->nmips : 0; }

rlevox_status rlevox_hscene_mip(const rlevox_hscene* s, int32_t m, rlevox_mip* out) {
  clear_error();
  if (!s || !out || m < 0 || m >= s->nmips)
    return fail(RLEVOX_ERR_INVALID_ARG, "hscene_mip: bad argument");
  const HostLevel& l = s->lv[m];
  out->sx = l.sx;
  out->sy = l.sy;
  out->sz = l.sz;
  out->offset_shift = l.shift;
  out->ptrmap = l.ptrmap.data();
  out->slabs = l.slabs.data();
  out->nslabs = l.slabs.size();
  return RLEVOX_OK;
}

void rlevox_hscene_free(rlevox_hscene* s) { delete s; }

// RLE4::save (Rle4.cpp:220-242): int32 nummaps, then per mip
// {int32 sx, sy, sz, slabs_size; uint16 slabs[slabs_size]}.
rlevox_status rlevox_rle4_save(const char* path, const rlevox_mip* mips, int32_t nmips) {
  clear_error();
  if (!path || !mips || nmips < 1 || nmips > RLEVOX_MAX_MIPS)
    return fail(RLEVOX_ERR_INVALID_ARG, "rle4_save: bad argument");
  for (int m = 0; m < nmips; m++)
    if (mips[m].offset_shift != 0)
      return fail(RLEVOX_ERR_INVALID_ARG, "rle4_save: mip %d has offset_shift %d; the format "
                  "stores contiguous columns only", m, mips[m].offset_shift);
  for (int m = 0; m < nmips; m++)
    if (mips[m].nslabs > 0x7fffffffull)
      return fail(RLEVOX_ERR_INVALID_ARG, "rle4_save: mip %d has %llu slab words; the format holds "
                  "at most 2^31-1", m, (unsigned long long)mips[m].nslabs);
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(RLEVOX_ERR_IO, "rle4_save: cannot open %s", path);
  bool ok = std::fwrite(&nmips, 4, 1, f) == 1;
  for (int m = 0; ok && m < nmips; m++) {
    const int32_t hdr[4] = {mips[m].sx, mips[m].sy, mips[m].sz, (int32_t)mips[m].nslabs};
    ok = std::fwrite(hdr, 4, 4, f) == 4 &&
         std::fwrite(mips[m].slabs, 2, mips[m].nslabs, f) == mips[m].nslabs;
  }
  if (std::fclose(f) != 0) ok = false;
  return ok ? RLEVOX_OK : fail(RLEVOX_ERR_IO, "rle4_save: write to %s failed", path);
}

// RLE4::load (Rle4.cpp:244-384) including the pointer-map rebuild.
rlevox_status rlevox_rle4_load(const char* path, rlevox_hscene** out) {
  clear_error();
  if (!path || !out) return fail(RLEVOX_ERR_INVALID_ARG, "rle4_load: bad argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(RLEVOX_ERR_IO, "rle4_load: cannot open %s", path);
  rlevox_hscene* s = new rlevox_hscene();
  int32_t n = 0;
  rlevox_status st = RLEVOX_OK;
  if (std::fread(&n, 4, 1, f) != 1 || n < 1 || n > RLEVOX_MAX_MIPS) {
    st = fail(RLEVOX_ERR_IO, "rle4_load: %s: bad mip count", path);
  }
  for (int m = 0; st == RLEVOX_OK && m < n; m++) {
    int32_t hdr[4];
    if (std::fread(hdr, 4, 4, f) != 4 || hdr[0] < 1 || hdr[1] < 1 || hdr[2] < 1 || hdr[3] < 2 ||
        (int64_t)hdr[0] * hdr[2] > (1ll << 30)) {
      st = fail(RLEVOX_ERR_IO, "rle4_load: %s: bad header of mip %d", path, m);
      break;
    }
    HostLevel& l = s->lv[m];
    l.sx = hdr[0];
    l.sy = hdr[1];
    l.sz = hdr[2];
    l.slabs.resize((size_t)hdr[3]);
    if (std::fread(l.slabs.data(), 2, l.slabs.size(), f) != l.slabs.size()) {
      st = fail(RLEVOX_ERR_IO, "rle4_load: %s: truncated mip %d", path, m);
      break;
    }
    l.ptrmap.resize((size_t)2 * l.sx * l.sz);
    st = build_ptrmap(l.slabs.data(), l.slabs.size(), l.sx, l.sz, 0, l.ptrmap.data());
    s->nmips = m + 1;
  }
  std::fclose(f);
  if (st != RLEVOX_OK) {
    delete s;
    return st;
  }
  *out = s;
  return RLEVOX_OK;
}

}  // extern "C"
