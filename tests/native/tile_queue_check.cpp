// Checks the tile queues of the batched post passes (csrc/tile_queue.h): for
// batches of 1 to ~10^6 tiles and 1 to 2048 workgroups (home queue b % 8),
// with the stripe and chunk sizes the build defines,
// every workgroup's walk is run to its end under seeded random interleavings
// of the requests (each request is one atomic add on the device, so any
// serialisation that keeps each walk's own order is a possible run).  Every
// tile must be handed out exactly once, no tile outside the batch, and each
// walk must end.  Also: the queue lengths sum to the batch and queue_tile
// maps the queues one to one onto the tiles.  Built and run by
// tests/test_native_checks.py; TILE_QUEUE_PLANT=1 makes back requests see
// one more taken than there were, a mistake the check must catch.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#define RLEVOX_HD
#include "../../rle-based-voxel-raycasting_amd/csrc/tile_queue.h"

#ifndef TILE_QUEUE_PLANT
#define TILE_QUEUE_PLANT 0
#endif

using namespace rlevox;

struct HostAdd {
  uint64_t operator()(uint64_t* p, uint64_t v) const {
    const uint64_t old = *p;
#if TILE_QUEUE_PLANT == 1
    // back requests see one more taken than there were
    *p += v;
    return (v >> 32) ? old + (1ull << 32) : old;
#else
    *p += v;
    return old;
#endif
  }
};

static long g_fail = 0;

static void fail(const char* what, long total, int groups, long a, long b) {
  if (g_fail++ < 10) printf("FAIL %s: total %ld groups %d (%ld, %ld)\n", what, total, groups, a, b);
}

static void check_layout(int64_t total) {
  std::vector<uint8_t> seen((size_t)total, 0);
  int64_t sum = 0;
  for (int j = 0; j < kTileQueues; j++) {
    const int64_t len = queue_len(total, j);
    sum += len;
    for (int64_t p = 0; p < len; p++) {
      const int64_t t = queue_tile(j, p);
      if (t < 0 || t >= total) { fail("queue_tile out of range", total, 0, j, p); return; }
      if (seen[(size_t)t]++) fail("queue_tile twice", total, 0, j, p);
    }
  }
  if (sum != total) fail("queue lengths", total, 0, sum, total);
}

static void check_walks(int64_t total, int groups, uint64_t seed) {
  std::mt19937_64 rng(seed);
  TileQueueWords q{};
  std::vector<TileWalk<HostAdd>> w((size_t)groups);
  std::vector<int> live;
  std::vector<int64_t> requests((size_t)groups, 0);
  for (int b = 0; b < groups; b++) {
    w[(size_t)b].start(&q, total, b % kTileQueues);
    live.push_back(b);
  }
  std::vector<uint32_t> seen((size_t)total, 0);
  // a few walks run ahead or lag: weights skew which walk moves next
  std::vector<double> weight((size_t)groups);
  for (auto& x : weight) x = 0.2 + std::uniform_real_distribution<double>(0, 1)(rng);
  const int64_t cap = total + 16;  // a walk makes at most len + 9 requests
  while (!live.empty()) {
    size_t i = rng() % live.size();
    if (std::uniform_real_distribution<double>(0, 1.2)(rng) > weight[(size_t)live[i]]) continue;
    const int b = live[i];
    const int64_t t = w[(size_t)b].take();
    if (++requests[(size_t)b] > cap) {
      fail("walk does not end", total, groups, b, requests[(size_t)b]);
      return;
    }
    if (t < 0) {
      live[i] = live.back();
      live.pop_back();
      continue;
    }
    if (t >= total) { fail("tile out of range", total, groups, b, t); return; }
    seen[(size_t)t]++;
  }
  for (int64_t t = 0; t < total; t++)
    if (seen[(size_t)t] != 1) { fail("tile taken != once", total, groups, t, seen[(size_t)t]); return; }
}

int main() {
  const int64_t totals[] = {1, 2, 7, kTileChunk + 1, kTileStripe - 1, kTileStripe, kTileStripe + 1,
                            8 * kTileStripe - 3, 8 * kTileStripe, 9 * kTileStripe + 5,
                            17 * kTileStripe + 1, 123457, 710400};
  const int groups[] = {1, 3, 8, 9, 64, 2048};
  long runs = 0;
  for (int64_t total : totals) {
    check_layout(total);
    for (int g : groups) {
      if (total > 200000 && g < 64) continue;  // long single walks add nothing
      for (uint64_t seed = 1; seed <= 3; seed++) {
        check_walks(total, g, seed * 7919 + (uint64_t)total + (uint64_t)g);
        runs++;
      }
    }
  }
  printf("tile queue walks %ld, stripe %d, chunk %d, failures %ld\n", runs, kTileStripe, kTileChunk, g_fail);
  return g_fail ? 1 : 0;
}
