// Tile queues of the batched post passes (compose, gap-fill).  Plain C++ so
// a host check can run the same code: the includer defines RLEVOX_HD as the
// function qualifiers it needs.
//
// A batch's pixel tiles run view by view, row by row.  They are dealt in
// stripes of kTileStripe consecutive tiles, stripe k to queue k % 8, and a
// queue hands them out in chunks of kTileChunk consecutive tiles.
// Workgroup b works queue b % 8 from the front; workgroups go round robin
// over the 8 XCDs, so a queue stays on one L2 and its tiles, whose borders
// overlap, are taken in order by workgroups that run at the same time.
// Once its own queue is empty, a workgroup takes chunks from the backs of
// the other queues.  A queue word counts the chunks taken from its front
// (low 32 bits) and its back (high 32 bits); every request adds one to
// either half, so the k-th request on a queue of n chunks sees k taken and
// holds a chunk exactly when k < n: the front requests take 0, 1, ... and
// the back requests n-1, n-2, ..., and every chunk goes once.  Each queue
// word has a 256-byte line of its own: requests to one address are served
// one at a time by the memory (a C4 batch's 710400 single-tile requests
// to one line took 4.5 ms).
#pragma once

#include <stdint.h>

namespace rlevox {

#ifndef RLEVOX_TILE_STRIPE
#define RLEVOX_TILE_STRIPE 1024
#endif
#ifndef RLEVOX_TILE_CHUNK
#define RLEVOX_TILE_CHUNK 4
#endif
constexpr int kTileStripe = RLEVOX_TILE_STRIPE;
constexpr int kTileChunk = RLEVOX_TILE_CHUNK;
constexpr int kTileQueues = 8;

// The queue words of one pass, one per 256-byte line.
struct TileQueueWords {
  struct alignas(256) Word {
    uint64_t v;
  } w[kTileQueues];
};

// Tiles in queue j: stripes j, j + 8, ... of `total` tiles.
RLEVOX_HD inline int64_t queue_len(int64_t total, int j) {
  const int64_t stripes = (total + kTileStripe - 1) / kTileStripe;
  if (j >= stripes) return 0;
  int64_t len = ((stripes - 1 - j) / kTileQueues + 1) * kTileStripe;
  if ((stripes - 1) % kTileQueues == j) len -= stripes * kTileStripe - total;  // partial last stripe
  return len;
}

// The batch tile at position p of queue j.
RLEVOX_HD inline int64_t queue_tile(int j, int64_t p) {
  return (j + kTileQueues * (p / kTileStripe)) * kTileStripe + p % kTileStripe;
}

// One workgroup's walk over the queues, run by one of its threads.  The home
// queue's next chunk is requested when a chunk starts (the request's result
// is read when the chunk is done); once the home queue is empty, each other
// queue's back is tried in turn until it is empty, so a walk makes at most
// (its chunks) + 9 requests and ends with every queue drained.
// Add(p, v) adds v to *p atomically and returns the old value.
template <class Add>
struct TileWalk {
  TileQueueWords* q;
  int64_t total, home_len;
  int64_t cur, end;  // tiles left of the current chunk: queue positions [cur, end)
  int home, from;    // from: 0 = home front, k = back of queue (home + k) % 8
  int cq;            // queue of the current chunk
  uint64_t pend;     // home-front request in flight

  RLEVOX_HD void start(TileQueueWords* q_, int64_t total_, int home_) {
    q = q_;
    total = total_;
    home = home_;
    home_len = queue_len(total, home);
    from = 0;
    cur = end = 0;
    cq = home;
    pend = Add()(&q->w[home].v, 1ull);
  }

  // The next tile, or -1 once every queue is empty.
  RLEVOX_HD int64_t take() {
    if (cur < end) return queue_tile(cq, cur++);
    if (from == 0) {
      const int64_t f = (int64_t)(pend & 0xffffffffu), b = (int64_t)(pend >> 32);
      if (f + b < (home_len + kTileChunk - 1) / kTileChunk) {
        pend = Add()(&q->w[home].v, 1ull);  // the next chunk, in flight
        cq = home;
        cur = f * kTileChunk;
        end = cur + kTileChunk < home_len ? cur + kTileChunk : home_len;
        return queue_tile(cq, cur++);
      }
      from = 1;
    }
    for (; from < kTileQueues; from++) {
      const int j = (home + from) & (kTileQueues - 1);
      const uint64_t old = Add()(&q->w[j].v, 1ull << 32);
      const int64_t f = (int64_t)(old & 0xffffffffu), b = (int64_t)(old >> 32);
      const int64_t len = queue_len(total, j), n = (len + kTileChunk - 1) / kTileChunk;
      if (f + b < n) {
        cq = j;
        cur = (n - 1 - b) * kTileChunk;
        end = cur + kTileChunk < len ? cur + kTileChunk : len;
        return queue_tile(cq, cur++);
      }
    }
    return -1;
  }
};

}  // namespace rlevox
