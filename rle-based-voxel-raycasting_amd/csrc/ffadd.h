// Closed form of repeated binary32 addition.
//
// The DDA of a ray line accumulates constant increments: the next x/z
// crossing distances (dds_dist0 += grad_dist0, Cuda_Render.h:402,411) and
// the crossing positions (isect += grad, :404-405,412-413).  ffadd(s, g, k)
// returns s after k rounds of s = RN(s + g) (round to nearest, ties to even)
// without doing the k additions, so the lanes of a wavefront can each take a
// different step of the same line and still see exactly the values the
// sequential walk produces.
//
// Within one binade [2^e, 2^(e+1)) every value is a multiple of u = ulp and
// s + g rounds to s + d*u with d fixed by g's remainder modulo u (d = q or
// q+1; on an exact half-ulp remainder the first step rounds to the even
// neighbour and every later step keeps the multiple even).  The walk stays
// on that grid while the exact sum is below 2^(e+1); the step that crosses
// is done with an ordinary addition and the next binade starts over.
//
// Preconditions: s and g finite, g != 0, and s == 0 or sign(s) == sign(g).
#pragma once

#include <stdint.h>

#include <cmath>

#ifdef __HIPCC__
#define RLX_HD __host__ __device__
#else
#define RLX_HD
#endif

namespace rlevox {

RLX_HD inline uint32_t f_bits(float f) { return __builtin_bit_cast(uint32_t, f); }
RLX_HD inline float f_from(uint32_t u) { return __builtin_bit_cast(float, u); }

RLX_HD inline float ffadd(float s, float g, int k) {
  if (k <= 0) return s;
  const bool neg = g < 0.0f;
  float a = neg ? -s : s;  // RN is symmetric: work with non-negative values
  const float b = neg ? -g : g;
  while (k > 0 && !(a >= b)) {  // small a: plain steps until a >= b
    a = a + b;
    --k;
  }
  const uint32_t bb = f_bits(b);
  const int eb = (int)(bb >> 23);
  const uint32_t mb = eb ? ((bb & 0x7fffffu) | 0x800000u) : (bb & 0x7fffffu);
  const int ebx = eb ? eb : 1;
  while (k > 0) {
    const uint32_t ab = f_bits(a);
    const int ea = (int)(ab >> 23);
    if (ea == 0 || ea >= 254) {  // subnormal or near overflow: plain step
      a = a + b;
      --k;
      continue;
    }
    const uint32_t m = (ab & 0x7fffffu) | 0x800000u;  // a / u
    const int sh = ea - ebx;                           // >= 0 because a >= b
    uint32_t q;
    int cls;  // 0: remainder below half an ulp, 1: exactly half, 2: above half
    if (sh == 0) {
      q = mb;
      cls = 0;
    } else if (sh >= 25) {
      q = 0;
      cls = 0;  // b < u/2
    } else {
      q = mb >> sh;
      const uint32_t r = mb & ((1u << sh) - 1u), half = 1u << (sh - 1);
      cls = r < half ? 0 : (r == half ? 1 : 2);
    }
    if (m + q >= (1u << 24)) {  // this step leaves the binade
      a = a + b;
      --k;
      continue;
    }
    uint32_t d1, d;
    if (cls == 0) {
      d1 = d = q;
    } else if (cls == 2) {
      d1 = d = q + 1;
    } else {
      d1 = ((m + q) & 1u) ? q + 1 : q;
      d = (q & 1u) ? q + 1 : q;
    }
    // steps j >= 1 stay in the binade while m + d1 + (j-1)*d + q < 2^24.
    // All quantities stay below 2^26 (take is bounded by the binade), so
    // 32-bit arithmetic is exact.
    uint32_t take;
    if (d == 0) {
      take = (uint32_t)k;  // absorbed: nothing changes after the first step
    } else {
      const uint32_t used = q + m + d1;  // < 2^26
      const uint32_t lim = used < (1u << 24) ? (1u << 24) - used : 0u;
      const uint32_t n = 1u + (lim + d - 1u) / d;
      take = (uint32_t)k < n ? (uint32_t)k : n;
    }
    const uint32_t mn = m + d1 + (take - 1u) * d;
    a = (mn >= (1u << 24)) ? f_from((uint32_t)(ea + 1) << 23)
                           : f_from(((uint32_t)ea << 23) | (mn & 0x7fffffu));
    k -= (int)take;
  }
  return neg ? -a : a;
}

// ------------------------------------------------- closed-form DDA terms

// One binade of s <- RN(s + g) for |s| >= |g| > 0: terms k = 0..n stay on
// the grid of ulp 2^(e-150): mantissa m for k = 0, m + d1 + (k-1)*d after.
struct Binade {
  uint32_t e, m, d1, d, n;  // n = steps that stay in the binade, capped at 64
};

RLX_HD inline Binade binade_of(float a, float b) {  // 0 < b <= a, both normal
  const uint32_t ab = f_bits(a), bb = f_bits(b);
  Binade r;
  r.e = ab >> 23;
  r.m = (ab & 0x7fffffu) | 0x800000u;
  const uint32_t mb = (bb & 0x7fffffu) | 0x800000u;
  const int sh = (int)r.e - (int)(bb >> 23);
  uint32_t q;
  int cls;
  if (sh == 0) {
    q = mb;
    cls = 0;
  } else if (sh >= 25) {
    q = 0;
    cls = 0;
  } else {
    q = mb >> sh;
    const uint32_t rem = mb & ((1u << sh) - 1u), half = 1u << (sh - 1);
    cls = rem < half ? 0 : (rem == half ? 1 : 2);
  }
  if (r.m + q >= (1u << 24)) {  // the very next step leaves the binade
    r.d1 = r.d = r.n = 0;
    return r;
  }
  if (cls == 0) {
    r.d1 = r.d = q;
  } else if (cls == 2) {
    r.d1 = r.d = q + 1;
  } else {
    r.d1 = ((r.m + q) & 1u) ? q + 1 : q;
    r.d = (q & 1u) ? q + 1 : q;
  }
  if (r.d == 0) {
    r.n = 64;
    return r;
  }
  const uint32_t used = q + r.m + r.d1;
  const uint32_t lim = used < (1u << 24) ? (1u << 24) - used : 0u;
  if (lim >= 63u * r.d) {
    r.n = 64;
    return r;
  }
  // n = 1 + ceil(lim / d) < 64: estimate, then correct exactly
  uint32_t c = (uint32_t)((float)lim / (float)r.d);
  while (c * r.d < lim) c++;
  while (c > 0 && (c - 1) * r.d >= lim) c--;
  r.n = 1 + c;
  return r;
}

RLX_HD inline float binade_term(const Binade& b, uint32_t k) {  // 1 <= k <= b.n
  const uint32_t m = b.m + b.d1 + (k - 1) * b.d;
  return m >= (1u << 24) ? f_from((b.e + 1) << 23) : f_from((b.e << 23) | (m & 0x7fffffu));
}

// Terms 0..63 of s <- RN(s + g) in at most two binades (valid when
// |s| >= 64|g|: 64 increments then at most double |s|).
struct SeqChunk {
  float s0, a2;
  Binade b1, b2;
  uint32_t k2;  // first term of the second binade (64: none)
  bool neg, ok;
};

RLX_HD inline SeqChunk seq_chunk(float s, float g) {
  SeqChunk c;
  c.s0 = s;
  c.neg = g < 0.0f;
  const float a = std::fabs(s), b = std::fabs(g);
  const uint32_t ab = f_bits(a), bb = f_bits(b);
  c.ok = std::isfinite(s) && std::isfinite(g) && g != 0.0f && (s == 0.0f || ((s < 0.0f) == c.neg)) &&
         (bb >> 23) >= 1u && (ab >> 23) < 250u && a >= 64.0f * b;
  if (!c.ok) return c;
  c.b1 = binade_of(a, b);
  c.k2 = 64;
  if (c.b1.n < 63) {
    const float v = c.b1.n ? binade_term(c.b1, c.b1.n) : a;
    c.a2 = v + b;  // the step that crosses into the next binade
    c.k2 = c.b1.n + 1;
    c.b2 = binade_of(c.a2, b);
    if (c.k2 + c.b2.n < 63) c.ok = false;  // a second crossing: not covered
  }
  return c;
}

RLX_HD inline float seq_term(const SeqChunk& c, uint32_t k) {
  if (k == 0) return c.s0;
  float v;
  if (k <= c.b1.n) v = binade_term(c.b1, k);
  else if (k == c.k2) v = c.a2;
  else v = binade_term(c.b2, k - c.k2);
  return c.neg ? -v : v;
}

}  // namespace rlevox
