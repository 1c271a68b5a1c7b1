// Integer quotients of the gap-fill from a reciprocal estimate, in plain
// C++ so a host check can run the same code with any reciprocal: the
// includer defines RLEVOX_HD as the function qualifiers it needs, and the
// kernels pass the hardware reciprocal (v_rcp_f32, at most 1 ulp off).
// Both functions are exact for any reciprocal within 2^-22 of 1/x
// relative, which covers 2 ulps.
#pragma once

// Nonzero only in the host check's planted builds: 1 drops div_small's
// half-step bias, 2 drops div_78720's upward correction.  The check must
// report mismatches for both.
#ifndef RLEVOX_DIV_PLANT
#define RLEVOX_DIV_PLANT 0
#endif

namespace rlevox {

// floor(num / n) for 0 <= num < 2^13 and 1 <= n <= 26: (num + 1/2) / n
// lies at least 1/(2n) >= 1/52 from an integer, and the product with the
// estimate is within 2^13 * 2^-21 = 1/256 of it.
template <class Rcp>
RLEVOX_HD inline int div_small(int num, int n, Rcp rcp) {
  const float bias = RLEVOX_DIV_PLANT == 1 ? 0.0f : 0.5f;
  return (int)(((float)num + bias) * rcp((float)n));
}

// floor(78720 / d) for 1 <= d < 2^31: the estimate's quotient is within
// 0.02 of 78720 / d (a value below 78721), so its truncation is off by at
// most one either way, and the integer products settle it (they stay
// below 2^18 for d <= 78720; above, the estimate is 0 and stays 0).
template <class Rcp>
RLEVOX_HD inline int div_78720(int d, Rcp rcp) {
  int r = (int)(78720.0f * rcp((float)d));
  if (RLEVOX_DIV_PLANT != 2) r += (r + 1) * d <= 78720 ? 1 : 0;
  r -= r * d > 78720 ? 1 : 0;
  return r;
}

}  // namespace rlevox
