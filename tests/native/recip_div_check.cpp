// Checks the gap-fill's integer quotients from a reciprocal estimate
// (csrc/recip_div.h) against integer division, with the estimate taken as
// the correctly rounded 1/x moved by -2..+2 ulps (the hardware reciprocal
// is at most 1 ulp off):
//   div_small: every num in [0, 8191] and n in [1, 26];
//   div_78720: every d in [1, 2^18] and 2 M seeded d up to 2^31 - 1.
// Built and run by tests/test_native_checks.py, also with the header's
// planted mistakes (RLEVOX_DIV_PLANT=1: no half-step bias; 2: no upward
// correction), which it must catch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#define RLEVOX_HD
#include "../../rle-based-voxel-raycasting_amd/csrc/recip_div.h"

struct Rcp {
  int ulps;
  float operator()(float x) const {
    float r = 1.0f / x;
    for (int i = 0; i < ulps; i++) r = std::nextafter(r, 1e30f);
    for (int i = 0; i < -ulps; i++) r = std::nextafter(r, 0.0f);
    return r;
  }
};

int main() {
  long bad_small = 0, bad_big = 0, cases = 0;
  for (int u = -2; u <= 2; u++) {
    const Rcp rcp{u};
    for (int n = 1; n <= 26; n++)
      for (int num = 0; num < 8192; num++) {
        cases++;
        if (rlevox::div_small(num, n, rcp) != num / n) {
          if (bad_small < 5) std::printf("div_small(%d, %d) ulps %d wrong\n", num, n, u);
          bad_small++;
        }
      }
    std::mt19937_64 rng(7);
    for (long i = 1; i <= (1 << 18) + 2000000; i++) {
      const int d = i <= (1 << 18) ? (int)i : (int)(1 + rng() % 2147483646u);
      cases++;
      if (rlevox::div_78720(d, rcp) != 78720 / d) {
        if (bad_big < 5) std::printf("div_78720(%d) ulps %d wrong\n", d, u);
        bad_big++;
      }
    }
  }
  std::printf("recip_div cases %ld, div_small mismatches %ld, div_78720 mismatches %ld\n", cases,
              bad_small, bad_big);
  return bad_small + bad_big ? 1 : 0;
}
