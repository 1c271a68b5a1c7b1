// Emulates the wavefront DDA chunk (64 lanes, lane k holding term k of each
// crossing sequence, steps ordered by merge ranks of the two distance
// sequences) and checks it against the sequential DDA of
// Cuda_Render.h:376-414, including the gradient doubling at mip switches and
// lines parallel to an axis.  Built and run by tests/test_native_checks.py.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>


struct St {
  float A, B, i0x, i0y, i1x, i1y, gA, gB, g0x, g0y, g1x, g1y;
  float dnow, px, py;
  int idx;
};
struct Step {
  float db, dn, pbx, pby;
  int ib;
};

static Step seq_step(St& s) {
  Step o;
  o.db = s.dnow; o.pbx = s.px; o.pby = s.py; o.ib = s.idx;
  if (s.B < s.A) {
    s.dnow = s.B; s.idx = 1; s.B += s.gB; s.px = s.i1x; s.py = s.i1y; s.i1x += s.g1x; s.i1y += s.g1y;
  } else {
    s.dnow = s.A; s.idx = 0; s.px = s.i0x; s.py = s.i0y; s.A += s.gA; s.i0x += s.g0x; s.i0y += s.g0y;
  }
  o.dn = s.dnow;
  return o;
}

static void chunk(St& s, int L, Step* out) {
  // the kernel's chunk: each sequence's terms 0..63 by its own additions
  float sA[64], sB[64], x0[64], y0[64], x1[64], y1[64];
  int slot[64];
  sA[0] = s.A; sB[0] = s.B; x0[0] = s.i0x; y0[0] = s.i0y; x1[0] = s.i1x; y1[0] = s.i1y;
  for (int k = 1; k < 64; k++) {
    sA[k] = sA[k - 1] + s.gA; sB[k] = sB[k - 1] + s.gB;
    x0[k] = x0[k - 1] + s.g0x; y0[k] = y0[k - 1] + s.g0y;
    x1[k] = x1[k - 1] + s.g1x; y1[k] = y1[k - 1] + s.g1y;
  }
  int nx = 0;
  for (int k = 0; k < 64; k++) {
    int lo = 0, hi = 64;
    while (lo < hi) { int m = (lo + hi) >> 1; if (sB[m] < sA[k]) lo = m + 1; else hi = m; }
    const int rx = k + lo;
    lo = 0; hi = 64;
    while (lo < hi) { int m = (lo + hi) >> 1; if (!(sB[k] < sA[m])) lo = m + 1; else hi = m; }
    const int rz = k + lo;
    if (rx < L) { slot[rx] = k; nx++; }
    if (rz < L) slot[rz] = 64 + k;
  }
  const int nz = L - nx;
  for (int t = 0; t < L; t++) {
    auto val = [&](int c, float& d, float& x, float& y, int& id) {
      if (c < 64) { d = sA[c]; x = x0[c]; y = y0[c]; id = 0; }
      else { d = sB[c - 64]; x = x1[c - 64]; y = y1[c - 64]; id = 1; }
    };
    float d, x, y; int id;
    val(slot[t], d, x, y, id);
    out[t].dn = d;
    if (t == 0) { out[t].db = s.dnow; out[t].pbx = s.px; out[t].pby = s.py; out[t].ib = s.idx; }
    else { float d2, x2, y2; int id2; val(slot[t - 1], d2, x2, y2, id2); out[t].db = d2; out[t].pbx = x2; out[t].pby = y2; out[t].ib = id2; }
  }
  { float d, x, y; int id;
    auto val = [&](int c) { if (c < 64) { d = sA[c]; x = x0[c]; y = y0[c]; id = 0; } else { d = sB[c - 64]; x = x1[c - 64]; y = y1[c - 64]; id = 1; } };
    val(slot[L - 1]); s.dnow = d; s.px = x; s.py = y; s.idx = id; }
  s.A = nx < 64 ? sA[nx] : sA[63] + s.gA;
  s.i0x = nx < 64 ? x0[nx] : x0[63] + s.g0x;
  s.i0y = nx < 64 ? y0[nx] : y0[63] + s.g0y;
  s.B = nz < 64 ? sB[nz] : sB[63] + s.gB;
  s.i1x = nz < 64 ? x1[nz] : x1[63] + s.g1x;
  s.i1y = nz < 64 ? y1[nz] : y1[63] + s.g1y;
}

static bool same(float a, float b) { return std::memcmp(&a, &b, 4) == 0; }

int main(int argc, char** argv) {
  const int lines = argc > 1 ? atoi(argv[1]) : 20000;
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long bad = 0, steps = 0;
  for (int ln = 0; ln < lines; ln++) {
    // a line's DDA setup as in Cuda_Render.h:270-304
    const double ang = U(rng) * 6.283185307179586;
    const float ddx = (float)std::cos(ang), ddy = (float)std::sin(ang);
    const float vpx = (float)(U(rng) * 20000.0), vpz = (float)(U(rng) * 20000.0);
    float fix_x = -1, fix_y = -1, sx = -1, sy = -1;
    float fx = vpx - (float)(int)vpx, fy = vpz - (float)(int)vpz;
    if (ddx >= 0) { fix_x = 0; sx = 1; fx = 1 - fx; }
    if (ddy >= 0) { fix_y = 0; sy = 1; fy = 1 - fy; }
    (void)fix_x; (void)fix_y;
    St s;
    s.g0y = ddy / std::fabs(ddx); s.g0x = sx; s.g1x = ddx / std::fabs(ddy); s.g1y = sy;
    s.i0x = s.g0x * fx; s.i0y = s.g0y * fx; s.i1x = s.g1x * fy; s.i1y = s.g1y * fy;
    s.gA = std::sqrt(s.g0x * s.g0x + s.g0y * s.g0y); s.gB = std::sqrt(s.g1x * s.g1x + s.g1y * s.g1y);
    s.A = std::sqrt(s.i0x * s.i0x + s.i0y * s.i0y); s.B = std::sqrt(s.i1x * s.i1x + s.i1y * s.i1y);
    s.dnow = 0; s.px = 0; s.py = 0; s.idx = 0;
    if (ln % 50 == 7) {  // exactly axis-parallel: one gradient is infinite
      s.g0y = 0.0f / 1.0f; s.i0y = 0; s.A = s.i0x >= 0 ? s.i0x : -s.i0x; s.gA = 1;
      s.g1x = INFINITY; s.i1x = INFINITY; s.B = INFINITY; s.gB = INFINITY; s.i1y = s.g1y * fy;
    }
    St q = s;
    const int total = 500 + (int)(U(rng) * 4000);
    int done = 0, mipstep = 600 + (int)(U(rng) * 700);
    while (done < total) {
      if (done >= mipstep) {  // gradients double at a mip switch
        for (St* p : {&s, &q}) { p->gA *= 2; p->gB *= 2; p->g0x *= 2; p->g0y *= 2; p->g1x *= 2; p->g1y *= 2; }
        mipstep += 300 + (int)(U(rng) * 700);
      }
      int L = 64;
      if (mipstep - done < L) L = mipstep - done;
      if (total - done < L) L = total - done;
      if ((int)(U(rng) * 10) == 0) L = 1 + (int)(U(rng) * 63);
      Step c[64];
      chunk(q, L, c);
      for (int t = 0; t < L; t++) {
        Step r = seq_step(s);
        steps++;
        if (!same(r.db, c[t].db) || !same(r.dn, c[t].dn) || !same(r.pbx, c[t].pbx) ||
            !same(r.pby, c[t].pby) || r.ib != c[t].ib) {
          if (bad < 5) std::printf("line %d step %d: seq (%a %a %a %a %d) chunk (%a %a %a %a %d)\n", ln, done + t,
                                   r.db, r.dn, r.pbx, r.pby, r.ib, c[t].db, c[t].dn, c[t].pbx, c[t].pby, c[t].ib);
          bad++;
        }
      }
      done += L;
    }
  }
  std::printf("lines %d steps %ld mismatches %ld\n", lines, steps, bad);
  return bad ? 1 : 0;
}
