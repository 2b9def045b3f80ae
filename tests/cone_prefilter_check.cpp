// Soundness of the correction's first-step pre-test (sbrsolver_amd/csrc/correction_prefilter.h)
// against the FP64 Moller-Trumbore test it stands in front of, on the CPU:
//   cone_prefilter_check <cases>
// Each case draws a triangle (sizes 1e-2 .. 1e3 m, any orientation, some slivers), an image source
// (far, near the triangle's plane, on an edge's line, exactly in the plane), a receiver box, and
// receivers in it: uniform, and placed just inside / outside the cone through the triangle's edges
// and vertices (relative offsets down to 1e-12), and at the image source itself (when the box
// holds it).
// Checks, for every receiver:
//   * rejected by cone_may_pass => tri_intersect(v1, e1, e2, I, C - I) fails;
//   * accepted => its x (and y) lies in cone_interval of a band holding it (the row ranges).
// Prints: cases receivers rejected passes_exact bad_reject bad_interval.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "correction_prefilter.h"

using namespace sbr;

static uint64_t g_state = 0x2545F4914F6CDD1Dull;
static uint64_t next_u64() {
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uni() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
static double uni(double a, double b) { return a + (b - a) * uni(); }
static double logu(double a, double b) { return std::exp(uni(std::log(a), std::log(b))); }
static V3 rand_dir() {
  for (;;) {
    V3 v = v3(uni(-1, 1), uni(-1, 1), uni(-1, 1));
    double n = vnorm(v);
    if (n > 0.1 && n <= 1.0) return v3(v.x / n, v.y / n, v.z / n);
  }
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 20000;
  long rx = 0, rej = 0, passes = 0, badRej = 0, badInt = 0;
  for (long i = 0; i < n; i++) {
    const double size = logu(1e-2, 1e3);
    const V3 v1 = v3(uni(-500, 500), uni(-500, 500), uni(0, 60));
    V3 e1 = vmul(rand_dir(), size * uni(0.05, 1.0));
    V3 e2 = vmul(rand_dir(), size * uni(0.05, 1.0));
    if (next_u64() % 8 == 0) e2 = vadd(vmul(e1, uni(-2, 2)), vmul(rand_dir(), size * 1e-4));   // sliver
    const V3 N = vcross(e1, e2);
    const V3 cen = vadd(v1, vmul(vadd(e1, e2), 1.0 / 3.0));
    V3 I;
    switch (next_u64() % 5) {
      case 0: I = vadd(cen, vmul(rand_dir(), logu(1.0, 5e3))); break;
      case 1: {   // near the plane
        const double nn = vnorm(N);
        I = vadd(vadd(cen, vmul(rand_dir(), logu(1.0, 1e3))), v3(0, 0, 0));
        const double d = vdot(vsub(I, v1), N) / nn;
        I = vsub(I, vmul(N, (d - uni(-1e-9, 1e-9) * size) / nn));
        break;
      }
      case 2: I = vadd(v1, vmul(e1, uni(-3, 3))); break;                       // on an edge's line
      case 3: I = vadd(v1, vadd(vmul(e1, uni(-1, 2)), vmul(e2, uni(-1, 2)))); break;   // in the plane
      default: I = vadd(v3(uni(-2e4, 2e4), uni(-2e4, 2e4), uni(-5e3, 5e3)), v3(0, 0, 0)); break;
    }
    // receiver box
    double bmin[3], bmax[3];
    const double h = logu(1.0, 600.0);
    const V3 bc = vadd(cen, vmul(rand_dir(), uni(0, 2 * h)));
    for (int a = 0; a < 3; a++) {
      bmin[a] = vget(bc, a) - h * uni(0.1, 1.0);
      bmax[a] = vget(bc, a) + h * uni(0.1, 1.0);
    }
    if (next_u64() % 4 == 0) bmax[2] = bmin[2];   // flat receiver grid
    if (next_u64() % 16 == 0) {   // a box that holds the image source
      for (int a = 0; a < 3; a++) {
        bmin[a] = std::fmin(bmin[a], vget(I, a));
        bmax[a] = std::fmax(bmax[a], vget(I, a));
      }
    }
    const ConeF c = cone_setup(I, v1, e1, e2, cone_reach(I, bmin, bmax));
    for (int k = 0; k < 64; k++) {
      V3 C;
      const int kind = (int)(next_u64() % 6);
      if (kind == 0) {
        C = v3(uni(bmin[0], bmax[0]), uni(bmin[1], bmax[1]), uni(bmin[2], bmax[2]));
      } else {
        // a point of the triangle's boundary (or a vertex), pushed off by a tiny relative amount,
        // then moved along the ray from I so that it falls in the box
        double u = uni(), v = uni();
        if (kind == 1) v = 0.0;
        if (kind == 2) u = 0.0;
        if (kind == 3) { u = uni(); v = 1.0 - u; }
        if (kind == 4) { u = (double)(next_u64() % 2); v = u > 0 ? 0.0 : (double)(next_u64() % 2); }
        if (u + v > 1.0) { u = 1.0 - u; v = 1.0 - v; }
        V3 P = vadd(v1, vadd(vmul(e1, u), vmul(e2, v)));
        P = vadd(P, vmul(rand_dir(), size * logu(1e-12, 1e-3) * (next_u64() % 2 ? 1.0 : 0.0)));
        const V3 d = vsub(P, I);
        const double t = uni(0.2, 3.0);
        C = vadd(I, vmul(d, t));
        for (int a = 0; a < 3; a++) {   // the box must hold the receiver: clamp (still a valid case)
          const double x = vget(C, a);
          const double cl = x < bmin[a] ? bmin[a] : (x > bmax[a] ? bmax[a] : x);
          if (a == 0) C.x = cl; else if (a == 1) C.y = cl; else C.z = cl;
        }
      }
      if (next_u64() % 200 == 0 && I.x >= bmin[0] && I.x <= bmax[0] && I.y >= bmin[1] && I.y <= bmax[1] &&
          I.z >= bmin[2] && I.z <= bmax[2])
        C = I;   // the receiver at the image source (dir = 0), when the box holds it
      rx++;
      V3 p;
      double dist;
      const bool exact = tri_intersect(v1, e1, e2, I, vsub(C, I), &p, &dist);
      passes += exact ? 1 : 0;
      const bool may = cone_may_pass(c, I, C);
      if (!may) {
        rej++;
        if (exact) {
          badRej++;
          if (badRej < 10) std::printf("bad reject: case %ld receiver %d kind %d\n", i, k, kind);
        }
        continue;
      }
      // accepted: inside the x interval of a band of rows holding it, and the y interval
      const double band[6] = {bmin[0], C.y - 1e-3 * h, bmin[2], bmax[0], C.y + 1e-3 * h, bmax[2]};
      double lo, hi;
      const bool okx = cone_interval(c, I, band, 0, &lo, &hi) && C.x >= lo && C.x <= hi;
      const double full[6] = {bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2]};
      double lo2, hi2;
      const bool oky = cone_interval(c, I, full, 1, &lo2, &hi2) && C.y >= lo2 && C.y <= hi2;
      if (!okx || !oky) {
        badInt++;
        if (badInt < 10) std::printf("bad interval: case %ld receiver %d kind %d\n", i, k, kind);
      }
    }
  }
  std::printf("%ld %ld %ld %ld %ld %ld\n", n, rx, rej, passes, badRej, badInt);
  return badRej == 0 && badInt == 0 ? 0 : 1;
}
