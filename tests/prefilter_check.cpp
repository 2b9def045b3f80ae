// Soundness of the single-precision traversal decisions (sbrsolver_amd/csrc/prefilter.h) against
// the reference's FP64 line / box slab test (box_line, BVHNode.h:64-86) and the FP64
// Moller-Trumbore test (tri_intersect, Triangle.cpp:80-125), on rays built to graze: aimed at box
// corners, edges and faces and at triangle vertices and edges, in scenes of extent 1 to 5000 m,
// with flat and thin boxes and axis-parallel directions.  Checks, for every case:
//   grown box (2 delta) missed          => the FP64 slab test fails
//   shrunk box (delta) crossed          => the FP64 slab test passes
//   FP64 triangle hit at t              => pre-test box (delta) kept at clip t (1 + 1e-6)
// and, for a box A enclosing the box (faces moved outward by random amounts, some by none), as an
// ancestor's box in the reference tree encloses its descendants' boxes:
//   shrunk box (delta) crossed          => the FP64 slab test of A passes (the wide traversal's
//                                          check of a winner's leaf stops at such a node)
//   FP64 triangle hit at t              => grown box (2 delta) of the box and of A kept at clip
//                                          t (1 + 1e-6) (the wide traversal reaches every hit)
// Prints "<cases> <left to FP64> <robust passes> <robust misses> <triangle hits> <violations>
// <ancestor implications checked>" and exits 1 on a violation.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "prefilter.h"

using namespace sbr;

int main(int argc, char** argv) {
  const long N = argc > 1 ? atol(argv[1]) : 400000;
  std::mt19937_64 g(20261203);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  long cases = 0, borderline = 0, bad = 0, robustPass = 0, robustMiss = 0, triHits = 0, ancestors = 0;
  const double scales[4] = {1.0, 50.0, 500.0, 5000.0};
  for (long it = 0; it < N; it++) {
    const double E0 = scales[it % 4], ext = E0 + 1.0, delta = prefilter_delta(ext);
    // box: random, sometimes flat or thin in one axis
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
      const double c = U(g) * E0 * 0.9, w = std::fabs(U(g)) * E0 * 0.1 * std::pow(10.0, -3.0 * std::fabs(U(g)));
      lo[a] = c - w;
      hi[a] = c + w;
    }
    if (it % 5 == 1) hi[it % 3] = lo[it % 3];
    // target on the box surface: a corner, an edge or a face point, then nudged by a tiny amount
    double p[3];
    for (int a = 0; a < 3; a++) {
      const int mode = (int)((it / 7 + a) % 3);
      p[a] = mode == 0 ? lo[a] : mode == 1 ? hi[a] : lo[a] + (hi[a] - lo[a]) * std::fabs(U(g));
      p[a] += (it % 3 == 0 ? 0.0 : U(g) * delta * std::pow(10.0, -6.0 * std::fabs(U(g))));
    }
    V3 o = v3(U(g) * 4 * ext, U(g) * 4 * ext, U(g) * 4 * ext);
    V3 d = vsub(v3(p[0], p[1], p[2]), o);
    if (it % 11 == 0) d.x = 0.0;
    if (it % 13 == 0) { d.y = 0.0; o.y = p[1]; }
    const double dn = std::sqrt(vdot(d, d));
    if (!(dn > 0)) continue;
    if (it % 2 == 0) d = vnormalized(d);
    const RayF f = rayf_pre(o, d, (float)ext);
    if (!f.on) continue;
    cases++;
    double tlo, thi;
    const bool exact = box_line(lo, hi, o, d, &tlo, &thi);
    float glo[3], ghi[3], slo[3], shi[3];
    box_grown(lo, hi, 2.0 * delta, glo, ghi);
    const bool inner = box_shrunk(lo, hi, delta, slo, shi);
    float e;
    const bool grownCrossed = box32_crossed(glo, ghi, f);
    const bool shrunkCrossed = inner && box32_crossed(slo, shi, f);
    if (!grownCrossed && exact) { bad++; if (bad < 10) printf("grown miss but FP64 pass: it %ld E %g\n", it, E0); }
    if (shrunkCrossed && !exact) { bad++; if (bad < 10) printf("shrunk hit but FP64 miss: it %ld E %g\n", it, E0); }
    if (grownCrossed && !shrunkCrossed) borderline++;
    robustPass += shrunkCrossed;
    robustMiss += !grownCrossed;
    // enclosing box
    double alo[3], ahi[3];
    for (int a = 0; a < 3; a++) {
      const double w = (hi[a] - lo[a]) + delta;
      alo[a] = (it + a) % 4 == 0 ? lo[a] : lo[a] - std::fabs(U(g)) * w * std::pow(10.0, -4.0 * std::fabs(U(g)));
      ahi[a] = (it + a) % 4 == 1 ? hi[a] : hi[a] + std::fabs(U(g)) * w * std::pow(10.0, -4.0 * std::fabs(U(g)));
    }
    if (shrunkCrossed) {
      ancestors++;
      double alo2, ahi2;
      if (!box_line(alo, ahi, o, d, &alo2, &ahi2)) {
        bad++;
        if (bad < 10) printf("shrunk hit but enclosing box's FP64 miss: it %ld E %g\n", it, E0);
      }
    }
    // a triangle on the box: corner p, plus two points on the box
    V3 a0 = v3(p[0], p[1], p[2]);
    V3 a1 = v3(lo[0] + (hi[0] - lo[0]) * std::fabs(U(g)), lo[1], hi[2]);
    V3 a2 = v3(hi[0], lo[1] + (hi[1] - lo[1]) * std::fabs(U(g)), lo[2]);
    V3 q;
    double dist, t;
    if (tri_intersect(a0, vsub(a1, a0), vsub(a2, a0), o, d, &q, &dist, &t)) {
      triHits++;
      double tl[3], th[3];
      for (int k = 0; k < 3; k++) {
        tl[k] = std::fmin(std::fmin(vget(a0, k), vget(a1, k)), vget(a2, k));
        th[k] = std::fmax(std::fmax(vget(a0, k), vget(a1, k)), vget(a2, k));
      }
      float blo[3], bhi[3];
      box_grown(tl, th, delta, blo, bhi);
      const float clip = std::nextafter((float)(t * (1.0 + 1e-6)), INFINITY);
      if (!box32_may_hit(blo, bhi, f, clip, &e)) { bad++; if (bad < 10) printf("triangle hit pre-tested away: it %ld\n", it); }
      float nlo[3], nhi[3];
      box_grown(lo, hi, 2.0 * delta, nlo, nhi);
      const bool inBox = box32_may_hit(nlo, nhi, f, clip, &e);
      box_grown(alo, ahi, 2.0 * delta, nlo, nhi);
      const bool inAnc = box32_may_hit(nlo, nhi, f, clip, &e);
      if (!inBox || !inAnc) { bad++; if (bad < 10) printf("triangle hit but a wide node box missed: it %ld\n", it); }
    }
  }
  printf("%ld %ld %ld %ld %ld %ld %ld\n", cases, borderline, robustPass, robustMiss, triHits, bad, ancestors);
  return bad ? 1 : 0;
}
