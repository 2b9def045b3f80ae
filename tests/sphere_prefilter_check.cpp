// Soundness of the single-precision reception-sphere pre-test (sbrsolver_amd/csrc/capture_prefilter.h)
// against the FP64 test, on the CPU:
//   sphere_prefilter_check <cases>
// Reciprocals and square roots are taken one ulp off at random (the device's instructions are
// within one ulp).  Cases: random segments and receivers over urban-sized boxes, long (1e10 m)
// and tiny segments, and receivers placed on the reception sphere's boundary (relative offsets
// down to 1e-9), near the segment's ends (t0 around 0 and 1) and on the segment itself.  Both
// models, rc over [1e-5, 0.5], base over [0, 5000].  Prints: cases decided_miss decided_hit
// undecided bad (a decided case whose FP64 result differs).
#define SBR_SPHERE_HOST_OPS 1
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "capture_prefilter.h"

using namespace sbr;

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uni() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
static double uni(double a, double b) { return a + (b - a) * uni(); }
static float jitter(float r) {
  switch (next_u64() % 3) {
    case 0: return std::nextafter(r, INFINITY);
    case 1: return std::nextafter(r, -INFINITY);
    default: return r;
  }
}
namespace sbr {
float pre_rcp(float x) { return jitter(1.0f / x); }
float pre_sqrt(float x) { return jitter(std::sqrt(x)); }
}  // namespace sbr

static V3 rand_point(double h) { return v3(uni(-h, h), uni(-h, h), uni(-h * 0.2, h * 0.4)); }
static V3 unit_perp(V3 d) {
  V3 a = std::fabs(d.x) < 0.6 * vnorm(d) ? v3(1, 0, 0) : v3(0, 1, 0);
  V3 p = vcross(d, a);
  const double n = vnorm(p);
  p = v3(p.x / n, p.y / n, p.z / n);
  V3 q = vcross(d, p);
  const double m = vnorm(q);
  q = v3(q.x / m, q.y / m, q.z / m);
  const double th = uni(0, 6.283185307179586);
  return v3(p.x * std::cos(th) + q.x * std::sin(th), p.y * std::cos(th) + q.y * std::sin(th),
            p.z * std::cos(th) + q.z * std::sin(th));
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  long miss = 0, hit = 0, undec = 0, bad = 0;
  for (long i = 0; i < n; i++) {
    const int model = (int)(next_u64() & 1);
    const double rc = std::exp(uni(std::log(1e-5), std::log(0.5)));
    const double base = (next_u64() % 8 == 0) ? 0.0 : uni(0, 5000);
    const int kind = (int)(next_u64() % 8);
    V3 S = rand_point(500), T;
    if (kind == 0) T = vadd(S, vmul(unit_perp(rand_point(1)), 1e10));          // long miss stub
    else if (kind == 1) T = vadd(S, vmul(unit_perp(rand_point(1)), uni(1e-6, 1e-3)));   // tiny
    else T = rand_point(500);
    const V3 D = vsub(T, S);
    V3 C;
    const int where = (int)(next_u64() % 6);
    if (where == 0) {
      C = rand_point(500);
    } else {
      // a point at parameter t off the segment by the radius there, times (1 + e)
      double t = uni(-0.3, 1.3);
      if (where == 2) t = uni(-1e-6, 1e-6);
      if (where == 3) t = 1.0 + uni(-1e-6, 1e-6);
      const double tc = t < 0 ? 0.0 : (t > 1 ? 1.0 : t);
      const V3 cp = vadd(S, vmul(D, tc));
      const double rad = (base + (model == 0 ? vnorm(cp) : vnorm(vmul(D, tc)))) * rc;
      const double e = (where == 4) ? uni(-1e-9, 1e-9) : std::exp(uni(std::log(1e-9), std::log(1e-1))) *
                                                            ((next_u64() & 1) ? 1.0 : -1.0);
      const double off = rad * (1.0 + e);
      C = vadd(vadd(S, vmul(D, t)), vmul(unit_perp(D), where == 5 ? 0.0 : off));
    }
    const SphereK k = sphere_k(rc, model);
    const SphereF f = sphere_f(S, T, base, k);
    const int p = sphere_pre(f, k, S, C);
    const bool ex = sphere_hit(S, T, C, base, model, rc);
    if (p == -1) miss++;
    else if (p == 1) hit++;
    else undec++;
    if ((p == -1 && ex) || (p == 1 && !ex)) {
      if (bad < 10)
        std::fprintf(stderr, "bad: pre %d exact %d model %d rc %.17g base %.17g S %.17g %.17g %.17g T %.17g %.17g %.17g C %.17g %.17g %.17g\n",
                     p, (int)ex, model, rc, base, S.x, S.y, S.z, T.x, T.y, T.z, C.x, C.y, C.z);
      bad++;
    }
  }
  std::printf("%ld %ld %ld %ld %ld\n", n, miss, hit, undec, bad);
  return 0;
}
