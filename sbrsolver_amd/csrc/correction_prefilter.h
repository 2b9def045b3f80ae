// A necessary condition of the image correction's first step, in single precision, host-callable so
// that its soundness can be checked on the CPU.
//
// The correction of a path of depth d >= 1 to a receiver C starts with Triangle::RayIntersects
// (Triangle.cpp:80-125, called from MultiPathCorrection, RTSolver.cpp:299-394) from the image source
// I of the path's last interaction along dir = C - I against that interaction's surface (v1, e1, e2).
// With s = I - v1, q = s x e1 and T0 = e2 . q (the Moller-Trumbore quantities; T0 does not depend on
// the receiver), the test computes a = e1 . (dir x e2), u = (s . (dir x e2)) / a, v = (dir . q) / a,
// t = T0 / a and passes iff |a| >= EPSILON, 0 <= u <= 1, v >= 0, u + v <= 1 and t > EPSILON.  Hence:
//   * T0 == 0 (or NaN): t is never above EPSILON, no receiver passes;
//   * otherwise, with sigma = sign(T0), a pass needs sign(a) = sigma, and then, in exact arithmetic
//     on the test's own FP64 inputs, dir . n_k >= -g W_k |dir| for the four planes
//         n_1 = sigma (e2 x e1)            (a, since e1 . (dir x e2) = dir . (e2 x e1))   W_1 = |e1||e2|
//         n_2 = sigma (e2 x s)             (u a)                                          W_2 = |e2||s|
//         n_3 = sigma q                    (v a)                                          W_3 = |s||e1|
//         n_4 = sigma (e2 x e1 - e2 x s - q)   ((1 - u - v) a)                            W_4 = W_1 + W_2 + W_3
//     where g ~ 1e-14 covers the FP64 rounding of the test (a few ulps of each triple product, the
//     reciprocal and the products by it, the sum u + v).
// The pre-test evaluates dir . n_k in single precision: dir = C - I taken in FP64 and rounded, n_k
// rounded (the conversion back to FP64 is exact, so the row ranges below use the same vectors).
// The computed value is within 8 u W_k |dir| of the exact one (u = 2^-24: rounding of both operands
// and the three-term dot), so a value below -M_k with M_k = 1e-5 W_k D, D >= |C - I| for every
// receiver of the grid, proves dir . n_k < -9.5e-6 W_k |dir|: the FP64 test fails.  Magnitudes that
// could overflow single precision (above 1e15) or underflow it (W_k D below 1e-20), and non-finite
// inputs, switch the pre-test off for the node (every receiver is then left to the exact test).
#pragma once

#include <cmath>

#include "sbrsolver/sbr_math.h"

namespace sbr {

struct ConeF {
  float n[4][3];   // plane normals n_k
  float m[4];      // margins M_k
  int state;       // 1: pre-test on, 0: off (undecided for every receiver), -1: no receiver passes
};

// The pre-test of image source I against the surface (v1, e1, e2); D: an upper bound of |C - I|
// over the receivers (see cone_reach).
SBR_HD ConeF cone_setup(V3 I, V3 v1, V3 e1, V3 e2, double D) {
  ConeF c;
  for (int k = 0; k < 4; k++) {
    c.n[k][0] = c.n[k][1] = c.n[k][2] = 0.0f;
    c.m[k] = 0.0f;
  }
  c.state = 0;
  const V3 s = vsub(I, v1);   // the test's own s and q (same operations)
  const V3 q = vcross(s, e1);
  const double T0 = vdot(e2, q);
  const double mag = fmax(fmax(fmax(vnorm(s), vnorm(e1)), fmax(vnorm(e2), vnorm(q))), D);
  if (!(mag < 1e15)) return c;   // overflow risk or a non-finite input: off
  if (!(T0 != 0.0)) {            // t = T0 / a is never above EPSILON
    c.state = -1;
    return c;
  }
  const double sg = T0 > 0.0 ? 1.0 : -1.0;
  const V3 N = vcross(e2, e1), A = vcross(e2, s);
  const V3 nk[4] = {smul(sg, N), smul(sg, A), smul(sg, q), smul(sg, vsub(vsub(N, A), q))};
  const double w1 = vnorm(e1) * vnorm(e2), w2 = vnorm(e2) * vnorm(s), w3 = vnorm(s) * vnorm(e1);
  const double W[4] = {w1, w2, w3, w1 + w2 + w3};
  for (int k = 0; k < 4; k++) {
    const double mk = 1e-5 * W[k] * D * (1.0 + 1e-6);
    if (!(mk >= 1e-20 && mk < 1e30)) return c;   // off
    c.n[k][0] = (float)nk[k].x;
    c.n[k][1] = (float)nk[k].y;
    c.n[k][2] = (float)nk[k].z;
    c.m[k] = (float)mk;
  }
  c.state = 1;
  return c;
}

// Upper bound of |C - I| over the points of the box [bmin, bmax].
SBR_HD double cone_reach(V3 I, const double* bmin, const double* bmax) {
  double d2 = 0.0;
  const double ii[3] = {I.x, I.y, I.z};
  for (int a = 0; a < 3; a++) {
    const double e = fmax(fabs(bmin[a] - ii[a]), fabs(bmax[a] - ii[a]));
    d2 += e * e;
  }
  return sqrt(d2) * (1.0 + 1e-6) + 1e-30;
}

// False when the first correction step provably fails for a receiver at C.
SBR_HD bool cone_may_pass(const ConeF& c, V3 I, V3 C) {
  if (c.state != 1) return c.state == 0;
  const float dx = (float)(C.x - I.x), dy = (float)(C.y - I.y), dz = (float)(C.z - I.z);
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float p = c.n[k][0] * dx + c.n[k][1] * dy + c.n[k][2] * dz;
    ok = ok && !(p < -c.m[k]);
  }
  return ok;
}

// The interval [*lo, *hi] of coordinate `ax` (0: x, 1: y) outside which no point of the box
// band[] = {x0, y0, z0, x1, y1, z1} (the band's extent along `ax` ignored) passes cone_may_pass.
// Computed in FP64 on the same (rounded) normals with 1.2 M_k, which covers the pre-test's own
// rounding (at most 0.05 M_k, since |C - I| <= D) with room for the FP64 rounding here.  Returns
// false when no point of the band can pass.
SBR_HD bool cone_interval(const ConeF& c, V3 I, const double* band, int ax, double* lo, double* hi) {
  *lo = -INFINITY;
  *hi = INFINITY;
  if (c.state == 0) return true;
  if (c.state < 0) return false;
  const double ii[3] = {I.x, I.y, I.z};
  for (int k = 0; k < 4; k++) {
    // n . (X - I) >= -1.2 M needs n_ax (x_ax - I_ax) >= -1.2 M - G, G the largest value of the other
    // two terms over the band
    double G = 0.0;
    for (int b = 0; b < 3; b++) {
      if (b == ax) continue;
      const double nb = (double)c.n[k][b];
      G += fmax(nb * (band[b] - ii[b]), nb * (band[3 + b] - ii[b]));
    }
    const double na = (double)c.n[k][ax];
    const double rhs = -1.2 * (double)c.m[k] - G;
    if (na > 0.0) {
      *lo = fmax(*lo, ii[ax] + rhs / na);
    } else if (na < 0.0) {
      *hi = fmin(*hi, ii[ax] + rhs / na);
    } else if (rhs > 0.0) {
      return false;
    }
  }
  // FP64 rounding of the bounds themselves: a relative 1e-9 and an absolute 1e-9 of the band
  const double pad = 1e-9 * (fabs(*lo < -1e300 ? 0.0 : *lo) + fabs(*hi > 1e300 ? 0.0 : *hi) + fabs(band[ax]) +
                             fabs(band[3 + ax]) + fabs(ii[ax])) + 1e-12;
  *lo -= pad;
  *hi += pad;
  return *lo <= *hi;
}

}  // namespace sbr
