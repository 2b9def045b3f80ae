// FP64 vector helpers shared by the host scene builder and the gfx950 kernels.
//
// Every operation that feeds a discrete decision of the trace (hit / miss, closest triangle,
// reception-sphere capture, image-correction success) is written here once, in the evaluation
// order of the reference's Eigen 3.4 build (common/VecMatDef.h): dot and squaredNorm reduce as
// (a0*b0 + a1*b1) + a2*b2 (the SSE2 packet path), cross is the textbook component form,
// normalized() divides each component by sqrt(squaredNorm).  All translation units that include
// this header are compiled with -ffp-contract=off so that no multiply-add is fused: with IEEE
// correctly rounded + - * / sqrt on both the CPU and gfx950, identical operation sequences give
// bit-identical results.
#pragma once

#include <cfloat>
#include <cmath>

#if defined(__HIPCC__)
#define SBR_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define SBR_HD inline
#endif

namespace sbr {

struct V3 {
  double x, y, z;
};

SBR_HD V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
SBR_HD V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
SBR_HD V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
SBR_HD V3 vmul(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }   // vector * scalar
SBR_HD V3 smul(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }   // scalar * vector
SBR_HD V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
SBR_HD double vdot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
SBR_HD double vsqnorm(V3 a) { return vdot(a, a); }
SBR_HD double vnorm(V3 a) { return sqrt(vsqnorm(a)); }
SBR_HD V3 vnormalized(V3 a) {
  double z = vsqnorm(a);
  if (z > 0.0) {
    double s = sqrt(z);
    return v3(a.x / s, a.y / s, a.z / s);
  }
  return a;
}
SBR_HD V3 vcross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
SBR_HD double vget(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// std::min / std::max as the reference calls them: NaN handling depends on argument order.
SBR_HD double smin(double a, double b) { return (b < a) ? b : a; }
SBR_HD double smax(double a, double b) { return (a < b) ? b : a; }

// Reflect(i, n) = i - (2n)(n.i)   (VecMatDef.h:51-54)
SBR_HD V3 vreflect(V3 i, V3 n) { return vsub(i, vmul(smul(2.0, n), vdot(n, i))); }

// Single-precision machine epsilon, the reference's EPSILON (Constants.h:13).
#define SBR_EPS ((double)FLT_EPSILON)

// Moller-Trumbore as Triangle::RayIntersects (Triangle.cpp:80-125).  Returns true with the hit
// point and distance on a hit with t > EPSILON.
SBR_HD bool tri_intersect(V3 v1, V3 edge1, V3 edge2, V3 org, V3 dir, V3* point, double* dist,
                          double* tparam = nullptr) {
  V3 h = vcross(dir, edge2);
  double a = vdot(edge1, h);
  if (a > -SBR_EPS && a < SBR_EPS) return false;
  double f = 1.0 / a;
  V3 s = vsub(org, v1);
  double u = f * vdot(s, h);
  if (u < 0.0 || u > 1.0) return false;
  V3 q = vcross(s, edge1);
  double v = f * vdot(dir, q);
  if (v < 0.0 || u + v > 1.0) return false;
  double t = f * vdot(edge2, q);
  if (t > SBR_EPS) {
    V3 p = vadd(org, vmul(dir, t));
    *point = p;
    *dist = vnorm(vsub(p, org));
    if (tparam) *tparam = t;
    return true;
  }
  return false;
}

// Line / box slab test of BVHNode::IntersectsAABB (BVHNode.h:64-86): infinite line, division per
// axis, std::min/std::max NaN semantics.  Also returns the final [tmin, tmax] interval.
SBR_HD bool box_line(const double* bmin, const double* bmax, V3 o, V3 d, double* tmin_out, double* tmax_out) {
  double t_max = INFINITY, t_min = -INFINITY;
  double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double t1 = (bmin[i] - oo[i]) / dd[i];
    double t2 = (bmax[i] - oo[i]) / dd[i];
    double tn = smin(t1, t2), tf = smax(t1, t2);
    t_max = smin(t_max, tf);
    t_min = smax(t_min, tn);
    if (t_min > t_max) return false;
  }
  *tmin_out = t_min;
  *tmax_out = t_max;
  return true;
}

}  // namespace sbr
