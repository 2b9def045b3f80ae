// The reception-sphere test of the capture (HitReceptionSphere, RTSolver.cpp:805-863, as RayCapture
// calls it, RTSolver.cpp:259-297) and a single-precision pre-test in front of it, host-callable so
// that the pre-test's soundness can be checked on the CPU.
//
// Exact test (FP64, the reference's operations): t0 = (T - S).(C - S) / |T - S|^2; no capture if
// t0 < 0; t0 clamped to 1; cp = S + (T - S) t0; radius = (base + |cp|) rc (model 0, |cp| from the
// origin) or (base + |(T - S) t0|) rc (otherwise); captured iff |C - cp| <= radius.
//
// Pre-test: the same quantities in single precision, relative to S, from inputs rounded once
// (C - S and T - S taken in FP64, then rounded).  With u = 2^-24, X = |C - S|_1, L = |T - S|:
//   * the computed t is within dt = 16 u (X / L + |t|) of t0 (dot product error 5.1 u L X, squared
//     length 5.1 u, reciprocal and product 3 u, one ulp reciprocals allowed);
//   * t < -dt proves t0 < 0 (no capture); t > dt proves t0 > 0; otherwise the exact test decides;
//   * the clamped parameter moves by at most dt, so |C - cp| and the radius move by at most
//     L dt and rc L dt; the roundings add at most 8 u (X + L) to the distance and
//     8 u rc (|S|_1 + L + base) to the radius (one-ulp square roots allowed);
//   * diff = distance - radius is decided when |diff| > M = 2 [(1 + rc) L dt + 16 u (X + L) +
//     16 u rc (|S|_1 + L + base)], a factor of two above the sum of those bounds, and far above the
//     FP64 test's own rounding (~1e-15 of the same magnitudes), so the FP64 decision is the same.
// Inputs whose magnitudes could overflow single precision squares (above 1e15) are left to the
// exact test, as are NaNs (every comparison with them fails).
#pragma once

#include <cmath>

#include "sbrsolver/sbr_math.h"

namespace sbr {

SBR_HD bool sphere_hit(V3 S, V3 T, V3 C, double base, int model, double rcoef) {
  V3 rayDir = vsub(T, S);
  double t0 = vdot(rayDir, vsub(C, S)) / vsqnorm(rayDir);
  if (t0 < 0) return false;
  if (t0 >= 1) t0 = 1;
  V3 cp = vadd(S, vmul(rayDir, t0));
  double unfolded = model == 0 ? base + vnorm(cp) : base + vnorm(vmul(rayDir, t0));
  double radius = unfolded * rcoef;
  double distance = vnorm(vsub(C, cp));
  return distance <= radius;
}

// Reciprocal and square root of the pre-test: the hardware's one-ulp instructions on the device.
// A host build may supply its own (SBR_SPHERE_HOST_OPS), e.g. to perturb results by an ulp.
#if defined(__HIP_DEVICE_COMPILE__)
SBR_HD float pre_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
SBR_HD float pre_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
#elif !defined(SBR_SPHERE_HOST_OPS)
SBR_HD float pre_rcp(float x) { return 1.0f / x; }
SBR_HD float pre_sqrt(float x) { return std::sqrt(x); }
#else
float pre_rcp(float x);
float pre_sqrt(float x);
#endif

// Per-segment part of the pre-test.
struct SphereF {
  float dx, dy, dz;   // T - S
  float l2inv;        // 1 / |T - S|^2
  float linv;         // 1 / |T - S|
  float sx, sy, sz;   // S
  float base;
  float mT, m0;       // margin = mX X + mT |t| + m0
};

// Receiver-independent constants of a transmitter's pre-tests.
struct SphereK {
  float rc;
  float mX;
  int model;
};

constexpr float kPreU = 5.9604645e-8f;   // 2^-24

SBR_HD SphereK sphere_k(double rc, int model) {
  SphereK k;
  k.rc = (float)rc;
  k.mX = 32.0f * kPreU * (2.0f + k.rc);
  k.model = model;
  if (!(rc >= 0.0 && rc < 1e6)) k.mX = NAN;
  return k;
}

SBR_HD SphereF sphere_f(V3 S, V3 T, double base, const SphereK& k) {
  SphereF f;
  f.dx = (float)(T.x - S.x);
  f.dy = (float)(T.y - S.y);
  f.dz = (float)(T.z - S.z);
  const float l2 = f.dx * f.dx + f.dy * f.dy + f.dz * f.dz;
  const float L = pre_sqrt(l2);
  f.l2inv = pre_rcp(l2);
  f.linv = pre_rcp(L);
  f.sx = (float)S.x;
  f.sy = (float)S.y;
  f.sz = (float)S.z;
  f.base = (float)base;
  const float s1 = fabsf(f.sx) + fabsf(f.sy) + fabsf(f.sz);
  f.mT = 32.0f * kPreU * (1.0f + k.rc) * L;
  f.m0 = 32.0f * kPreU * (L + k.rc * (s1 + L + f.base)) + 1e-30f;
  if (!(s1 < 1e15f && L < 1e15f && f.base >= 0.0f && f.base < 1e15f)) f.m0 = NAN;
  return f;
}

// -1: no capture, 1: capture (both proven), 0: undecided (the exact test decides).
SBR_HD int sphere_pre(const SphereF& f, const SphereK& k, V3 S, V3 C) {
  const float cx = (float)(C.x - S.x), cy = (float)(C.y - S.y), cz = (float)(C.z - S.z);
  const float X = fabsf(cx) + fabsf(cy) + fabsf(cz);
  if (!(X < 1e15f)) return 0;
  const float t = (f.dx * cx + f.dy * cy + f.dz * cz) * f.l2inv;
  const float at = fabsf(t);
  const float dt = 16.0f * kPreU * (X * f.linv + at);
  if (t < -dt) return -1;
  if (!(t > dt)) return 0;
  const float tc = fminf(t, 1.0f);
  const float px = f.dx * tc, py = f.dy * tc, pz = f.dz * tc;
  const float vx = cx - px, vy = cy - py, vz = cz - pz;
  const float dist = pre_sqrt(vx * vx + vy * vy + vz * vz);
  float nrm;
  if (k.model == 0) {
    const float ax = f.sx + px, ay = f.sy + py, az = f.sz + pz;
    nrm = pre_sqrt(ax * ax + ay * ay + az * az);
  } else {
    nrm = pre_sqrt(px * px + py * py + pz * pz);
  }
  const float diff = dist - (f.base + nrm) * k.rc;
  const float M = k.mX * X + f.mT * at + f.m0;
  if (diff > M) return -1;
  if (diff < -M) return 1;
  return 0;
}

}  // namespace sbr
