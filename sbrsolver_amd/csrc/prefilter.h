// Single-precision ray / box tests of the traversal, host-callable so that their soundness can be
// checked on the CPU.  Error budget: with E = the scene extent (max |coordinate| + 1), ray origins
// with |coordinates| <= 4E and directions whose largest component lies in [1e-3, 1e3], the computed
// slab endpoints of a box are within ~1.5e-6 E (in arc length) of those of the exact ray; boxes are
// padded by delta = 1e-3 + 4e-6 E at upload, which covers that with a margin of two.
#pragma once

#include <cmath>

#include "sbrsolver/sbr_math.h"

namespace sbr {

// Single-precision pre-test in front of each ray-triangle test: the triangle's box, rounded outward
// and padded by delta (set at upload), is tested against the ray in single precision.  For rays
// with |origin coordinates| <= 4E (E = scene extent) and a direction whose largest component lies
// in [1e-3, 1e3], every parameter t at which the ray meets the exact box lies inside the computed
// interval, so a triangle is skipped only when the exact test misses it (interval empty, or wholly
// at t < 0) or when every hit on it lies beyond the best hit so far (entry > t_best (1 + 1e-6)).
// Other rays take the exact test on every triangle.
struct RayF {
  float ix, iy, iz;     // 1 / direction (components clamped away from 0)
  float nx, ny, nz;     // -origin / direction
  bool on;
};

SBR_HD RayF rayf_pre(V3 o, V3 d, float ext) {
  RayF f;
  const double dm = fmax(fabs(d.x), fmax(fabs(d.y), fabs(d.z)));
  const double lim = 4.0 * (double)ext;
  f.on = fabs(o.x) <= lim && fabs(o.y) <= lim && fabs(o.z) <= lim && dm >= 1e-3 && dm <= 1e3;
  auto inv = [](double c) {
    float x = (float)c;
    if (fabsf(x) < 1e-20f) x = x < 0.0f ? -1e-20f : 1e-20f;
    return 1.0f / x;
  };
  f.ix = inv(d.x);
  f.iy = inv(d.y);
  f.iz = inv(d.z);
  f.nx = -(float)o.x * f.ix;
  f.ny = -(float)o.y * f.iy;
  f.nz = -(float)o.z * f.iz;
  return f;
}

// True if what lies behind the padded box [lo, hi] may still hold a hit at t <= clip; *entry is
// then a lower bound of the t of every such hit.  (Bounds given per coordinate.)
SBR_HD bool box32_may_hit(float lx, float ly, float lz, float hx, float hy, float hz, const RayF& f, float clip,
                          float* entry) {
  const float ax = fmaf(lx, f.ix, f.nx), bx = fmaf(hx, f.ix, f.nx);
  const float ay = fmaf(ly, f.iy, f.ny), by = fmaf(hy, f.iy, f.ny);
  const float az = fmaf(lz, f.iz, f.nz), bz = fmaf(hz, f.iz, f.nz);
  const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
  const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
  *entry = tn;
  return tn <= tf && tf >= 0.0f && tn <= clip;
}

SBR_HD bool box32_may_hit(const float* lo, const float* hi, const RayF& f, float clip, float* entry) {
  return box32_may_hit(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], f, clip, entry);
}

// Non-empty computed interval of the box [lo, hi] (no clipping).
SBR_HD bool box32_crossed(const float* lo, const float* hi, const RayF& f) {
  const float ax = fmaf(lo[0], f.ix, f.nx), bx = fmaf(hi[0], f.ix, f.nx);
  const float ay = fmaf(lo[1], f.iy, f.ny), by = fmaf(hi[1], f.iy, f.ny);
  const float az = fmaf(lo[2], f.iz, f.nz), bz = fmaf(hi[2], f.iz, f.nz);
  const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
  const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
  return tn <= tf;
}

// ---- boxes, built on the host at upload ----------------------------------------------------------
// Padding for a scene whose coordinates are bounded by ext - 1 in magnitude.
inline double prefilter_delta(double ext) { return 1e-3 + 4e-6 * ext; }

// The exact box [lo, hi] grown by g per side and rounded outward to single precision.
inline void box_grown(const double* lo, const double* hi, double g, float* flo, float* fhi) {
  for (int a = 0; a < 3; a++) {
    flo[a] = std::nextafter((float)(lo[a] - g), -INFINITY);
    fhi[a] = std::nextafter((float)(hi[a] + g), INFINITY);
  }
}

// The exact box shrunk by s per side and rounded inward; false if nothing is left.
inline bool box_shrunk(const double* lo, const double* hi, double s, float* flo, float* fhi) {
  bool ok = true;
  for (int a = 0; a < 3; a++) {
    flo[a] = std::nextafter((float)(lo[a] + s), INFINITY);
    fhi[a] = std::nextafter((float)(hi[a] - s), -INFINITY);
    if (!(flo[a] <= fhi[a])) ok = false;
  }
  return ok;
}

}  // namespace sbr
