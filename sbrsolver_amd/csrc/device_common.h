// Device helpers of the gfx950 engine: BVH closest hit, reception sphere, image correction,
// complex arithmetic and the per-path field.  Compiled with -ffp-contract=off (see sbr_math.h).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "engine.h"

namespace sbr {

__device__ __forceinline__ V3 ld3(const double* p) { return v3(p[0], p[1], p[2]); }

// ---------------------------------------------------------------------------------------------
// Closest hit over the BVH: BVH<Triangle>::RayIntersects (BVH/BVH.h:276-328).
//
// The reference visits, in DFS order, every node whose box the infinite line crosses
// (BVHNode.h:64-86) and keeps the first triangle with strictly smaller distance, i.e. the
// minimum over (distance, DFS position) among the triangles of visited leaves.  Triangles are
// stored in DFS leaf order, so the position is the triangle index and the tie-break is explicit.
// EXACT=true walks every crossed box, as the reference does.  EXACT=false adds two conservative
// cuts that never drop a triangle that could win: boxes wholly behind the origin (a hit needs
// t > EPSILON) and boxes whose entry parameter exceeds the best hit's parameter by a relative
// margin of 1e-7; children are visited nearest entry first so the cut bites early.  A box is
// only entered when the reference's own slab test passes, so the reference's grazing misses are
// reproduced too.
//
// The slab test runs with a per-ray reciprocal direction (one multiply instead of a divide per
// slab).  The reciprocal product differs from the reference's quotient by at most ~2 ulp and has
// the same inf / NaN pattern (both come only from zero direction components), so the pass / fail
// decision can differ only when the entry and exit parameters agree to ~1e-15; inside a band of
// 1e-13 relative the test is redone with the reference's divisions.
// ---------------------------------------------------------------------------------------------
struct Hit {
  int idx;       // DFS triangle index, INT_MAX if none
  double dist;
  double t;
  int nodes;     // BVH nodes entered
  int tests;     // ray-triangle tests
};

// Hit point of a closest hit: the same operations tri_intersect applies, so the value is
// bit-identical to the point it computed (kept out of the traversal loop to save registers).
__device__ __forceinline__ V3 hit_point(V3 o, V3 d, const Hit& h) {
  return h.idx == INT_MAX ? v3(0, 0, 0) : vadd(o, vmul(d, h.t));
}

struct RayPre {
  V3 o, d, inv;
  bool safe;     // no subnormal direction component (reciprocal exact in inf / NaN pattern)
};

__device__ __forceinline__ RayPre ray_pre(V3 o, V3 d) {
  RayPre r;
  r.o = o;
  r.d = d;
  r.inv = v3(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
  const double tiny = 1e-290;
  r.safe = !((d.x != 0.0 && fabs(d.x) < tiny) || (d.y != 0.0 && fabs(d.y) < tiny) || (d.z != 0.0 && fabs(d.z) < tiny));
  return r;
}

__device__ __forceinline__ double fin_abs(double x) { return isfinite(x) ? fabs(x) : 0.0; }

__device__ __forceinline__ bool box_test(const double* bmin, const double* bmax, const RayPre& r, double* lo,
                                         double* hi) {
  if (r.safe) {
    double t_max = INFINITY, t_min = -INFINITY;
    const double oo[3] = {r.o.x, r.o.y, r.o.z}, iv[3] = {r.inv.x, r.inv.y, r.inv.z};
#pragma unroll
    for (int i = 0; i < 3; i++) {
      double t1 = (bmin[i] - oo[i]) * iv[i];
      double t2 = (bmax[i] - oo[i]) * iv[i];
      t_max = smin(t_max, smax(t1, t2));
      t_min = smax(t_min, smin(t1, t2));
    }
    double m = 1e-13 * (fin_abs(t_min) + fin_abs(t_max));
    if (t_min > t_max + m) return false;
    if (t_min < t_max - m) {
      *lo = t_min;
      *hi = t_max;
      return true;
    }
  }
  return box_line(bmin, bmax, r.o, r.d, lo, hi);
}

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

__device__ __forceinline__ RayF rayf_pre(V3 o, V3 d, float ext) {
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

// True if the triangle behind box b may still hold a hit at t <= clip.
__device__ __forceinline__ bool box32_may_hit(const DBox32& b, const RayF& f, float clip) {
  const float ax = fmaf(b.lo[0], f.ix, f.nx), bx = fmaf(b.hi[0], f.ix, f.nx);
  const float ay = fmaf(b.lo[1], f.iy, f.ny), by = fmaf(b.hi[1], f.iy, f.ny);
  const float az = fmaf(b.lo[2], f.iz, f.nz), bz = fmaf(b.hi[2], f.iz, f.nz);
  const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
  const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
  return tn <= tf && tf >= 0.0f && tn <= clip;
}

template <bool EXACT>
__device__ __forceinline__ bool keep_box(double lo, double hi, const Hit& h) {
  if (EXACT) return true;
  if (hi < -1e-9) return false;
  if (h.idx != INT_MAX && lo > h.t * (1.0 + 1e-7) + 1e-7) return false;
  return true;
}

// Traversal stack: entry e of the calling thread is at stkN[e * stride] / stkT[e * stride]
// (stride = threads per block, so a wave's entries are contiguous).
template <bool EXACT>
__device__ __forceinline__ bool closest_hit(const DNode* __restrict__ nodes, const DTri* __restrict__ tris,
                                            const DBox32* __restrict__ tbox, float ext, V3 o, V3 d, Hit& h, int* stkN,
                                            float* stkT, int stride) {
  h.idx = INT_MAX;
  h.dist = INFINITY;
  h.t = INFINITY;
  h.nodes = 0;
  h.tests = 0;
  const RayPre r = ray_pre(o, d);
  const RayF f = rayf_pre(o, d, ext);
  float clip = INFINITY;
  double lo, hi;
  if (!box_test(nodes[0].bmin, nodes[0].bmax, r, &lo, &hi) || !keep_box<EXACT>(lo, hi, h)) return false;
  int sp = 0;
  int cur = 0;
  while (true) {
    const DNode nd = nodes[cur];
    h.nodes++;
    if (nd.count > 0) {
      for (int k = nd.a; k < nd.a + nd.count; k++) {
        if (f.on && !box32_may_hit(tbox[k], f, clip)) continue;
        h.tests++;
        const DTri& tr = tris[k];
        V3 p;
        double dist, t;
        if (tri_intersect(ld3(tr.v1), ld3(tr.e1), ld3(tr.e2), o, d, &p, &dist, &t)) {
          if (dist < h.dist || (dist == h.dist && k < h.idx)) {
            h.idx = k;
            h.dist = dist;
            h.t = t;
            clip = __double2float_ru(t * (1.0 + 1e-6));
          }
        }
      }
    } else {
      double la, ha, lb, hb;
      bool ina = box_test(nodes[nd.a].bmin, nodes[nd.a].bmax, r, &la, &ha) && keep_box<EXACT>(la, ha, h);
      bool inb = box_test(nodes[nd.b].bmin, nodes[nd.b].bmax, r, &lb, &hb) && keep_box<EXACT>(lb, hb, h);
      if (ina && inb) {
        int nearN = nd.a, farN = nd.b;
        double farT = lb;
        if (!EXACT && lb < la) { nearN = nd.b; farN = nd.a; farT = la; }
        stkN[sp * stride] = farN;
        stkT[sp * stride] = __double2float_rd(farT);
        sp++;
        cur = nearN;
        continue;
      }
      if (ina || inb) {
        cur = ina ? nd.a : nd.b;
        continue;
      }
    }
    bool popped = false;
    while (sp > 0) {
      --sp;
      if (!EXACT && h.idx != INT_MAX && (double)stkT[sp * stride] > h.t * (1.0 + 1e-7) + 1e-7) continue;
      cur = stkN[sp * stride];
      popped = true;
      break;
    }
    if (!popped) break;
  }
  return h.idx != INT_MAX;
}

// ---------------------------------------------------------------------------------------------
// Reception sphere: HitReceptionSphere (RTSolver.cpp:805-863), and the alpha*L/sqrt(3) model.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ bool sphere_hit(V3 S, V3 T, V3 C, double plArg, double trueLen, int model, double rcoef) {
  V3 rayDir = vsub(T, S);
  double t0 = vdot(rayDir, vsub(C, S)) / vsqnorm(rayDir);
  if (t0 < 0) return false;
  if (t0 >= 1) t0 = 1;
  V3 cp = vadd(S, vmul(rayDir, t0));
  double unfolded = model == 0 ? plArg + vnorm(cp) : trueLen + vnorm(vmul(rayDir, t0));
  double radius = unfolded * rcoef;
  double distance = vnorm(vsub(C, cp));
  return distance <= radius;
}

// ImagePoint (RTSolver.cpp:910-924)
__device__ __forceinline__ V3 image_point(const DSurf& s, V3 src) {
  double dist = vdot(ld3(s.n), vsub(src, ld3(s.v1)));
  return vsub(src, smul(2 * dist, ld3(s.n)));
}

// ---------------------------------------------------------------------------------------------
// Complex FP64 arithmetic for the field (std::complex<double> semantics up to last-ulp rounding).
// ---------------------------------------------------------------------------------------------
struct cx {
  double r, i;
};
__device__ __forceinline__ cx mkc(double r, double i) { cx z; z.r = r; z.i = i; return z; }
__device__ __forceinline__ cx cadd(cx a, cx b) { return mkc(a.r + b.r, a.i + b.i); }
__device__ __forceinline__ cx csub(cx a, cx b) { return mkc(a.r - b.r, a.i - b.i); }
__device__ __forceinline__ cx cmul(cx a, cx b) { return mkc(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r); }
__device__ __forceinline__ cx cscale(cx a, double s) { return mkc(a.r * s, a.i * s); }
__device__ __forceinline__ cx cneg(cx a) { return mkc(-a.r, -a.i); }
// Smith's division.
__device__ __forceinline__ cx cdiv(cx a, cx b) {
  if (fabs(b.r) >= fabs(b.i)) {
    double q = b.i / b.r, den = b.r + b.i * q;
    return mkc((a.r + a.i * q) / den, (a.i - a.r * q) / den);
  }
  double q = b.r / b.i, den = b.r * q + b.i;
  return mkc((a.r * q + a.i) / den, (a.i * q - a.r) / den);
}
__device__ __forceinline__ double cabs_(cx a) { return hypot(a.r, a.i); }
// Principal square root.
__device__ __forceinline__ cx csqrt_(cx z) {
  if (z.r == 0.0 && z.i == 0.0) return mkc(0.0, z.i);
  double m = cabs_(z);
  if (z.r >= 0.0) {
    double t = sqrt(0.5 * (m + z.r));
    return mkc(t, z.i / (2.0 * t));
  }
  double t = sqrt(0.5 * (m - z.r));
  return mkc(fabs(z.i) / (2.0 * t), copysign(t, z.i));
}
__device__ __forceinline__ cx cexp_(cx z) {
  double e = exp(z.r);
  return mkc(e * cos(z.i), e * sin(z.i));
}
__device__ __forceinline__ cx csin_(cx z) { return mkc(sin(z.r) * cosh(z.i), cos(z.r) * sinh(z.i)); }
__device__ __forceinline__ cx ccos_(cx z) { return mkc(cos(z.r) * cosh(z.i), -sin(z.r) * sinh(z.i)); }
__device__ __forceinline__ cx clog_(cx z) { return mkc(log(cabs_(z)), atan2(z.i, z.r)); }
// asin(z) = -i log(i z + sqrt(1 - z^2)); for real z in [-1, 1] the real routine is used.
__device__ __forceinline__ cx casin_(cx z) {
  if (z.i == 0.0 && fabs(z.r) <= 1.0) return mkc(asin(z.r), z.i);
  cx w = csqrt_(csub(mkc(1.0, 0.0), cmul(z, z)));
  cx u = cadd(mkc(-z.i, z.r), w);
  cx l = clog_(u);
  return mkc(l.i, -l.r);
}

struct V3C {
  cx x, y, z;
};

// Physical constants as the reference's MSVC build evaluates them (Constants.h:6-22).
#define SBR_PI 3.141592653589793
#define SBR_C0 2.99792458e8

// AngleBetween (VecMatDef.h:56-77), result in radians.
__device__ __forceinline__ double angle_between_rad(V3 a, V3 b) {
  a = vnormalized(a);
  b = vnormalized(b);
  double ratio = vdot(a, b);
  if (ratio < 0) return SBR_PI - 2.0 * asin(vnorm(vsub(vneg(a), b)) / 2.0);
  return 2.0 * asin(vnorm(vsub(a, b)) / 2.0);
}

// rows-as-given 3x3 (row-major) times vector, reduction order (m0*v0 + m1*v1) + m2*v2.
__device__ __forceinline__ V3 mat_vec(const double* m, V3 v) {
  return v3((m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[3] * v.x + m[4] * v.y) + m[5] * v.z,
            (m[6] * v.x + m[7] * v.y) + m[8] * v.z);
}
__device__ __forceinline__ cx rdot(double a, double b, double c, cx x, cx y, cx z) {
  return cadd(cadd(cscale(x, a), cscale(y, b)), cscale(z, c));
}
__device__ __forceinline__ V3C mat_vec_c(const double* m, const V3C& v) {
  V3C r;
  r.x = rdot(m[0], m[1], m[2], v.x, v.y, v.z);
  r.y = rdot(m[3], m[4], m[5], v.x, v.y, v.z);
  r.z = rdot(m[6], m[7], m[8], v.x, v.y, v.z);
  return r;
}
__device__ __forceinline__ V3C matT_vec_c(const double* m, const V3C& v) {   // m^T v
  V3C r;
  r.x = rdot(m[0], m[3], m[6], v.x, v.y, v.z);
  r.y = rdot(m[1], m[4], m[7], v.x, v.y, v.z);
  r.z = rdot(m[2], m[5], m[8], v.x, v.y, v.z);
  return r;
}

// TE/TM coefficients (FieldCompute.cpp:237-302) for incidence thetaI from air onto a material
// of relative permittivity relPerm and conductivity sigma; returns the two factors applied to
// (E_theta, E_phi): reflection (refl=true) or transmission.
__device__ __forceinline__ void fresnel(bool refl, double thetaI, double relPerm, double sigma, double freq,
                                        bool infWall, cx* fTM, cx* fTE) {
  const double U0 = 4 * SBR_PI * 1e-7;
  const double E0 = 1.0 / (U0 * SBR_C0 * SBR_C0);
  const double ETA0 = U0 * SBR_C0;
  cx et = mkc(relPerm, -1 * sigma / (E0 * 2 * SBR_PI * freq));
  cx ni = csqrt_(mkc(1.0, 0.0));
  cx nt = csqrt_(et);
  cx ti = mkc(thetaI, 0.0);
  cx tt = casin_(cmul(cdiv(ni, nt), csin_(ti)));
  cx etai = cdiv(mkc(ETA0, 0.0), ni), etat = cdiv(mkc(ETA0, 0.0), nt);
  cx ci = ccos_(ti), ct = ccos_(tt), si = csin_(ti);
  cx gtm = cdiv(cadd(cmul(cneg(etat), ct), cmul(etai, ci)), cadd(cmul(etat, ct), cmul(etai, ci)));
  cx a = cdiv(etat, ct), b = cdiv(etai, ci);
  cx gte = cneg(cdiv(csub(a, b), cadd(a, b)));
  if (infWall) {
    if (refl) { *fTM = gtm; *fTE = gte; }
    else { *fTM = mkc(0, 0); *fTE = mkc(0, 0); }
    return;
  }
  // Slab of width 0 (the reference passes width = 0): q = 0, phase factor 1.
  const double lamda = SBR_C0 / freq;
  const double width = 0.0;
  cx q = cmul(mkc(2 * SBR_PI * width / lamda, 0.0), csqrt_(csub(nt, cmul(si, si))));
  cx ph = cexp_(cmul(mkc(0.0, -2.0), q));
  cx g2[2] = {gtm, gte};
  cx outR[2], outT[2];
  for (int k = 0; k < 2; k++) {
    cx g = g2[k];
    cx den = csub(mkc(1.0, 0.0), cmul(cmul(g, g), ph));
    outR[k] = cdiv(cmul(g, csub(mkc(1.0, 0.0), ph)), den);
    outT[k] = cdiv(cmul(csub(mkc(1.0, 0.0), cmul(g, g)), ph), den);
  }
  if (refl) { *fTM = outR[0]; *fTE = outR[1]; }
  else { *fTM = outT[0]; *fTE = outT[1]; }
}

}  // namespace sbr
