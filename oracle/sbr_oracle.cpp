// TEST INFRASTRUCTURE ONLY -- never linked into, loaded by or called from the product library.
//
// FP64 CPU oracle for SBRSolver's shoot-bounce-reflect trace-and-field step.  It restates the
// reference's algorithm function by function so that the HIP path can be checked against it:
// same control flow, same FP64 operation order (Eigen 3.4 SSE2 reduction order for dot /
// squaredNorm, no FMA contraction), same quirks.  Every function cites the reference file:line it
// follows (paths relative to SBRSolver_rev2/).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this library, through
// the C interface at the bottom of the file (ctypes).
//
// Pinning (what checks this restatement, besides reading it against the source):
//   * STL read + vertex dedup + coplanar ids: data/output/Faces.vtk, Edges.vtk (ground.stl).
//   * icosahedron order and directions: data/output/Icosahedron.vtk (level 3, 642 rays).
//   * trace / capture / image correction / dedup / field for a 1-leaf BVH at R=1:
//     data/output/RayPath.vtk, RxLocations.vtk, ElectricField.csv (ground.stl, tess 2).
//   * BVH closest hit vs brute force over all triangles, Fresnel closed forms at normal
//     incidence, image-point reflection, direct field closed form: tests/test_oracle_*.py.
//   * Tabulated transmitter pattern (antennaType 3, an extension: the reference reads
//     TXPattern.dat, Preprocessor.cpp:3-35, but never uses it): bilinear interpolation pinned by
//     tables sampled from functions it reproduces exactly, and by the Hertzian dipole at grid angles
//     (tests/test_pattern.py).
//   * Multi-bounce (R>=2), transmission, oblique lossy Fresnel, half-wave / isotropic patterns:
//     parity unpinned by reference output (no reference vector covers them); they rest on this
//     restatement and on the invariant / closed-form tests.

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <queue>
#include <sstream>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------------------------
// L0 math (common/VecMatDef.h, common/Constants.h).  Vec3 = 3 contiguous doubles (Eigen layout).
// ---------------------------------------------------------------------------------------------
// MSVC has long double == double, so EIGEN_PI-based constants are double expressions there.
static const double kPI = 3.141592653589793;                  // Constants.h:19 (EIGEN_PI as double)
static const double kTWOPI = 2 * kPI;                          // Constants.h:18
static const double kC0 = 2.99792458e8;                        // Constants.h:6
static const double kU0 = 4 * kPI * 1e-7;                     // Constants.h:7
static const double kE0 = 1.0 / (kU0 * kC0 * kC0);             // Constants.h:8
static const double kETA0 = kU0 * kC0;                         // Constants.h:9
static const double kEPS = (double)std::numeric_limits<float>::epsilon();  // Constants.h:13
static const double kINF = std::numeric_limits<double>::infinity();       // Constants.h:16
typedef std::complex<double> cd;
static const cd kJ(0.0, 1.0);                                  // Constants.h:22

// Tabulated transmitter pattern (no reference counterpart beyond the GainInfo rows it reads,
// DataStructures.h:32-37): grid theta_i = i * 180 / (nt - 1), phi_j = j * 360 / np degrees, per point
// the coefficients j sqrt(60 G) e^{j psi} of the theta and phi fields (G from dB, psi in degrees,
// FieldCompute.cpp:114-116), interpolated bilinearly in degrees, periodic in phi.
static std::vector<cd> gPatTheta, gPatPhi;
static int gPatNT = 0, gPatNP = 0;

static void SetTablePattern(const float* gi, int nt, int np) {
  gPatTheta.assign((size_t)nt * np, cd(0, 0));
  gPatPhi.assign((size_t)nt * np, cd(0, 0));
  for (size_t k = 0; k < (size_t)nt * np; k++) {
    for (int h = 0; h < 2; h++) {
      double gdb = gi[4 * k + 2 * h], psi = gi[4 * k + 2 * h + 1] * (kPI / 180.0);
      cd c = kJ * std::sqrt(60.0 * std::pow(10.0, gdb / 10.0)) * cd(std::cos(psi), std::sin(psi));
      (h == 0 ? gPatTheta : gPatPhi)[k] = c;
    }
  }
  gPatNT = nt;
  gPatNP = np;
}

static void TablePattern(double theta, double phi, cd* eth, cd* eph) {
  double u = theta * (180.0 / kPI) * (gPatNT - 1) / 180.0;
  int i0 = std::min(std::max((int)std::floor(u), 0), gPatNT - 2);
  double a = u - i0;
  double v = phi * (180.0 / kPI) * gPatNP / 360.0;
  double fv = std::floor(v), b = v - fv;
  int j0 = (((int)fv % gPatNP) + gPatNP) % gPatNP, j1 = (j0 + 1) % gPatNP;
  size_t k00 = (size_t)i0 * gPatNP + j0, k01 = (size_t)i0 * gPatNP + j1;
  size_t k10 = (size_t)(i0 + 1) * gPatNP + j0, k11 = (size_t)(i0 + 1) * gPatNP + j1;
  double w00 = (1 - a) * (1 - b), w01 = (1 - a) * b, w10 = a * (1 - b), w11 = a * b;
  *eth = w00 * gPatTheta[k00] + w01 * gPatTheta[k01] + w10 * gPatTheta[k10] + w11 * gPatTheta[k11];
  *eph = w00 * gPatPhi[k00] + w01 * gPatPhi[k01] + w10 * gPatPhi[k10] + w11 * gPatPhi[k11];
}

struct V3 {
  double x, y, z;
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline V3 mk(double x, double y, double z) { V3 r = {x, y, z}; return r; }
static inline V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 mul(V3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline V3 smul(double s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
// Eigen 3.4 redux of a 3-vector with SSE2 packets: (a0*b0 + a1*b1) + a2*b2.
static inline double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline double sqnorm(V3 a) { return dot(a, a); }
static inline double norm(V3 a) { return std::sqrt(sqnorm(a)); }
static inline V3 normalized(V3 a) {
  double z = sqnorm(a);
  if (z > 0.0) { double s = std::sqrt(z); return mk(a.x / s, a.y / s, a.z / s); }
  return a;
}
static inline V3 cross(V3 a, V3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline bool eq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
// std::min / std::max semantics (NaN handling follows the argument order).
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

struct M3 { double m[9]; };   // row-major storage (values, not Eigen's memory order)
static inline M3 ident() { M3 r = {{1, 0, 0, 0, 1, 0, 0, 0, 1}}; return r; }
static inline M3 transpose(const M3& a) {
  M3 r;
  for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++) r.m[i * 3 + k] = a.m[k * 3 + i];
  return r;
}
static inline M3 matmul(const M3& a, const M3& b) {
  M3 r;
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++)
      r.m[i * 3 + k] = (a.m[i * 3 + 0] * b.m[0 * 3 + k] + a.m[i * 3 + 1] * b.m[1 * 3 + k]) +
                       a.m[i * 3 + 2] * b.m[2 * 3 + k];
  return r;
}
static inline V3 mv(const M3& a, V3 v) {
  return mk((a.m[0] * v.x + a.m[1] * v.y) + a.m[2] * v.z, (a.m[3] * v.x + a.m[4] * v.y) + a.m[5] * v.z,
            (a.m[6] * v.x + a.m[7] * v.y) + a.m[8] * v.z);
}
struct V3c { cd x, y, z; };
static inline V3c mvc(const M3& a, const V3c& v) {
  V3c r;
  r.x = (a.m[0] * v.x + a.m[1] * v.y) + a.m[2] * v.z;
  r.y = (a.m[3] * v.x + a.m[4] * v.y) + a.m[5] * v.z;
  r.z = (a.m[6] * v.x + a.m[7] * v.y) + a.m[8] * v.z;
  return r;
}

// VecMatDef.h:51-54
static inline V3 Reflect(V3 i, V3 n) { return sub(i, mul(smul(2.0, n), dot(n, i))); }

// VecMatDef.h:56-77
static double AngleBetween(V3 v1, V3 v2, bool alreadyNormalized, bool inRadian) {
  if (!alreadyNormalized) { v1 = normalized(v1); v2 = normalized(v2); }
  double ratio = dot(v1, v2);
  double theta;
  if (ratio < 0) theta = kPI - 2.0 * std::asin(norm(sub(neg(v1), v2)) / 2.0);
  else theta = 2.0 * std::asin(norm(sub(v1, v2)) / 2.0);
  return inRadian ? theta : (theta * (180.0 / kPI));
}
// VecMatDef.h:79-83, 101-109: Rotate(v, old, new) = (new * old^T) * v
static inline M3 RotationMatrix(const M3& oldSys, const M3& newSys) { return matmul(newSys, transpose(oldSys)); }
// VecMatDef.h:85-99
static M3 TransformationMatrix(double theta, double phi, bool sph2Cart) {
  double st = std::sin(theta), sp = std::sin(phi), ct = std::cos(theta), cp = std::cos(phi);
  M3 t = {{st * cp, ct * cp, -1 * sp, st * sp, ct * sp, cp, ct, -1 * st, 0.0}};
  return sph2Cart ? t : transpose(t);
}
// VecMatDef.h:132-155
static V3 CartesianToSpherical(V3 v) {
  V3 r;
  r.x = norm(v);
  if (v.x == 0.0 && v.y == 0.0 && v.z == 0.0) r.y = kPI / 2;
  else r.y = std::atan2(std::sqrt(v.x * v.x + v.y * v.y), v.z);
  r.z = std::atan2(v.y, v.x);
  if (r.z < 0) r.z += 2 * kPI;
  return r;
}

// ---------------------------------------------------------------------------------------------
// L1 geometry: Triangle (Triangle.cpp), BVH (BVH/BVH.h, BVH/BVHNode.h), Icosahedron.cpp
// ---------------------------------------------------------------------------------------------
struct HitInfo {   // common/DataStructures.h:9-17
  V3 normal = {0, 0, 0}, pointIntersect = {0, 0, 0};
  int materialId = -1, triangleId = -1, coplanarId = -1;
  double distance = kINF;
};

struct Tri {
  V3 v1, v2, v3, norm;
  M3 coordSys;
  double bmin[3], bmax[3];
  V3 center;
  int triangleId = -1, coplanarId = -1, materialId = 0;

  // Triangle.cpp:28-35 (constructor with file normal), :37-70
  void init(V3 a, V3 b, V3 c, V3 n, int id) {
    v1 = a; v2 = b; v3 = c; norm = n; triangleId = id;
    bmin[0] = std::min({v1.x, v2.x, v3.x}); bmin[1] = std::min({v1.y, v2.y, v3.y});
    bmin[2] = std::min({v1.z, v2.z, v3.z});
    bmax[0] = std::max({v1.x, v2.x, v3.x}); bmax[1] = std::max({v1.y, v2.y, v3.y});
    bmax[2] = std::max({v1.z, v2.z, v3.z});
    center = mk((bmin[0] + bmax[0]) / 2.0, (bmin[1] + bmax[1]) / 2.0, (bmin[2] + bmax[2]) / 2.0);
    V3 zw = norm, yw = normalized(sub(v2, v1)), xw = cross(yw, zw);
    M3 s = {{xw.x, xw.y, xw.z, yw.x, yw.y, yw.z, zw.x, zw.y, zw.z}};
    coordSys = s;
  }

  // Triangle.cpp:80-125 (Moller-Trumbore, EPSILON = FLT_EPSILON)
  bool RayIntersects(V3 rayOrg, V3 rayDir, HitInfo& info) const {
    V3 edge1 = sub(v2, v1), edge2 = sub(v3, v1);
    V3 h = cross(rayDir, edge2);
    double a = dot(edge1, h);
    if (a > -kEPS && a < kEPS) return false;
    double f = 1.0 / a;
    V3 s = sub(rayOrg, v1);
    double u = f * dot(s, h);
    if (u < 0.0 || u > 1.0) return false;
    V3 q = cross(s, edge1);
    double v = f * dot(rayDir, q);
    if (v < 0.0 || u + v > 1.0) return false;
    double t = f * dot(edge2, q);
    if (t > kEPS) {
      info.pointIntersect = add(rayOrg, mul(rayDir, t));
      info.distance = orc::norm(sub(info.pointIntersect, rayOrg));
      info.normal = norm;
      info.materialId = materialId;
      info.triangleId = triangleId;
      info.coplanarId = coplanarId;
      return true;
    }
    return false;
  }
};

struct BNode {
  double bmin[3], bmax[3];
  int child[2] = {-1, -1};
  std::vector<Tri*> objects;
  // BVHNode.h:54-62
  BNode() {
    for (int a = 0; a < 3; a++) { bmin[a] = FLT_MAX; bmax[a] = -FLT_MAX; }
  }
  // BVHNode.h:64-86 (infinite line, std::min/std::max NaN semantics)
  bool IntersectsAABB(V3 o, V3 d) const {
    double t_max = kINF, t_min = -kINF;
    for (int i = 0; i < 3; i++) {
      double t1 = (bmin[i] - o[i]) / d[i];
      double t2 = (bmax[i] - o[i]) / d[i];
      double t_near = smin(t1, t2), t_far = smax(t1, t2);
      t_max = smin(t_max, t_far);
      t_min = smax(t_min, t_near);
      if (t_min > t_max) return false;
    }
    return true;
  }
  // The same slab test, also returning the final [t_min, t_max] (counting traversal only).
  bool LineInterval(V3 o, V3 d, double* lo, double* hi) const {
    double t_max = kINF, t_min = -kINF;
    for (int i = 0; i < 3; i++) {
      double t1 = (bmin[i] - o[i]) / d[i];
      double t2 = (bmax[i] - o[i]) / d[i];
      t_max = smin(t_max, smax(t1, t2));
      t_min = smax(t_min, smin(t1, t2));
      if (t_min > t_max) return false;
    }
    *lo = t_min;
    *hi = t_max;
    return true;
  }
};

struct BVH {
  std::vector<BNode> nodes;
  long long nodeVisits = 0;  // counting mode (not in the reference)
  long long canonNodes = 0, canonTris = 0;
  BVH() { nodes.resize(1); }  // BVH.h:39-43

  // BVH.h:247-267
  static double SurfaceArea(const double* bmin, const double* bmax) {
    if (bmax[0] <= bmin[0] || bmax[1] <= bmin[1] || bmax[2] <= bmin[2]) return 0.0;
    const double dx = bmax[0] - bmin[0], dy = bmax[1] - bmin[1], dz = bmax[2] - bmin[2];
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }
  static void Merge(const double* amin, const double* amax, const double* bmin, const double* bmax,
                    double* rmin, double* rmax) {  // BVH.h:241-249
    for (int i = 0; i < 3; i++) { rmin[i] = std::min(amin[i], bmin[i]); rmax[i] = std::max(amax[i], bmax[i]); }
  }

  // BVH.h:135-230 (full-sweep SAH; sorts `objects` in place -- the caller's vector at the root)
  void Build(std::vector<Tri*>& objects, int nodeIdx) {
    {
      BNode& node = nodes[nodeIdx];
      for (const Tri* o : objects) Merge(o->bmin, o->bmax, node.bmin, node.bmax, node.bmin, node.bmax);
    }
    const double T_tri = 1.0, T_aabb = 1.0;
    double bestCost = T_tri * objects.size();
    int bestAxis = -1, bestSplit = -1;
    const double SA_root = SurfaceArea(nodes[nodeIdx].bmin, nodes[nodeIdx].bmax);
    const int n = (int)objects.size();
    for (int axis = 0; axis < 3; axis++) {
      std::sort(objects.begin(), objects.end(),
                [axis](const Tri* a, const Tri* b) { return a->center[axis] < b->center[axis]; });
      double lmin[3], lmax[3];
      for (int a = 0; a < 3; a++) { lmin[a] = FLT_MAX; lmax[a] = -FLT_MAX; }
      std::vector<double> v1SA(n + 1, std::numeric_limits<double>::max());
      v1SA[0] = SurfaceArea(lmin, lmax);
      for (int i = 0; i < n; i++) {
        Merge(lmin, lmax, objects[i]->bmin, objects[i]->bmax, lmin, lmax);
        v1SA[i] = SurfaceArea(lmin, lmax);
      }
      double rmin[3], rmax[3];
      for (int a = 0; a < 3; a++) { rmin[a] = FLT_MAX; rmax[a] = -FLT_MAX; }
      std::vector<double> v2SA(n + 1, std::numeric_limits<double>::max());
      v2SA[n] = SurfaceArea(rmin, rmax);
      for (int i = n - 1; i >= 0; i--) {
        Merge(rmin, rmax, objects[i]->bmin, objects[i]->bmax, rmin, rmax);
        v2SA[i] = SurfaceArea(rmin, rmax);
      }
      for (int i = 1; i < n - 1; i++) {
        const int j = n - i;
        const double cost = 2 * T_aabb + (i * v1SA[i] + j * v2SA[i]) * T_tri / SA_root;
        if (cost < bestCost) { bestCost = cost; bestAxis = axis; bestSplit = i; }
      }
    }
    if (bestAxis == -1) {
      nodes[nodeIdx].child[0] = nodes[nodeIdx].child[1] = -1;   // BVH.h:269-274
      nodes[nodeIdx].objects = objects;
    } else {
      std::sort(objects.begin(), objects.end(),
                [bestAxis](const Tri* a, const Tri* b) { return a->center[bestAxis] < b->center[bestAxis]; });
      nodes[nodeIdx].child[0] = (int)nodes.size();
      nodes.push_back(BNode());
      nodes[nodeIdx].child[1] = (int)nodes.size();
      nodes.push_back(BNode());
      std::vector<Tri*> left(objects.begin(), objects.begin() + bestSplit);
      std::vector<Tri*> right(objects.begin() + bestSplit, objects.end());
      Build(left, nodes[nodeIdx].child[0]);
      Build(right, nodes[nodeIdx].child[1]);
    }
  }

  // Counting mode (not in the reference): a canonical closest-first, t-clipped traversal of the
  // same tree -- boxes entered nearest entry first, boxes wholly behind the origin or entered beyond
  // the best hit skipped -- counting boxes entered and triangles tested.  It defines the work of one
  // bounce for the algorithmic byte count; its hit is not used.
  void CanonicalCount(V3 o, V3 d) {
    double best = kINF, lo, hi;
    if (!nodes[0].LineInterval(o, d, &lo, &hi) || hi < 0.0) return;
    std::vector<std::pair<int, double>> st{{0, lo}};
    while (!st.empty()) {
      const std::pair<int, double> top = st.back();
      st.pop_back();
      if (top.second > best) continue;
      canonNodes++;
      const BNode& n = nodes[top.first];
      if (n.child[0] == -1) {
        for (Tri* obj : n.objects) {
          canonTris++;
          HitInfo cur;
          if (obj->RayIntersects(o, d, cur) && cur.distance < best) best = cur.distance;
        }
        continue;
      }
      double l0, h0, l1, h1;
      const bool in0 = nodes[n.child[0]].LineInterval(o, d, &l0, &h0) && h0 >= 0.0 && l0 <= best;
      const bool in1 = nodes[n.child[1]].LineInterval(o, d, &l1, &h1) && h1 >= 0.0 && l1 <= best;
      if (in0 && in1) {
        if (l1 < l0) { st.push_back({n.child[0], l0}); st.push_back({n.child[1], l1}); }
        else { st.push_back({n.child[1], l1}); st.push_back({n.child[0], l0}); }
      } else if (in0) {
        st.push_back({n.child[0], l0});
      } else if (in1) {
        st.push_back({n.child[1], l1});
      }
    }
  }

  // BVH.h:276-328 (recursive DFS, children[0] first, strict '<' replacement)
  bool RayIntersects(int index, V3 rayOrg, V3 rayDir, HitInfo& info) {
    nodeVisits++;
    if (nodes[index].IntersectsAABB(rayOrg, rayDir)) {
      if (nodes[index].child[0] != -1) {
        bool childResult = false;
        for (int i = 0; i < 2; i++) {
          bool r = RayIntersects(nodes[index].child[i], rayOrg, rayDir, info);
          if (r) childResult = r;
        }
        if (childResult) return childResult;
      } else {
        bool result = false;
        for (Tri* obj : nodes[index].objects) {
          HitInfo cur;
          if (obj->RayIntersects(rayOrg, rayDir, cur)) {
            if (cur.distance < info.distance) {
              result = true;
              info = cur;
            }
          }
        }
        if (result) return result;
      }
    }
    return false;
  }
};

// Icosahedron.cpp:3-21
static int vertex_for_edge(std::map<std::pair<int, int>, int>& lookup, std::vector<V3>& vertices, int first,
                           int second) {
  std::pair<int, int> key(first, second);
  if (key.first > key.second) std::swap(key.first, key.second);
  auto ins = lookup.insert({key, (int)vertices.size()});
  if (ins.second) {
    V3 p = normalized(add(vertices[first], vertices[second]));
    vertices.push_back(p);
  }
  return ins.first->second;
}
struct I3 { int a, b, c; };
// Icosahedron.cpp:23-41
static std::vector<I3> Subdivide(std::vector<V3>& vertices, const std::vector<I3>& tris) {
  std::map<std::pair<int, int>, int> lookup;
  std::vector<I3> result;
  result.reserve(tris.size() * 4);
  for (const I3& e : tris) {
    int m0 = vertex_for_edge(lookup, vertices, e.a, e.b);
    int m1 = vertex_for_edge(lookup, vertices, e.b, e.c);
    int m2 = vertex_for_edge(lookup, vertices, e.c, e.a);
    result.push_back({e.a, m0, m2});
    result.push_back({e.b, m1, m0});
    result.push_back({e.c, m2, m1});
    result.push_back({m0, m1, m2});
  }
  return result;
}
// Icosahedron.cpp:43-78, :80-110 (vertex mode: directions = normalized vertices, in vertex order)
static std::vector<V3> GenerateRaysOnIcosahedron(int tess) {
  const double X = .525731112119133606, Z = .850650808352039932, N = 0.0;
  std::vector<V3> v = {mk(-X, N, Z), mk(X, N, Z), mk(-X, N, -Z), mk(X, N, -Z), mk(N, Z, X), mk(N, Z, -X),
                       mk(N, -Z, X), mk(N, -Z, -X), mk(Z, X, N), mk(-Z, X, N), mk(Z, -X, N), mk(-Z, -X, N)};
  std::vector<I3> t = {{0, 4, 1}, {0, 9, 4}, {9, 5, 4}, {4, 5, 8}, {4, 8, 1}, {8, 10, 1}, {8, 3, 10},
                       {5, 3, 8}, {5, 2, 3}, {2, 7, 3}, {7, 10, 3}, {7, 6, 10}, {7, 11, 6}, {11, 0, 6},
                       {0, 1, 6}, {6, 1, 10}, {9, 0, 11}, {9, 11, 2}, {9, 2, 5}, {7, 2, 11}};
  for (int i = 0; i < tess; ++i) t = Subdivide(v, t);
  std::vector<V3> out(v.size());
  for (size_t i = 0; i < v.size(); i++) out[i] = normalized(v[i]);
  return out;
}

// ---------------------------------------------------------------------------------------------
// L2 scene preprocessing: StlReader.h:448-700 (read + RemoveDoubles), Preprocessor.cpp:69-246
// ---------------------------------------------------------------------------------------------
struct CWI { float d[3]; unsigned index; };
static bool cwi_less(const CWI& a, const CWI& c) {   // StlReader.h:431-436
  return (a.d[0] < c.d[0]) || (a.d[0] == c.d[0] && a.d[1] < c.d[1]) ||
         (a.d[0] == c.d[0] && a.d[1] == c.d[1] && a.d[2] < c.d[2]);
}
static bool cwi_ne(const CWI& a, const CWI& c) { return a.d[0] != c.d[0] || a.d[1] != c.d[1] || a.d[2] != c.d[2]; }

// StlReader.h:448-508
static void RemoveDoubles(std::vector<float>& uniq, std::vector<unsigned>& tris, std::vector<CWI>& cw) {
  std::sort(cw.begin(), cw.end(), cwi_less);
  unsigned numUnique = 1;
  for (size_t i = 1; i < cw.size(); ++i) if (cwi_ne(cw[i], cw[i - 1])) ++numUnique;
  uniq.resize(numUnique * 3);
  std::vector<unsigned> newIndex(cw.size());
  unsigned cur = 0;
  newIndex[0] = 0;   // as written; index 0 of the *sorted* array's slot
  for (int i = 0; i < 3; ++i) uniq[i] = cw[0].d[i];
  for (size_t i = 1; i < cw.size(); ++i) {
    const CWI c = cw[i];
    if (cwi_ne(c, cw[i - 1])) {
      ++cur;
      for (int j = 0; j < 3; ++j) uniq[cur * 3 + j] = cw[i].d[j];
    }
    newIndex[c.index] = cur;
  }
  unsigned nu = 0;
  for (unsigned i = 0; i < tris.size(); i += 3) {
    unsigned ni[3];
    for (int j = 0; j < 3; ++j) ni[j] = newIndex[tris[i + j]];
    if (ni[0] != ni[1] && ni[0] != ni[2] && ni[1] != ni[2]) {
      for (int j = 0; j < 3; ++j) tris[nu + j] = ni[j];
      nu += 3;
    }
  }
  if (nu < tris.size()) tris.resize(nu);
}

static bool StlHasAscii(const std::string& fn) {   // StlReader.h:704-717
  std::ifstream in(fn);
  if (!in) return false;
  char chars[256];
  in.read(chars, 256);
  std::string buf(chars, in.gcount());
  std::transform(buf.begin(), buf.end(), buf.begin(), ::tolower);
  return buf.find("solid") != std::string::npos && buf.find("\n") != std::string::npos &&
         buf.find("facet") != std::string::npos && buf.find("normal") != std::string::npos;
}

static bool ReadStl(const std::string& fn, std::vector<float>& coords, std::vector<float>& normals,
                    std::vector<unsigned>& tris) {
  std::vector<CWI> cw;
  coords.clear(); normals.clear(); tris.clear();
  if (StlHasAscii(fn)) {   // StlReader.h:527-634
    std::ifstream in(fn);
    if (!in) return false;
    std::string line;
    size_t nfv = 0;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::vector<std::string> tok;
      std::string s;
      while (ls >> s) tok.push_back(s);
      if (tok.empty()) continue;
      if (tok[0] == "vertex") {
        if (tok.size() < 4) return false;
        CWI c;
        for (int i = 0; i < 3; i++) c.d[i] = (float)atof(tok[i + 1].c_str());
        c.index = (unsigned)cw.size();
        cw.push_back(c);
        ++nfv;
      } else if (tok[0] == "facet") {
        if (tok.size() < 5 || tok[1] != "normal") return false;
        for (int i = 0; i < 3; i++) normals.push_back((float)atof(tok[i + 2].c_str()));
        nfv = 0;
      } else if (tok[0] == "endfacet") {
        if (nfv != 3) return false;
        tris.push_back((unsigned)cw.size() - 3); tris.push_back((unsigned)cw.size() - 2);
        tris.push_back((unsigned)cw.size() - 1);
      }
    }
  } else {   // StlReader.h:637-700
    std::ifstream in(fn, std::ios::binary);
    if (!in) return false;
    char hdr[80];
    in.read(hdr, 80);
    unsigned nt = 0;
    in.read((char*)&nt, 4);
    if (!in) return false;
    for (unsigned t = 0; t < nt; ++t) {
      float d[12];
      in.read((char*)d, 48);
      if (!in) return false;
      for (int i = 0; i < 3; ++i) normals.push_back(d[i]);
      for (int iv = 1; iv < 4; ++iv) {
        CWI c;
        for (int i = 0; i < 3; ++i) c.d[i] = d[iv * 3 + i];
        c.index = (unsigned)cw.size();
        cw.push_back(c);
      }
      tris.push_back((unsigned)cw.size() - 3); tris.push_back((unsigned)cw.size() - 2);
      tris.push_back((unsigned)cw.size() - 1);
      char add[2];
      in.read(add, 2);
      if (!in) return false;
    }
  }
  if (cw.empty()) return false;
  RemoveDoubles(coords, tris, cw);
  return true;
}

struct Scene {
  std::vector<Tri*> owned;      // by triangleId
  std::vector<Tri*> mesh;       // caller's vector: sorted in place by ConstructBVH (Q1)
  BVH bvh;
  std::vector<float> coords;
  std::vector<unsigned> tris;
  std::vector<std::pair<std::pair<int, int>, int>> edges;   // (v*3 pair, connection count), map order
  ~Scene() { for (Tri* t : owned) delete t; }
};

// Preprocessor.cpp:116-190 (one edge)
static void InsertEdge(std::pair<int, int> e, size_t triId, const std::vector<Tri*>& tris,
                       std::map<std::pair<int, int>, std::vector<int>>& edgeMap, std::vector<std::vector<int>>& adj) {
  auto it = edgeMap.find(e);
  if (it != edgeMap.end()) {
    for (const int& i : it->second) {
      if (eq(tris[i]->norm, tris[triId]->norm) || eq(tris[i]->norm, neg(tris[triId]->norm))) {
        adj[i].push_back((int)triId);
        adj[triId].push_back(i);
      }
    }
    it->second.push_back((int)triId);
  } else {
    edgeMap[e] = std::vector<int>();
    edgeMap[e].push_back((int)triId);
  }
}

// Preprocessor.cpp:69-114 + 192-246.  `mats` (optional) sets Triangle::materialId per triangle.
static Scene* BuildScene(std::vector<float>& coords, std::vector<float>& normals, std::vector<unsigned>& tris,
                         const int* mats) {
  Scene* sc = new Scene();
  std::map<std::pair<int, int>, std::vector<int>> edgeMap;
  const size_t numTris = tris.size() / 3;
  std::vector<std::vector<int>> adj(numTris);
  for (size_t it = 0; it < numTris; ++it) {
    size_t a = 3 * tris[3 * it], b = 3 * tris[3 * it + 1], c = 3 * tris[3 * it + 2];
    V3 v1 = mk(coords[a], coords[a + 1], coords[a + 2]);
    V3 v2 = mk(coords[b], coords[b + 1], coords[b + 2]);
    V3 v3 = mk(coords[c], coords[c + 1], coords[c + 2]);
    V3 n = mk(normals[3 * it], normals[3 * it + 1], normals[3 * it + 2]);
    Tri* t = new Tri();
    t->init(v1, v2, v3, n, (int)it);
    if (mats) t->materialId = mats[it];
    sc->owned.push_back(t);
    std::pair<int, int> e1 = a < b ? std::make_pair((int)a, (int)b) : std::make_pair((int)b, (int)a);
    std::pair<int, int> e2 = b < c ? std::make_pair((int)b, (int)c) : std::make_pair((int)c, (int)b);
    std::pair<int, int> e3 = a < c ? std::make_pair((int)a, (int)c) : std::make_pair((int)c, (int)a);
    InsertEdge(e1, it, sc->owned, edgeMap, adj);
    InsertEdge(e2, it, sc->owned, edgeMap, adj);
    InsertEdge(e3, it, sc->owned, edgeMap, adj);
  }
  // BfsCoplanarSurface / BFSUtil
  std::vector<bool> visited(numTris, false);
  int cop = 0;
  for (size_t u0 = 0; u0 < numTris; u0++) {
    if (visited[u0]) continue;
    std::queue<int> q;
    visited[u0] = true;
    q.push((int)u0);
    while (!q.empty()) {
      int u = q.front();
      sc->owned[u]->coplanarId = cop;
      q.pop();
      for (size_t i = 0; i != adj[u].size(); ++i)
        if (!visited[adj[u][i]]) { visited[adj[u][i]] = true; q.push(adj[u][i]); }
    }
    cop++;
  }
  for (auto& kv : edgeMap) sc->edges.push_back(std::make_pair(kv.first, (int)kv.second.size()));
  sc->coords = coords;
  sc->tris = tris;
  sc->mesh = sc->owned;
  sc->bvh.Build(sc->mesh, 0);   // SBRSolver_rev2.cpp:68 -- sorts mesh in place
  return sc;
}

// ---------------------------------------------------------------------------------------------
// L3 tracing core: Ray.cpp, RTSolver.cpp
// ---------------------------------------------------------------------------------------------
struct Ray {   // Ray.h:15-27
  V3 sourcePoint = {kINF, kINF, kINF}, targetPoint = {kINF, kINF, kINF};
  int penetrationMaterialId = -1, reflectionMaterialId = -1, hitSurfaceID = -1, hitCoplanarId = -1;
  double angleFromSurfaceNormal = kINF, pathLength = kINF;
  bool captured = false;
  double trueLength = 0.0;   // extension: geometric length from Tx to sourcePoint (config reception model)
};
struct PTN {   // Ray.h:38-43
  Ray ray;
  PTN* childTransmit = nullptr;
  PTN* childReflect = nullptr;
};
static PTN* NewNode(V3 s, V3 t, int pen, int refl, int surf, int cop, double pl, double ang) {   // Ray.cpp:5-30
  PTN* r = new PTN;
  r->ray.sourcePoint = s; r->ray.targetPoint = t;
  r->ray.penetrationMaterialId = pen; r->ray.reflectionMaterialId = refl;
  r->ray.hitSurfaceID = surf; r->ray.hitCoplanarId = cop;
  r->ray.angleFromSurfaceNormal = ang; r->ray.pathLength = pl;
  return r;
}
static PTN* DeleteChildNodes(PTN* n) {   // Ray.cpp:32-48
  if (n) {
    if (n->childTransmit) n->childTransmit = DeleteChildNodes(n->childTransmit);
    if (n->childReflect) n->childReflect = DeleteChildNodes(n->childReflect);
    delete n;
  }
  return nullptr;
}
static PTN* CloneNode(PTN* n) {   // Ray.cpp:66-86
  if (!n) return nullptr;
  PTN* r = new PTN;
  r->ray = n->ray;
  r->childTransmit = r->childReflect = nullptr;
  return r;
}
static void CloneTree(PTN* o, PTN* c) {   // Ray.cpp:50-64
  if (o) {
    c->childTransmit = CloneNode(o->childTransmit);
    CloneTree(o->childTransmit, c->childTransmit);
    c->childReflect = CloneNode(o->childReflect);
    CloneTree(o->childReflect, c->childReflect);
  }
}
typedef std::vector<Ray> Path;
// Ray.cpp:136-182 (BFS, transmit child before reflect child)
static void TreePaths(PTN* root, std::vector<Path>& out) {
  if (!root) return;
  std::queue<std::pair<std::pair<Path, bool>, PTN*>> q;
  Path nv;
  nv.push_back(root->ray);
  q.push(std::make_pair(std::make_pair(nv, false), root));
  while (!q.empty()) {
    auto p = q.front();
    q.pop();
    PTN* n = p.second;
    if (n->childTransmit || n->childReflect) {
      if (n->childTransmit) {
        bool term = p.first.second || n->childTransmit->ray.captured;
        Path v(p.first.first);
        v.push_back(n->childTransmit->ray);
        q.push(std::make_pair(std::make_pair(v, term), n->childTransmit));
      }
      if (n->childReflect) {
        bool term = p.first.second || n->childReflect->ray.captured;
        Path v(p.first.first);
        v.push_back(n->childReflect->ray);
        q.push(std::make_pair(std::make_pair(v, term), n->childReflect));
      }
    } else if (p.first.second || n->ray.captured) {
      out.push_back(p.first.first);
    }
  }
}

struct Params {
  double tx[3];
  double txCoordSys[9];   // row-major
  int maxR, maxT, maxTotal;   // maxTotal <= 0: no total-depth cap (reference)
  double freq, pt;
  int useFresnel;
  int tess;
  int antennaType;        // reference hard-codes 1 (FieldCompute.cpp:117)
  int receptionModel;     // 0: reference (RTSolver.cpp:838-841), 1: alpha*L/sqrt(3), alpha = 69deg/2^n
  int launchStride, launchOffset;
  int q1MeshIndexing;     // 1: mesh[hitSurfaceID] on the BVH-sorted vector (reference), 0: by id
  int countCanonical = 0; // 1: also run BVH::CanonicalCount per bounce (work count only)
  double wallWidth = 0;   // > 0 with useFresnel: slab walls (extension, see SlabCoeffs)
  int typedDedup = 0;     // 1: duplicate removal also keyed by the segments' interaction types (extension)
  int checkOcclusion = 0; // 1: corrected multi-bounce paths must be unobstructed (extension, Unobstructed)
  int longMiss = 0;       // 1: a miss after a bounce extends 1e10 m, as at the root (extension)
  int physicalField = 0;  // 1: s/p field transport in the plane of incidence (extension, FieldForPath)
};

struct Tracer {
  Scene* sc;
  Params p;
  const double* rx;
  int nrx;
  std::vector<double> mats;   // 6 per material (MaterialProperties layout)
  double rcoef = 0.0;
  int txTess = 0;

  const Tri& MeshTri(int id) const { return p.q1MeshIndexing ? *sc->mesh[id] : *sc->owned[id]; }

  // RTSolver.cpp:128-257 (RayLaunch, recursive)
  void RayLaunch(BVH& bvh, PTN* node, V3 src, V3 dir, int transMat, int refMat, int rc, int tc, double totalPL,
                 bool isRoot, double lastAngle, double trueLen, long long& bounces) {
    if (rc > p.maxR && tc > p.maxT) return;
    bounces++;
    HitInfo hit;
    bool hasHit = bvh.RayIntersects(0, src, dir, hit);
    if (p.countCanonical) bvh.CanonicalCount(src, dir);
    if (hasHit) {
      double angle = AngleBetween(hit.normal, dir, false, false);
      PTN* nn = node;
      if (!isRoot) {
        PTN* c = NewNode(src, hit.pointIntersect, transMat, refMat, hit.triangleId, hit.coplanarId,
                         totalPL + hit.distance, lastAngle);
        c->ray.trueLength = trueLen;
        if (refMat >= 0) node->childReflect = c; else node->childTransmit = c;
        nn = c;
      } else {
        node->ray.sourcePoint = src; node->ray.targetPoint = hit.pointIntersect;
        node->ray.penetrationMaterialId = transMat; node->ray.reflectionMaterialId = refMat;
        node->ray.hitSurfaceID = hit.triangleId; node->ray.hitCoplanarId = hit.coplanarId;
        node->ray.pathLength = hit.distance; node->ray.angleFromSurfaceNormal = lastAngle;
        node->ray.trueLength = 0.0;
      }
      double childTrue = trueLen + norm(sub(hit.pointIntersect, src));
      bool depthOk = p.maxTotal <= 0 || (rc + tc) < p.maxTotal;
      if (rc < p.maxR && depthOk) {
        V3 nd = normalized(Reflect(dir, hit.normal));
        RayLaunch(bvh, nn, hit.pointIntersect, nd, -1, hit.materialId, rc + 1, tc, totalPL + hit.distance, false,
                  angle, childTrue, bounces);
      }
      if (tc < p.maxT && depthOk) {
        RayLaunch(bvh, nn, hit.pointIntersect, dir, hit.materialId, -1, rc, tc + 1, totalPL + hit.distance, false,
                  angle, childTrue, bounces);
      }
    } else {
      if (rc > 0 || tc > 0) {
        PTN* c = NewNode(src, add(src, mul(dir, p.longMiss ? 1E10 : 1E2)), transMat, refMat, -1, -1, kINF, lastAngle);
        c->ray.trueLength = trueLen;
        if (refMat >= 0) node->childReflect = c;
        else if (transMat >= 0) node->childTransmit = c;
        else delete c;
      }
      if (isRoot) {
        node->ray.sourcePoint = src;
        node->ray.targetPoint = add(src, mul(dir, 1E10));
        node->ray.reflectionMaterialId = -1; node->ray.penetrationMaterialId = -1;
        node->ray.hitSurfaceID = -1; node->ray.hitCoplanarId = -1;
        node->ray.pathLength = kINF;
        node->ray.trueLength = 0.0;
      }
    }
  }

  // RTSolver.cpp:805-863 (model 0) and the alpha*L/sqrt(3) model (model 1)
  bool HitReceptionSphere(V3 S, V3 T, V3 C, double pathLength, double trueLen) const {
    V3 rayDir = sub(T, S);
    double t0 = dot(rayDir, sub(C, S)) / sqnorm(rayDir);
    if (t0 < 0) return false;
    else if (t0 >= 1) t0 = 1;
    V3 cp = add(S, mul(rayDir, t0));
    double unfolded;
    if (p.receptionModel == 0) unfolded = pathLength + norm(cp);
    else unfolded = trueLen + norm(mul(rayDir, t0));
    double radius = unfolded * rcoef;
    double distance = norm(sub(C, cp));
    return distance <= radius;
  }

  // RTSolver.cpp:259-297
  void RayCapture(PTN* n, V3 C, double totalPL) {
    bool captured = HitReceptionSphere(n->ray.sourcePoint, n->ray.targetPoint, C, totalPL, n->ray.trueLength);
    if (!n->childTransmit && !n->childReflect && !captured) return;
    if (captured) {
      n->ray.captured = true;
      n->ray.hitSurfaceID = -1;
      n->ray.hitCoplanarId = -1;
      n->childTransmit = DeleteChildNodes(n->childTransmit);
      n->childReflect = DeleteChildNodes(n->childReflect);
    }
    if (n->childTransmit) RayCapture(n->childTransmit, C, totalPL + n->ray.pathLength);
    if (n->childReflect) RayCapture(n->childReflect, C, totalPL + n->ray.pathLength);
  }

  // RTSolver.cpp:910-924
  static V3 ImagePoint(const Tri& t, V3 src) {
    double dist = dot(t.norm, sub(src, t.v1));
    return sub(src, smul(2 * dist, t.norm));
  }
  // RTSolver.cpp:396-433
  std::vector<std::pair<int, V3>> GetImageSources(const Path& path) const {
    V3 txp = mk(p.tx[0], p.tx[1], p.tx[2]);
    V3 cur = txp;
    std::vector<std::pair<int, V3>> r{std::make_pair(-1, txp)};
    for (size_t i = 0; i + 1 < path.size(); i++) {
      int surf = path[i].hitSurfaceID;
      if (surf < 0) return r;
      if (path[i + 1].reflectionMaterialId >= 0) {
        cur = ImagePoint(MeshTri(surf), cur);
        r.push_back(std::make_pair(surf, cur));
      } else if (path[i + 1].penetrationMaterialId >= 0) {
        r.push_back(std::make_pair(surf, cur));
      } else {
        return r;
      }
    }
    return r;
  }
  // RTSolver.cpp:330-373
  bool MultiPathCorrection(Path& mp, V3 rxp) const {
    std::vector<std::pair<int, V3>> img = GetImageSources(mp);
    HitInfo result;
    V3 target = rxp;
    for (int i = (int)mp.size() - 1; i >= 0; i--) {
      int surf = img.at(i).first;   // .at(): out_of_range would escape, as in the reference
      V3 src = img.at(i).second;
      V3 d = sub(target, src);
      if (surf < 0) {
        mp[i].sourcePoint = src;
        mp[i].targetPoint = target;
        mp[i].pathLength = norm(sub(target, result.pointIntersect));
        mp[i].angleFromSurfaceNormal = kINF;
      } else {
        const Tri& t = MeshTri(surf);
        bool hh = t.RayIntersects(src, d, result);
        if (!hh) return false;
        mp[i].sourcePoint = result.pointIntersect;
        mp[i].targetPoint = target;
        mp[i].pathLength = norm(sub(target, result.pointIntersect));
        mp[i].angleFromSurfaceNormal = AngleBetween(t.norm, sub(target, result.pointIntersect), false, false);
        target = result.pointIntersect;
      }
    }
    return true;
  }
  // RTSolver.cpp:375-394
  // Extension (checkOcclusion): a corrected multi-bounce path is dropped when the closest hit along
  // one of its segments (source point towards target point, the line of sight's test) lies nearer
  // than the segment's end; a hit at the end is the next surface or lies beyond the receiver.
  bool Unobstructed(BVH& bvh, const Path& mp) const {
    for (const Ray& r : mp) {
      V3 d = sub(r.targetPoint, r.sourcePoint);
      HitInfo h;
      if (bvh.RayIntersects(0, r.sourcePoint, d, h) && h.distance < norm(d) * (1.0 - 1e-7)) return false;
    }
    return true;
  }

  bool DirectPathCorrection(BVH& bvh, Path& dp, V3 rxp) const {
    if (dp.size() != 1) return false;
    V3 txp = mk(p.tx[0], p.tx[1], p.tx[2]);
    HitInfo result;
    V3 d = sub(rxp, txp);
    bool hh = bvh.RayIntersects(0, txp, d, result);
    if (!hh || norm(d) < result.distance) {
      dp[0].sourcePoint = txp;
      dp[0].targetPoint = rxp;
      dp[0].pathLength = norm(sub(dp[0].targetPoint, dp[0].sourcePoint));
      return true;
    }
    return false;
  }

  // FieldCompute.cpp:153-183
  V3c Pattern(int type, double theta, double phi, double pt) const {
    V3c e = {cd(0, 0), cd(0, 0), cd(0, 0)};
    if (type == 3) {   // tabulated pattern (extension, see TablePattern)
      cd eth, eph;
      TablePattern(theta, phi, &eth, &eph);
      e.y = std::sqrt(pt) * eth;
      e.z = std::sqrt(pt) * eph;
      return e;
    }
    if (type == 0) { e.x = cd(0, std::sqrt(60 * pt)); e.y = 0; e.z = 0; }
    if (type == 1) { e.x = 0; e.y = cd(0, std::sqrt(60 * pt * 1.5) * std::sin(theta)); e.z = 0; }
    else if (type == 2) {
      e.x = 0;
      e.y = cd(0, std::sqrt(60 * pt) * (std::cos(kPI * std::cos(theta) / 2.0) / std::sin(theta)));
      e.z = 0;
    }
    return e;
  }
  // FieldCompute.cpp:294-302
  static cd TransAngle(double thetaI, cd epsI, cd epsT) {
    cd ti(thetaI, 0);
    cd ni = std::sqrt(epsI), nt = std::sqrt(epsT);
    return std::asin((ni / nt) * std::sin(ti));
  }
  // FieldCompute.cpp:237-292
  static void Coeffs(bool te, cd ti, cd tt, cd ei, cd et, double lamda, double width, bool infWall, cd& ref, cd& tran) {
    cd ni = std::sqrt(ei), nt = std::sqrt(et);
    cd etai = kETA0 / ni, etat = kETA0 / nt;
    cd ci = std::cos(ti), ct = std::cos(tt), si = std::sin(ti);
    cd g;
    if (te) g = -((etat / ct) - (etai / ci)) / ((etat / ct) + (etai / ci));
    else g = (-etat * ct + etai * ci) / (etat * ct + etai * ci);
    if (infWall) { ref = g; tran = 0; }
    else {
      cd q = (kTWOPI * width / lamda) * std::sqrt(nt - (si * si));
      cd ph = std::exp(-kJ * 2.0 * q);
      cd den = 1.0 - (g * g * ph);
      ref = (g * (1.0 - ph)) / den;
      tran = ((1.0 - g * g) * ph) / den;
    }
  }
  // Slab of thickness width (extension; ITU-R P.2040 eqs. for a dielectric slab in air): the
  // reference's slab branch (above, never reached with width > 0 by the reference) uses
  // sqrt(n_t - sin^2) for the phase thickness and e^{-2jq} on the transmitted wave; here
  // q = (2 pi d / lambda) sqrt(eps_t - sin^2 theta_i), R = G (1 - e^{-2jq}) / (1 - G^2 e^{-2jq}),
  // T = (1 - G^2) e^{-jq} / (1 - G^2 e^{-2jq}), with G the half-space coefficient above.
  static void SlabCoeffs(bool te, cd ti, cd tt, cd ei, cd et, double lamda, double width, cd& ref, cd& tran) {
    cd g, unusedT;
    Coeffs(te, ti, tt, ei, et, lamda, 0.0, true, g, unusedT);
    cd si = std::sin(ti);
    cd q = (kTWOPI * width / lamda) * std::sqrt(et - si * si);
    cd e1 = std::exp(-kJ * q), e2 = e1 * e1;
    cd den = 1.0 - g * g * e2;
    ref = g * (1.0 - e2) / den;
    tran = (1.0 - g * g) * e1 / den;
  }
  // FieldCompute.cpp:185-235
  V3c RefTrans(bool refl, const V3c& e, double relPerm, double sigma, double freq, double thetaI, double width,
               bool infWall) const {
    cd et(relPerm, -1 * sigma / (kE0 * 2 * kPI * freq));
    double lamda = kC0 / freq;
    cd ei(1, 0);
    cd tt = TransAngle(thetaI, ei, et);
    cd rTE, rTM, tTE, tTM;
    if (infWall && p.wallWidth > 0) {
      SlabCoeffs(false, cd(thetaI, 0), tt, ei, et, lamda, p.wallWidth, rTM, tTM);
      SlabCoeffs(true, cd(thetaI, 0), tt, ei, et, lamda, p.wallWidth, rTE, tTE);
    } else {
      Coeffs(false, cd(thetaI, 0), tt, ei, et, lamda, width, infWall, rTM, tTM);
      Coeffs(true, cd(thetaI, 0), tt, ei, et, lamda, width, infWall, rTE, tTE);
    }
    V3c r;
    r.x = e.x;
    if (refl) { r.y = e.y * rTM; r.z = e.z * rTE; }
    else { r.y = e.y * tTM; r.z = e.z * tTE; }
    return r;
  }
  // FieldCompute.cpp:71-151
  V3c FieldForPath(const Path& path) const {
    double lamda = kC0 / p.freq;
    double k = (2 * kPI) / lamda;
    M3 G = ident(), cur = G;
    M3 next = {{p.txCoordSys[0], p.txCoordSys[1], p.txCoordSys[2], p.txCoordSys[3], p.txCoordSys[4],
                p.txCoordSys[5], p.txCoordSys[6], p.txCoordSys[7], p.txCoordSys[8]}};
    V3c inc = {0, 0, 0}, tot = {0, 0, 0};
    double L = 0.0;
    for (size_t i = 0; i < path.size(); i++) {
      const Ray& ray = path[i];
      V3 vg = sub(ray.targetPoint, ray.sourcePoint);
      V3 vl = mv(RotationMatrix(G, next), vg);
      V3 sph = CartesianToSpherical(vl);
      double r = sph.x, theta = sph.y, phi = sph.z;
      L += r;
      cur = next;
      if (ray.hitSurfaceID == -1) next = G;
      else next = MeshTri(ray.hitSurfaceID).coordSys;
      if (p.physicalField && i > 0) {
        // Extension: the global field split at the interaction into its components along s (normal
        // to the plane of incidence) and p (in it: s x direction, before and after), with the
        // Fresnel coefficients of the incidence angle in [0, 90] degrees (the reference's TE
        // coefficient is minus the usual s one for a reflection, its TM one the usual p one).
        V3 din = normalized(sub(path[i - 1].targetPoint, path[i - 1].sourcePoint)), dout = normalized(vg);
        V3 nrm = normalized(mk(cur.m[6], cur.m[7], cur.m[8]));
        double cth = std::fabs(dot(din, nrm));
        if (cth > 1.0) cth = 1.0;
        V3 s = cross(din, nrm);
        if (norm(s) < 1e-12) {
          V3 ax = std::fabs(din.x) <= std::fabs(din.y) && std::fabs(din.x) <= std::fabs(din.z) ? mk(1, 0, 0)
                  : (std::fabs(din.y) <= std::fabs(din.z) ? mk(0, 1, 0) : mk(0, 0, 1));
          s = cross(din, ax);
        }
        s = normalized(s);
        V3 pin = cross(s, din), pout = cross(s, dout);
        bool refl = ray.reflectionMaterialId >= 0;
        const double* m = &mats[6 * (refl ? ray.reflectionMaterialId : ray.penetrationMaterialId)];
        V3c unit = {cd(0, 0), cd(1, 0), cd(1, 0)};
        V3c c = RefTrans(refl, unit, m[3], m[5], p.freq, std::acos(cth), 0, p.useFresnel != 0);
        cd cs = refl ? -c.z : c.z, cp = c.y;
        cd es = tot.x * s.x + tot.y * s.y + tot.z * s.z, ep = tot.x * pin.x + tot.y * pin.y + tot.z * pin.z;
        cd as = cs * es, ap = cp * ep;
        tot.x = as * s.x + ap * pout.x;
        tot.y = as * s.y + ap * pout.y;
        tot.z = as * s.z + ap * pout.z;
        continue;
      }
      if (i == 0) inc = Pattern(p.antennaType, theta, phi, p.pt);
      else {
        inc = mvc(RotationMatrix(cur, G), tot);
        inc = mvc(TransformationMatrix(theta, phi, false), inc);
      }
      V3 n2 = mk(cur.m[6], cur.m[7], cur.m[8]);
      if (ray.reflectionMaterialId >= 0) {
        double ang = AngleBetween(vg, n2, false, true);
        const double* m = &mats[6 * ray.reflectionMaterialId];
        tot = RefTrans(true, inc, m[3], m[5], p.freq, ang, 0, p.useFresnel != 0);
      } else if (ray.penetrationMaterialId >= 0) {
        double ang = AngleBetween(vg, n2, false, true);
        const double* m = &mats[6 * ray.penetrationMaterialId];
        tot = RefTrans(false, inc, m[3], m[5], p.freq, ang, 0, p.useFresnel != 0);
      } else {
        tot = inc;
      }
      tot = mvc(TransformationMatrix(theta, phi, true), tot);
      tot = mvc(RotationMatrix(cur, G), tot);
    }
    cd ph = std::exp(-kJ * k * L);
    tot.x = tot.x * ph / L; tot.y = tot.y * ph / L; tot.z = tot.z * ph / L;
    return tot;
  }
};

struct Result {
  long long bounces = 0;
  long long nodeVisits = 0;
  long long canonNodes = 0, canonTris = 0;   // counting traversal (countCanonical)
  long long captureRecords = 0;
  std::vector<double> fields;   // 6 per receiver
  // kept paths, rx ascending, launch ascending, BFS order
  std::vector<int> prx, plaunch, pnseg, poff;
  std::vector<int> stri, scop, stype;
  std::vector<double> ssrc, stgt;
  std::vector<double> pfield;   // 6 per kept path
  double seconds = 0.0;
};

// RTSolver.cpp:52-126 (ExecuteRayTracing), restructured so that only captured paths are stored:
// the per-(launch, receiver) path lists are kept per receiver instead of as an N_rays x N_rx array.
static Result* Trace(Scene* sc, const Params& p, const double* rx, int nrx, const double* mats, int nmat,
                     int nthreads) {
  auto t0 = std::chrono::steady_clock::now();
  Result* res = new Result();
  Tracer tr;
  tr.sc = sc; tr.p = p; tr.rx = rx; tr.nrx = nrx;
  tr.mats.assign(mats, mats + 6 * nmat);
  tr.txTess = p.tess + 1;   // RTSolver.cpp:70
  if (p.receptionModel == 0) {
    double beta = 1.0 / (3.0 * tr.txTess) * std::acos(-1.0 / std::sqrt(5.0));
    tr.rcoef = std::tan(beta);
  } else {
    double alpha = (69.0 * kPI / 180.0) / std::ldexp(1.0, p.tess);
    tr.rcoef = alpha / std::sqrt(3.0);
  }
  std::vector<V3> dirs = GenerateRaysOnIcosahedron(p.tess);
  const int nl = (int)dirs.size();
  const int stride = p.launchStride > 0 ? p.launchStride : 1;
  std::vector<int> launches;
  for (int i = 0; i < nl; i++) if (i % stride == p.launchOffset) launches.push_back(i);
  const int nsel = (int)launches.size();
  // per selected launch: list of (rx, paths)
  std::vector<std::vector<std::pair<int, std::vector<Path>>>> perLaunch(nsel);
  std::vector<long long> bl(nsel, 0), nv(nsel, 0), cn(nsel, 0), ct(nsel, 0);
  V3 txp = mk(p.tx[0], p.tx[1], p.tx[2]);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    BVH bvh = sc->bvh;   // per-thread copy: only the counter is mutated
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int li = 0; li < nsel; li++) {
      int i = launches[li];
      PTN root;
      long long b = 0;
      long long v0 = bvh.nodeVisits, c0 = bvh.canonNodes, t0c = bvh.canonTris;
      tr.RayLaunch(bvh, &root, txp, dirs[i], -1, -1, 0, 0, 0, true, kINF, 0.0, b);
      bl[li] = b;
      for (int j = 0; j < nrx; j++) {
        PTN* cr = CloneNode(&root);
        CloneTree(&root, cr);
        tr.RayCapture(cr, mk(rx[3 * j], rx[3 * j + 1], rx[3 * j + 2]), 0.0);
        std::vector<Path> paths;
        TreePaths(cr, paths);
        if (!paths.empty()) perLaunch[li].push_back(std::make_pair(j, std::move(paths)));
        cr = DeleteChildNodes(cr);
      }
      DeleteChildNodes(root.childTransmit);
      DeleteChildNodes(root.childReflect);
      nv[li] = bvh.nodeVisits - v0;
      cn[li] = bvh.canonNodes - c0;
      ct[li] = bvh.canonTris - t0c;
    }
  }
  for (int li = 0; li < nsel; li++) {
    res->bounces += bl[li];
    res->nodeVisits += nv[li];
    res->canonNodes += cn[li];
    res->canonTris += ct[li];
  }
  // per receiver: (launch, path) in launch order
  std::vector<std::vector<std::pair<int, Path>>> perRx(nrx);
  for (int li = 0; li < nsel; li++)
    for (auto& e : perLaunch[li])
      for (auto& pth : e.second) { perRx[e.first].push_back(std::make_pair(launches[li], pth)); res->captureRecords++; }
  perLaunch.clear();
  res->fields.assign(6 * (size_t)nrx, 0.0);
  std::vector<std::vector<std::pair<int, Path>>> kept(nrx);
  std::vector<std::vector<V3c>> keptField(nrx);
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    BVH bvh = sc->bvh;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
    for (int k = 0; k < nrx; k++) {
      V3 rxp = mk(rx[3 * k], rx[3 * k + 1], rx[3 * k + 2]);
      // ImagePathCorrection (RTSolver.cpp:299-328)
      std::vector<std::pair<int, Path>> surv;
      for (auto& lp : perRx[k]) {
        Path pth = lp.second;
        bool ok = pth.size() == 1 ? tr.DirectPathCorrection(bvh, pth, rxp) : tr.MultiPathCorrection(pth, rxp);
        if (ok && p.checkOcclusion && pth.size() > 1) ok = tr.Unobstructed(bvh, pth);
        if (ok) surv.push_back(std::make_pair(lp.first, pth));
      }
      // RemoveDuplicatePath (RTSolver.cpp:435-459): first (launch, BFS) occurrence of a coplanar key wins
      std::map<std::vector<int>, int> seen;
      for (auto& lp : surv) {
        std::vector<int> key;
        for (const Ray& r : lp.second) key.push_back(r.hitCoplanarId);
        if (p.typedDedup)   // extension: then how each segment left its start surface
          for (const Ray& r : lp.second)
            key.push_back(r.reflectionMaterialId >= 0 ? 1 : (r.penetrationMaterialId >= 0 ? 2 : 0));
        if (seen.count(key)) continue;
        seen.insert(std::make_pair(key, 0));
        kept[k].push_back(lp);
      }
      // FieldAtReceiver (FieldCompute.cpp:11-25)
      V3c tot = {0, 0, 0};
      for (auto& lp : kept[k]) {
        V3c f = tr.FieldForPath(lp.second);
        keptField[k].push_back(f);
        tot.x += f.x; tot.y += f.y; tot.z += f.z;
      }
      double* o = &res->fields[6 * (size_t)k];
      o[0] = tot.x.real(); o[1] = tot.x.imag(); o[2] = tot.y.real(); o[3] = tot.y.imag();
      o[4] = tot.z.real(); o[5] = tot.z.imag();
    }
  }
  for (int k = 0; k < nrx; k++) {
    for (const V3c& f : keptField[k]) {
      double v[6] = {f.x.real(), f.x.imag(), f.y.real(), f.y.imag(), f.z.real(), f.z.imag()};
      res->pfield.insert(res->pfield.end(), v, v + 6);
    }
    for (auto& lp : kept[k]) {
      res->prx.push_back(k);
      res->plaunch.push_back(lp.first);
      res->pnseg.push_back((int)lp.second.size());
      res->poff.push_back((int)res->stri.size());
      for (const Ray& r : lp.second) {
        res->stri.push_back(r.hitSurfaceID);
        res->scop.push_back(r.hitCoplanarId);
        res->stype.push_back(r.reflectionMaterialId >= 0 ? 1 : (r.penetrationMaterialId >= 0 ? 2 : 0));
        res->ssrc.push_back(r.sourcePoint.x); res->ssrc.push_back(r.sourcePoint.y); res->ssrc.push_back(r.sourcePoint.z);
        res->stgt.push_back(r.targetPoint.x); res->stgt.push_back(r.targetPoint.y); res->stgt.push_back(r.targetPoint.z);
      }
    }
  }
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

// Interaction-tree dump (BFS per launch, T before R): one row per RayLaunch call.
struct TreeDump {
  std::vector<int> launch, depth, tri, cop, type;
  std::vector<long long> code;
  std::vector<double> src, tgt, pl;
};
static TreeDump* DumpTrees(Scene* sc, const Params& p) {
  TreeDump* d = new TreeDump();
  Tracer tr;
  tr.sc = sc; tr.p = p;
  std::vector<V3> dirs = GenerateRaysOnIcosahedron(p.tess);
  const int stride = p.launchStride > 0 ? p.launchStride : 1;
  V3 txp = mk(p.tx[0], p.tx[1], p.tx[2]);
  BVH bvh = sc->bvh;
  for (int i = 0; i < (int)dirs.size(); i++) {
    if (i % stride != p.launchOffset) continue;
    PTN root;
    long long b = 0;
    tr.RayLaunch(bvh, &root, txp, dirs[i], -1, -1, 0, 0, 0, true, kINF, 0.0, b);
    std::queue<std::pair<PTN*, std::pair<int, long long>>> q;
    q.push(std::make_pair(&root, std::make_pair(0, 0LL)));
    while (!q.empty()) {
      auto e = q.front();
      q.pop();
      PTN* n = e.first;
      d->launch.push_back(i); d->depth.push_back(e.second.first); d->code.push_back(e.second.second);
      d->tri.push_back(n->ray.hitSurfaceID); d->cop.push_back(n->ray.hitCoplanarId);
      d->type.push_back(n->ray.reflectionMaterialId >= 0 ? 1 : (n->ray.penetrationMaterialId >= 0 ? 2 : 0));
      d->src.push_back(n->ray.sourcePoint.x); d->src.push_back(n->ray.sourcePoint.y); d->src.push_back(n->ray.sourcePoint.z);
      d->tgt.push_back(n->ray.targetPoint.x); d->tgt.push_back(n->ray.targetPoint.y); d->tgt.push_back(n->ray.targetPoint.z);
      d->pl.push_back(n->ray.pathLength);
      if (n->childTransmit) q.push(std::make_pair(n->childTransmit, std::make_pair(e.second.first + 1, e.second.second * 2)));
      if (n->childReflect) q.push(std::make_pair(n->childReflect, std::make_pair(e.second.first + 1, e.second.second * 2 + 1)));
    }
    DeleteChildNodes(root.childTransmit);
    DeleteChildNodes(root.childReflect);
  }
  return d;
}

}  // namespace orc

// =============================================================================================
// C interface (ctypes)
// =============================================================================================
using namespace orc;
// memcpy for output arrays that may be empty (a null pointer with a zero size is undefined
// behaviour for memcpy itself).
static void copy_out(void* dst, const void* src, size_t n) {
  if (n > 0) memcpy(dst, src, n);
}

extern "C" {

typedef struct {
  double tx[3];
  double txCoordSys[9];
  int maxR, maxT, maxTotal;
  double freq, pt;
  int useFresnel, tess, antennaType, receptionModel, launchStride, launchOffset, q1MeshIndexing;
  int countCanonical;
  double wallWidth;
  int typedDedup;
  int checkOcclusion;
  int longMiss;
  int physicalField;
} orc_params;

void* orc_scene_from_stl(const char* path) {
  std::vector<float> c, n;
  std::vector<unsigned> t;
  if (!ReadStl(path, c, n, t)) return nullptr;
  return BuildScene(c, n, t, nullptr);
}

// Triangle soup as a binary STL carries it: 9 floats of vertices and 3 floats of normal per
// facet (StlReader.h:651-672); optional per-triangle material ids (indexed like the normals).
void* orc_scene_from_soup(const float* coords9, const float* normals3, const int* mats, int ntri) {
  std::vector<CWI> cw;
  std::vector<float> c, n(normals3, normals3 + 3 * (size_t)ntri);
  std::vector<unsigned> t;
  for (int i = 0; i < ntri; i++) {
    for (int v = 0; v < 3; v++) {
      CWI x;
      for (int k = 0; k < 3; k++) x.d[k] = coords9[9 * (size_t)i + 3 * v + k];
      x.index = (unsigned)cw.size();
      cw.push_back(x);
    }
    t.push_back((unsigned)cw.size() - 3); t.push_back((unsigned)cw.size() - 2); t.push_back((unsigned)cw.size() - 1);
  }
  if (cw.empty()) return nullptr;
  RemoveDoubles(c, t, cw);
  return BuildScene(c, n, t, mats);
}

// A scene over given triangles and a given BVH (not built here): triangles by id (vertices,
// file normal, coplanar and material ids), the mesh order (meshTriId[k] = id at position k), and the
// BVH as nodes in DFS order with node 0 the root: box, and either children (a, b) or a leaf's
// triangle range [a, a + b) of the DFS triangle list triDfs.  Used to run the reference's traversal
// and capture over a tree the library built on the device.  Returns nullptr on malformed input.
void* orc_scene_from_arrays(int ntri, const double* verts9, const double* norm3, const int* cop, const int* mat,
                            const int* meshTriId, int nnode, const double* box6, const int* a, const int* b,
                            const int* leaf, const int* triDfs) {
  if (ntri <= 0 || nnode <= 0) return nullptr;
  Scene* sc = new Scene();
  for (int i = 0; i < ntri; i++) {
    const double* v = verts9 + 9 * (size_t)i;
    Tri* t = new Tri();
    t->init(mk(v[0], v[1], v[2]), mk(v[3], v[4], v[5]), mk(v[6], v[7], v[8]),
            mk(norm3[3 * i], norm3[3 * i + 1], norm3[3 * i + 2]), i);
    t->coplanarId = cop[i];
    t->materialId = mat[i];
    sc->owned.push_back(t);
  }
  for (int k = 0; k < ntri; k++) {
    if (meshTriId[k] < 0 || meshTriId[k] >= ntri) { delete sc; return nullptr; }
    sc->mesh.push_back(sc->owned[meshTriId[k]]);
  }
  sc->bvh.nodes.assign(nnode, BNode());
  for (int k = 0; k < nnode; k++) {
    BNode& n = sc->bvh.nodes[k];
    for (int q = 0; q < 3; q++) { n.bmin[q] = box6[6 * k + q]; n.bmax[q] = box6[6 * k + 3 + q]; }
    if (leaf[k]) {
      if (a[k] < 0 || b[k] <= 0 || a[k] + b[k] > ntri) { delete sc; return nullptr; }
      for (int j = a[k]; j < a[k] + b[k]; j++) n.objects.push_back(sc->owned[triDfs[j]]);
    } else {
      if (a[k] <= k || b[k] <= k || a[k] >= nnode || b[k] >= nnode) { delete sc; return nullptr; }
      n.child[0] = a[k];
      n.child[1] = b[k];
    }
  }
  return sc;
}

void orc_scene_free(void* s) { delete (Scene*)s; }
int orc_scene_ntri(void* s) { return (int)((Scene*)s)->owned.size(); }
int orc_scene_nnodes(void* s) { return (int)((Scene*)s)->bvh.nodes.size(); }
int orc_scene_nverts(void* s) { return (int)((Scene*)s)->coords.size() / 3; }
int orc_scene_nedges(void* s) { return (int)((Scene*)s)->edges.size(); }

// Per triangle id: coplanar id, material id, vertices (9), normal (3); mesh order (Q1 permutation):
// meshTriId[k] = triangleId of the triangle at position k of the BVH-sorted caller vector.
void orc_scene_triangles(void* sp, int* coplanar, int* material, double* verts9, double* norm3, int* meshTriId) {
  Scene* s = (Scene*)sp;
  for (size_t i = 0; i < s->owned.size(); i++) {
    Tri* t = s->owned[i];
    coplanar[i] = t->coplanarId;
    material[i] = t->materialId;
    double v[9] = {t->v1.x, t->v1.y, t->v1.z, t->v2.x, t->v2.y, t->v2.z, t->v3.x, t->v3.y, t->v3.z};
    copy_out(verts9 + 9 * i, v, sizeof v);
    norm3[3 * i] = t->norm.x; norm3[3 * i + 1] = t->norm.y; norm3[3 * i + 2] = t->norm.z;
    meshTriId[i] = s->mesh[i]->triangleId;
  }
}
// Unique vertex coordinates and triangle indices after RemoveDoubles (Faces.vtk content).
void orc_scene_geometry(void* sp, float* coords, unsigned* tris) {
  Scene* s = (Scene*)sp;
  copy_out(coords, s->coords.data(), s->coords.size() * sizeof(float));
  copy_out(tris, s->tris.data(), s->tris.size() * sizeof(unsigned));
}
// Edge map in key order: (vertex index pair, connection count) (Edges.vtk content).
void orc_scene_edges(void* sp, int* pairs, int* counts) {
  Scene* s = (Scene*)sp;
  for (size_t i = 0; i < s->edges.size(); i++) {
    pairs[2 * i] = s->edges[i].first.first / 3;
    pairs[2 * i + 1] = s->edges[i].first.second / 3;
    counts[i] = s->edges[i].second;
  }
}
// BVH nodes in creation order: box (6), children (2), leaf object count, and the DFS leaf order of
// triangle ids (children[0] first) in leafTri (capacity ntri).
void orc_scene_bvh(void* sp, double* box6, int* child2, int* nobj, int* leafTri) {
  Scene* s = (Scene*)sp;
  for (size_t i = 0; i < s->bvh.nodes.size(); i++) {
    const BNode& n = s->bvh.nodes[i];
    for (int a = 0; a < 3; a++) { box6[6 * i + a] = n.bmin[a]; box6[6 * i + 3 + a] = n.bmax[a]; }
    child2[2 * i] = n.child[0]; child2[2 * i + 1] = n.child[1];
    nobj[i] = (int)n.objects.size();
  }
  int k = 0;
  std::vector<int> st{0};
  while (!st.empty()) {
    int id = st.back();
    st.pop_back();
    const BNode& n = s->bvh.nodes[id];
    if (n.child[0] == -1) { for (Tri* t : n.objects) leafTri[k++] = t->triangleId; }
    else { st.push_back(n.child[1]); st.push_back(n.child[0]); }
  }
}

// Closest hit of one ray (reference traversal): returns triangle id or -1; point and distance out.
int orc_closest_hit(void* sp, const double* org, const double* dir, double* point, double* dist) {
  Scene* s = (Scene*)sp;
  HitInfo h;
  bool hh = s->bvh.RayIntersects(0, mk(org[0], org[1], org[2]), mk(dir[0], dir[1], dir[2]), h);
  if (!hh) return -1;
  point[0] = h.pointIntersect.x; point[1] = h.pointIntersect.y; point[2] = h.pointIntersect.z;
  *dist = h.distance;
  return h.triangleId;
}
// Brute-force closest hit over all triangles in triangle-id order (strict '<').
int orc_closest_hit_brute(void* sp, const double* org, const double* dir, double* point, double* dist) {
  Scene* s = (Scene*)sp;
  HitInfo best;
  bool any = false;
  for (Tri* t : s->owned) {
    HitInfo h;
    if (t->RayIntersects(mk(org[0], org[1], org[2]), mk(dir[0], dir[1], dir[2]), h) && h.distance < best.distance) {
      best = h;
      any = true;
    }
  }
  if (!any) return -1;
  point[0] = best.pointIntersect.x; point[1] = best.pointIntersect.y; point[2] = best.pointIntersect.z;
  *dist = best.distance;
  return best.triangleId;
}

long long orc_icosahedron(int tess, double* out, long long cap) {
  std::vector<V3> d = GenerateRaysOnIcosahedron(tess);
  long long n = (long long)d.size();
  for (long long i = 0; i < n && i < cap; i++) { out[3 * i] = d[i].x; out[3 * i + 1] = d[i].y; out[3 * i + 2] = d[i].z; }
  return n;
}

void* orc_trace(void* sp, const orc_params* pp, const double* rx, int nrx, const double* mats, int nmat, int nthreads) {
  Params p;
  copy_out(p.tx, pp->tx, sizeof p.tx);
  copy_out(p.txCoordSys, pp->txCoordSys, sizeof p.txCoordSys);
  p.maxR = pp->maxR; p.maxT = pp->maxT; p.maxTotal = pp->maxTotal;
  p.freq = pp->freq; p.pt = pp->pt; p.useFresnel = pp->useFresnel; p.tess = pp->tess;
  p.antennaType = pp->antennaType; p.receptionModel = pp->receptionModel;
  p.launchStride = pp->launchStride; p.launchOffset = pp->launchOffset; p.q1MeshIndexing = pp->q1MeshIndexing;
  p.countCanonical = pp->countCanonical;
  p.wallWidth = pp->wallWidth;
  p.typedDedup = pp->typedDedup;
  p.checkOcclusion = pp->checkOcclusion;
  p.longMiss = pp->longMiss;
  p.physicalField = pp->physicalField;
  return Trace((Scene*)sp, p, rx, nrx, mats, nmat, nthreads);
}
void orc_result_free(void* r) { delete (Result*)r; }
long long orc_result_bounces(void* r) { return ((Result*)r)->bounces; }
long long orc_result_node_visits(void* r) { return ((Result*)r)->nodeVisits; }
long long orc_result_canon_nodes(void* r) { return ((Result*)r)->canonNodes; }
long long orc_result_canon_tris(void* r) { return ((Result*)r)->canonTris; }
long long orc_result_capture_records(void* r) { return ((Result*)r)->captureRecords; }
double orc_result_seconds(void* r) { return ((Result*)r)->seconds; }
void orc_result_fields(void* r, double* out) {
  Result* x = (Result*)r;
  copy_out(out, x->fields.data(), x->fields.size() * sizeof(double));
}
int orc_result_npaths(void* r) { return (int)((Result*)r)->prx.size(); }
int orc_result_nsegs(void* r) { return (int)((Result*)r)->stri.size(); }
void orc_result_paths(void* r, int* prx, int* plaunch, int* pnseg, int* poff, int* stri, int* scop, int* stype,
                      double* ssrc, double* stgt) {
  Result* x = (Result*)r;
  size_t np = x->prx.size(), ns = x->stri.size();
  copy_out(prx, x->prx.data(), np * sizeof(int)); copy_out(plaunch, x->plaunch.data(), np * sizeof(int));
  copy_out(pnseg, x->pnseg.data(), np * sizeof(int)); copy_out(poff, x->poff.data(), np * sizeof(int));
  copy_out(stri, x->stri.data(), ns * sizeof(int)); copy_out(scop, x->scop.data(), ns * sizeof(int));
  copy_out(stype, x->stype.data(), ns * sizeof(int));
  copy_out(ssrc, x->ssrc.data(), 3 * ns * sizeof(double)); copy_out(stgt, x->stgt.data(), 3 * ns * sizeof(double));
}

void orc_result_path_fields(void* r, double* out) {
  Result* x = (Result*)r;
  copy_out(out, x->pfield.data(), x->pfield.size() * sizeof(double));
}

void* orc_trees(void* sp, const orc_params* pp) {
  Params p;
  copy_out(p.tx, pp->tx, sizeof p.tx);
  copy_out(p.txCoordSys, pp->txCoordSys, sizeof p.txCoordSys);
  p.maxR = pp->maxR; p.maxT = pp->maxT; p.maxTotal = pp->maxTotal;
  p.freq = pp->freq; p.pt = pp->pt; p.useFresnel = pp->useFresnel; p.tess = pp->tess;
  p.antennaType = pp->antennaType; p.receptionModel = pp->receptionModel;
  p.launchStride = pp->launchStride; p.launchOffset = pp->launchOffset; p.q1MeshIndexing = pp->q1MeshIndexing;
  p.wallWidth = pp->wallWidth;
  p.typedDedup = pp->typedDedup;
  p.checkOcclusion = pp->checkOcclusion;
  p.longMiss = pp->longMiss;
  p.physicalField = pp->physicalField;
  return DumpTrees((Scene*)sp, p);
}
int orc_trees_n(void* d) { return (int)((TreeDump*)d)->launch.size(); }
void orc_trees_get(void* dp, int* launch, int* depth, long long* code, int* tri, int* cop, int* type, double* src,
                   double* tgt, double* pl) {
  TreeDump* d = (TreeDump*)dp;
  size_t n = d->launch.size();
  copy_out(launch, d->launch.data(), n * sizeof(int)); copy_out(depth, d->depth.data(), n * sizeof(int));
  copy_out(code, d->code.data(), n * sizeof(long long)); copy_out(tri, d->tri.data(), n * sizeof(int));
  copy_out(cop, d->cop.data(), n * sizeof(int)); copy_out(type, d->type.data(), n * sizeof(int));
  copy_out(src, d->src.data(), 3 * n * sizeof(double)); copy_out(tgt, d->tgt.data(), 3 * n * sizeof(double));
  copy_out(pl, d->pl.data(), n * sizeof(double));
}
void orc_trees_free(void* d) { delete (TreeDump*)d; }

// Tabulated pattern for antennaType 3 (rows as in sbr_set_tx_pattern); not thread-safe against a
// running trace.
int orc_set_pattern(const float* gain_info, int n_theta, int n_phi) {
  if (!gain_info || n_theta < 2 || n_phi < 1) return 1;
  orc::SetTablePattern(gain_info, n_theta, n_phi);
  return 0;
}
// Field of the tabulated pattern at (theta, phi) radians for power pt: out = theta re, im, phi re, im.
void orc_pattern_eval(double theta, double phi, double pt, double* out) {
  std::complex<double> a, b;
  orc::TablePattern(theta, phi, &a, &b);
  a *= std::sqrt(pt);
  b *= std::sqrt(pt);
  out[0] = a.real(); out[1] = a.imag(); out[2] = b.real(); out[3] = b.imag();
}

// Fresnel coefficients for tests (FieldCompute.cpp:237-302): out = [rTM, tTM, rTE, tTE] as re,im pairs.
void orc_fresnel(double thetaI, double relPerm, double sigma, double freq, double width, int infWall, double* out) {
  cd et(relPerm, -1 * sigma / (kE0 * 2 * kPI * freq));
  cd ei(1, 0);
  double lamda = kC0 / freq;
  cd tt = Tracer::TransAngle(thetaI, ei, et);
  cd rTE, rTM, tTE, tTM;
  if (infWall == 2) {   // slab of the given width (extension)
    Tracer::SlabCoeffs(false, cd(thetaI, 0), tt, ei, et, lamda, width, rTM, tTM);
    Tracer::SlabCoeffs(true, cd(thetaI, 0), tt, ei, et, lamda, width, rTE, tTE);
  } else {
    Tracer::Coeffs(false, cd(thetaI, 0), tt, ei, et, lamda, width, infWall != 0, rTM, tTM);
    Tracer::Coeffs(true, cd(thetaI, 0), tt, ei, et, lamda, width, infWall != 0, rTE, tTE);
  }
  cd v[4] = {rTM, tTM, rTE, tTE};
  for (int i = 0; i < 4; i++) { out[2 * i] = v[i].real(); out[2 * i + 1] = v[i].imag(); }
}

}  // extern "C"
