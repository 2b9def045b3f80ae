// Triangle record of the reference call surface (Triangle.h:8-49, Triangle.cpp:28-125).
#pragma once

#include <algorithm>
#include <cfloat>

#include "VecMatDef.h"

class Triangle {
 public:
  Triangle() {
    for (int a = 0; a < 3; a++) { bbox(0, a) = FLT_MAX; bbox(1, a) = -FLT_MAX; }
  }
  // v1, v2, v3, file normal, triangleId (Triangle.cpp:28-35)
  Triangle(Vec3 a, Vec3 b, Vec3 c, Vec3 n, int id) : v1(a), v2(b), v3(c), norm(n), triangleId(id) {
    bbox = findBbox();
    center = findCenter();
    coordSys = findCoordSys();
  }

  Vec3 v1, v2, v3;
  Vec3 norm;       // the normal given with the facet, not recomputed
  Mat3 coordSys;   // rows: x = y cross n, y = normalized(v2 - v1), z = n
  Mat23 bbox;
  Vec3 center;
  int triangleId = -1;
  int coplanarId = -1;
  int materialId = 0;

  Mat3 findCoordSys() const {
    sbr::V3 zw = norm.v();
    sbr::V3 yw = sbr::vnormalized(sbr::vsub(v2.v(), v1.v()));
    sbr::V3 xw = sbr::vcross(yw, zw);
    Mat3 s;
    s(0, 0) = xw.x; s(0, 1) = xw.y; s(0, 2) = xw.z;
    s(1, 0) = yw.x; s(1, 1) = yw.y; s(1, 2) = yw.z;
    s(2, 0) = zw.x; s(2, 1) = zw.y; s(2, 2) = zw.z;
    return s;
  }
  Mat23 findBbox() const {
    Mat23 b;
    for (int a = 0; a < 3; a++) {
      b(0, a) = std::min({v1(a), v2(a), v3(a)});
      b(1, a) = std::max({v1(a), v2(a), v3(a)});
    }
    return b;
  }
  Vec3 findCenter() const {
    Vec3 c;
    for (int a = 0; a < 3; a++)
      c(a) = (std::min({v1(a), v2(a), v3(a)}) + std::max({v1(a), v2(a), v3(a)})) / 2.0;
    return c;
  }
  bool RayIntersects(const Vec3& rayOrg, const Vec3& rayDir, HitInfo& info) const {
    sbr::V3 p;
    double dist;
    if (!sbr::tri_intersect(v1.v(), sbr::vsub(v2.v(), v1.v()), sbr::vsub(v3.v(), v1.v()), rayOrg.v(), rayDir.v(),
                            &p, &dist))
      return false;
    info.pointIntersect = Vec3::of(p);
    info.distance = dist;
    info.normal = norm;
    info.materialId = materialId;
    info.triangelId = triangleId;
    info.coplanarId = coplanarId;
    return true;
  }
};
