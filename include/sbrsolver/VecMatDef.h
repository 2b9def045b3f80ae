// Value types of the reference's call surface (common/VecMatDef.h:11-49, DataStructures.h:9-37)
// without Eigen: Vec3 is 3 contiguous doubles and Mat3 9 doubles in column-major order, the memory
// layout of Eigen::Vector3d / Eigen::Matrix3d, so arrays of them can be handed to the C-ABI as is.
#pragma once

#include <complex>
#include <cstdint>
#include <limits>
#include <vector>

#include "sbr_math.h"

typedef std::complex<double> cdouble;

struct Vec3 {
  double d[3];
  Vec3() : d{0, 0, 0} {}
  Vec3(double x, double y, double z) : d{x, y, z} {}
  double& operator()(int i) { return d[i]; }
  double operator()(int i) const { return d[i]; }
  double& operator[](int i) { return d[i]; }
  double operator[](int i) const { return d[i]; }
  double x() const { return d[0]; }
  double y() const { return d[1]; }
  double z() const { return d[2]; }
  sbr::V3 v() const { return sbr::v3(d[0], d[1], d[2]); }
  static Vec3 of(sbr::V3 a) { return Vec3(a.x, a.y, a.z); }
  bool operator==(const Vec3& o) const { return d[0] == o.d[0] && d[1] == o.d[1] && d[2] == o.d[2]; }
  Vec3 operator-() const { return Vec3(-d[0], -d[1], -d[2]); }
  double norm() const { return sbr::vnorm(v()); }
};
typedef std::vector<Vec3> VecVec3;

struct Vec3c {
  cdouble d[3];
  Vec3c() : d{0.0, 0.0, 0.0} {}
  Vec3c(cdouble x, cdouble y, cdouble z) : d{x, y, z} {}
  cdouble& operator()(int i) { return d[i]; }
  cdouble operator()(int i) const { return d[i]; }
  cdouble x() const { return d[0]; }
  cdouble y() const { return d[1]; }
  cdouble z() const { return d[2]; }
  double norm() const {
    return std::sqrt((std::norm(d[0]) + std::norm(d[1])) + std::norm(d[2]));
  }
};

struct Mat3 {
  double m[9];  // column-major
  Mat3() : m{0, 0, 0, 0, 0, 0, 0, 0, 0} {}
  static Mat3 Identity() {
    Mat3 r;
    r(0, 0) = r(1, 1) = r(2, 2) = 1.0;
    return r;
  }
  // Row-wise brace construction as Eigen's Matrix3d{{..},{..},{..}}.
  Mat3(std::initializer_list<std::initializer_list<double>> rows) : Mat3() {
    int r = 0;
    for (auto& row : rows) {
      int c = 0;
      for (double v : row) { (*this)(r, c++) = v; }
      r++;
    }
  }
  double& operator()(int r, int c) { return m[c * 3 + r]; }
  double operator()(int r, int c) const { return m[c * 3 + r]; }
  void to_row_major(double* out) const {
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) out[r * 3 + c] = (*this)(r, c);
  }
};

struct Mat23 {
  double m[6];  // column-major, row 0 = min, row 1 = max
  double& operator()(int r, int c) { return m[c * 2 + r]; }
  double operator()(int r, int c) const { return m[c * 2 + r]; }
};

struct Idx2 {
  int d[2] = {-1, -1};
  int& operator[](int i) { return d[i]; }
  int operator[](int i) const { return d[i]; }
};

// common/DataStructures.h:9-17
struct HitInfo {
  Vec3 normal;
  Vec3 pointIntersect;
  int materialId = -1;
  int triangelId = -1;
  int coplanarId = -1;
  double distance = std::numeric_limits<double>::infinity();
};

// common/DataStructures.h:20-28
struct MaterialProperties {
  double frequency;
  double transmissionLoss;
  double reflectionLoss;
  double relPermittivityRe;
  double relPermittivityIm;
  double relConductivity;
};
using Materials = std::vector<MaterialProperties>;
