// SAH bounding-volume hierarchy of the reference call surface (BVH/BVH.h:13-37, BVHNode.h:8-25).
//
// The build is the reference's full-sweep SAH (BVH/BVH.h:135-230): for each axis the object list is
// sorted by centre, prefix / suffix box areas are swept, and split i (1 <= i <= n-2) costs
// 2 + (i * A(objects[0..i]) + (n-i) * A(objects[i..n-1])) / A(node); a leaf is made when no split
// beats n.  Like the reference, ConstructBVH sorts the caller's vector in place; the hot path then
// indexes that vector by triangleId, which the device reproduces through mesh_tri_id.  The
// sequence of std::sort calls is kept because the final order of equal-centre objects decides both
// the tree and that permutation.
#pragma once

#include <algorithm>
#include <cfloat>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "VecMatDef.h"

template <typename T>
class BVHNode {
 public:
  BVHNode() { EmptyAABB(bbox); }
  Mat23 bbox;
  Idx2 children;
  std::vector<T*> objects;

  static void EmptyAABB(Mat23& b) {
    for (int a = 0; a < 3; a++) { b(0, a) = FLT_MAX; b(1, a) = -FLT_MAX; }
  }
  // The box's eight corners, x varying fastest, then y, then z (BVHNode.h:39-51).
  VecVec3 CornerPoints() const {
    VecVec3 v(8);
    for (int k = 0; k < 8; k++) v[k] = Vec3(bbox(k & 1, 0), bbox((k >> 1) & 1, 1), bbox((k >> 2) & 1, 2));
    return v;
  }
  bool IntersectsAABB(const Vec3& o, const Vec3& d) const {
    double bmin[3] = {bbox(0, 0), bbox(0, 1), bbox(0, 2)}, bmax[3] = {bbox(1, 0), bbox(1, 1), bbox(1, 2)};
    double a, b;
    return sbr::box_line(bmin, bmax, o.v(), d.v(), &a, &b);
  }
};

template <typename T>
class BVH {
 public:
  BVH() : T_tri(1.0), T_aabb(1.0) { nodes.resize(1); }
  const double T_tri;
  const double T_aabb;
  std::vector<BVHNode<T>> nodes;

  void ConstructBVH(std::vector<T*>& objects) { Build(objects, 0); }

  // Every node's box as a VTK hexahedron with its id, children and object count (BVH/BVH.h:57-133).
  void SaveAsVtk(const std::string& fname) const {
    std::ofstream ofs(fname);
    const size_t n = nodes.size();
    ofs << "# vtk DataFile Version 2.0" << std::endl;
    ofs << "3D model of boudary volume hierarchy" << std::endl;   // the reference's title, typo included
    ofs << "ASCII" << std::endl;
    ofs << "DATASET UNSTRUCTURED_GRID" << std::endl;
    ofs << std::endl;
    ofs << "POINTS " << 8 * n << " float" << std::endl;
    for (const BVHNode<T>& node : nodes)
      for (const Vec3& v : node.CornerPoints()) ofs << " " << v.x() << " " << v.y() << " " << v.z() << std::endl;
    ofs << "CELLS " << n << " " << 9 * n << std::endl;
    for (size_t i = 0; i < n; i++) {
      ofs << " 8";
      for (size_t j = 0; j < 8; j++) ofs << " " << 8 * i + j;
      ofs << std::endl;
    }
    ofs << "CELL_TYPES " << n << std::endl;
    for (size_t i = 0; i < n; i++) ofs << " 11" << std::endl;
    ofs << "CELL_DATA " << (int)n << std::endl;
    ofs << "FIELD FieldData 4" << std::endl;
    ofs << "BoxID 1 " << (int)n << " int" << std::endl;
    for (size_t i = 0; i < n; i++) ofs << " " << i << std::endl;
    ofs << "Child0 1 " << (int)n << " int" << std::endl;
    for (size_t i = 0; i < n; i++) ofs << " " << nodes[i].children[0] << std::endl;
    ofs << "Child1 1 " << (int)n << " int" << std::endl;
    for (size_t i = 0; i < n; i++) ofs << " " << nodes[i].children[1] << std::endl;
    ofs << "NumFaces 1 " << (int)n << " int" << std::endl;
    for (size_t i = 0; i < n; i++) ofs << " " << nodes[i].objects.size() << std::endl;
  }

  // Host closest hit in the reference's traversal order (DFS, children[0] first, strict '<').
  bool RayIntersects(int index, const Vec3& o, const Vec3& d, HitInfo& info) const {
    const BVHNode<T>& n = nodes[index];
    if (!n.IntersectsAABB(o, d)) return false;
    bool any = false;
    if (n.children[0] != -1) {
      for (int i = 0; i < 2; i++)
        if (RayIntersects(n.children[i], o, d, info)) any = true;
    } else {
      for (T* obj : n.objects) {
        HitInfo cur;
        if (obj->RayIntersects(o, d, cur) && cur.distance < info.distance) {
          info = cur;
          any = true;
        }
      }
    }
    return any;
  }

 private:
  static double Area(const Mat23& b) {
    if (b(1, 0) <= b(0, 0) || b(1, 1) <= b(0, 1) || b(1, 2) <= b(0, 2)) return 0.0;
    const double dx = b(1, 0) - b(0, 0), dy = b(1, 1) - b(0, 1), dz = b(1, 2) - b(0, 2);
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }
  // result = merge(first, second) with std::min(first, second) / std::max(first, second); the
  // argument order matters for signed zeros, so both call sites keep the reference's order.
  static void Merge(const Mat23& first, const Mat23& second, Mat23& result) {
    for (int a = 0; a < 3; a++) {
      result(0, a) = std::min(first(0, a), second(0, a));
      result(1, a) = std::max(first(1, a), second(1, a));
    }
  }
  static void SortAxis(std::vector<T*>& v, int axis) {
    std::sort(v.begin(), v.end(), [axis](const T* a, const T* b) { return a->center[axis] < b->center[axis]; });
  }

  void Build(std::vector<T*>& objects, int nodeIdx) {
    for (const T* o : objects) Merge(o->bbox, nodes[nodeIdx].bbox, nodes[nodeIdx].bbox);
    const int n = (int)objects.size();
    const double areaNode = Area(nodes[nodeIdx].bbox);
    double bestCost = T_tri * n;
    int bestAxis = -1, bestSplit = -1;
    std::vector<double> prefix(n + 1), suffix(n + 1);
    for (int axis = 0; axis < 3; axis++) {
      SortAxis(objects, axis);
      Mat23 acc;
      BVHNode<T>::EmptyAABB(acc);
      for (int i = 0; i < n; i++) { Merge(acc, objects[i]->bbox, acc); prefix[i] = Area(acc); }   // i+1 objects
      BVHNode<T>::EmptyAABB(acc);
      for (int i = n - 1; i >= 0; i--) { Merge(acc, objects[i]->bbox, acc); suffix[i] = Area(acc); }
      for (int i = 1; i < n - 1; i++) {
        const double cost = 2 * T_aabb + (i * prefix[i] + (n - i) * suffix[i]) * T_tri / areaNode;
        if (cost < bestCost) { bestCost = cost; bestAxis = axis; bestSplit = i; }
      }
    }
    if (bestAxis == -1) {
      nodes[nodeIdx].children[0] = nodes[nodeIdx].children[1] = -1;
      nodes[nodeIdx].objects = objects;
      return;
    }
    SortAxis(objects, bestAxis);
    const int c0 = (int)nodes.size();
    nodes[nodeIdx].children[0] = c0;
    nodes[nodeIdx].children[1] = c0 + 1;
    nodes.emplace_back();
    nodes.emplace_back();
    std::vector<T*> left(objects.begin(), objects.begin() + bestSplit);
    std::vector<T*> right(objects.begin() + bestSplit, objects.end());
    Build(left, c0);
    Build(right, c0 + 1);
  }
};
