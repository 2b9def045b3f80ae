// Ray segments, path trees and path lists of the reference call surface (Ray.h:15-67, Ray.cpp).
//
// A Ray is one segment of a path: where it starts and ends, the material of the surface it leaves
// (by reflection or by transmission), the surface it ends on, its path length and the angle to the
// normal it left with.  PathTreeNode is the node of the binary interaction tree of one launch
// (transmitted child, reflected child); Paths(root) lists the root-to-node segment chains of the
// nodes a receiver captured, breadth first with the transmitted child before the reflected one.
// These are host data structures: the library's trace keeps its trees on the device and hands
// back the kept paths as Paths (RTSolver), so callers that walk or build trees themselves keep the
// reference's functions.
#pragma once

#include <atomic>
#include <limits>
#include <vector>

#include "VecMatDef.h"

extern std::atomic<int> RayGlobalId;   // id of the next node newPathTreeNode makes (Ray.cpp:3)

struct Ray {
  Vec3 sourcePoint = Vec3(std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
                          std::numeric_limits<double>::infinity());
  Vec3 targetPoint = Vec3(std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
                          std::numeric_limits<double>::infinity());
  int penetrationMaterialId = -1;  // >= 0: the segment leaves its start surface by transmission
  int reflectionMaterialId = -1;   // >= 0: ... by reflection
  int hitSurfaceID = -1;           // surface the segment ends on (-1: a receiver, or a miss)
  int hitCoplanarId = -1;
  double angleFromSurfaceNormal = std::numeric_limits<double>::infinity();   // degrees
  double pathLength = std::numeric_limits<double>::infinity();
  bool captured = false;
  int id = -1;
};

struct PathTreeNode {
  Ray ray;
  PathTreeNode* childTransmit = nullptr;
  PathTreeNode* childReflect = nullptr;
};

// Ray.cpp:5-30: a node with the next RayGlobalId.
PathTreeNode* newPathTreeNode(Vec3 sourcePoint, Vec3 targetPoint, int penetrationMaterialId, int reflectionMaterialId,
                              int surfaceId, int coplanarId, double pathLength, double angleFromSurfaceNormal);
// Ray.cpp:32-48: deletes the node and its subtree, returns nullptr.
PathTreeNode* DeleteChildNodes(PathTreeNode* node);
// Ray.cpp:66-86: a copy of one node without children (nullptr for nullptr).
PathTreeNode* CloneNode(PathTreeNode* node);
// Ray.cpp:50-64: copies orgTree's subtrees under cloneTree.
void CloneTree(PathTreeNode* orgTree, PathTreeNode* cloneTree);

class Paths {
 public:
  std::vector<std::vector<Ray>> rayPaths;
  Paths() = default;
  // Ray.cpp:136-182: breadth first over the tree; a chain ends at a captured node (or a leaf below
  // one) and is listed if a node on it was captured.
  explicit Paths(PathTreeNode* rootNode);
};
