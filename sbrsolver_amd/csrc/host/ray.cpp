// Path-tree helpers of the reference call surface (Ray.cpp:3-182).
#include "sbrsolver/Ray.h"

#include <deque>
#include <utility>

std::atomic<int> RayGlobalId{0};

PathTreeNode* newPathTreeNode(Vec3 sourcePoint, Vec3 targetPoint, int penetrationMaterialId, int reflectionMaterialId,
                              int surfaceId, int coplanarId, double pathLength, double angleFromSurfaceNormal) {
  PathTreeNode* n = new PathTreeNode;
  n->ray.id = RayGlobalId++;
  n->ray.sourcePoint = sourcePoint;
  n->ray.targetPoint = targetPoint;
  n->ray.penetrationMaterialId = penetrationMaterialId;
  n->ray.reflectionMaterialId = reflectionMaterialId;
  n->ray.hitSurfaceID = surfaceId;
  n->ray.hitCoplanarId = coplanarId;
  n->ray.pathLength = pathLength;
  n->ray.angleFromSurfaceNormal = angleFromSurfaceNormal;
  return n;
}

PathTreeNode* DeleteChildNodes(PathTreeNode* node) {
  // iterative, so that deep trees do not exhaust the stack
  std::vector<PathTreeNode*> st;
  if (node) st.push_back(node);
  while (!st.empty()) {
    PathTreeNode* n = st.back();
    st.pop_back();
    if (n->childTransmit) st.push_back(n->childTransmit);
    if (n->childReflect) st.push_back(n->childReflect);
    delete n;
  }
  return nullptr;
}

PathTreeNode* CloneNode(PathTreeNode* node) {
  if (!node) return nullptr;
  PathTreeNode* n = new PathTreeNode;
  n->ray = node->ray;   // every field, the id and the captured flag included
  return n;
}

void CloneTree(PathTreeNode* orgTree, PathTreeNode* cloneTree) {
  std::vector<std::pair<PathTreeNode*, PathTreeNode*>> st;
  if (orgTree && cloneTree) st.emplace_back(orgTree, cloneTree);
  while (!st.empty()) {
    auto [o, c] = st.back();
    st.pop_back();
    c->childTransmit = CloneNode(o->childTransmit);
    c->childReflect = CloneNode(o->childReflect);
    if (o->childTransmit) st.emplace_back(o->childTransmit, c->childTransmit);
    if (o->childReflect) st.emplace_back(o->childReflect, c->childReflect);
  }
}

Paths::Paths(PathTreeNode* rootNode) {
  if (!rootNode) return;
  // (segments so far, a node on the chain was captured, node)
  struct Item {
    std::vector<Ray> chain;
    bool done;
    PathTreeNode* node;
  };
  std::deque<Item> q;
  q.push_back(Item{{rootNode->ray}, false, rootNode});
  while (!q.empty()) {
    Item it = std::move(q.front());
    q.pop_front();
    PathTreeNode* n = it.node;
    if (n->childTransmit || n->childReflect) {
      for (PathTreeNode* c : {n->childTransmit, n->childReflect}) {
        if (!c) continue;
        Item next{it.chain, it.done || c->ray.captured, c};
        next.chain.push_back(c->ray);
        q.push_back(std::move(next));
      }
    } else if (it.done || n->ray.captured) {
      rayPaths.push_back(std::move(it.chain));
    }
  }
}
