// Host-only checks of the reference call surface compiled against include/ (no GPU):
//   surface_check <ground.stl> <out dir> <receiver file>
// 1. Preprocessor::StlToGeometry(..., saveEdges, saveFaces, dir) writes Edges.vtk / Faces.vtk, and
//    BVH<Triangle>::SaveAsVtk writes BVH.vtk (compared with the reference's files by the caller).
// 2. Path trees: newPathTreeNode numbering, CloneNode / CloneTree copies, Paths(root) breadth-first
//    chains ending at captured nodes, DeleteChildNodes.
// 3. Preprocessor::ReadLocationFile on a receiver file as the reference ships it.
// Prints one "key value..." line per fact.
#include <cstdio>
#include <string>
#include <vector>

#include "sbrsolver/BVH.h"
#include "sbrsolver/FieldCompute.h"
#include "sbrsolver/Preprocessor.h"
#include "sbrsolver/RTSolver.h"
#include "sbrsolver/Ray.h"

static PathTreeNode* node(double x, int refl, int pen, int surf) {
  return newPathTreeNode(Vec3(x, 0, 0), Vec3(x + 1, 0, 0), pen, refl, surf, surf, x, 0.0);
}

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const std::string stl = argv[1], dir = std::string(argv[2]) + "/";
  std::vector<Triangle*> mesh;
  const bool ok = Preprocessor::StlToGeometry(stl, mesh, true, true, dir);
  BVH<Triangle> bvh;
  bvh.ConstructBVH(mesh);
  bvh.SaveAsVtk(dir + "BVH.vtk");
  std::printf("stl %d %zu %zu\n", ok ? 1 : 0, mesh.size(), bvh.nodes.size());

  // A tree: root -> reflect A (captured) -> reflect B; root -> transmit C -> reflect D (captured),
  // transmit E.  Paths(root) walks breadth first, transmit before reflect, and lists a chain when it
  // reaches a leaf with a captured node on the way: [root, C, D], then [root, A, B] (the walk goes on
  // below a captured node; RayCapture deletes those subtrees before Paths sees them).
  RayGlobalId = 0;
  PathTreeNode* root = new PathTreeNode;
  root->ray.id = RayGlobalId++;
  root->ray.sourcePoint = Vec3(0, 0, 0);
  PathTreeNode* a = node(1, 0, -1, 10);
  PathTreeNode* b = node(2, 0, -1, 11);
  PathTreeNode* c = node(3, -1, 0, 12);
  PathTreeNode* d = node(4, 0, -1, 13);
  PathTreeNode* e = node(5, -1, 0, 14);
  root->childReflect = a;
  a->childReflect = b;
  root->childTransmit = c;
  c->childReflect = d;
  c->childTransmit = e;
  a->ray.captured = true;
  d->ray.captured = true;
  std::printf("ids %d %d %d %d %d %d next %d\n", root->ray.id, a->ray.id, b->ray.id, c->ray.id, d->ray.id, e->ray.id,
              (int)RayGlobalId);
  PathTreeNode* copy = CloneNode(root);
  CloneTree(root, copy);
  Paths p(copy);
  std::printf("paths %zu", p.rayPaths.size());
  for (const auto& chain : p.rayPaths) {
    std::printf(" [");
    for (const Ray& r : chain) std::printf(" %d", r.hitSurfaceID);
    std::printf(" ]");
  }
  std::printf("\n");
  std::printf("clone %d %d %d\n", copy->childTransmit->childReflect->ray.captured ? 1 : 0,
              copy->childReflect->ray.id, copy->childTransmit != c ? 1 : 0);
  copy = DeleteChildNodes(copy);
  root = DeleteChildNodes(root);
  std::printf("deleted %d %d\n", copy == nullptr ? 1 : 0, root == nullptr ? 1 : 0);
  Paths none(nullptr);
  std::printf("empty %zu\n", none.rayPaths.size());

  VecVec3 rx;
  const bool rok = Preprocessor::ReadLocationFile(argv[3], rx);
  std::printf("rx %d %zu", rok ? 1 : 0, rx.size());
  for (size_t i = 0; i < rx.size() && i < 3; i++) std::printf(" %g %g %g", rx[i].x(), rx[i].y(), rx[i].z());
  std::printf("\n");

  // FieldCompute without a device: an empty path list gives zero fields and no device call.
  Paths** none2 = new Paths*[1];
  none2[0] = new Paths[1];
  MaterialProperties mats[1] = {{1.8e9, 0, 0, 1.0, 0.0, 1e6}};
  FieldCompute fc(none2, mesh, rx, 1, mats, Mat3::Identity(), 1.8e9, 1.0, true);
  Vec3c f = fc.FieldAtReceiver(0);
  std::printf("fc %g %d\n", f.norm(), fc.LastStatus());
  delete[] none2[0];
  delete[] none2;
  for (Triangle* t : mesh) delete t;
  return 0;
}
