// Driver for an AddressSanitizer / UndefinedBehaviorSanitizer build of the host library parts that
// need no device (sbrsolver_amd/csrc/host/preprocessor.cpp, ray.cpp, include/sbrsolver/BVH.h):
//   sanitize_host <stl dir> <rx dir> <out dir>
// Reads the reference's STL files (ASCII and binary) with and without the Edges / Faces writers,
// builds the SAH BVH, runs closest hits, writes BVH.vtk, converts a soup (with a degenerate
// facet), reads the receiver files and a pattern file, generates receiver planes and lines, and
// builds, clones, flattens and deletes path trees.  Any sanitizer report aborts with a non-zero
// status.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "sbrsolver/BVH.h"
#include "sbrsolver/Preprocessor.h"
#include "sbrsolver/Ray.h"

static PathTreeNode* node(double x, int refl, int pen, int surf) {
  return newPathTreeNode(Vec3(x, 0, 0), Vec3(x + 1, 0, 0), pen, refl, surf, surf, x, 0.0);
}

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const std::string stl = std::string(argv[1]) + "/", rxd = std::string(argv[2]) + "/",
                    out = std::string(argv[3]) + "/";
  const char* files[] = {"ground.stl", "corner.stl", "two_planes.stl", "office.stl", "geometry_test.stl",
                         "geometry_test2.stl", "cube.stl", "divide.stl", "tetrahedron.stl", "two_tetrahedron.stl",
                         "triangle.stl", "icosphere_at_origin.stl", "two_icospheres.stl", "menger_sponge.stl",
                         "bahen.stl"};
  for (const char* f : files) {
    std::vector<Triangle*> mesh;
    const bool save = std::string(f) == "ground.stl";   // the writers (compared with the reference's files)
    if (!Preprocessor::StlToGeometry(stl + f, mesh, save, save, out)) {
      std::printf("unreadable %s\n", f);
      return 3;
    }
    BVH<Triangle> bvh;
    bvh.ConstructBVH(mesh);
    bvh.SaveAsVtk(out + "BVH.vtk");
    int hits = 0;
    for (int k = 0; k < 100; k++) {
      const Vec3 o(0.01 * k, 0.02 * k, 0.5);
      const double dz = std::sin(0.7 * k) - 0.2, dn = std::sqrt(1.0 + dz * dz);
      const Vec3 d(std::cos(0.3 * k) / dn, std::sin(0.3 * k) / dn, dz / dn);
      HitInfo info;
      hits += bvh.RayIntersects(0, o, d, info) ? 1 : 0;
    }
    std::printf("%s %zu %zu %d\n", f, mesh.size(), bvh.nodes.size(), hits);
    for (Triangle* t : mesh) delete t;
  }
  // soup with a degenerate facet and material ids
  std::vector<float> c9 = {0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, 2, 2, 2, 2, 2, 2, 3, 3, 3};
  std::vector<float> n3 = {0, 0, 1, 0, 0, 1, 0, 0, 1};
  std::vector<int32_t> mats = {0, 1, 0};
  std::vector<Triangle*> soup;
  const bool sok = Preprocessor::SoupToGeometry(c9.data(), n3.data(), mats.data(), 3, soup);
  std::printf("soup %d %zu\n", sok ? 1 : 0, soup.size());
  for (Triangle* t : soup) delete t;
  // receivers, pattern, generated receiver sets
  const char* rxf[] = {"RX_Corner.dat", "RX_Ground.dat", "RX_Office.dat", "RX_TwoPlanes.dat"};
  for (const char* f : rxf) {
    VecVec3 rx;
    const bool ok = Preprocessor::ReadLocationFile(rxd + f, rx);
    std::printf("%s %d %zu\n", f, ok ? 1 : 0, rx.size());
  }
  {
    std::ofstream pf(out + "TXPattern.dat");
    pf << "theta phi gt pt gp pp\n";
    for (int t = 0; t <= 180; t += 45)
      for (int ph = 0; ph < 360; ph += 90) pf << t << " " << ph << " -3 10 -20 0\n";
    pf << "1 2 3\nx y z a b c\n";
  }
  GainMap gm;
  const bool pok = Preprocessor::ReadPatternFile(out + "TXPattern.dat", gm);
  std::printf("pattern %d %zu\n", pok ? 1 : 0, gm.size());
  VecVec3 plane, line;
  Preprocessor::GenerateRxPlane(-5, -5, 5, 5, 1.5, 0.5, plane);
  Preprocessor::GenerateRXLine(Vec3(0, 0, 0), Vec3(3, 4, 0), 0.25, line);
  std::printf("generated %zu %zu\n", plane.size(), line.size());
  // path trees
  RayGlobalId = 0;
  PathTreeNode* root = new PathTreeNode;
  root->ray.id = RayGlobalId++;
  PathTreeNode* a = node(1, 0, -1, 10);
  PathTreeNode* b = node(2, 0, -1, 11);
  PathTreeNode* c = node(3, -1, 0, 12);
  root->childReflect = a;
  a->childReflect = b;
  root->childTransmit = c;
  b->ray.captured = true;
  c->ray.captured = true;
  PathTreeNode* copy = CloneNode(root);
  CloneTree(root, copy);
  Paths p(copy);
  std::printf("paths %zu\n", p.rayPaths.size());
  copy = DeleteChildNodes(copy);
  root = DeleteChildNodes(root);
  return copy == nullptr && root == nullptr ? 0 : 5;
}
