// The reference call surface on the GPU, compiled against include/:
//   fieldcompute_check <office.stl> <out dir>
// 1. RTSolver::ExecuteRayTracing on the office scene (two materials, reflections and a
//    transmission), then FieldCompute(solver.RayPaths(), mesh, rx, launches, materials, ...) gives
//    FieldAtReceiver(k) == solver.Fields()[k] bit for bit, and FieldAtAllReceivers() the same.
// 2. The kept paths' Ray fields: the last segment captured, path lengths and angles as the
//    corrections leave them, reference node ids (options.rayIds).
// 3. RefCoeffTest / TransCoeffTest write their sweeps (checked by the caller).
// 4. A mesh edited in place (same vector, same BVH object rebuilt) is traced anew, and restoring it
//    gives the first result again.
// Prints one "key value..." line per fact.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "sbrsolver/BVH.h"
#include "sbrsolver/FieldCompute.h"
#include "sbrsolver/Preprocessor.h"
#include "sbrsolver/RTSolver.h"

static bool same_bits(const Vec3c& a, const Vec3c& b) {
  return std::memcmp(&a, &b, sizeof(Vec3c)) == 0;
}

static void rebuild(std::vector<Triangle*>& mesh, BVH<Triangle>& bvh, double scale) {
  for (Triangle* t : mesh) {
    for (Vec3* v : {&t->v1, &t->v2, &t->v3})
      for (int a = 0; a < 3; a++) (*v)(a) *= scale;
    t->bbox = t->findBbox();
    t->center = t->findCenter();
    t->coordSys = t->findCoordSys();
  }
  bvh.nodes.clear();
  bvh.nodes.resize(1);
  bvh.ConstructBVH(mesh);
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string out = std::string(argv[2]) + "/";
  std::vector<Triangle*> mesh;
  if (!Preprocessor::StlToGeometry(argv[1], mesh)) return 3;
  for (Triangle* t : mesh) t->materialId = t->triangleId % 2;
  BVH<Triangle> bvh;
  bvh.ConstructBVH(mesh);
  MaterialProperties mats[2] = {{1.8e9, -1, -1, 2.5, 0.3, 0.03}, {1.8e9, -1, -1, 6.0, 0.05, 0.005}};
  VecVec3 rx;
  Preprocessor::GenerateRxPlane(0.3, 0.3, 11.7, 8.7, 1.2, 0.6, rx);
  const Vec3 tx(0.835938, 4.53906, 2.5);
  const Mat3 txcs{{0, 1, 0}, {0, 0, 1}, {1, 0, 0}};

  RTSolver solver;
  solver.options.rayIds = 1;
  solver.Init(mats, 2, bvh);
  const int n = solver.ExecuteRayTracing(tx, txcs, 2, 1, 1.8e9, 1.0, true, rx, mesh, 3);
  std::printf("trace %d %zu %d\n", n, rx.size(), solver.RayPaths() != nullptr ? 1 : 0);
  if (n < 0) {
    std::printf("error %s\n", solver.LastError());
    return 4;
  }
  const std::vector<Vec3c> fields = solver.Fields();

  FieldCompute fc(solver.RayPaths(), mesh, rx, solver.LaunchCount(), mats, txcs, 1.8e9, 1.0, true);
  int equal = 0, nonzero = 0;
  for (int k = 0; k < (int)rx.size(); k++) {
    const Vec3c f = fc.FieldAtReceiver(k);
    equal += same_bits(f, fields[k]) ? 1 : 0;
    nonzero += f.norm() > 0 ? 1 : 0;
  }
  const std::vector<Vec3c> all = fc.FieldAtAllReceivers();
  int equalAll = 0;
  for (size_t k = 0; k < all.size(); k++) equalAll += same_bits(all[k], fields[k]) ? 1 : 0;
  std::printf("fieldcompute %d %d %d %d\n", equal, equalAll, nonzero, fc.LastStatus());

  // Ray fields of the kept paths
  int paths = 0, multi = 0, capOk = 0, lenOk = 0, angOk = 0, idsOk = 0, types = 0;
  std::set<std::tuple<int, int, int>> seenIds;
  for (const auto& perRx : solver.ReceiverPaths())
    for (const auto& lp : perRx) {
      const std::vector<Ray>& p = lp.second;
      paths++;
      bool cap = true, len = true, ang = true, ids = true;
      for (size_t s = 0; s < p.size(); s++) {
        const Ray& r = p[s];
        cap = cap && (r.captured == (s + 1 == p.size()));
        const double d = std::sqrt(std::pow(r.targetPoint.x() - r.sourcePoint.x(), 2) +
                                   std::pow(r.targetPoint.y() - r.sourcePoint.y(), 2) +
                                   std::pow(r.targetPoint.z() - r.sourcePoint.z(), 2));
        if (p.size() > 1 && s == 0) len = len && r.pathLength == 0.0;
        else len = len && std::fabs(r.pathLength - d) <= 1e-12 * (1 + d);
        if (p.size() > 1 && s > 0) ang = ang && r.angleFromSurfaceNormal >= 0 && r.angleFromSurfaceNormal <= 180;
        else ang = ang && std::isinf(r.angleFromSurfaceNormal);
        ids = ids && r.id >= 0 && (s == 0 || r.id > p[s - 1].id);
        if (r.reflectionMaterialId >= 0 || r.penetrationMaterialId >= 0) types |= r.reflectionMaterialId >= 0 ? 1 : 2;
      }
      multi += p.size() > 1 ? 1 : 0;
      capOk += cap;
      lenOk += len;
      angOk += ang;
      idsOk += ids;
    }
  std::printf("rays %d %d %d %d %d %d %d\n", paths, multi, capOk, lenOk, angOk, idsOk, types);

  fc.RefCoeffTest(91, 2.4e9, cdouble(1.0, 0.0), cdouble(5.31, -0.0326 / (2 * 3.141592653589793 * 2.4e9 * 8.854187817e-12)),
                  out + "ref_coeff.csv");
  const int st1 = fc.LastStatus();
  fc.TransCoeffTest(91, 2.4e9, cdouble(2.0, 0.0), cdouble(6.0, -0.5), out + "trans_coeff.csv");
  std::printf("coeff %d %d\n", st1, fc.LastStatus());

  // the same mesh vector and BVH object, edited in place
  rebuild(mesh, bvh, 1.05);
  const int n2 = solver.ExecuteRayTracing(tx, txcs, 2, 1, 1.8e9, 1.0, true, rx, mesh, 3);
  int changed = 0;
  for (size_t k = 0; k < rx.size(); k++) changed += same_bits(solver.Fields()[k], fields[k]) ? 0 : 1;
  rebuild(mesh, bvh, 1.0 / 1.05);
  std::vector<Triangle*> fresh;
  Preprocessor::StlToGeometry(argv[1], fresh);
  for (Triangle* t : fresh) t->materialId = t->triangleId % 2;
  // restore the exact original coordinates (the round trip through 1.05 is not exact)
  for (Triangle* t : mesh) {
    *t = *fresh[t->triangleId];
  }
  bvh.nodes.clear();
  bvh.nodes.resize(1);
  std::vector<Triangle*> order(mesh.size());
  for (Triangle* t : mesh) order[t->triangleId] = t;
  mesh = order;   // the STL order, as before the first build
  bvh.ConstructBVH(mesh);
  const int n3 = solver.ExecuteRayTracing(tx, txcs, 2, 1, 1.8e9, 1.0, true, rx, mesh, 3);
  int restored = 0;
  for (size_t k = 0; k < rx.size(); k++) restored += same_bits(solver.Fields()[k], fields[k]) ? 1 : 0;
  std::printf("inplace %d %d %d %d\n", n2, changed, n3, restored);
  for (Triangle* t : fresh) delete t;
  for (Triangle* t : mesh) delete t;
  return 0;
}
