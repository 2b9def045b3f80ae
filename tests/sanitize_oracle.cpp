// Driver for an AddressSanitizer / UndefinedBehaviorSanitizer build of the FP64 oracle
// (oracle/sbr_oracle.cpp, linked in directly):
//   sanitize_oracle <stl dir>
// Builds the scenes of the reference's STL files and a synthetic soup, traces each at a small level
// in both reception models and with the physical-mode switches, dumps trees, runs closest hits
// (BVH and brute force), the icosahedron, the tabulated pattern and the Fresnel sweep.  Prints one
// line per scene; any sanitizer report aborts with a non-zero status.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

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
void* orc_scene_from_stl(const char* path);
void* orc_scene_from_soup(const float* coords9, const float* normals3, const int* mats, int ntri);
void orc_scene_free(void* s);
int orc_scene_ntri(void* s);
int orc_scene_nnodes(void* s);
void orc_scene_triangles(void* sp, int* coplanar, int* material, double* verts9, double* norm3, int* meshTriId);
int orc_closest_hit(void* sp, const double* org, const double* dir, double* point, double* dist);
int orc_closest_hit_brute(void* sp, const double* org, const double* dir, double* point, double* dist);
long long orc_icosahedron(int tess, double* out, long long cap);
void* orc_trace(void* sp, const orc_params* pp, const double* rx, int nrx, const double* mats, int nmat, int nthreads);
void orc_result_free(void* r);
long long orc_result_bounces(void* r);
int orc_result_npaths(void* r);
int orc_result_nsegs(void* r);
void orc_result_fields(void* r, double* out);
void orc_result_paths(void* r, int* prx, int* plaunch, int* pnseg, int* poff, int* stri, int* scop, int* stype,
                      double* ssrc, double* stgt);
void orc_result_path_fields(void* r, double* out);
void* orc_trees(void* sp, const orc_params* pp);
int orc_trees_n(void* d);
void orc_trees_get(void* dp, int* launch, int* depth, long long* code, int* tri, int* cop, int* type, double* src,
                   double* tgt, double* pl);
void orc_trees_free(void* d);
int orc_set_pattern(const float* gain_info, int n_theta, int n_phi);
void orc_fresnel(double thetaI, double relPerm, double sigma, double freq, double width, int infWall, double* out);
}

static orc_params base_params(const double* tx, int tess, int maxR, int maxT) {
  orc_params p{};
  for (int a = 0; a < 3; a++) p.tx[a] = tx[a];
  for (int k = 0; k < 9; k++) p.txCoordSys[k] = (k % 4 == 0) ? 1.0 : 0.0;
  p.maxR = maxR;
  p.maxT = maxT;
  p.freq = 1.8e9;
  p.pt = 1.0;
  p.useFresnel = 1;
  p.tess = tess;
  p.antennaType = 2;
  p.launchStride = 1;
  p.q1MeshIndexing = 1;
  return p;
}

// Trace, read every output array back, and dump the trees.
static long long run(void* sc, const orc_params& p, const std::vector<double>& rx, const double* mats, int nmat,
                     int threads) {
  void* r = orc_trace(sc, &p, rx.data(), (int)(rx.size() / 3), mats, nmat, threads);
  if (!r) return -1;
  const int np = orc_result_npaths(r), ns = orc_result_nsegs(r);
  std::vector<int> prx(np), pl(np), pn(np), po(np), st(ns), sc2(ns), sy(ns);
  std::vector<double> ss(3 * (size_t)ns), sg(3 * (size_t)ns), f(6 * rx.size() / 3), pf(6 * (size_t)np);
  orc_result_paths(r, prx.data(), pl.data(), pn.data(), po.data(), st.data(), sc2.data(), sy.data(), ss.data(),
                   sg.data());
  orc_result_fields(r, f.data());
  orc_result_path_fields(r, pf.data());
  const long long b = orc_result_bounces(r);
  orc_result_free(r);
  void* t = orc_trees(sc, &p);
  const int n = orc_trees_n(t);
  std::vector<int> la(n), de(n), tr(n), co(n), ty(n);
  std::vector<long long> cd(n);
  std::vector<double> s3(3 * (size_t)n), t3(3 * (size_t)n), plen(n);
  orc_trees_get(t, la.data(), de.data(), cd.data(), tr.data(), co.data(), ty.data(), s3.data(), t3.data(), plen.data());
  orc_trees_free(t);
  return b + np;
}

static void exercise(const char* label, void* sc, const double* tx) {
  const int nt = orc_scene_ntri(sc);
  std::vector<int> cop(nt), mat(nt), mid(nt);
  std::vector<double> v9(9 * (size_t)nt), n3(3 * (size_t)nt);
  orc_scene_triangles(sc, cop.data(), mat.data(), v9.data(), n3.data(), mid.data());
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (size_t i = 0; i < v9.size(); i++) {
    lo[i % 3] = std::fmin(lo[i % 3], v9[i]);
    hi[i % 3] = std::fmax(hi[i % 3], v9[i]);
  }
  std::vector<double> rx;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      rx.push_back(lo[0] + (hi[0] - lo[0]) * (i + 0.5) / 4);
      rx.push_back(lo[1] + (hi[1] - lo[1]) * (j + 0.5) / 4);
      rx.push_back(0.5 * (lo[2] + hi[2]) + 0.1);
    }
  const double mats[12] = {1.8e9, -1, -1, 2.5, 0.3, 0.03, 1.8e9, -1, -1, 6.0, 0.05, 0.005};
  long long acc = 0;
  orc_params p = base_params(tx, 2, 3, 1);
  acc += run(sc, p, rx, mats, 2, 1);
  p.receptionModel = 1;
  acc += run(sc, p, rx, mats, 2, 2);
  p.q1MeshIndexing = 0;
  p.typedDedup = 1;
  p.checkOcclusion = 1;
  p.longMiss = 1;
  p.physicalField = 1;
  p.wallWidth = 0.2;
  p.maxTotal = 3;
  acc += run(sc, p, rx, mats, 2, 1);
  p = base_params(tx, 1, 2, 0);
  p.antennaType = 3;
  acc += run(sc, p, rx, mats, 2, 1);
  p.antennaType = 0;
  p.useFresnel = 0;
  p.launchStride = 3;
  p.launchOffset = 2;
  acc += run(sc, p, rx, mats, 2, 1);
  // closest hits from inside the scene's box, BVH and brute force
  int agree = 0;
  for (int k = 0; k < 200; k++) {
    const double o[3] = {lo[0] + (hi[0] - lo[0]) * ((k * 37) % 101) / 100.0,
                         lo[1] + (hi[1] - lo[1]) * ((k * 53) % 103) / 102.0, 0.5 * (lo[2] + hi[2])};
    const double d[3] = {std::cos(0.1 * k), std::sin(0.1 * k), std::sin(0.37 * k)};
    double p1[3], p2[3], d1 = 0, d2 = 0;
    const int h1 = orc_closest_hit(sc, o, d, p1, &d1), h2 = orc_closest_hit_brute(sc, o, d, p2, &d2);
    agree += (h1 == h2) ? 1 : 0;
  }
  std::printf("%s %d %d %lld %d\n", label, nt, orc_scene_nnodes(sc), acc, agree);
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = std::string(argv[1]) + "/";
  // tabulated pattern (antennaType 3): 3 x 4 grid
  std::vector<float> pat;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 4; j++) {
      const float row[6] = {90.0f * i, 90.0f * j, -3.0f * j, 10.0f * i, -20.0f, 0.0f};
      pat.insert(pat.end(), row, row + 6);
    }
  if (orc_set_pattern(pat.data(), 3, 4) != 0) return 3;
  struct S {
    const char* f;
    double tx[3];
  } files[] = {{"ground.stl", {0, 10, 2}},        {"corner.stl", {0, 0, 2.5}},
               {"two_planes.stl", {0, 0, 1}},     {"office.stl", {0.835938, 4.53906, 2.5}},
               {"geometry_test2.stl", {0.1, 0.2, 0.5}}, {"cube.stl", {0.1, 0.2, 0.3}},
               {"menger_sponge.stl", {0.5, 0.5, 1.5}}};
  for (const S& s : files) {
    void* sc = orc_scene_from_stl((dir + s.f).c_str());
    if (!sc) {
      std::printf("unreadable %s\n", s.f);
      return 4;
    }
    exercise(s.f, sc, s.tx);
    orc_scene_free(sc);
  }
  // a closed box room as a soup (12 facets), with a degenerate facet and per-facet materials
  const float X = 10, Y = 8, Z = 3;
  const float q[6][4][3] = {{{0, 0, 0}, {X, 0, 0}, {X, Y, 0}, {0, Y, 0}}, {{0, 0, Z}, {X, 0, Z}, {X, Y, Z}, {0, Y, Z}},
                            {{0, 0, 0}, {X, 0, 0}, {X, 0, Z}, {0, 0, Z}}, {{0, Y, 0}, {X, Y, 0}, {X, Y, Z}, {0, Y, Z}},
                            {{0, 0, 0}, {0, Y, 0}, {0, Y, Z}, {0, 0, Z}}, {{X, 0, 0}, {X, Y, 0}, {X, Y, Z}, {X, 0, Z}}};
  std::vector<float> c9, n3;
  std::vector<int> m;
  for (int f = 0; f < 6; f++)
    for (int t = 0; t < 2; t++) {
      const int idx[2][3] = {{0, 1, 2}, {0, 2, 3}};
      for (int v = 0; v < 3; v++)
        for (int a = 0; a < 3; a++) c9.push_back(q[f][idx[t][v]][a]);
      n3.insert(n3.end(), {0.0f, 0.0f, 1.0f});
      m.push_back(f % 2);
    }
  for (int v = 0; v < 3; v++)
    for (int a = 0; a < 3; a++) c9.push_back(1.0f);   // degenerate facet (dropped)
  n3.insert(n3.end(), {0.0f, 0.0f, 1.0f});
  m.push_back(0);
  void* room = orc_scene_from_soup(c9.data(), n3.data(), m.data(), (int)m.size());
  const double tx[3] = {2, 2, 1.5};
  exercise("box_room", room, tx);
  orc_scene_free(room);
  std::vector<double> ico(3 * (10 * 16 + 2));
  const long long nv = orc_icosahedron(2, ico.data(), (long long)ico.size() / 3);
  double fr[8];
  for (int k = 0; k <= 90; k += 5) orc_fresnel(k * 3.141592653589793 / 180.0, 5.31, 0.0326, 2.4e9, 0.2, k % 3, fr);
  std::printf("ico %lld\n", nv);
  return 0;
}
