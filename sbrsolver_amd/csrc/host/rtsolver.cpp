// RTSolver on the gfx950 library: the reference's call surface and output writers
// (RTSolver.cpp:1-803), with the trace delegated to the C-ABI.
#include "sbrsolver/RTSolver.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <tuple>

#include "sbrsolver/SceneDesc.h"

RTSolver::RTSolver() {}

// AngleBetween (VecMatDef.h:56-77) in degrees: both vectors normalized, the angle from the chord
// length, on the obtuse side from the chord to the reversed vector.
double RTSolver::AngleBetweenDeg(sbr::V3 a, sbr::V3 b) {
  a = sbr::vnormalized(a);
  b = sbr::vnormalized(b);
  const double kPi = 3.141592653589793;
  const double th = sbr::vdot(a, b) < 0 ? kPi - 2.0 * std::asin(sbr::vnorm(sbr::vsub(sbr::vneg(a), b)) / 2.0)
                                         : 2.0 * std::asin(sbr::vnorm(sbr::vsub(a, b)) / 2.0);
  return th * (180.0 / kPi);
}

RTSolver::~RTSolver() {
  Cleanup();
  if (_ctx) sbr_destroy(_ctx);
  _ctx = nullptr;
}

void RTSolver::FreeRayPaths() {
  for (int i = 0; i < _rayPathRows; i++) delete[] _rayPaths[i];
  delete[] _rayPaths;
  _rayPaths = nullptr;
  _rayPathRows = 0;
}

void RTSolver::Cleanup() {
  FreeRayPaths();
  delete[] _materialProperties;
  _materialProperties = nullptr;
  _materialCount = 0;
  _shootRayList.clear();
  _efield.clear();
  _kept.clear();
  _isSceneInitialized = false;
  _uploadedFingerprint = 0;   // the next trace uploads the scene again
}

namespace {

// FNV-1a over the bytes of the flattened scene: the device copy is reused only while the triangles,
// their frames, ids and materials and the BVH are unchanged (the reference reads the live objects on
// every trace, so an edit in place, a rebuilt BVH or a new mesh at the old address must be seen).
uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

template <typename T>
uint64_t fnv1a_vec(uint64_t h, const std::vector<T>& v) {
  const uint64_t n = v.size();
  h = fnv1a(h, &n, sizeof n);
  return v.empty() ? h : fnv1a(h, v.data(), v.size() * sizeof(T));
}

uint64_t scene_fingerprint(const SceneArrays& sa) {
  uint64_t h = 0xcbf29ce484222325ull;
  h = fnv1a_vec(h, sa.tri_v);
  h = fnv1a_vec(h, sa.tri_n);
  h = fnv1a_vec(h, sa.tri_cs);
  h = fnv1a_vec(h, sa.tri_cop);
  h = fnv1a_vec(h, sa.tri_mat);
  h = fnv1a_vec(h, sa.mesh_tri_id);
  h = fnv1a_vec(h, sa.node_box);
  h = fnv1a_vec(h, sa.node_child);
  h = fnv1a_vec(h, sa.node_first);
  h = fnv1a_vec(h, sa.node_count);
  h = fnv1a_vec(h, sa.leaf_tri);
  return h | 1;   // 0 means "nothing uploaded"
}

}  // namespace

const char* RTSolver::LastError() const { return _ctx ? sbr_last_error(_ctx) : "no context"; }

// The map must hold a full grid: theta 0..180 degrees and phi 0..360 - dphi degrees, each evenly
// spaced, every (theta, phi) pair present.
int RTSolver::SetTxPattern(const GainMap& pattern) {
  std::vector<float> th, ph;
  for (const auto& kv : pattern) {
    th.push_back(kv.first.first);
    ph.push_back(kv.first.second);
  }
  auto uniq = [](std::vector<float>& v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
  uniq(th);
  uniq(ph);
  const int nt = (int)th.size(), np = (int)ph.size();
  if (nt < 2 || np < 1 || (size_t)nt * np != pattern.size() || th.front() != 0.0f || th.back() != 180.0f ||
      ph.front() != 0.0f)
    return SBR_E_ARG;
  for (int i = 0; i < nt; i++)
    if (std::fabs(th[i] - 180.0 * i / (nt - 1)) > 1e-3) return SBR_E_ARG;
  for (int j = 0; j < np; j++)
    if (std::fabs(ph[j] - 360.0 * j / np) > 1e-3) return SBR_E_ARG;
  std::vector<float> rows(4 * (size_t)nt * np);
  for (int i = 0; i < nt; i++)
    for (int j = 0; j < np; j++) {
      const GainInfo& g = pattern.at(ThetaPhiAngle{th[i], ph[j]});
      std::copy(g.data, g.data + 4, rows.begin() + 4 * ((size_t)i * np + j));
    }
  _pattern.swap(rows);
  _patternRows = nt;
  _patternCols = np;
  _patternVersion++;
  options.antennaType = 3;
  return SBR_OK;
}

int RTSolver::EnsureContext() {
  if (_ctx) return SBR_OK;
  return sbr_create(options.device, &_ctx);
}

// RTSolver.cpp:13-22
int RTSolver::Init(MaterialProperties materialProperties[], int materialPropertiesCount, BVH<Triangle>& triangles) {
  Cleanup();
  _materialProperties = new MaterialProperties[materialPropertiesCount];
  std::copy(materialProperties, materialProperties + materialPropertiesCount, _materialProperties);
  _materialCount = materialPropertiesCount;
  _bvh = &triangles;
  _isSceneInitialized = true;
  return 0;
}

// RTSolver.cpp:52-126: returns N_rays * N_rx as the reference does; a negative value is an sbr
// status (the reference has none).
int RTSolver::ExecuteRayTracing(Vec3 sourcePoint, const Mat3& txCoordSys, int maxReflectionCount,
                                int maxTransmissionCount, double frequency, double pt, bool useFresnelCoeff,
                                VecVec3& receivers, std::vector<Triangle*>& triangleMesh, int txTesslation) {
  if (!_isSceneInitialized || !_bvh) return SBR_E_STATE;
  int rc = EnsureContext();
  if (rc) return rc;
  {
    SceneArrays sa;
    if (!sa.build(triangleMesh, *_bvh)) return SBR_E_ARG;
    const uint64_t fp = scene_fingerprint(sa);
    if (fp != _uploadedFingerprint) {
      _uploadedFingerprint = 0;
      rc = sbr_scene_upload(_ctx, &sa.desc);
      if (rc) return rc;
      _uploadedFingerprint = fp;
    }
  }
  rc = sbr_materials(_ctx, reinterpret_cast<const sbr_material*>(_materialProperties), _materialCount);
  if (rc) return rc;
  if (_patternRows > 0 && _patternSent != _patternVersion) {
    rc = sbr_set_tx_pattern(_ctx, _pattern.data(), _patternRows, _patternCols);
    if (rc) return rc;
    _patternSent = _patternVersion;
  }
  sbr_trace_params p{};
  for (int a = 0; a < 3; a++) p.tx[a] = sourcePoint(a);
  txCoordSys.to_row_major(p.txCoordSys);
  p.maxR = maxReflectionCount;
  p.maxT = maxTransmissionCount;
  p.maxTotal = options.maxTotal;
  p.frequency = frequency;
  p.pt = pt;
  p.useFresnel = useFresnelCoeff ? 1 : 0;
  p.tess = txTesslation;
  p.antennaType = options.antennaType;
  p.wallWidth = options.wallWidth;
  p.typedDedup = options.typedDedup;
  p.checkOcclusion = options.checkOcclusion;
  p.longMiss = options.longMiss;
  p.physicalField = options.physicalField;
  p.receptionModel = options.receptionModel;
  p.launchStride = 1;
  p.launchOffset = 0;
  p.q1MeshIndexing = options.q1MeshIndexing;
  _receivers = &receivers;
  _receiverCount = (int)receivers.size();
  _txTesslation = txTesslation + 1;
  rc = sbr_trace(_ctx, &p, reinterpret_cast<const double*>(receivers.data()), _receiverCount);
  if (rc) return rc;
  sbr_result_stats(_ctx, &_stats);
  _pathsCount = (int)_stats.launches;
  std::vector<double> f(6 * (size_t)_receiverCount);
  rc = sbr_result_fields(_ctx, f.data(), _receiverCount);
  if (rc) return rc;
  _efield.assign(_receiverCount, Vec3c());
  for (int r = 0; r < _receiverCount; r++)
    for (int c = 0; c < 3; c++) _efield[r](c) = cdouble(f[6 * r + 2 * c], f[6 * r + 2 * c + 1]);
  int64_t np = 0, ns = 0;
  rc = sbr_result_path_counts(_ctx, &np, &ns);
  if (rc) return rc;
  std::vector<int32_t> prx(np), pl(np), pn(np), stri(ns), scop(ns), stype(ns);
  std::vector<int64_t> poff(np);
  std::vector<double> ssrc(3 * ns), stgt(3 * ns);
  if (np > 0) {
    rc = sbr_result_paths(_ctx, prx.data(), pl.data(), pn.data(), poff.data(), stri.data(), scop.data(), stype.data(),
                          ssrc.data(), stgt.data());
    if (rc) return rc;
  }
  std::vector<int> matOfId(triangleMesh.size(), 0);
  for (const Triangle* t : triangleMesh)
    if (t->triangleId >= 0 && t->triangleId < (int)matOfId.size()) matOfId[t->triangleId] = t->materialId;
  std::vector<int64_t> nodeIds;
  if (options.rayIds && np > 0) {
    rc = NumberTreeNodes(prx, pl, pn, poff, stri, stype, nodeIds);
    if (rc) return rc;
  }
  _kept.assign(_receiverCount, {});
  const double inf = std::numeric_limits<double>::infinity();
  for (int64_t k = 0; k < np; k++) {
    std::vector<Ray> path(pn[k]);
    for (int s = 0; s < pn[k]; s++) {
      const int64_t q = poff[k] + s;
      Ray& r = path[s];
      r.sourcePoint = Vec3(ssrc[3 * q], ssrc[3 * q + 1], ssrc[3 * q + 2]);
      r.targetPoint = Vec3(stgt[3 * q], stgt[3 * q + 1], stgt[3 * q + 2]);
      r.hitSurfaceID = stri[q];
      r.hitCoplanarId = scop[q];
      const int leftMat = s > 0 && stri[q - 1] >= 0 ? matOfId[stri[q - 1]] : -1;
      if (stype[q] == 1) r.reflectionMaterialId = leftMat;
      if (stype[q] == 2) r.penetrationMaterialId = leftMat;
      // Lengths and angles as the corrections leave them: a direct path gets its length
      // (RTSolver.cpp:384); in a corrected multi-bounce path each later segment gets its length and
      // the angle (degrees) between it and the normal of mesh[its start surface] (RTSolver.cpp:
      // 353-357), and the first one the length from the first bounce point to itself, 0 (:348).
      const sbr::V3 seg = sbr::vsub(r.targetPoint.v(), r.sourcePoint.v());
      if (pn[k] == 1) {
        r.pathLength = sbr::vnorm(seg);
      } else if (s == 0) {
        r.pathLength = 0.0;
      } else {
        r.pathLength = sbr::vnorm(seg);
        const int surf = stri[q - 1];
        if (surf >= 0 && surf < (int)triangleMesh.size()) r.angleFromSurfaceNormal = AngleBetweenDeg(triangleMesh[surf]->norm.v(), seg);
      }
      if (pn[k] == 1 || s == 0) r.angleFromSurfaceNormal = inf;
      r.captured = s == pn[k] - 1;   // the capturing node; no ancestor captured (RTSolver.cpp:259-297)
      if (!nodeIds.empty()) r.id = (int)nodeIds[q];
    }
    _kept[prx[k]].push_back(std::make_pair((int)pl[k], std::move(path)));
  }
  FreeRayPaths();
  if ((int64_t)_pathsCount * _receiverCount <= options.rayPathArrayLimit && _pathsCount > 0) {
    _rayPaths = new Paths*[_pathsCount]();
    _rayPathRows = _pathsCount;
    for (int i = 0; i < _pathsCount; i++) _rayPaths[i] = new Paths[_receiverCount];
    for (int r = 0; r < _receiverCount; r++)
      for (auto& lp : _kept[r]) _rayPaths[lp.first][r].rayPaths.push_back(lp.second);
  }
  _shootRayList.clear();
  return _pathsCount * _receiverCount;
}

// Reference node ids of the kept paths' segments: the id of a node is its position in the creation
// order of all tree nodes of the trace, launch by launch, each tree depth first with the reflected
// child before the transmitted one (RTSolver.cpp:82-84, 186-220; Ray.cpp:18).  The trees are read
// back from the device; within a launch a node's rank in that order follows from its branch code.
int RTSolver::NumberTreeNodes(const std::vector<int32_t>& prx, const std::vector<int32_t>& pl,
                              const std::vector<int32_t>& pn, const std::vector<int64_t>& poff,
                              const std::vector<int32_t>& stri, const std::vector<int32_t>& stype,
                              std::vector<int64_t>& ids) {
  (void)prx;
  (void)stri;
  int64_t n = 0;
  int rc = sbr_result_tree_count(_ctx, &n);
  if (rc) return rc;
  std::vector<int32_t> launch(n), depth(n), tri(n), cop(n), type(n);
  std::vector<int64_t> code(n);
  std::vector<double> src(3 * n), tgt(3 * n), plen(n);
  if (n > 0) {
    rc = sbr_result_trees(_ctx, launch.data(), depth.data(), code.data(), tri.data(), cop.data(), type.data(),
                          src.data(), tgt.data(), plen.data());
    if (rc) return rc;
  }
  // Depth-first rank with the reflected branch first: compare branch strings bit by bit from the
  // first level, reflect (1) before transmit (0), a prefix before its extensions.
  auto dfs_less = [&](int64_t a, int64_t b) {
    const int da = depth[a], db = depth[b], m = std::min(da, db);
    for (int l = 0; l < m; l++) {
      const int ba = (int)((code[a] >> (da - 1 - l)) & 1), bb = (int)((code[b] >> (db - 1 - l)) & 1);
      if (ba != bb) return ba > bb;
    }
    return da < db;
  };
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return launch[a] != launch[b] ? launch[a] < launch[b] : dfs_less(a, b);
  });
  // (launch, depth, code) -> global id
  std::map<std::tuple<int32_t, int32_t, int64_t>, int64_t> idOf;
  for (int64_t r = 0; r < n; r++) {
    const int64_t i = order[r];
    idOf.emplace(std::make_tuple(launch[i], depth[i], code[i]), r);
  }
  ids.assign(stype.size(), -1);
  for (size_t k = 0; k < pn.size(); k++) {
    int64_t c = 0;
    for (int s = 0; s < pn[k]; s++) {
      const int64_t q = poff[k] + s;
      if (s > 0) c = (c << 1) | (stype[q] == 1 ? 1 : 0);
      auto it = idOf.find(std::make_tuple(pl[k], s, c));
      if (it != idOf.end()) ids[q] = it->second;
    }
  }
  return SBR_OK;
}

namespace {

void vtk_header(std::ofstream& ofs, const char* title) {
  ofs << "# vtk DataFile Version 2.0" << std::endl;
  ofs << title << std::endl;
  ofs << "ASCII" << std::endl;
  ofs << "DATASET POLYDATA" << std::endl;
  ofs << std::endl;
}

double vec_norm(const Vec3c& e) {
  return std::sqrt((std::norm(e(0)) + std::norm(e(1))) + std::norm(e(2)));
}

// Sum of the three components' arguments, in radians: the degree conversion the reference writes
// is an expression whose result is discarded (RTSolver.cpp:596, 676).
double phase_sum(const Vec3c& e) {
  double p = std::arg(e(0)) + std::arg(e(1)) + std::arg(e(2));
  return std::isfinite(p) ? p : 0.0;
}

}  // namespace

// RTSolver.cpp:510-610 with PathsToVector (:768-803); paths in (launch, receiver) order.
bool RTSolver::SavePathsAsVtk(std::string fname) {
  std::ofstream ofs(fname);
  vtk_header(ofs, "Ray Path Visualization");
  struct Item { int launch, rx; const std::vector<Ray>* path; };
  std::vector<Item> items;
  for (int r = 0; r < (int)_kept.size(); r++)
    for (auto& lp : _kept[r]) items.push_back({lp.first, r, &lp.second});
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
    return a.launch != b.launch ? a.launch < b.launch : a.rx < b.rx;
  });
  std::vector<Vec3> vertex;
  std::vector<std::vector<int>> lineIdx;
  int totalLines = 0;
  for (const Item& it : items) {
    std::vector<int> lines;
    for (size_t j = 0; j < it.path->size(); j++) {
      if (j == 0) {
        vertex.push_back((*it.path)[j].sourcePoint);
        lines.push_back((int)vertex.size() - 1);
      }
      vertex.push_back((*it.path)[j].targetPoint);
      lines.push_back((int)vertex.size() - 1);
    }
    totalLines += (int)lines.size() + 1;
    lineIdx.push_back(lines);
  }
  ofs << "POINTS " << vertex.size() << " float" << std::endl;
  for (const Vec3& v : vertex) ofs << " " << v.x() << " " << v.y() << " " << v.z() << std::endl;
  ofs << "LINES " << lineIdx.size() << " " << totalLines << std::endl;
  for (auto& l : lineIdx) {
    ofs << " " << l.size() << " ";
    for (int q : l) ofs << q << " ";
    ofs << std::endl;
  }
  const size_t n = items.size();
  ofs << "CELL_DATA " << n << std::endl;
  ofs << "FIELD FieldData 5" << std::endl;
  ofs << "LaunchID 1 " << n << " int" << std::endl;
  for (const Item& it : items) ofs << " " << it.launch << std::endl;
  ofs << "ReceiverID 1 " << n << " int" << std::endl;
  for (const Item& it : items) ofs << " " << it.rx << std::endl;
  ofs << "Eabs 1 " << n << " double" << std::endl;
  for (const Item& it : items) {
    double e = vec_norm(_efield[it.rx]);
    ofs << " " << (std::isfinite(e) ? e : 0.0) << std::endl;
  }
  ofs << "EdB 1 " << n << " double" << std::endl;
  for (const Item& it : items) {
    double e = vec_norm(_efield[it.rx]);
    double edb = std::isfinite(e) ? 20 * std::log10(e * 1e6) : -300.0;
    ofs << " " << (std::isfinite(edb) ? edb : -300.0) << std::endl;
  }
  ofs << "Phase 1 " << n << " double" << std::endl;
  for (const Item& it : items) ofs << " " << phase_sum(_efield[it.rx]) << std::endl;
  return true;
}

// RTSolver.cpp:612-690
bool RTSolver::SaveReceiversAsVtk(std::string fname) {
  std::ofstream ofs(fname);
  vtk_header(ofs, "Ray Path Visualization");
  const VecVec3& rx = *_receivers;
  ofs << "POINTS " << _receiverCount << " float" << std::endl;
  for (int i = 0; i < _receiverCount; i++) ofs << " " << rx[i].x() << " " << rx[i].y() << " " << rx[i].z() << std::endl;
  ofs << "VERTICES 1 " << _receiverCount + 1 << std::endl;
  ofs << _receiverCount << std::endl;
  for (int i = 0; i < _receiverCount; i++) {
    if ((i % 5) == 4) ofs << " " << i << std::endl;
    else ofs << " " << i;
  }
  ofs << std::endl;
  ofs << "POINT_DATA " << _receiverCount << std::endl;
  ofs << "FIELD FieldData 4" << std::endl;
  ofs << "ReceiverID 1 " << _receiverCount << " int" << std::endl;
  for (int i = 0; i < _receiverCount; i++) ofs << " " << i << std::endl;
  ofs << "Eabs 1 " << _receiverCount << " double" << std::endl;
  for (int i = 0; i < _receiverCount; i++) {
    double e = vec_norm(_efield[i]);
    ofs << " " << (std::isfinite(e) ? e : 0.0) << std::endl;
  }
  ofs << "EdB 1 " << _receiverCount << " double" << std::endl;
  for (int i = 0; i < _receiverCount; i++) {
    double e = vec_norm(_efield[i]);
    e = std::isfinite(e) ? e : 0.0;
    ofs << " " << (e > 0.0 ? 20 * std::log10(vec_norm(_efield[i]) * 1e6) : -300.0) << std::endl;
  }
  ofs << "Phase 1 " << _receiverCount << " double" << std::endl;
  for (int i = 0; i < _receiverCount; i++) ofs << " " << phase_sum(_efield[i]) << std::endl;
  return true;
}

// RTSolver.cpp:692-717 (showpos is set on the stream in the first row and stays set).
bool RTSolver::SaveFieldAsCsv(std::string fname) {
  std::ofstream ofs(fname);
  const VecVec3& rx = *_receivers;
  ofs << "X,Y,Z,Ex,Ey,Ez" << std::endl;
  for (int i = 0; i < _receiverCount; i++) {
    ofs << rx[i].x() << "," << rx[i].y() << "," << rx[i].z() << ",";
    ofs << std::showpos << _efield[i].x().real() << _efield[i].x().imag() << "i" << "," << _efield[i].y().real()
        << _efield[i].y().imag() << "i" << "," << _efield[i].z().real() << _efield[i].z().imag() << "i" << std::endl;
  }
  return true;
}

// RTSolver.cpp:719-766
bool RTSolver::SaveIcosahedronAsVtk(std::string fname, Vec3 rayOrg, int tessellation) {
  if (EnsureContext()) return false;
  int64_t n = 10 * (int64_t(1) << (2 * tessellation)) + 2;
  std::vector<double> d(3 * n);
  if (sbr_generate_rays(_ctx, tessellation, d.data(), n, &n)) return false;
  std::ofstream ofs(fname);
  vtk_header(ofs, "TX Icosahedron Visualization");
  ofs << "POINTS " << n + 1 << " float" << std::endl;
  ofs << " " << rayOrg.x() << " " << rayOrg.y() << " " << rayOrg.z() << std::endl;
  for (int64_t i = 0; i < n; i++)
    ofs << " " << rayOrg.x() + 1.0 * d[3 * i] << " " << rayOrg.y() + 1.0 * d[3 * i + 1] << " "
        << rayOrg.z() + 1.0 * d[3 * i + 2] << std::endl;
  ofs << "LINES " << n << " " << 3 * n << std::endl;
  for (int64_t i = 1; i < n + 1; i++) ofs << " 2" << " 0 " << i << std::endl;
  ofs << "CELL_DATA " << n << std::endl;
  ofs << "FIELD FieldData 1" << std::endl;
  ofs << "LaunchID 1 " << n << " int" << std::endl;
  for (int64_t i = 0; i < n; i++) ofs << " " << i << std::endl;
  return true;
}

// RTSolver.cpp:479-508 (kept paths only: the device does not keep uncaptured tree branches here).
void RTSolver::CmdLineDebug() {
  for (int j = 0; j < (int)_kept.size(); j++)
    for (auto& lp : _kept[j]) {
      std::cout << "Launch ID: " << lp.first << "\tReceiver ID: " << j << std::endl;
      for (const Ray& r : lp.second)
        std::cout << "\t\t Source Point: " << r.sourcePoint.x() << " " << r.sourcePoint.y() << " " << r.sourcePoint.z()
                  << "  Target Point: " << r.targetPoint.x() << " " << r.targetPoint.y() << " " << r.targetPoint.z()
                  << "  Hit Triangle ID: " << r.hitSurfaceID << std::endl;
    }
}
