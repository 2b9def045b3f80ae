// RTSolver on the gfx950 library: the reference's call surface and output writers
// (RTSolver.cpp:1-803), with the trace delegated to the C-ABI.
#include "sbrsolver/RTSolver.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>

#include "sbrsolver/SceneDesc.h"

RTSolver::RTSolver() {}

RTSolver::~RTSolver() {
  Cleanup();
  if (_ctx) sbr_destroy(_ctx);
  _ctx = nullptr;
}

void RTSolver::Cleanup() {
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
  _kept.assign(_receiverCount, {});
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
    }
    _kept[prx[k]].push_back(std::make_pair((int)pl[k], std::move(path)));
  }
  _shootRayList.clear();
  return _pathsCount * _receiverCount;
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
