// FieldCompute on the gfx950 library (FieldCompute.h:9-94, FieldCompute.cpp:11-69): paths are
// flattened into segment arrays and handed to sbr_path_fields; receiver sums are taken on the host in
// the reference's order.
#include "sbrsolver/FieldCompute.h"

#include <cmath>
#include <fstream>
#include <limits>
#include <unordered_map>

FieldCompute::FieldCompute() {}

FieldCompute::FieldCompute(Paths** rayPaths, std::vector<Triangle*>& triangleMesh, VecVec3& receivers, int pathsCount,
                           MaterialProperties* materials, Mat3 txCoordSys, double frequency, double pt,
                           bool useFresnelCoeff)
    : _rayPaths(rayPaths),
      _triangleMesh(&triangleMesh),
      _pathsCount(pathsCount),
      _receivers(&receivers),
      _materials(materials),
      _txCoordSys(txCoordSys),
      _frequency(frequency),
      _txPower(pt),
      _useFresnelCoeff(useFresnelCoeff) {}

FieldCompute::~FieldCompute() {
  if (_ctx) sbr_destroy(_ctx);
  _ctx = nullptr;
}

const char* FieldCompute::LastError() const { return _ctx ? sbr_last_error(_ctx) : "no context"; }

int FieldCompute::EnsureContext() {
  if (_ctx) return SBR_OK;
  return sbr_create(device, &_ctx);
}

void FieldCompute::PathFields(const std::vector<const std::vector<Ray>*>& paths, std::vector<Vec3c>& out) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  out.assign(paths.size(), Vec3c(cdouble(nan, nan), cdouble(nan, nan), cdouble(nan, nan)));
  if (paths.empty()) {
    _status = SBR_OK;
    return;
  }
  _status = EnsureContext();
  if (_status) return;
  // Segment arrays; surfaces are renumbered to the ones the paths name, whose frames are read from
  // the mesh vector by hitSurfaceID (FieldCompute.cpp:107).
  std::vector<int64_t> off;
  std::vector<int32_t> cnt, surf, type, mat;
  std::vector<double> src, tgt, frames;
  std::unordered_map<int, int32_t> surfIndex;
  int maxMat = -1;
  off.reserve(paths.size());
  cnt.reserve(paths.size());
  for (const std::vector<Ray>* p : paths) {
    off.push_back((int64_t)type.size());
    cnt.push_back((int32_t)p->size());
    for (const Ray& r : *p) {
      for (int a = 0; a < 3; a++) {
        src.push_back(r.sourcePoint(a));
        tgt.push_back(r.targetPoint(a));
      }
      int32_t s = -1;
      if (r.hitSurfaceID != -1) {
        auto it = surfIndex.find(r.hitSurfaceID);
        if (it == surfIndex.end()) {
          if (!_triangleMesh || r.hitSurfaceID < 0 || r.hitSurfaceID >= (int)_triangleMesh->size()) {
            _status = SBR_E_ARG;
            return;
          }
          const int32_t k = (int32_t)surfIndex.size();
          it = surfIndex.emplace(r.hitSurfaceID, k).first;
          double cs[9];
          (*_triangleMesh)[r.hitSurfaceID]->coordSys.to_row_major(cs);
          frames.insert(frames.end(), cs, cs + 9);
        }
        s = it->second;
      }
      surf.push_back(s);
      // the interaction is the one the segment left its start surface by (FieldCompute.cpp:126-145)
      if (r.reflectionMaterialId >= 0) {
        type.push_back(1);
        mat.push_back(r.reflectionMaterialId);
      } else if (r.penetrationMaterialId >= 0) {
        type.push_back(2);
        mat.push_back(r.penetrationMaterialId);
      } else {
        type.push_back(0);
        mat.push_back(0);
      }
      if (type.back() != 0) maxMat = std::max(maxMat, mat.back());
    }
  }
  if (maxMat >= 0 && !_materials) {
    _status = SBR_E_ARG;
    return;
  }
  // the materials the paths name: the reference indexes _materials by those ids
  const int nMats = maxMat + 1;
  sbr_field_params fp{};
  _txCoordSys.to_row_major(fp.txCoordSys);
  fp.frequency = _frequency;
  fp.pt = _txPower;
  fp.useFresnel = _useFresnelCoeff ? 1 : 0;
  fp.antennaType = antennaType;
  std::vector<double> f(6 * paths.size());
  _status = sbr_path_fields(_ctx, &fp, reinterpret_cast<const sbr_material*>(_materials), nMats, (int64_t)paths.size(),
                            off.data(), cnt.data(), src.data(), tgt.data(), surf.data(), type.data(), mat.data(),
                            (int32_t)surfIndex.size(), frames.empty() ? nullptr : frames.data(), f.data());
  if (_status) return;
  for (size_t k = 0; k < paths.size(); k++)
    for (int c = 0; c < 3; c++) out[k](c) = cdouble(f[6 * k + 2 * c], f[6 * k + 2 * c + 1]);
}

Vec3c FieldCompute::FieldForPath(const std::vector<Ray>& path) {
  std::vector<Vec3c> f;
  PathFields({&path}, f);
  return f[0];
}

// FieldCompute.cpp:11-25
Vec3c FieldCompute::FieldAtReceiver(int receiverId) {
  std::vector<const std::vector<Ray>*> paths;
  for (int i = 0; i < _pathsCount; i++)
    for (const std::vector<Ray>& p : _rayPaths[i][receiverId].rayPaths) paths.push_back(&p);
  std::vector<Vec3c> f;
  PathFields(paths, f);
  Vec3c total(cdouble(0, 0), cdouble(0, 0), cdouble(0, 0));
  for (const Vec3c& e : f)
    for (int c = 0; c < 3; c++) total(c) += e(c);
  return total;
}

std::vector<Vec3c> FieldCompute::FieldAtAllReceivers() {
  const int nrx = _receivers ? (int)_receivers->size() : 0;
  std::vector<const std::vector<Ray>*> paths;
  std::vector<int> owner;
  for (int k = 0; k < nrx; k++)
    for (int i = 0; i < _pathsCount; i++)
      for (const std::vector<Ray>& p : _rayPaths[i][k].rayPaths) {
        paths.push_back(&p);
        owner.push_back(k);
      }
  std::vector<Vec3c> f;
  PathFields(paths, f);
  std::vector<Vec3c> total(nrx, Vec3c(cdouble(0, 0), cdouble(0, 0), cdouble(0, 0)));
  for (size_t q = 0; q < f.size(); q++)
    for (int c = 0; c < 3; c++) total[owner[q]](c) += f[q](c);
  return total;
}

void FieldCompute::CoeffSweep(int numPts, double freq, cdouble epsilon1, cdouble epsilon2, std::string fname,
                              bool reflection) {
  _status = EnsureContext();
  if (_status) return;
  if (numPts <= 0) {
    _status = SBR_E_ARG;
    return;
  }
  std::vector<double> theta(numPts), c(8 * (size_t)numPts);
  const double e1[2] = {epsilon1.real(), epsilon1.imag()}, e2[2] = {epsilon2.real(), epsilon2.imag()};
  _status = sbr_fresnel_sweep(_ctx, numPts, freq, e1, e2, 0.0, 1, theta.data(), c.data());
  if (_status) return;
  std::ofstream ofs(fname);
  if (reflection) ofs << "theta_i,refTM_mag,refTM_phase,refTE_mag, refTE_phase" << std::endl;
  else ofs << "theta_i,tranTM_mag,tranTM_phase, tranTE_mag, tranTE_phase" << std::endl;
  const double kPi = 3.141592653589793;
  for (int i = 0; i < numPts; i++) {
    // rTM, tTM, rTE, tTE as (re, im)
    const cdouble tm(c[8 * i + (reflection ? 0 : 2)], c[8 * i + (reflection ? 1 : 3)]);
    const cdouble te(c[8 * i + (reflection ? 4 : 6)], c[8 * i + (reflection ? 5 : 7)]);
    ofs << theta[i] * (180.0 / kPi) << "," << std::abs(tm) << "," << std::arg(tm) << ",";
    ofs << std::abs(te) << "," << std::arg(te) << std::endl;
  }
}

// FieldCompute.cpp:27-47
void FieldCompute::RefCoeffTest(int numPts, double freq, cdouble epsilon1, cdouble epsilon2, std::string fname) {
  CoeffSweep(numPts, freq, epsilon1, epsilon2, fname, true);
}

// FieldCompute.cpp:49-69
void FieldCompute::TransCoeffTest(int numPts, double freq, cdouble epsilon1, cdouble epsilon2, std::string fname) {
  CoeffSweep(numPts, freq, epsilon1, epsilon2, fname, false);
}
