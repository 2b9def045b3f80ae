// FieldCompute of the reference call surface (FieldCompute.h:9-94) on the gfx950 library.
//
// Same constructor and methods: the object borrows the per-(launch, receiver) path lists, the mesh
// vector (surfaces addressed by hitSurfaceID through it, as the reference does), the receivers and
// the materials.  FieldAtReceiver(k) sums FieldForPath over launch i ascending and then path order
// (FieldCompute.cpp:11-25); each path's field is computed on the device by the same transport the
// trace uses (sbr_path_fields), and the sum is taken in that order, so for the paths an RTSolver kept
// it equals RTSolver::Fields() bit for bit.  FieldAtReceiver reads the paths of receiver k when it is
// called, as the reference does (RTSolver::ExecuteRayTracing corrects and de-duplicates receiver k's
// paths just before asking for its field).  RefCoeffTest / TransCoeffTest write the coefficient
// sweeps of FieldCompute.cpp:27-69, computed on the device (sbr_fresnel_sweep).
//
// Errors: the reference has none to report; a failed device call yields NaN fields and a CSV that is
// not written, with the status in LastStatus() and the message in LastError().
#pragma once

#include <string>
#include <vector>

#include "../sbr_capi.h"
#include "Ray.h"
#include "Triangle.h"
#include "VecMatDef.h"

class FieldCompute {
 public:
  FieldCompute();
  FieldCompute(Paths** rayPaths, std::vector<Triangle*>& triangleMesh, VecVec3& receivers, int pathsCount,
               MaterialProperties* materials, Mat3 txCoordSys, double frequency, double pt, bool useFresnelCoeff);
  ~FieldCompute();
  FieldCompute(const FieldCompute&) = delete;
  FieldCompute& operator=(const FieldCompute&) = delete;

  Vec3c FieldAtReceiver(int receiverId);
  void RefCoeffTest(int numPts, double freq, cdouble epsilon1, cdouble epsilon2, std::string fname);
  void TransCoeffTest(int numPts, double freq, cdouble epsilon1, cdouble epsilon2, std::string fname);

  // ---- extensions (no reference counterpart)
  int antennaType = 1;   // FieldCompute.cpp:117 hard-codes 1 (Hertzian dipole); 0 isotropic, 2 half-wave
  int device = 0;        // HIP device of the context the object creates on first use
  // Every receiver's field in one device call; equal to FieldAtReceiver(k) for each k.
  std::vector<Vec3c> FieldAtAllReceivers();
  int LastStatus() const { return _status; }
  const char* LastError() const;

 protected:
  Paths** _rayPaths = nullptr;
  std::vector<Triangle*>* _triangleMesh = nullptr;
  int _pathsCount = 0;
  VecVec3* _receivers = nullptr;
  MaterialProperties* _materials = nullptr;
  Mat3 _txCoordSys;
  double _frequency = 0.0;
  double _txPower = 0.0;
  bool _useFresnelCoeff = true;

  // FieldCompute.cpp:71-151 for one path (one device call).
  Vec3c FieldForPath(const std::vector<Ray>& path);

 private:
  sbr_ctx* _ctx = nullptr;
  int _status = 0;
  int EnsureContext();
  // Fields of the given paths, in order, into out (NaN on failure).
  void PathFields(const std::vector<const std::vector<Ray>*>& paths, std::vector<Vec3c>& out);
  void CoeffSweep(int numPts, double freq, cdouble epsilon1, cdouble epsilon2, std::string fname, bool reflection);
};
