// RTSolver: the reference's tracing entry point (RTSolver.h:18-83) on the gfx950 library.
//
// Same call surface: Init(materials, count, bvh) keeps a non-owning BVH pointer and copies the
// materials; ExecuteRayTracing(...) takes the same arguments, keeps non-owning pointers to the
// receivers and the mesh, and returns N_rays * N_rx like the reference (RTSolver.cpp:52-126); the
// Save* writers produce the reference's file formats (RTSolver.cpp:510-766).  The trace itself runs
// through the C-ABI (sbr_capi.h).  Each trace flattens the live mesh and BVH, as the reference reads
// them on every call; the device copy is replaced whenever their content hash changes and stays
// resident otherwise.  Extensions: result getters, trace options, batch tracing.
#pragma once

#include <string>
#include <vector>

#include "../sbr_capi.h"
#include "BVH.h"
#include "Preprocessor.h"
#include "Ray.h"
#include "Triangle.h"
#include "VecMatDef.h"

struct RTSolverOptions {
  int antennaType = 1;        // FieldCompute.cpp:117 hard-codes 1 (Hertzian dipole); 3: SetTxPattern's table
  int receptionModel = 0;     // 0: reference sphere (RTSolver.cpp:838-841); 1: alpha*L/sqrt(3)
  int maxTotal = 0;           // cap on reflections + transmissions, 0 = none (reference)
  int q1MeshIndexing = 1;     // 1: surface data of mesh[hitSurfaceID] (reference); 0: by id
  double wallWidth = 0;       // > 0: slab walls of this thickness in m (physical mode); 0: reference walls
  int typedDedup = 0;         // 1: duplicate removal keyed by coplanar ids and interaction types
  int checkOcclusion = 0;     // 1: drop corrected paths with a scene triangle on a segment
  int longMiss = 0;           // 1: misses after a bounce extend 1e10 m instead of a 100 m stub
  int physicalField = 0;      // 1: s/p field transport in the plane of incidence
  // Every correction of the reference's quirks at once (wallWidth, a thickness, is left as it is).
  void SetPhysical() {
    q1MeshIndexing = 0;
    receptionModel = 1;
    typedDedup = checkOcclusion = longMiss = physicalField = 1;
  }
  int device = 0;
  // _rayPaths (the reference's Paths[launches][receivers], RTSolver.h:45) is filled after a trace
  // when launches * receivers is at most this (24 bytes per entry; the reference's array at the
  // urban canyon's 2.6 M launches x 10 k receivers would be 629 GB); otherwise it stays null.
  int64_t rayPathArrayLimit = int64_t(1) << 22;
  // 1: number the kept paths' segments with the reference's Ray::id (the global creation order of
  // tree nodes, reflected child first, RTSolver.cpp:82-84, Ray.cpp:18); this reads every tree node
  // of the trace back from the device.  0: id stays -1.
  int rayIds = 0;
};

class RTSolver {
 public:
  RTSolver();
  virtual ~RTSolver();
  int Init(MaterialProperties materialProperties[], int materialPropertiesCount, BVH<Triangle>& triangles);
  void Cleanup();
  int ExecuteRayTracing(Vec3 sourcePoint, const Mat3& txCoordSys, int maxReflectionCount, int maxTransmissionCount,
                        double frequency, double pt, bool useFresnelCoeff, VecVec3& receivers,
                        std::vector<Triangle*>& triangleMesh, int txTesslation);
  bool SavePathsAsVtk(std::string fname);
  bool SaveReceiversAsVtk(std::string fname);
  bool SaveFieldAsCsv(std::string fname);
  bool SaveIcosahedronAsVtk(std::string fname, Vec3 rayOrg, int tessellation);
  void CmdLineDebug();

  // ---- extensions (no reference counterpart)
  RTSolverOptions options;
  const std::vector<Vec3c>& Fields() const { return _efield; }
  // Kept paths per receiver: (launch id, segments), launch ascending then path order.
  const std::vector<std::vector<std::pair<int, std::vector<Ray>>>>& ReceiverPaths() const { return _kept; }
  // The kept paths as the reference's array (see RTSolverOptions::rayPathArrayLimit), for
  // FieldCompute(RayPaths(), mesh, receivers, launches, ...); null when not filled.
  Paths** RayPaths() const { return _rayPaths; }
  int LaunchCount() const { return _pathsCount; }
  const sbr_stats& Stats() const { return _stats; }
  const char* LastError() const;
  // Transmitter pattern from a TXPattern.dat map (Preprocessor::ReadPatternFile), used by the
  // following traces (sets options.antennaType = 3).  SBR_E_ARG unless the map is a full, evenly
  // spaced (theta, phi) grid.
  int SetTxPattern(const GainMap& pattern);

 protected:
  MaterialProperties* _materialProperties = nullptr;
  int _materialCount = 0;
  BVH<Triangle>* _bvh = nullptr;
  VecVec3* _receivers = nullptr;
  std::vector<Vec3> _shootRayList;
  std::vector<Vec3c> _efield;
  std::vector<std::vector<std::pair<int, std::vector<Ray>>>> _kept;
  Paths** _rayPaths = nullptr;   // [_pathsCount][_receiverCount] when filled
  int _rayPathRows = 0;
  sbr_stats _stats{};
  bool _isSceneInitialized = false;
  int _txTesslation = 0;
  int _pathsCount = 0;
  int _receiverCount = 0;

 private:
  sbr_ctx* _ctx = nullptr;
  std::vector<float> _pattern;   // GainInfo rows, theta-major
  int _patternRows = 0, _patternCols = 0, _patternVersion = 0, _patternSent = -1;
  uint64_t _uploadedFingerprint = 0;   // content hash of the scene on the device, 0: none
  int EnsureContext();
  void FreeRayPaths();
  int NumberTreeNodes(const std::vector<int32_t>& prx, const std::vector<int32_t>& pl, const std::vector<int32_t>& pn,
                      const std::vector<int64_t>& poff, const std::vector<int32_t>& stri,
                      const std::vector<int32_t>& stype, std::vector<int64_t>& ids);
  static double AngleBetweenDeg(sbr::V3 a, sbr::V3 b);
};
