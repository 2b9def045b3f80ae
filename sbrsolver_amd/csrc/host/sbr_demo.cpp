// The reference's demo driver (SBRSolver_rev2.cpp:18-144) on the gfx950 library: ground scene,
// 3000 receivers on a line, Tx (0, 0, 5), icosahedron level 2, one reflection, 1.8 GHz, metal.
// Usage: sbr_demo <ground.stl> <output dir>
#include <iostream>
#include <string>

#include "sbrsolver/BVH.h"
#include "sbrsolver/Preprocessor.h"
#include "sbrsolver/RTSolver.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::cerr << "usage: sbr_demo <ground.stl> <output dir>" << std::endl;
    return 2;
  }
  const std::string stl = argv[1], out = std::string(argv[2]) + "/";
  VecVec3 rxLocation;
  std::vector<Triangle*> triangle_mesh;
  BVH<Triangle> bvh;
  Preprocessor::GenerateRXLine(Vec3(0, 0, 3), Vec3(0, 3000, 3), 1, rxLocation);
  if (!Preprocessor::StlToGeometry(stl, triangle_mesh, false, false, out)) {
    std::cerr << "cannot read " << stl << std::endl;
    return 1;
  }
  bvh.ConstructBVH(triangle_mesh);
  MaterialProperties materials[1];
  materials[0].frequency = 1.8e9;
  materials[0].reflectionLoss = 0.05;
  materials[0].transmissionLoss = 221.13;
  materials[0].relConductivity = 1e6;
  materials[0].relPermittivityRe = 1;
  materials[0].relPermittivityIm = 0;
  Mat3 txCoordSys{{0, 1, 0}, {0, 0, 1}, {1, 0, 0}};
  RTSolver* rayTracer = new RTSolver();
  rayTracer->Init(materials, 1, bvh);
  int total = rayTracer->ExecuteRayTracing(Vec3(0, 0, 5), txCoordSys, 1, 0, 1.8e9, 1, true, rxLocation, triangle_mesh, 2);
  if (total < 0) {
    std::cerr << "trace failed: " << rayTracer->LastError() << std::endl;
    return 1;
  }
  rayTracer->SavePathsAsVtk(out + "RayPath.vtk");
  rayTracer->SaveReceiversAsVtk(out + "RxLocations.vtk");
  rayTracer->SaveFieldAsCsv(out + "ElectricField.csv");
  rayTracer->SaveIcosahedronAsVtk(out + "Icosahedron.vtk", Vec3(0, 0, 5), 3);
  std::cout << "launches x receivers = " << total << ", bounces " << rayTracer->Stats().bounces << std::endl;
  delete rayTracer;
  for (Triangle* t : triangle_mesh) delete t;
  return 0;
}
