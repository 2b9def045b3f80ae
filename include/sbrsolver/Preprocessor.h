// Host scene preparation of the reference call surface (Preprocessor.h:18-76).
#pragma once

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "Triangle.h"
#include "VecMatDef.h"

// Antenna gain information (DataStructures.h:32-37): theta gain dB, theta phase deg, phi gain dB,
// phi phase deg, keyed by (theta, phi) in degrees.
struct GainInfo {
  float data[4];
};
using ThetaPhiAngle = std::pair<float, float>;
using GainMap = std::map<ThetaPhiAngle, GainInfo>;

namespace Preprocessor {

// TXPattern.dat reader (Preprocessor.cpp:3-35): every line of exactly six numbers (theta, phi,
// theta gain, theta phase, phi gain, phi phase) becomes one entry; other lines (the header) are
// skipped, as are lines whose six tokens are not all numbers (the reference's std::stof would
// throw there).  Returns false when the file cannot be opened.
bool ReadPatternFile(std::string fileName, GainMap& output);

// STL (ASCII or binary) -> triangles with the file normals and coplanar ids
// (Preprocessor.cpp:69-114).  saveEdges / saveFaces write dataDir + "Edges.vtk" / "Faces.vtk"
// (the dedup'd vertices with the edge map / the triangles with their ids).  Returns false when the
// file cannot be read.
bool StlToGeometry(std::string fileName, std::vector<Triangle*>& output, bool saveEdges = false,
                   bool saveFaces = false, std::string dataDir = "");

// Edge map of the scene (key: the two vertex offsets into coords, 3 per vertex; value: the
// triangles sharing the edge) as VTK lines with their connection counts (Preprocessor.cpp:290-351).
bool SaveEdgesAsVtk(std::string fileName, const std::map<std::pair<int, int>, std::vector<int>>& edgeMap,
                    const std::vector<float>& coords);
// Triangles as VTK polygons with FaceID, CoplanarID and MaterialID (Preprocessor.cpp:354-421).
bool SaveFacesAsVtk(std::string fileName, const std::vector<Triangle*>& triangles, const std::vector<float>& coords,
                    const std::vector<unsigned int> tris);

// Facet arrays as a binary STL holds them (9 vertex floats + 3 normal floats per facet), with
// optional per-facet material ids; same vertex dedup, normal indexing and coplanar grouping.
bool SoupToGeometry(const float* coords9, const float* normals3, const int32_t* materialIds, int nFacets,
                    std::vector<Triangle*>& output);

// Raw STL parse with exact vertex dedup (the stl_reader contract, StlReader.h:448-700).
bool ReadStl(const std::string& fileName, std::vector<float>& coords, std::vector<float>& normals,
             std::vector<unsigned>& tris);

bool ReadLocationFile(std::string fileName, VecVec3& output);                       // :37-67
void GenerateRxPlane(double xMin, double yMin, double xMax, double yMax, double height, double resolution,
                     VecVec3& output);                                                 // :254-268
void GenerateRXLine(Vec3 start, Vec3 end, double resolution, VecVec3& output);         // :270-284

}  // namespace Preprocessor
