// Host scene preparation: STL parse, exact vertex dedup, triangles with file normals, coplanar ids.
// Behaviour follows Preprocessor.cpp:37-284 and the stl_reader contract (StlReader.h:448-700).
#include "sbrsolver/Preprocessor.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>

namespace {

// Unique vertex table: every facet corner gets the index of its coordinate triple in the
// lexicographically sorted list of distinct triples; facets that lose a corner to a duplicate
// are dropped (StlReader.h:448-508).
void dedup_vertices(const std::vector<float>& corners, std::vector<float>& coords, std::vector<unsigned>& tris) {
  const size_t n = corners.size() / 3;
  std::vector<unsigned> order(n);
  std::iota(order.begin(), order.end(), 0u);
  auto less = [&](unsigned a, unsigned b) {
    const float* p = &corners[3 * a];
    const float* q = &corners[3 * b];
    return (p[0] < q[0]) || (p[0] == q[0] && p[1] < q[1]) || (p[0] == q[0] && p[1] == q[1] && p[2] < q[2]);
  };
  std::sort(order.begin(), order.end(), less);
  std::vector<unsigned> remap(n);
  coords.clear();
  for (size_t i = 0; i < n; i++) {
    const float* p = &corners[3 * order[i]];
    bool fresh = i == 0;
    if (!fresh) {
      const float* q = &corners[3 * order[i - 1]];
      fresh = p[0] != q[0] || p[1] != q[1] || p[2] != q[2];
    }
    if (fresh) coords.insert(coords.end(), p, p + 3);
    remap[order[i]] = (unsigned)(coords.size() / 3 - 1);
  }
  std::vector<unsigned> kept;
  for (size_t f = 0; f + 2 < n; f += 3) {
    unsigned a = remap[f], b = remap[f + 1], c = remap[f + 2];
    if (a != b && a != c && b != c) { kept.push_back(a); kept.push_back(b); kept.push_back(c); }
  }
  tris.swap(kept);
}

bool looks_ascii(const std::string& fn) {
  std::ifstream in(fn);
  if (!in) return false;
  char buf[256];
  in.read(buf, sizeof buf);
  std::string head(buf, (size_t)in.gcount());
  for (char& ch : head) ch = (char)std::tolower((unsigned char)ch);
  return head.find("solid") != std::string::npos && head.find('\n') != std::string::npos &&
         head.find("facet") != std::string::npos && head.find("normal") != std::string::npos;
}

// Triangles from deduplicated geometry.  As in Preprocessor.cpp:87-101 the normal (and here the
// material id) of triangle k is read from facet slot k of the input, i.e. after dropped facets.
bool make_triangles(const std::vector<float>& coords, const std::vector<float>& normals,
                    const std::vector<unsigned>& tris, const int32_t* mats, std::vector<Triangle*>& out,
                    std::map<std::pair<int, int>, std::vector<int>>* edgesOut = nullptr) {
  const size_t nt = tris.size() / 3;
  if (normals.size() < 3 * nt) return false;
  std::map<std::pair<int, int>, std::vector<int>> edges;
  std::vector<std::vector<int>> adj(nt);
  const size_t base = out.size();
  for (size_t t = 0; t < nt; t++) {
    size_t i0 = 3 * (size_t)tris[3 * t], i1 = 3 * (size_t)tris[3 * t + 1], i2 = 3 * (size_t)tris[3 * t + 2];
    Vec3 a(coords[i0], coords[i0 + 1], coords[i0 + 2]);
    Vec3 b(coords[i1], coords[i1 + 1], coords[i1 + 2]);
    Vec3 c(coords[i2], coords[i2 + 1], coords[i2 + 2]);
    Vec3 nrm(normals[3 * t], normals[3 * t + 1], normals[3 * t + 2]);
    Triangle* tri = new Triangle(a, b, c, nrm, (int)t);
    if (mats) tri->materialId = mats[t];
    out.push_back(tri);
    // Two triangles are coplanar neighbours when they share an edge and their stored normals are
    // equal or opposite (Preprocessor.cpp:116-190).
    const size_t corner[3][2] = {{i0, i1}, {i1, i2}, {i0, i2}};
    for (auto& e : corner) {
      std::pair<int, int> key((int)std::min(e[0], e[1]), (int)std::max(e[0], e[1]));
      std::vector<int>& users = edges[key];
      for (int u : users) {
        const Vec3& nu = out[base + u]->norm;
        if (nu == nrm || nu == -nrm) { adj[u].push_back((int)t); adj[t].push_back(u); }
      }
      users.push_back((int)t);
    }
  }
  // Coplanar id = index of the connected component, numbered by its lowest triangle.
  std::vector<char> seen(nt, 0);
  int cop = 0;
  std::vector<int> queue;
  for (size_t s = 0; s < nt; s++) {
    if (seen[s]) continue;
    queue.assign(1, (int)s);
    seen[s] = 1;
    for (size_t h = 0; h < queue.size(); h++) {
      int u = queue[h];
      out[base + u]->coplanarId = cop;
      for (int w : adj[u])
        if (!seen[w]) { seen[w] = 1; queue.push_back(w); }
    }
    cop++;
  }
  if (edgesOut) edgesOut->swap(edges);
  return true;
}

}  // namespace

namespace Preprocessor {

bool ReadStl(const std::string& fn, std::vector<float>& coords, std::vector<float>& normals,
             std::vector<unsigned>& tris) {
  std::vector<float> corners;
  normals.clear();
  if (looks_ascii(fn)) {
    std::ifstream in(fn);
    if (!in) return false;
    std::string line, tok;
    int inFacet = 0;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::vector<std::string> w;
      while (ls >> tok) w.push_back(tok);
      if (w.empty()) continue;
      if (w[0] == "facet") {
        if (w.size() < 5 || w[1] != "normal") return false;
        for (int i = 0; i < 3; i++) normals.push_back((float)std::atof(w[2 + i].c_str()));
        inFacet = 0;
      } else if (w[0] == "vertex") {
        if (w.size() < 4) return false;
        for (int i = 0; i < 3; i++) corners.push_back((float)std::atof(w[1 + i].c_str()));
        inFacet++;
      } else if (w[0] == "endfacet") {
        if (inFacet != 3) return false;
      }
    }
  } else {
    std::ifstream in(fn, std::ios::binary);
    if (!in) return false;
    char header[80];
    uint32_t count = 0;
    in.read(header, 80);
    in.read((char*)&count, 4);
    if (!in) return false;
    for (uint32_t f = 0; f < count; f++) {
      float rec[12];
      char attr[2];
      in.read((char*)rec, sizeof rec);
      in.read(attr, 2);
      if (!in) return false;
      normals.insert(normals.end(), rec, rec + 3);
      corners.insert(corners.end(), rec + 3, rec + 12);
    }
  }
  if (corners.empty()) return false;
  dedup_vertices(corners, coords, tris);
  return true;
}

bool StlToGeometry(std::string fileName, std::vector<Triangle*>& output, bool saveEdges, bool saveFaces,
                   std::string dataDir) {
  std::vector<float> coords, normals;
  std::vector<unsigned> tris;
  if (!ReadStl(fileName, coords, normals, tris)) return false;
  const size_t first = output.size();
  std::map<std::pair<int, int>, std::vector<int>> edges;
  if (!make_triangles(coords, normals, tris, nullptr, output, saveEdges ? &edges : nullptr)) return false;
  if (saveEdges) SaveEdgesAsVtk(dataDir + "Edges.vtk", edges, coords);
  if (saveFaces) {
    const std::vector<Triangle*> mine(output.begin() + first, output.end());
    SaveFacesAsVtk(dataDir + "Faces.vtk", mine, coords, tris);
  }
  return true;
}

bool SaveEdgesAsVtk(std::string fileName, const std::map<std::pair<int, int>, std::vector<int>>& edgeMap,
                    const std::vector<float>& coords) {
  std::ofstream ofs(fileName);
  ofs << "# vtk DataFile Version 2.0" << std::endl;
  ofs << "3D model of edge geometry" << std::endl;
  ofs << "ASCII" << std::endl;
  ofs << "DATASET POLYDATA" << std::endl;
  ofs << std::endl;
  ofs << "POINTS " << int(coords.size() / 3) << " float" << std::endl;
  // one point per line; the last line runs into the LINES header, as the reference writes it
  for (size_t i = 0; i < coords.size(); i++) {
    if (i > 0 && i % 3 == 0) ofs << std::endl;
    ofs << " " << coords[i];
  }
  ofs << "LINES " << edgeMap.size() << " " << 3 * edgeMap.size() << std::endl;
  for (const auto& kv : edgeMap) ofs << " 2 " << kv.first.first / 3 << " " << kv.first.second / 3 << std::endl;
  ofs << "CELL_DATA " << edgeMap.size() << std::endl;
  ofs << "FIELD FieldData 2" << std::endl;
  ofs << "EdgeID 1 " << edgeMap.size() << " int" << std::endl;
  for (size_t i = 0; i < edgeMap.size(); i++) ofs << " " << i << std::endl;
  ofs << "Connection 1 " << edgeMap.size() << " int" << std::endl;
  for (const auto& kv : edgeMap) ofs << " " << int(kv.second.size()) << std::endl;
  return (bool)ofs;
}

bool SaveFacesAsVtk(std::string fileName, const std::vector<Triangle*>& triangles, const std::vector<float>& coords,
                    const std::vector<unsigned int> tris) {
  std::ofstream ofs(fileName);
  ofs << "# vtk DataFile Version 2.0" << std::endl;
  ofs << "3D model of face geometry" << std::endl;
  ofs << "ASCII" << std::endl;
  ofs << "DATASET POLYDATA" << std::endl;
  ofs << std::endl;
  ofs << "POINTS " << int(coords.size() / 3) << " float" << std::endl;
  for (size_t i = 0; i < coords.size(); i++) {
    if (i > 0 && i % 3 == 0) ofs << std::endl;
    ofs << " " << coords[i];
  }
  ofs << std::endl;
  const int numTris = int(tris.size() / 3);
  ofs << "POLYGONS " << numTris << " " << 4 * numTris << std::endl;
  for (int i = 0; i < numTris; i++)
    ofs << " 3 " << tris[3 * i] << " " << tris[3 * i + 1] << " " << tris[3 * i + 2] << std::endl;
  ofs << "CELL_DATA " << numTris << std::endl;
  ofs << "FIELD FieldData 3" << std::endl;
  ofs << "FaceID 1 " << triangles.size() << " int" << std::endl;
  for (const Triangle* t : triangles) ofs << " " << t->triangleId << std::endl;
  ofs << "CoplanarID 1 " << triangles.size() << " int" << std::endl;
  for (const Triangle* t : triangles) ofs << " " << t->coplanarId << std::endl;
  ofs << "MaterialID 1 " << triangles.size() << " int" << std::endl;
  for (const Triangle* t : triangles) ofs << " " << t->materialId << std::endl;
  return (bool)ofs;
}

bool SoupToGeometry(const float* coords9, const float* normals3, const int32_t* mats, int nFacets,
                    std::vector<Triangle*>& output) {
  if (nFacets <= 0) return false;
  std::vector<float> corners(coords9, coords9 + 9 * (size_t)nFacets), normals(normals3, normals3 + 3 * (size_t)nFacets);
  std::vector<float> coords;
  std::vector<unsigned> tris;
  dedup_vertices(corners, coords, tris);
  return make_triangles(coords, normals, tris, mats, output);
}

bool ReadPatternFile(std::string fileName, GainMap& output) {
  std::ifstream in(fileName);
  if (!in) return false;
  std::string line, tok;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::vector<std::string> w;
    while (ls >> tok) w.push_back(tok);
    if (w.size() != 6) continue;
    float v[6];
    bool ok = true;
    for (int i = 0; i < 6 && ok; i++) {
      char* end = nullptr;
      v[i] = std::strtof(w[i].c_str(), &end);
      ok = end && *end == 0 && end != w[i].c_str();
    }
    if (!ok) continue;
    GainInfo g{};
    for (int i = 0; i < 4; i++) g.data[i] = v[i + 2];
    output.insert({ThetaPhiAngle{v[0], v[1]}, g});   // first entry of an angle pair wins, as std::map::insert
  }
  return true;
}

bool ReadLocationFile(std::string fileName, VecVec3& output) {
  std::ifstream in(fileName);
  if (!in) return false;
  std::string line, tok;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::vector<std::string> w;
    while (ls >> tok) w.push_back(tok);
    if (w.size() != 3) continue;
    char* end = nullptr;
    Vec3 p;
    bool ok = true;
    for (int i = 0; i < 3; i++) {
      p(i) = std::strtod(w[i].c_str(), &end);
      ok = ok && end && *end == 0;
    }
    if (ok) output.push_back(p);
  }
  return true;
}

void GenerateRxPlane(double xMin, double yMin, double xMax, double yMax, double height, double resolution,
                     VecVec3& output) {
  int nx = int((xMax - xMin) / resolution), ny = int((yMax - yMin) / resolution);
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j <= ny; j++) output.emplace_back(xMin + i * resolution, yMin + j * resolution, height);
}

void GenerateRXLine(Vec3 start, Vec3 end, double resolution, VecVec3& output) {
  sbr::V3 d = sbr::vsub(end.v(), start.v());
  int numPoints = (int)std::floor(sbr::vnorm(d) / resolution);
  d = sbr::vnormalized(d);
  for (int i = 0; i < numPoints; i++) output.push_back(Vec3::of(sbr::vadd(start.v(), sbr::smul(i * resolution, d))));
}

}  // namespace Preprocessor
