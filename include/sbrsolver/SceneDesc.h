// Builds the sbr_scene_desc arrays from the reference-style scene objects: the caller's mesh vector
// (after BVH<Triangle>::ConstructBVH sorted it) and the BVH itself.
#pragma once

#include <cstdint>
#include <vector>

#include "../sbr_capi.h"
#include "BVH.h"
#include "Triangle.h"

struct SceneArrays {
  std::vector<double> tri_v, tri_n, tri_cs, node_box;
  std::vector<int32_t> tri_cop, tri_mat, mesh_tri_id, node_child, node_first, node_count, leaf_tri;
  sbr_scene_desc desc{};

  // `mesh` is the vector the reference hands to ExecuteRayTracing: triangle k of it is what
  // mesh[hitSurfaceID = k] names.  Triangles are addressed by their triangleId, which must be a
  // permutation of 0..n-1.
  bool build(const std::vector<Triangle*>& mesh, const BVH<Triangle>& bvh) {
    const int n = (int)mesh.size();
    std::vector<const Triangle*> byId(n, nullptr);
    for (const Triangle* t : mesh) {
      if (!t || t->triangleId < 0 || t->triangleId >= n || byId[t->triangleId]) return false;
      byId[t->triangleId] = t;
    }
    tri_v.resize(9 * (size_t)n);
    tri_n.resize(3 * (size_t)n);
    tri_cs.resize(9 * (size_t)n);
    tri_cop.resize(n);
    tri_mat.resize(n);
    mesh_tri_id.resize(n);
    for (int i = 0; i < n; i++) {
      const Triangle* t = byId[i];
      for (int a = 0; a < 3; a++) {
        tri_v[9 * i + a] = t->v1(a);
        tri_v[9 * i + 3 + a] = t->v2(a);
        tri_v[9 * i + 6 + a] = t->v3(a);
        tri_n[3 * i + a] = t->norm(a);
      }
      t->coordSys.to_row_major(&tri_cs[9 * (size_t)i]);
      tri_cop[i] = t->coplanarId;
      tri_mat[i] = t->materialId;
      mesh_tri_id[i] = mesh[i]->triangleId;
    }
    const int nn = (int)bvh.nodes.size();
    node_box.resize(6 * (size_t)nn);
    node_child.resize(2 * (size_t)nn);
    node_first.resize(nn);
    node_count.resize(nn);
    leaf_tri.clear();
    for (int k = 0; k < nn; k++) {
      const BVHNode<Triangle>& b = bvh.nodes[k];
      for (int a = 0; a < 3; a++) {
        node_box[6 * k + a] = b.bbox(0, a);
        node_box[6 * k + 3 + a] = b.bbox(1, a);
      }
      node_child[2 * k] = b.children[0];
      node_child[2 * k + 1] = b.children[1];
      node_first[k] = (int32_t)leaf_tri.size();
      node_count[k] = (int32_t)b.objects.size();
      for (const Triangle* t : b.objects) leaf_tri.push_back(t->triangleId);
    }
    desc.n_tri = n;
    desc.tri_v = tri_v.data();
    desc.tri_n = tri_n.data();
    desc.tri_cs = tri_cs.data();
    desc.tri_cop = tri_cop.data();
    desc.tri_mat = tri_mat.data();
    desc.mesh_tri_id = mesh_tri_id.data();
    desc.n_node = nn;
    desc.node_box = node_box.data();
    desc.node_child = node_child.data();
    desc.node_first = node_first.data();
    desc.node_count = node_count.data();
    desc.n_leaf_tri = (int32_t)leaf_tri.size();
    desc.leaf_tri = leaf_tri.data();
    return true;
  }
};
