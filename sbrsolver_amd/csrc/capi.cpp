// C-ABI entry points (include/sbr_capi.h) and the scene upload into the device layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <new>

#include "engine.h"
#include "prefilter.h"
#include "sbr_capi.h"
#include "sbrsolver/BVH.h"
#include "sbrsolver/Preprocessor.h"
#include "sbrsolver/SceneDesc.h"


namespace sbr {

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail(SBR_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static int put(DBuf<T>& b, const std::vector<T>& h, hipStream_t s) {
  b.ensure(std::max<size_t>(h.size(), 1));
  if (!b.p) return -1;
  if (!h.empty() && hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess)
    return -1;
  return 0;
}

// Flattens the BVH into depth-first order (children[0] subtree first), so that the DFS position of a
// triangle is the reference's visiting order, and builds the surface tables addressed by
// hitSurfaceID (through the mesh vector, and by id).
int Engine::upload(const sbr_scene_desc* d) {
  // A failed upload leaves no scene: device buffers are replaced one by one below, so after an error
  // they may hold parts of two scenes, and no trace may run on them.
  haveScene = false;
  haveResult = false;
  if (!d || d->n_tri <= 0 || d->n_node <= 0 || !d->tri_v || !d->tri_n || !d->tri_cs || !d->tri_cop || !d->tri_mat ||
      !d->mesh_tri_id || !d->node_box || !d->node_child || !d->node_first || !d->node_count || !d->leaf_tri)
    return fail(SBR_E_ARG, "scene: missing arrays");
  const int n = d->n_tri, nn = d->n_node;
  for (int i = 0; i < n; i++)
    if (d->mesh_tri_id[i] < 0 || d->mesh_tri_id[i] >= n) return fail(SBR_E_ARG, "scene: mesh_tri_id out of range");
  std::vector<DNode> nodesOut;
  std::vector<DTri> trisOut;
  std::vector<double> triN;
  nodesOut.reserve(nn);
  // Iterative pre-order walk; `slot` remembers where a parent expects its second child index.
  struct Item { int node; int parentSlot; int depth; };
  std::vector<Item> st{{0, -1, 1}};
  int height = 0;
  int visited = 0;
  while (!st.empty()) {
    Item it = st.back();
    st.pop_back();
    if (it.node < 0 || it.node >= nn || ++visited > nn) return fail(SBR_E_ARG, "scene: malformed BVH");
    height = std::max(height, it.depth);
    const int me = (int)nodesOut.size();
    if (it.parentSlot >= 0) nodesOut[it.parentSlot].b = me;
    DNode dn{};
    for (int a = 0; a < 3; a++) { dn.bmin[a] = d->node_box[6 * it.node + a]; dn.bmax[a] = d->node_box[6 * it.node + 3 + a]; }
    const int c0 = d->node_child[2 * it.node], c1 = d->node_child[2 * it.node + 1];
    if (c0 == -1) {
      const int first = d->node_first[it.node], cnt = d->node_count[it.node];
      if (cnt <= 0 || first < 0 || first + cnt > d->n_leaf_tri) return fail(SBR_E_ARG, "scene: bad leaf");
      dn.a = (int32_t)trisOut.size();
      dn.count = cnt;
      dn.b = -1;
      for (int k = 0; k < cnt; k++) {
        int id = d->leaf_tri[first + k];
        if (id < 0 || id >= n) return fail(SBR_E_ARG, "scene: leaf triangle id out of range");
        const double* v = d->tri_v + 9 * (size_t)id;
        DTri t{};
        for (int a = 0; a < 3; a++) {
          t.v1[a] = v[a];
          t.e1[a] = v[3 + a] - v[a];
          t.e2[a] = v[6 + a] - v[a];
        }
        t.id = id;
        t.cop = d->tri_cop[id];
        t.mat = d->tri_mat[id];
        trisOut.push_back(t);
        for (int a = 0; a < 3; a++) triN.push_back(d->tri_n[3 * (size_t)id + a]);
      }
      nodesOut.push_back(dn);
    } else {
      dn.a = me + 1;
      dn.count = 0;
      nodesOut.push_back(dn);
      st.push_back({c1, me, it.depth + 1});
      st.push_back({c0, -1, it.depth + 1});
    }
  }
  if (height > kMaxStack) return fail(SBR_E_LIMIT, "BVH taller than the traversal stack");
  // Single-precision pre-test boxes: exact FP64 bounds rounded outward and padded by delta.  With
  // E = max |coordinate| + 1 and ray origins with |coordinates| <= 4E, the single-precision
  // slab endpoints of a ray are off by at most ~1e-6 E in arc length (origin and direction rounding
  // 8e-7 E, fused slab products 6e-7 E), so delta = 1e-3 + 4e-6 E keeps every point of the exact box
  // strictly inside the computed interval.
  double ext = 0.0;
  for (int i = 0; i < 9 * n; i++) ext = std::max(ext, std::fabs(d->tri_v[i]));
  ext += 1.0;
  const double delta = prefilter_delta(ext);
  auto box32 = [&](const double* lo, const double* hi) {
    DBox32 b{};
    box_grown(lo, hi, delta, b.lo, b.hi);
    return b;
  };
  std::vector<DBox32> tb32(trisOut.size());
  for (size_t k = 0; k < trisOut.size(); k++) {
    double lo[3], hi[3];
    const double* tv = d->tri_v + 9 * (size_t)trisOut[k].id;
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min({tv[a], tv[3 + a], tv[6 + a]});
      hi[a] = std::max({tv[a], tv[3 + a], tv[6 + a]});
    }
    tb32[k] = box32(lo, hi);
  }
  // Leaf BVHs of the large leaves: median splits on the longest axis of the triangles' pre-test box
  // centres, leaves of at most kSubLeafSize triangles, node boxes the union of the pre-test boxes.
  std::vector<DSub> subs;
  std::vector<int32_t> subTri;
  int subInternal = 0;
  {
    std::vector<int32_t> ids;
    std::function<void(int, int, int, int)> fill = [&](int node, int lo, int hi, int level) {
      DSub dsub{};
      for (int a = 0; a < 3; a++) { dsub.lo[a] = INFINITY; dsub.hi[a] = -INFINITY; }
      float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int i = lo; i < hi; i++) {
        const DBox32& b = tb32[ids[i]];
        for (int a = 0; a < 3; a++) {
          dsub.lo[a] = std::min(dsub.lo[a], b.lo[a]);
          dsub.hi[a] = std::max(dsub.hi[a], b.hi[a]);
          const float c = 0.5f * (b.lo[a] + b.hi[a]);
          clo[a] = std::min(clo[a], c);
          chi[a] = std::max(chi[a], c);
        }
      }
      if (hi - lo <= kSubLeafSize) {
        dsub.first = (int32_t)subTri.size();
        dsub.count = hi - lo;
        subTri.insert(subTri.end(), ids.begin() + lo, ids.begin() + hi);
        subs[node] = dsub;
        return;
      }
      subInternal = std::max(subInternal, level + 1);
      int axis = 0;
      for (int a = 1; a < 3; a++)
        if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
      const int mid = (lo + hi) / 2;
      std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi, [&](int32_t x, int32_t y) {
        const float cx = tb32[x].lo[axis] + tb32[x].hi[axis], cy = tb32[y].lo[axis] + tb32[y].hi[axis];
        return cx != cy ? cx < cy : x < y;
      });
      const int c = (int)subs.size();
      subs.resize(c + 2);
      dsub.first = c;
      dsub.count = 0;
      subs[node] = dsub;
      fill(c, lo, mid, level + 1);
      fill(c + 1, mid, hi, level + 1);
    };
    for (auto& nd : nodesOut) {
      nd.sub = -1;
      if (nd.count > kSubLeaf) {
        ids.resize(nd.count);
        for (int i = 0; i < nd.count; i++) ids[i] = nd.a + i;
        nd.sub = (int32_t)subs.size();
        subs.resize(subs.size() + 1);
        fill(nd.sub, 0, nd.count, 0);
      }
    }
  }
  std::vector<DNodeF> nf(nodesOut.size());
  for (size_t k = 0; k < nodesOut.size(); k++) {
    const DNode& n = nodesOut[k];
    if (n.count > 0) {   // a leaf's own slot holds its triangle range: a, b, count, sub
      nf[k].ch[0] = DChildF{n.a, n.b};
      nf[k].ch[1] = DChildF{n.count, n.sub};
      continue;
    }
    const int32_t ch[2] = {n.a, n.b};
    for (int c = 0; c < 2; c++) {
      const DNode& cn = nodesOut[ch[c]];
      box_grown(cn.bmin, cn.bmax, 2.0 * delta, &nf[k].box[c][0], &nf[k].box[c][3]);
      const bool inner = box_shrunk(cn.bmin, cn.bmax, delta, &nf[k].box[c][6], &nf[k].box[c][9]);
      nf[k].ch[c].item = ch[c] | (cn.count > 0 ? kLeafBit : 0);
      nf[k].ch[c].inner = inner ? 1 : 0;
    }
  }
  // Wide traversal tree over the reference tree (DNode4), its leaves, and per reference-tree node the
  // data of the winner check (DVer).
  std::vector<DNode4> n4;
  std::vector<DLeaf> leaves;
  std::vector<DVer> ver(nodesOut.size());
  int wideStack = 0;
  {
    std::vector<int32_t> leafSlot(nodesOut.size(), -1);
    for (size_t k = 0; k < nodesOut.size(); k++) {
      ver[k].parent = -1;
      if (nodesOut[k].count > 0) {
        if (nodesOut[k].sub < 0) {   // a leaf with a leaf BVH becomes a wide subtree (buildSub)
          leafSlot[k] = (int32_t)leaves.size();
          leaves.push_back(DLeaf{nodesOut[k].a, nodesOut[k].count, 0, 0});
        }
        for (int i = 0; i < nodesOut[k].count; i++) trisOut[nodesOut[k].a + i].leaf = (int32_t)k;
      }
    }
    // The check of a winner walks up from its leaf and stops at the first node whose shrunk box the
    // line crosses; that this implies the tests of all ancestors needs every box to contain its
    // children's boxes, which a description normally has.  Without it every node on the way is tested.
    bool nested = true;
    for (size_t k = 0; k < nodesOut.size(); k++) {
      const DNode& nd = nodesOut[k];
      if (nd.count > 0) continue;
      for (int c : {nd.a, nd.b}) {
        ver[c].parent = (int32_t)k;
        for (int a = 0; a < 3; a++)
          if (!(nodesOut[c].bmin[a] >= nd.bmin[a] && nodesOut[c].bmax[a] <= nd.bmax[a])) nested = false;
      }
    }
    for (size_t k = 0; k < nodesOut.size(); k++)
      ver[k].shrunk = (box_shrunk(nodesOut[k].bmin, nodesOut[k].bmax, delta, ver[k].lo, ver[k].hi) && nested) ? 1 : 0;
    auto area = [&](int k) {
      const DNode& x = nodesOut[k];
      const double e0 = x.bmax[0] - x.bmin[0], e1 = x.bmax[1] - x.bmin[1], e2 = x.bmax[2] - x.bmin[2];
      return e0 * e1 + e1 * e2 + e2 * e0;
    };
    auto subArea = [&](int q) {
      const DSub& x = subs[q];
      const double e0 = (double)x.hi[0] - x.lo[0], e1 = (double)x.hi[1] - x.lo[1], e2 = (double)x.hi[2] - x.lo[2];
      return e0 * e1 + e1 * e2 + e2 * e0;
    };
    // buildSub(q, deferred): wide subtree over the inner leaf-BVH node q, collapsed four-wide as
    // below; its boxes are the leaf BVH's (unions of the triangles' padded pre-test boxes, which a
    // hit triangle's computed interval enters, and so does any box holding them), its leaves list
    // their triangles in subTri.
    std::function<int(int, int)> buildSub = [&](int q, int deferred) {
      int c[kWide] = {subs[q].first, subs[q].first + 1};
      int m = 2;
      while (m < kWide) {
        int best = -1;
        double ba = -1.0;
        for (int i = 0; i < m; i++)
          if (subs[c[i]].count == 0 && subArea(c[i]) > ba) {
            best = i;
            ba = subArea(c[i]);
          }
        if (best < 0) break;
        const int x = c[best];
        for (int j = m; j > best + 1; j--) c[j] = c[j - 1];
        c[best] = subs[x].first;
        c[best + 1] = subs[x].first + 1;
        m++;
      }
      const int me = (int)n4.size();
      n4.emplace_back();
      wideStack = std::max(wideStack, deferred + m - 1);
      for (int i = 0; i < kWide; i++) {
        float lo[3], hi[3];
        int32_t item = -1;
        if (i < m) {
          const DSub sq = subs[c[i]];
          for (int a = 0; a < 3; a++) {
            lo[a] = sq.lo[a];
            hi[a] = sq.hi[a];
          }
          if (sq.count > 0) {
            item = kLeafBit | (int32_t)leaves.size();
            leaves.push_back(DLeaf{sq.first, sq.count, 1, 0});
          } else {
            item = buildSub(c[i], deferred + m - 1);
          }
        } else {
          for (int a = 0; a < 3; a++) lo[a] = hi[a] = std::nanf("");
        }
        for (int a = 0; a < 3; a++) {
          n4[me].lo[a][i] = lo[a];
          n4[me].hi[a][i] = hi[a];
        }
        n4[me].item[i] = item;
      }
      return me;
    };
    // build(k, deferred): wide node over reference node k (a leaf only at the root); deferred = the
    // siblings the traversal may hold on its stack above this node.
    std::function<int(int, int)> build = [&](int k, int deferred) {
      int c[kWide] = {k};
      int m = 1;
      if (nodesOut[k].count == 0) {
        c[0] = nodesOut[k].a;
        c[1] = nodesOut[k].b;
        m = 2;
      }
      while (m < kWide) {
        int best = -1;
        double ba = -1.0;
        for (int i = 0; i < m; i++)
          if (nodesOut[c[i]].count == 0 && area(c[i]) > ba) {
            best = i;
            ba = area(c[i]);
          }
        if (best < 0) break;
        const int x = c[best];
        for (int j = m; j > best + 1; j--) c[j] = c[j - 1];
        c[best] = nodesOut[x].a;
        c[best + 1] = nodesOut[x].b;
        m++;
      }
      const int me = (int)n4.size();
      n4.emplace_back();
      wideStack = std::max(wideStack, deferred + m - 1);
      for (int i = 0; i < kWide; i++) {
        float lo[3], hi[3];
        int32_t item = -1;
        if (i < m) {
          const DNode& cn = nodesOut[c[i]];
          box_grown(cn.bmin, cn.bmax, 2.0 * delta, lo, hi);
          if (cn.count == 0)
            item = build(c[i], deferred + m - 1);
          else
            item = cn.sub >= 0 ? buildSub(cn.sub, deferred + m - 1) : (kLeafBit | leafSlot[c[i]]);
        } else {
          for (int a = 0; a < 3; a++) lo[a] = hi[a] = std::nanf("");
        }
        for (int a = 0; a < 3; a++) {
          n4[me].lo[a][i] = lo[a];
          n4[me].hi[a][i] = hi[a];
        }
        n4[me].item[i] = item;
      }
      return me;
    };
    build(0, 0);
  }
  auto surf_of = [&](int id) {
    DSurf s{};
    const double* v = d->tri_v + 9 * (size_t)id;
    for (int a = 0; a < 3; a++) {
      s.v1[a] = v[a];
      s.e1[a] = v[3 + a] - v[a];
      s.e2[a] = v[6 + a] - v[a];
      s.n[a] = d->tri_n[3 * (size_t)id + a];
    }
    for (int k = 0; k < 9; k++) s.cs[k] = d->tri_cs[9 * (size_t)id + k];
    return s;
  };
  std::vector<DSurf> sq(n), si(n);
  for (int t = 0; t < n; t++) {
    sq[t] = surf_of(d->mesh_tri_id[t]);
    si[t] = surf_of(t);
  }
  if (device >= 0) {
    HIPCHK(hipSetDevice(device));
    if (put(dNodes, nodesOut, stream) || put(dTris, trisOut, stream) || put(dTriN, triN, stream) ||
        put(dSurfQ1, sq, stream) || put(dSurfId, si, stream) ||
        put(dTriBox32, tb32, stream) || put(dSub, subs, stream) || put(dSubTri, subTri, stream) ||
        put(dNodesF, nf, stream) || put(dNodes4, n4, stream) || put(dLeaves, leaves, stream) || put(dVer, ver, stream))
      return fail(SBR_E_NOMEM, "scene upload");
    HIPCHK(hipStreamSynchronize(stream));
  }
  hNodes.swap(nodesOut);
  hTris.swap(trisOut);
  hTriN.swap(triN);
  hSurfQ1.swap(sq);
  hSurfId.swap(si);
  hCop.assign(d->tri_cop, d->tri_cop + n);
  hMat.assign(d->tri_mat, d->tri_mat + n);
  hMeshTriId.assign(d->mesh_tri_id, d->mesh_tri_id + n);
  hVerts.assign(d->tri_v, d->tri_v + 9 * (size_t)n);
  hNorms.assign(d->tri_n, d->tri_n + 3 * (size_t)n);
  ntri = n;
  nnode = (int)hNodes.size();   // nodes reachable from the root (a description may hold others)
  treeHeight = height;
  hTriBox32.swap(tb32);
  hNodesF.swap(nf);
  hSub.swap(subs);
  hSubTri.swap(subTri);
  subHeight = subInternal;
  sceneExtent = (float)ext;
  boxDelta = (float)delta;
  sceneBytes = hNodesF.size() * sizeof(DNodeF);   // LDS-staged part (reference-tree traversal)
  hNodes4.swap(n4);
  hLeaves.swap(leaves);
  hVer.swap(ver);
  this->wideStack = wideStack;
  wideBytes = hNodes4.size() * sizeof(DNode4) + hLeaves.size() * sizeof(DLeaf);   // LDS-staged part (wide)
  ncop = 0;
  for (int t = 0; t < n; t++) ncop = std::max(ncop, hCop[t] + 1);
  haveScene = true;
  haveResult = false;
  return SBR_OK;
}

int Engine::clone_scene(const Engine& s) {
  haveScene = false;
  haveResult = false;
  if (!s.haveScene) return fail(SBR_E_STATE, "source has no scene");
  if (s.mats.empty()) return fail(SBR_E_STATE, "source has no materials");
  hNodes = s.hNodes;
  hTris = s.hTris;
  hTriN = s.hTriN;
  hSurfQ1 = s.hSurfQ1;
  hSurfId = s.hSurfId;
  hCop = s.hCop;
  hMat = s.hMat;
  hMeshTriId = s.hMeshTriId;
  hVerts = s.hVerts;
  hNorms = s.hNorms;
  hTriBox32 = s.hTriBox32;
  hNodesF = s.hNodesF;
  hSub = s.hSub;
  hSubTri = s.hSubTri;
  subHeight = s.subHeight;
  ntri = s.ntri;
  nnode = s.nnode;
  ncop = s.ncop;
  treeHeight = s.treeHeight;
  sceneBytes = s.sceneBytes;
  hNodes4 = s.hNodes4;
  hLeaves = s.hLeaves;
  hVer = s.hVer;
  wideStack = s.wideStack;
  wideBytes = s.wideBytes;
  sceneExtent = s.sceneExtent;
  boxDelta = s.boxDelta;
  if (device >= 0) {
    HIPCHK(hipSetDevice(device));
    if (put(dNodes, hNodes, stream) || put(dTris, hTris, stream) || put(dTriN, hTriN, stream) ||
        put(dSurfQ1, hSurfQ1, stream) || put(dSurfId, hSurfId, stream) || put(dTriBox32, hTriBox32, stream) ||
        put(dSub, hSub, stream) || put(dSubTri, hSubTri, stream) || put(dNodesF, hNodesF, stream) ||
        put(dNodes4, hNodes4, stream) || put(dLeaves, hLeaves, stream) || put(dVer, hVer, stream))
      return fail(SBR_E_NOMEM, "scene upload");
    HIPCHK(hipStreamSynchronize(stream));
  }
  haveScene = true;
  haveResult = false;
  if (s.patNT > 0) {
    const int rc = set_pattern(s.hPattern.data(), s.patNT, s.patNP);
    if (rc != SBR_OK) return rc;
  }
  return set_materials(reinterpret_cast<const sbr_material*>(s.mats.data()), (int)(s.mats.size() / 6));
}

}  // namespace sbr

namespace {

int from_triangles(sbr_ctx* ctx, std::vector<Triangle*>& mesh) {
  BVH<Triangle> bvh;
  bvh.ConstructBVH(mesh);   // sorts mesh in place, as SBRSolver_rev2.cpp:67-68 does
  SceneArrays sa;
  int rc;
  if (!sa.build(mesh, bvh)) rc = ctx->eng.fail(SBR_E_INTERNAL, "scene flattening failed");
  else rc = ctx->eng.upload(&sa.desc);
  for (Triangle* t : mesh) delete t;
  mesh.clear();
  return rc;
}

template <typename F>
int guarded(sbr_ctx* ctx, F&& f) {
  if (!ctx) return SBR_E_ARG;
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return ctx->eng.fail(SBR_E_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return ctx->eng.fail(SBR_E_INTERNAL, e.what());
  } catch (...) {
    return ctx->eng.fail(SBR_E_INTERNAL, "unknown exception");
  }
}

}  // namespace

extern "C" {

int sbr_version(void) { return 1; }

int sbr_create(int device, sbr_ctx** out) {
  if (!out) return SBR_E_ARG;
  *out = nullptr;
  sbr_ctx* c = new (std::nothrow) sbr_ctx();
  if (!c) return SBR_E_NOMEM;
  int rc = c->eng.init(device);
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return SBR_OK;
}

int sbr_destroy(sbr_ctx* ctx) {
  if (!ctx) return SBR_E_ARG;
  if (ctx->eng.device >= 0) (void)hipSetDevice(ctx->eng.device);
  delete ctx;
  return SBR_OK;
}

const char* sbr_last_error(const sbr_ctx* ctx) { return ctx ? ctx->eng.err.c_str() : "null context"; }

int sbr_scene_from_stl(sbr_ctx* ctx, const char* path) {
  return guarded(ctx, [&] {
    if (!path) return ctx->eng.fail(SBR_E_ARG, "null path");
    std::vector<Triangle*> mesh;
    if (!Preprocessor::StlToGeometry(path, mesh)) {
      for (Triangle* t : mesh) delete t;
      return ctx->eng.fail(SBR_E_IO, std::string("cannot read STL file ") + path);
    }
    return from_triangles(ctx, mesh);
  });
}

int sbr_scene_from_soup(sbr_ctx* ctx, const float* coords9, const float* normals3, const int32_t* mats, int32_t n) {
  return guarded(ctx, [&] {
    if (!coords9 || !normals3 || n <= 0) return ctx->eng.fail(SBR_E_ARG, "soup: empty input");
    std::vector<Triangle*> mesh;
    if (!Preprocessor::SoupToGeometry(coords9, normals3, mats, n, mesh)) {
      for (Triangle* t : mesh) delete t;
      return ctx->eng.fail(SBR_E_ARG, "soup: no valid triangle");
    }
    return from_triangles(ctx, mesh);
  });
}

int sbr_scene_upload(sbr_ctx* ctx, const sbr_scene_desc* desc) {
  return guarded(ctx, [&] { return ctx->eng.upload(desc); });
}

int sbr_scene_from_soup_device(sbr_ctx* ctx, const float* coords9, const float* normals3, const int32_t* material_ids,
                               int32_t n_facets, double* ms_out) {
  return guarded(ctx, [&] { return ctx->eng.ingest_soup(coords9, normals3, material_ids, n_facets, ms_out); });
}

int sbr_scene_info(sbr_ctx* ctx, int32_t* n_tri, int32_t* n_node, int32_t* n_coplanar) {
  return guarded(ctx, [&] {
    if (!ctx->eng.haveScene) return ctx->eng.fail(SBR_E_STATE, "no scene");
    if (n_tri) *n_tri = ctx->eng.ntri;
    if (n_node) *n_node = ctx->eng.nnode;
    if (n_coplanar) *n_coplanar = ctx->eng.ncop;
    return SBR_OK;
  });
}

int sbr_scene_triangles(sbr_ctx* ctx, int32_t* cop, int32_t* mat, double* verts9, double* norm3, int32_t* mesh) {
  return guarded(ctx, [&] {
    sbr::Engine& e = ctx->eng;
    if (!e.haveScene) return e.fail(SBR_E_STATE, "no scene");
    const size_t n = e.ntri;
    if (cop) memcpy(cop, e.hCop.data(), n * sizeof(int32_t));
    if (mat) memcpy(mat, e.hMat.data(), n * sizeof(int32_t));
    if (verts9) memcpy(verts9, e.hVerts.data(), 9 * n * sizeof(double));
    if (norm3) memcpy(norm3, e.hNorms.data(), 3 * n * sizeof(double));
    if (mesh) memcpy(mesh, e.hMeshTriId.data(), n * sizeof(int32_t));
    return SBR_OK;
  });
}

int sbr_scene_bvh_dfs(sbr_ctx* ctx, double* box, int32_t* a, int32_t* b, int32_t* leaf, int32_t* triId) {
  return guarded(ctx, [&] {
    sbr::Engine& e = ctx->eng;
    if (!e.haveScene) return e.fail(SBR_E_STATE, "no scene");
    for (size_t k = 0; k < e.hNodes.size(); k++) {
      const sbr::DNode& d = e.hNodes[k];
      for (int q = 0; q < 3; q++) { box[6 * k + q] = d.bmin[q]; box[6 * k + 3 + q] = d.bmax[q]; }
      a[k] = d.a;
      b[k] = d.count > 0 ? d.count : d.b;
      leaf[k] = d.count > 0 ? 1 : 0;
    }
    for (size_t k = 0; k < e.hTris.size(); k++) triId[k] = e.hTris[k].id;
    return SBR_OK;
  });
}

int sbr_materials(sbr_ctx* ctx, const sbr_material* mats, int32_t count) {
  return guarded(ctx, [&] { return ctx->eng.set_materials(mats, count); });
}

int sbr_set_tx_pattern(sbr_ctx* ctx, const float* gain_info, int32_t n_theta, int32_t n_phi) {
  return guarded(ctx, [&] { return ctx->eng.set_pattern(gain_info, n_theta, n_phi); });
}

int sbr_set_tuning(sbr_ctx* ctx, const sbr_tuning* t) {
  return guarded(ctx, [&] {
    if (!t || t->lds_scene < -1 || t->lds_scene > 1 || t->blocks_per_cu < 0 || t->blocks_per_cu > 64 ||
        t->window_nodes < 0 || t->window_nodes > (int64_t(1) << 30) || t->stack_lds < 0 || t->stack_lds > 64 ||
        t->oversub < 0 || t->oversub > 64 || t->table_entries < 0 || t->table_entries > (int64_t(1) << 32))
      return ctx->eng.fail(SBR_E_ARG, "tuning out of range");
    ctx->eng.tune.ldsScene = t->lds_scene;
    ctx->eng.tune.blocksPerCU = t->blocks_per_cu;
    ctx->eng.tune.windowNodes = t->window_nodes;
    ctx->eng.tune.stackLds = t->stack_lds;
    ctx->eng.tune.oversub = t->oversub;
    ctx->eng.tune.tableEntries = t->table_entries;
    return SBR_OK;
  });
}

int sbr_autotune(sbr_ctx* ctx, const sbr_trace_params* params, const double* rx, int32_t n_rx, sbr_tuning* chosen) {
  return guarded(ctx, [&] {
    sbr::Engine& e = ctx->eng;
    if (!params || !chosen) return e.fail(SBR_E_ARG, "autotune: null argument");
    if (e.device < 0) return e.fail(SBR_E_STATE, "host-only context has no device");
    sbr_trace_params p = *params;
    const int64_t launches = 10 * (int64_t(1) << (2 * std::max(0, p.tess))) + 2;
    p.launchStride = (int32_t)std::max<int64_t>(1, launches >> 18);
    p.launchOffset = 0;
    p.countTraversal = 0;
    const sbr::Tuning saved = e.tune;
    sbr::Tuning best = saved;
    double bestMs = 1e300;
    // kernel time of the sample under tuning t (first run warms buffers and caches); <0: error
    auto timed = [&](const sbr::Tuning& t, double* ms) {
      e.tune = t;
      *ms = 0;
      for (int rep = 0; rep < 3; rep++) {
        int rc = e.trace(&p, 1, rx, n_rx);
        if (rc) return rc;
        if (rep > 0) *ms += e.stats.ms_k_trace;
      }
      return (int)SBR_OK;
    };
    // 1. scene staging x blocks per CU; the first trial is the automatic choice, which also
    //    reports the occupancy maximum
    std::vector<sbr::Tuning> trials{{-1, 0, saved.windowNodes, 0, 0, saved.tableEntries}};
    for (int lds = 0; lds <= 1; lds++)
      for (int bpc = 1; bpc <= 8; bpc++) trials.push_back({lds, bpc, saved.windowNodes, 0, 0, saved.tableEntries});
    int maxBpc = 8;
    for (size_t k = 0; k < trials.size(); k++) {
      if (k > 0 && trials[k].blocksPerCU > maxBpc) continue;
      double ms = 0;
      int rc = timed(trials[k], &ms);
      if (rc) {
        e.tune = saved;
        return rc;
      }
      if (k == 0) {   // the automatic choice, recorded as the configuration it resolved to
        maxBpc = std::max(1, (int)e.stats.trace_blocks_per_cu);
        trials[0].ldsScene = (int32_t)e.stats.trace_lds;
        trials[0].blocksPerCU = maxBpc;
      }
      if (k > 0 && e.stats.trace_lds != trials[k].ldsScene) continue;   // staging did not fit
      if (ms < bestMs) {
        bestMs = ms;
        best = trials[k];
      }
    }
    // 2. for that configuration: LDS stack entries (the rest in device memory), then the grid's
    //    blocks per resident block
    for (int sl : {2, 4, 8, 16}) {
      sbr::Tuning t = best;
      t.stackLds = sl;
      double ms = 0;
      int rc = timed(t, &ms);
      if (rc) {
        e.tune = saved;
        return rc;
      }
      if (e.stats.trace_lds != t.ldsScene) continue;
      if (ms < bestMs) {
        bestMs = ms;
        best = t;
      }
    }
    for (int os : {1, 2, 4, 16}) {
      sbr::Tuning t = best;
      t.oversub = os;
      double ms = 0;
      int rc = timed(t, &ms);
      if (rc) {
        e.tune = saved;
        return rc;
      }
      if (ms < bestMs) {
        bestMs = ms;
        best = t;
      }
    }
    e.tune = best;
    chosen->lds_scene = best.ldsScene;
    chosen->blocks_per_cu = best.blocksPerCU;
    chosen->window_nodes = best.windowNodes;
    chosen->stack_lds = best.stackLds;
    chosen->oversub = best.oversub;
    return SBR_OK;
  });
}

int sbr_get_tuning(sbr_ctx* ctx, sbr_tuning* t) {
  return guarded(ctx, [&] {
    if (!t) return ctx->eng.fail(SBR_E_ARG, "null tuning");
    t->lds_scene = ctx->eng.tune.ldsScene;
    t->blocks_per_cu = ctx->eng.tune.blocksPerCU;
    t->window_nodes = ctx->eng.tune.windowNodes;
    t->stack_lds = ctx->eng.tune.stackLds;
    t->oversub = ctx->eng.tune.oversub;
    t->table_entries = ctx->eng.tune.tableEntries;
    return SBR_OK;
  });
}

int sbr_generate_rays(sbr_ctx* ctx, int32_t tess, double* out, int64_t cap, int64_t* n_out) {
  return guarded(ctx, [&] {
    sbr::Engine& e = ctx->eng;
    if (e.device < 0) return e.fail(SBR_E_STATE, "host-only context has no device");
    if (hipSetDevice(e.device) != hipSuccess) return e.fail(SBR_E_HIP, "hipSetDevice");
    int64_t n = 0;
    const double* rays = nullptr;
    int rc = e.rays_for(tess, &rays, &n);
    if (rc) return rc;
    if (n_out) *n_out = n;
    if (out && cap > 0) {
      int64_t k = std::min(cap, n);
      if (hipMemcpy(out, rays, 3 * sizeof(double) * k, hipMemcpyDeviceToHost) != hipSuccess)
        return e.fail(SBR_E_HIP, "copy rays");
    }
    return SBR_OK;
  });
}

int sbr_closest_hit(sbr_ctx* ctx, const double* o, const double* d, int32_t n, int32_t exact, int32_t* tri, double* pt,
                    double* dist) {
  return guarded(ctx, [&] {
    if (n < 0 || (n > 0 && (!o || !d || !tri || !pt || !dist)) || exact < 0 || exact > 2)
      return ctx->eng.fail(SBR_E_ARG, "closest_hit: bad args");
    return ctx->eng.closest_hit(o, d, n, exact, tri, pt, dist);
  });
}

int sbr_closest_hit_noalloc(sbr_ctx* ctx, const double* o, const double* d, int32_t n, int32_t exact, int32_t* tri,
                            double* pt, double* dist) {
  return guarded(ctx, [&] {
    if (n < 0 || (n > 0 && (!o || !d || !tri || !pt || !dist)) || exact < 0 || exact > 2)
      return ctx->eng.fail(SBR_E_ARG, "closest_hit_noalloc: bad args");
    return ctx->eng.closest_hit_noalloc(o, d, n, exact, tri, pt, dist);
  });
}

int sbr_trace(sbr_ctx* ctx, const sbr_trace_params* p, const double* rx, int32_t n_rx) {
  return guarded(ctx, [&] { return ctx->eng.trace(p, 1, rx, n_rx); });
}

int sbr_path_fields(sbr_ctx* ctx, const sbr_field_params* p, const sbr_material* mats, int32_t n_mats, int64_t n_paths,
                    const int64_t* seg_off, const int32_t* seg_count, const double* seg_src, const double* seg_tgt,
                    const int32_t* seg_surf, const int32_t* seg_type, const int32_t* seg_mat, int32_t n_surf,
                    const double* surf_cs, double* fields_out) {
  return guarded(ctx, [&] {
    return ctx->eng.path_fields_given(p, mats, n_mats, n_paths, seg_off, seg_count, seg_src, seg_tgt, seg_surf,
                                      seg_type, seg_mat, n_surf, surf_cs, fields_out);
  });
}

int sbr_fresnel_sweep(sbr_ctx* ctx, int32_t n, double frequency, const double* eps_i, const double* eps_t, double width,
                      int32_t inf_wall, double* theta_out, double* coeff_out) {
  return guarded(ctx, [&] {
    return ctx->eng.fresnel_sweep(n, frequency, eps_i, eps_t, width, inf_wall, theta_out, coeff_out);
  });
}

int sbr_trace_batch(sbr_ctx* ctx, const sbr_trace_params* p, int32_t n_tx, const double* rx, int32_t n_rx) {
  return guarded(ctx, [&] { return ctx->eng.trace(p, n_tx, rx, n_rx); });
}

int sbr_result_stats(sbr_ctx* ctx, sbr_stats* out) {
  return guarded(ctx, [&] {
    if (!ctx->eng.haveResult) return ctx->eng.fail(SBR_E_STATE, "no trace result");
    if (!out) return ctx->eng.fail(SBR_E_ARG, "null output");
    *out = ctx->eng.stats;
    return SBR_OK;
  });
}

int sbr_result_fields(sbr_ctx* ctx, double* out, int32_t n) {
  return guarded(ctx, [&] {
    sbr::Engine& e = ctx->eng;
    if (!e.haveResult) return e.fail(SBR_E_STATE, "no trace result");
    if ((int64_t)n != (int64_t)e.nrx * e.ntx || (!out && n > 0))
      return e.fail(SBR_E_ARG, "fields: count must be transmitters x receivers");
    if (n > 0) memcpy(out, e.field_host(), 6 * sizeof(double) * (size_t)n);
    return SBR_OK;
  });
}

int sbr_result_path_counts(sbr_ctx* ctx, int64_t* n_paths, int64_t* n_segs) {
  return guarded(ctx, [&] {
    int64_t a = 0, b = 0;
    int rc = ctx->eng.paths(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &a,
                            &b, true);
    if (n_paths) *n_paths = a;
    if (n_segs) *n_segs = b;
    return rc;
  });
}

int sbr_result_paths(sbr_ctx* ctx, int32_t* prx, int32_t* plaunch, int32_t* pnseg, int64_t* poff, int32_t* stri,
                     int32_t* scop, int32_t* stype, double* ssrc, double* stgt) {
  return guarded(ctx, [&] {
    if (!prx || !plaunch || !pnseg || !poff || !stri || !scop || !stype || !ssrc || !stgt)
      return ctx->eng.fail(SBR_E_ARG, "paths: null output");
    int64_t a = 0, b = 0;
    return ctx->eng.paths(nullptr, prx, plaunch, pnseg, poff, stri, scop, stype, ssrc, stgt, &a, &b, false);
  });
}

int sbr_result_path_tx(sbr_ctx* ctx, int32_t* out, int64_t n_paths) {
  return guarded(ctx, [&] { return ctx->eng.path_tx(out, n_paths); });
}

int sbr_result_path_fields(sbr_ctx* ctx, double* out, int64_t n_paths) {
  return guarded(ctx, [&] { return ctx->eng.path_fields(out, n_paths); });
}

int sbr_result_tree_count(sbr_ctx* ctx, int64_t* n) {
  return guarded(ctx, [&] {
    if (!ctx->eng.haveResult) return ctx->eng.fail(SBR_E_STATE, "no trace result");
    if (ctx->eng.multiWindow)
      return ctx->eng.fail(SBR_E_STATE, "interaction trees are kept for single-window traces only");
    if (n) *n = ctx->eng.nodes;
    return SBR_OK;
  });
}

int sbr_result_trees(sbr_ctx* ctx, int32_t* launch, int32_t* depth, int64_t* code, int32_t* tri, int32_t* cop,
                     int32_t* type, double* src, double* tgt, double* pl) {
  return guarded(ctx, [&] {
    if (!launch || !depth || !code || !tri || !cop || !type || !src || !tgt || !pl)
      return ctx->eng.fail(SBR_E_ARG, "trees: null output");
    return ctx->eng.trees(launch, depth, code, tri, cop, type, src, tgt, pl);
  });
}

}  // extern "C"
