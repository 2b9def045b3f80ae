// Scene ingestion on the device, for scenes whose host build is slow (the reference's full-sweep SAH
// over a million triangles takes seconds): the triangle set of Preprocessor::StlToGeometry
// (Preprocessor.cpp:69-114 with the stl_reader vertex dedup, StlReader.h:448-508, and the coplanar
// grouping, Preprocessor.cpp:116-246) and a linear BVH, built with sorts, scans and a union-find on
// the device.  The result is handed to Engine::upload as a scene description, as a host-built scene is.
//
// Triangles: the facet corners are sorted by (x, y, z) and numbered in that order (one vertex per
// distinct coordinate triple; -0 is taken as +0), facets with two equal corners are dropped, and
// triangle t takes its vertices from the t-th kept facet and its normal and material from facet slot
// t, as the host preprocessor does.  Coplanar ids: triangles sharing an edge with equal or opposite
// normals are joined (lock-free union-find, every root linked under the smaller one), and components
// are numbered in the order of their smallest triangle, which is the host's breadth-first numbering.
// Both are identical to the host build (tests/test_gpu_ingest.py).
//
// BVH: triangle centres (the centre of the FP64 bounding box, as Triangle::findCenter) quantised to
// 21 bits per axis over the centres' bounds and interleaved into 63-bit Morton codes; triangles sorted
// by (code, triangle id); Karras' radix tree over that order (ties of equal codes broken by position);
// subtrees of at most 4 triangles collapsed into leaves; node boxes the exact FP64 min / max of their
// triangles' vertices, built bottom up.  The mesh vector is taken in triangle-id order (nothing sorts
// it), so the reference's mesh[hitSurfaceID] indexing is the identity here.  A trace over this tree
// is the reference's algorithm over another tree: the closest hit is the minimum over (distance,
// DFS position) of the triangles whose leaves the FP64 slab tests admit, as for the SAH tree; the
// oracle runs the same traversal over the exported tree (sbr_scene_bvh_dfs) for parity.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cmath>
#include <cstring>

#include "engine.h"

namespace sbr {

#define IHIPCHK(x)                                                                    \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail(SBR_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define IENSURE(buf, cnt)                                                                        \
  do {                                                                                           \
    (buf).ensure((size_t)(cnt));                                                                 \
    if (!(buf).p) return fail(SBR_E_NOMEM, std::string("device allocation failed: ") + #buf);  \
  } while (0)

static inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

constexpr int kIngestLeaf = 4;   // triangles per collapsed leaf at most

// Unsigned key whose order is the float order (-0 taken as +0; no NaN in scene coordinates).
__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t u = __float_as_uint(f == 0.0f ? 0.0f : f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_ing_iota(int32_t* __restrict__ a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (int32_t)i;
}

// key[i] = coordinate `axis` of corner perm[i]
__global__ void k_ing_corner_key(const float* __restrict__ c, const int32_t* __restrict__ perm, int64_t K, int axis,
                                 uint32_t* __restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < K) key[i] = fkey(c[3 * (int64_t)perm[i] + axis]);
}

// flag[i] = 1 where sorted corner i starts a new coordinate triple
__global__ void k_ing_unique(const float* __restrict__ c, const int32_t* __restrict__ perm, int64_t K,
                             int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  int f = 1;
  if (i > 0) {
    const float* p = c + 3 * (int64_t)perm[i];
    const float* q = c + 3 * (int64_t)perm[i - 1];
    f = (fkey(p[0]) != fkey(q[0]) || fkey(p[1]) != fkey(q[1]) || fkey(p[2]) != fkey(q[2])) ? 1 : 0;
  }
  flag[i] = f;
}

// vertex index of every corner; the vertex table in sorted order
__global__ void k_ing_vertices(const float* __restrict__ c, const int32_t* __restrict__ perm, const int32_t* __restrict__ flag,
                               const int32_t* __restrict__ scan, int64_t K, int32_t* __restrict__ vid,
                               double* __restrict__ verts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  const int32_t v = scan[i] + flag[i] - 1;
  vid[perm[i]] = v;
  if (flag[i]) {
    const float* p = c + 3 * (int64_t)perm[i];
    for (int a = 0; a < 3; a++) verts[3 * (int64_t)v + a] = (double)(p[a] == 0.0f ? 0.0f : p[a]);
  }
}

__global__ void k_ing_keep(const int32_t* __restrict__ vid, int64_t nf, int32_t* __restrict__ keep) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int32_t a = vid[3 * f], b = vid[3 * f + 1], c = vid[3 * f + 2];
  keep[f] = (a != b && a != c && b != c) ? 1 : 0;
}

// Triangle t (the t-th kept facet): vertices, the normal and material of facet slot t, frame, box,
// centre; its three edges (vertex pairs, smaller first) in the host's order (01, 12, 02).
__global__ void k_ing_triangles(const int32_t* __restrict__ vid, const int32_t* __restrict__ keep,
                                const int32_t* __restrict__ keepScan, int64_t nf, const double* __restrict__ verts,
                                const float* __restrict__ normals, const int32_t* __restrict__ mats,
                                double* __restrict__ triV, double* __restrict__ triN, double* __restrict__ triCs,
                                int32_t* __restrict__ triMat, double* __restrict__ cen,
                                unsigned long long* __restrict__ edgeKey, int32_t* __restrict__ edgeVal) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf || !keep[f]) return;
  const int64_t t = keepScan[f];
  const int32_t iv[3] = {vid[3 * f], vid[3 * f + 1], vid[3 * f + 2]};
  V3 p[3];
  for (int k = 0; k < 3; k++) {
    p[k] = v3(verts[3 * (int64_t)iv[k]], verts[3 * (int64_t)iv[k] + 1], verts[3 * (int64_t)iv[k] + 2]);
    triV[9 * t + 3 * k] = p[k].x;
    triV[9 * t + 3 * k + 1] = p[k].y;
    triV[9 * t + 3 * k + 2] = p[k].z;
  }
  const V3 n = v3((double)normals[3 * t], (double)normals[3 * t + 1], (double)normals[3 * t + 2]);
  triN[3 * t] = n.x;
  triN[3 * t + 1] = n.y;
  triN[3 * t + 2] = n.z;
  triMat[t] = mats ? mats[t] : 0;
  // Triangle::findCoordSys: rows (y x n, normalized(v2 - v1), n)
  const V3 yw = vnormalized(vsub(p[1], p[0]));
  const V3 xw = vcross(yw, n);
  const double cs[9] = {xw.x, xw.y, xw.z, yw.x, yw.y, yw.z, n.x, n.y, n.z};
  for (int k = 0; k < 9; k++) triCs[9 * t + k] = cs[k];
  for (int a = 0; a < 3; a++) {
    const double lo = fmin(fmin(vget(p[0], a), vget(p[1], a)), vget(p[2], a));
    const double hi = fmax(fmax(vget(p[0], a), vget(p[1], a)), vget(p[2], a));
    cen[3 * t + a] = (lo + hi) / 2.0;
  }
  const int pairs[3][2] = {{0, 1}, {1, 2}, {0, 2}};
  for (int e = 0; e < 3; e++) {
    const uint32_t a = (uint32_t)iv[pairs[e][0]], b = (uint32_t)iv[pairs[e][1]];
    edgeKey[3 * t + e] = ((unsigned long long)min(a, b) << 32) | (unsigned long long)max(a, b);
    edgeVal[3 * t + e] = (int32_t)t;
  }
}

__device__ __forceinline__ int uf_find(int32_t* par, int32_t x) {
  while (true) {
    const int32_t p = __atomic_load_n(&par[x], __ATOMIC_RELAXED);
    if (p == x) return x;
    const int32_t g = __atomic_load_n(&par[p], __ATOMIC_RELAXED);
    if (g == p) return p;
    __atomic_store_n(&par[x], g, __ATOMIC_RELAXED);   // path halving: g is an ancestor of x
    x = g;
  }
}

__device__ __forceinline__ void uf_union(int32_t* par, int32_t a, int32_t b) {
  while (true) {
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b) return;
    if (a < b) {
      const int32_t t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(&par[a], a, b) == a) return;   // the larger root goes under the smaller one
  }
}

__device__ __forceinline__ bool same_or_opposite(const double* n, int32_t u, int32_t w) {
  const double* a = n + 3 * (int64_t)u;
  const double* b = n + 3 * (int64_t)w;
  return (a[0] == b[0] && a[1] == b[1] && a[2] == b[2]) || (a[0] == -b[0] && a[1] == -b[1] && a[2] == -b[2]);
}

// Sorted edges: every triangle is joined with the earlier users of its edge whose normal is equal
// or opposite (Preprocessor.cpp:116-190).
__global__ void k_ing_edge_union(const unsigned long long* __restrict__ key, const int32_t* __restrict__ val,
                                 int64_t E, const double* __restrict__ triN, int32_t* par) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E) return;
  const unsigned long long k = key[i];
  for (int64_t j = i - 1; j >= 0 && key[j] == k; j--)
    if (same_or_opposite(triN, val[i], val[j])) uf_union(par, val[i], val[j]);
}

__global__ void k_ing_roots(int32_t* par, int64_t n, int32_t* __restrict__ label, int32_t* __restrict__ isRoot) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t r = uf_find(par, (int32_t)i);
  label[i] = r;
  isRoot[i] = r == (int32_t)i ? 1 : 0;
}

__global__ void k_ing_cop(const int32_t* __restrict__ label, const int32_t* __restrict__ rootScan, int64_t n,
                          int32_t* __restrict__ cop) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cop[i] = rootScan[label[i]];
}

// Centre bounds: one block per 1024 triangles, block minima / maxima, then a final block.
__global__ void k_ing_bounds(const double* __restrict__ cen, int64_t n, double* __restrict__ part) {
  __shared__ double s[6][256];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x; i < n && i < ((int64_t)blockIdx.x + 1) * 1024; i += 256)
    for (int a = 0; a < 3; a++) {
      lo[a] = fmin(lo[a], cen[3 * i + a]);
      hi[a] = fmax(hi[a], cen[3 * i + a]);
    }
  for (int a = 0; a < 3; a++) { s[a][threadIdx.x] = lo[a]; s[3 + a][threadIdx.x] = hi[a]; }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; a++) {
        s[a][threadIdx.x] = fmin(s[a][threadIdx.x], s[a][threadIdx.x + w]);
        s[3 + a][threadIdx.x] = fmax(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 6; a++) part[6 * blockIdx.x + a] = s[a][0];
}

__global__ void k_ing_bounds_final(const double* __restrict__ part, int64_t nb, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int64_t b = 0; b < nb; b++)
    for (int a = 0; a < 3; a++) {
      r[a] = fmin(r[a], part[6 * b + a]);
      r[3 + a] = fmax(r[3 + a], part[6 * b + 3 + a]);
    }
  for (int a = 0; a < 6; a++) out[a] = r[a];
}

__device__ __forceinline__ unsigned long long spread21(unsigned long long x) {
  x &= 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ void k_ing_morton(const double* __restrict__ cen, int64_t n, const double* __restrict__ bounds,
                             unsigned long long* __restrict__ code, int32_t* __restrict__ ids) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  unsigned long long q[3];
  for (int a = 0; a < 3; a++) {
    const double lo = bounds[a], ext = bounds[3 + a] - bounds[a];
    double u = ext > 0.0 ? (cen[3 * t + a] - lo) / ext * 2097151.0 : 0.0;
    u = fmin(fmax(u, 0.0), 2097151.0);
    q[a] = (unsigned long long)u;
  }
  code[t] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
  ids[t] = (int32_t)t;
}

// Common prefix length of sorted keys i and j (codes, then positions for equal codes); -1 outside.
__device__ __forceinline__ int delta(const unsigned long long* code, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const unsigned long long a = code[i], b = code[j];
  if (a != b) return __clzll((long long)(a ^ b));
  return 64 + __clzll((long long)((unsigned long long)i ^ (unsigned long long)j));
}

// Karras (2012) radix tree: internal node i of n - 1; node ids: internal 0 .. n-2, leaf j = n - 1 + j.
__global__ void k_ing_karras(const unsigned long long* __restrict__ code, int64_t n, int32_t* __restrict__ child,
                             int32_t* __restrict__ parent, int32_t* __restrict__ rlo, int32_t* __restrict__ rhi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = delta(code, n, i, i + 1) > delta(code, n, i, i - 1) ? 1 : -1;
  const int dmin = delta(code, n, i, i - d);
  int64_t lmax = 2;
  while (delta(code, n, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(code, n, i, i + (l + t) * d) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = delta(code, n, i, j);
  int64_t s = 0;
  for (int64_t div = 2;; div *= 2) {
    const int64_t t = (l + div - 1) / div;
    if (delta(code, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const int64_t g = i + s * d + (d < 0 ? -1 : 0);
  const int64_t lo = d > 0 ? i : j, hi = d > 0 ? j : i;
  const int32_t left = lo == g ? (int32_t)(n - 1 + g) : (int32_t)g;
  const int32_t right = hi == g + 1 ? (int32_t)(n - 1 + g + 1) : (int32_t)(g + 1);
  child[2 * i] = left;
  child[2 * i + 1] = right;
  parent[left] = (int32_t)i;
  parent[right] = (int32_t)i;
  rlo[i] = (int32_t)lo;
  rhi[i] = (int32_t)hi;
}

__global__ void k_ing_leaf_ranges(int64_t n, int32_t* __restrict__ rlo, int32_t* __restrict__ rhi) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  rlo[n - 1 + j] = (int32_t)j;
  rhi[n - 1 + j] = (int32_t)j;
}

// A node of at most kIngestLeaf triangles whose parent holds more (or the root) is a leaf: its box is
// its triangles' box; its ancestors' boxes are built bottom up (the second child to arrive merges).
__global__ void k_ing_boxes(int64_t n, const int32_t* __restrict__ parent, const int32_t* __restrict__ child,
                            const int32_t* __restrict__ rlo, const int32_t* __restrict__ rhi,
                            const int32_t* __restrict__ sortedIds, const double* __restrict__ triV,
                            double* box, int32_t* visits, int32_t* __restrict__ isLeaf) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= 2 * n - 1) return;
  const int32_t size = rhi[x] - rlo[x] + 1;
  const bool leaf = size <= kIngestLeaf && (x == 0 || rhi[parent[x]] - rlo[parent[x]] + 1 > kIngestLeaf);
  isLeaf[x] = leaf ? 1 : 0;
  if (!leaf) return;
  double b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int32_t k = rlo[x]; k <= rhi[x]; k++) {
    const double* v = triV + 9 * (int64_t)sortedIds[k];
    for (int c = 0; c < 3; c++)
      for (int a = 0; a < 3; a++) {
        b[a] = fmin(b[a], v[3 * c + a]);
        b[3 + a] = fmax(b[3 + a], v[3 * c + a]);
      }
  }
  for (int a = 0; a < 6; a++) box[6 * x + a] = b[a];
  int64_t cur = x;
  while (cur != 0) {
    const int32_t p = parent[cur];
    __threadfence();
    if (atomicAdd(&visits[p], 1) == 0) return;   // the sibling is not done yet: it continues
    __threadfence();
    const int32_t c0 = child[2 * p], c1 = child[2 * p + 1];
    // the sibling's box was written by another thread before its fence and counter increment:
    // read it coherently (agent-scope atomic loads, not a possibly stale first-level cache line)
    auto ld = [&](int64_t k) {
      return __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<unsigned long long*>(box + k),
                                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    };
    for (int a = 0; a < 3; a++) {
      box[6 * (int64_t)p + a] = fmin(ld(6 * (int64_t)c0 + a), ld(6 * (int64_t)c1 + a));
      box[6 * (int64_t)p + 3 + a] = fmax(ld(6 * (int64_t)c0 + 3 + a), ld(6 * (int64_t)c1 + 3 + a));
    }
    cur = p;
  }
}

// Scene-description node arrays over all 2n - 1 radix-tree nodes (nodes under a leaf are never
// reached from the root).
__global__ void k_ing_desc(int64_t n, const int32_t* __restrict__ child, const int32_t* __restrict__ rlo,
                           const int32_t* __restrict__ rhi, const int32_t* __restrict__ isLeaf,
                           int32_t* __restrict__ nodeChild, int32_t* __restrict__ nodeFirst,
                           int32_t* __restrict__ nodeCount) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= 2 * n - 1) return;
  const bool leaf = isLeaf[x] || x >= n - 1;
  nodeChild[2 * x] = leaf ? -1 : child[2 * x];
  nodeChild[2 * x + 1] = leaf ? -1 : child[2 * x + 1];
  nodeFirst[x] = rlo[x];
  nodeCount[x] = rhi[x] - rlo[x] + 1;
}

int Engine::ingest_soup(const float* coords9, const float* normals3, const int32_t* matIds, int nf, double* msOut) {
  // A failed ingestion leaves no scene, as a failed upload does.
  haveScene = false;
  haveResult = false;
  if (device < 0) return fail(SBR_E_STATE, "host-only context has no device");
  if (!coords9 || !normals3 || nf <= 0) return fail(SBR_E_ARG, "ingest: missing facets");
  if ((int64_t)nf * 3 >= (int64_t(1) << 31)) return fail(SBR_E_LIMIT, "ingest: too many facets");
  const auto t0 = std::chrono::steady_clock::now();
  IHIPCHK(hipSetDevice(device));
  const int64_t K = 3 * (int64_t)nf;
  DBuf<float> dC, dN;
  DBuf<int32_t> dM, dPerm, dPerm2, dFlag, dScan, dVid, dKeep, dKeepScan;
  DBuf<uint32_t> dKey, dKey2;
  DBuf<double> dVerts;
  DBuf<uint8_t> dTmp;
  IENSURE(dC, 9 * (size_t)nf);
  IENSURE(dN, 3 * (size_t)nf);
  if (matIds) IENSURE(dM, nf);
  IENSURE(dPerm, K);
  IENSURE(dPerm2, K);
  IENSURE(dKey, K);
  IENSURE(dKey2, K);
  IENSURE(dFlag, K);
  IENSURE(dScan, K);
  IENSURE(dVid, K);
  IHIPCHK(hipMemcpyAsync(dC.p, coords9, 9 * sizeof(float) * (size_t)nf, hipMemcpyHostToDevice, stream));
  IHIPCHK(hipMemcpyAsync(dN.p, normals3, 3 * sizeof(float) * (size_t)nf, hipMemcpyHostToDevice, stream));
  if (matIds) IHIPCHK(hipMemcpyAsync(dM.p, matIds, sizeof(int32_t) * (size_t)nf, hipMemcpyHostToDevice, stream));

  auto tmp_for = [&](size_t bytes) -> int {
    dTmp.ensure(std::max<size_t>(bytes, 1));
    return dTmp.p ? SBR_OK : fail(SBR_E_NOMEM, "ingest: temporary storage");
  };
  // 1. corners sorted by (x, y, z): stable key sorts by z, then y, then x
  k_ing_iota<<<nblk(K), 256, 0, stream>>>(dPerm.p, K);
  for (int axis = 2; axis >= 0; axis--) {
    k_ing_corner_key<<<nblk(K), 256, 0, stream>>>(dC.p, dPerm.p, K, axis, dKey.p);
    size_t tb = 0;
    IHIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, dKey.p, dKey2.p, dPerm.p, dPerm2.p, (int)K, 0, 32, stream));
    if (tmp_for(tb)) return SBR_E_NOMEM;
    IHIPCHK(hipcub::DeviceRadixSort::SortPairs(dTmp.p, tb, dKey.p, dKey2.p, dPerm.p, dPerm2.p, (int)K, 0, 32, stream));
    std::swap(dPerm.p, dPerm2.p);
    std::swap(dPerm.n, dPerm2.n);
  }
  // 2. distinct triples -> vertex ids
  k_ing_unique<<<nblk(K), 256, 0, stream>>>(dC.p, dPerm.p, K, dFlag.p);
  {
    size_t tb = 0;
    IHIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, dFlag.p, dScan.p, (int)K, stream));
    if (tmp_for(tb)) return SBR_E_NOMEM;
    IHIPCHK(hipcub::DeviceScan::ExclusiveSum(dTmp.p, tb, dFlag.p, dScan.p, (int)K, stream));
  }
  int32_t lastScan = 0, lastFlag = 0;
  IHIPCHK(hipMemcpyAsync(&lastScan, dScan.p + K - 1, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(&lastFlag, dFlag.p + K - 1, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipStreamSynchronize(stream));
  const int64_t nv = (int64_t)lastScan + lastFlag;
  IENSURE(dVerts, 3 * (size_t)nv);
  k_ing_vertices<<<nblk(K), 256, 0, stream>>>(dC.p, dPerm.p, dFlag.p, dScan.p, K, dVid.p, dVerts.p);
  // 3. facets with three distinct vertices -> triangles
  IENSURE(dKeep, nf);
  IENSURE(dKeepScan, nf);
  k_ing_keep<<<nblk(nf), 256, 0, stream>>>(dVid.p, nf, dKeep.p);
  {
    size_t tb = 0;
    IHIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, dKeep.p, dKeepScan.p, nf, stream));
    if (tmp_for(tb)) return SBR_E_NOMEM;
    IHIPCHK(hipcub::DeviceScan::ExclusiveSum(dTmp.p, tb, dKeep.p, dKeepScan.p, nf, stream));
  }
  int32_t lastKeepScan = 0, lastKeep = 0;
  IHIPCHK(hipMemcpyAsync(&lastKeepScan, dKeepScan.p + nf - 1, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(&lastKeep, dKeep.p + nf - 1, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipStreamSynchronize(stream));
  const int64_t n = (int64_t)lastKeepScan + lastKeep;
  if (n <= 0) return fail(SBR_E_ARG, "ingest: no triangle left after vertex dedup");
  DBuf<double> dTriV, dTriN, dTriCs, dCen, dBox, dBounds, dPart;
  DBuf<int32_t> dTriMat, dEdgeVal, dEdgeVal2, dPar, dLabel, dIsRoot, dRootScan, dCop;
  DBuf<unsigned long long> dEdgeKey, dEdgeKey2;
  IENSURE(dTriV, 9 * (size_t)n);
  IENSURE(dTriN, 3 * (size_t)n);
  IENSURE(dTriCs, 9 * (size_t)n);
  IENSURE(dTriMat, n);
  IENSURE(dCen, 3 * (size_t)n);
  IENSURE(dEdgeKey, 3 * (size_t)n);
  IENSURE(dEdgeKey2, 3 * (size_t)n);
  IENSURE(dEdgeVal, 3 * (size_t)n);
  IENSURE(dEdgeVal2, 3 * (size_t)n);
  k_ing_triangles<<<nblk(nf), 256, 0, stream>>>(dVid.p, dKeep.p, dKeepScan.p, nf, dVerts.p, dN.p,
                                                matIds ? dM.p : nullptr, dTriV.p, dTriN.p, dTriCs.p, dTriMat.p, dCen.p,
                                                dEdgeKey.p, dEdgeVal.p);
  IHIPCHK(hipGetLastError());
  // 4. coplanar ids: edges sorted by vertex pair (stable: users in triangle order), union-find
  const int64_t E = 3 * n;
  {
    size_t tb = 0;
    IHIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, dEdgeKey.p, dEdgeKey2.p, dEdgeVal.p, dEdgeVal2.p, (int)E, 0,
                                               64, stream));
    if (tmp_for(tb)) return SBR_E_NOMEM;
    IHIPCHK(hipcub::DeviceRadixSort::SortPairs(dTmp.p, tb, dEdgeKey.p, dEdgeKey2.p, dEdgeVal.p, dEdgeVal2.p, (int)E, 0,
                                               64, stream));
  }
  IENSURE(dPar, n);
  IENSURE(dLabel, n);
  IENSURE(dIsRoot, n);
  IENSURE(dRootScan, n);
  IENSURE(dCop, n);
  k_ing_iota<<<nblk(n), 256, 0, stream>>>(dPar.p, n);
  k_ing_edge_union<<<nblk(E), 256, 0, stream>>>(dEdgeKey2.p, dEdgeVal2.p, E, dTriN.p, dPar.p);
  k_ing_roots<<<nblk(n), 256, 0, stream>>>(dPar.p, n, dLabel.p, dIsRoot.p);
  {
    size_t tb = 0;
    IHIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, dIsRoot.p, dRootScan.p, (int)n, stream));
    if (tmp_for(tb)) return SBR_E_NOMEM;
    IHIPCHK(hipcub::DeviceScan::ExclusiveSum(dTmp.p, tb, dIsRoot.p, dRootScan.p, (int)n, stream));
  }
  k_ing_cop<<<nblk(n), 256, 0, stream>>>(dLabel.p, dRootScan.p, n, dCop.p);
  // 5. linear BVH
  const int64_t nb = (n + 1023) / 1024;
  IENSURE(dPart, 6 * (size_t)nb);
  IENSURE(dBounds, 6);
  k_ing_bounds<<<(unsigned)nb, 256, 0, stream>>>(dCen.p, n, dPart.p);
  k_ing_bounds_final<<<1, 64, 0, stream>>>(dPart.p, nb, dBounds.p);
  DBuf<unsigned long long> dCode, dCode2;
  DBuf<int32_t> dIds, dIds2, dChild, dParent, dLo, dHi, dVisits, dIsLeaf, dNodeChild, dNodeFirst, dNodeCount;
  const int64_t nn = 2 * n - 1;
  IENSURE(dCode, n);
  IENSURE(dCode2, n);
  IENSURE(dIds, n);
  IENSURE(dIds2, n);
  k_ing_morton<<<nblk(n), 256, 0, stream>>>(dCen.p, n, dBounds.p, dCode.p, dIds.p);
  {
    size_t tb = 0;
    IHIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, dCode.p, dCode2.p, dIds.p, dIds2.p, (int)n, 0, 63, stream));
    if (tmp_for(tb)) return SBR_E_NOMEM;
    IHIPCHK(hipcub::DeviceRadixSort::SortPairs(dTmp.p, tb, dCode.p, dCode2.p, dIds.p, dIds2.p, (int)n, 0, 63, stream));
  }
  IENSURE(dChild, 2 * (size_t)std::max<int64_t>(n - 1, 1));
  IENSURE(dParent, nn);
  IENSURE(dLo, nn);
  IENSURE(dHi, nn);
  IENSURE(dVisits, nn);
  IENSURE(dIsLeaf, nn);
  IENSURE(dBox, 6 * (size_t)nn);
  IENSURE(dNodeChild, 2 * (size_t)nn);
  IENSURE(dNodeFirst, nn);
  IENSURE(dNodeCount, nn);
  IHIPCHK(hipMemsetAsync(dVisits.p, 0, sizeof(int32_t) * (size_t)nn, stream));
  IHIPCHK(hipMemsetAsync(dParent.p, 0xFF, sizeof(int32_t) * (size_t)nn, stream));
  if (n > 1) k_ing_karras<<<nblk(n - 1), 256, 0, stream>>>(dCode2.p, n, dChild.p, dParent.p, dLo.p, dHi.p);
  k_ing_leaf_ranges<<<nblk(n), 256, 0, stream>>>(n, dLo.p, dHi.p);
  k_ing_boxes<<<nblk(nn), 256, 0, stream>>>(n, dParent.p, dChild.p, dLo.p, dHi.p, dIds2.p, dTriV.p, dBox.p, dVisits.p,
                                            dIsLeaf.p);
  k_ing_desc<<<nblk(nn), 256, 0, stream>>>(n, dChild.p, dLo.p, dHi.p, dIsLeaf.p, dNodeChild.p, dNodeFirst.p,
                                           dNodeCount.p);
  IHIPCHK(hipGetLastError());
  // 6. the description on the host, and the usual upload (DFS flattening, device layouts)
  std::vector<double> triV(9 * n), triN(3 * n), triCs(9 * n), box(6 * nn);
  std::vector<int32_t> cop(n), mat(n), mesh(n), nodeChild(2 * nn), nodeFirst(nn), nodeCount(nn), leafTri(n);
  IHIPCHK(hipMemcpyAsync(triV.data(), dTriV.p, sizeof(double) * triV.size(), hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(triN.data(), dTriN.p, sizeof(double) * triN.size(), hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(triCs.data(), dTriCs.p, sizeof(double) * triCs.size(), hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(cop.data(), dCop.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(mat.data(), dTriMat.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(box.data(), dBox.p, sizeof(double) * box.size(), hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(nodeChild.data(), dNodeChild.p, sizeof(int32_t) * nodeChild.size(), hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(nodeFirst.data(), dNodeFirst.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(nodeCount.data(), dNodeCount.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipMemcpyAsync(leafTri.data(), dIds2.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
  IHIPCHK(hipStreamSynchronize(stream));
  for (int64_t i = 0; i < n; i++) mesh[i] = (int32_t)i;
  const double msDevice = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  sbr_scene_desc d{};
  d.n_tri = (int32_t)n;
  d.tri_v = triV.data();
  d.tri_n = triN.data();
  d.tri_cs = triCs.data();
  d.tri_cop = cop.data();
  d.tri_mat = mat.data();
  d.mesh_tri_id = mesh.data();
  d.n_node = (int32_t)nn;
  d.node_box = box.data();
  d.node_child = nodeChild.data();
  d.node_first = nodeFirst.data();
  d.node_count = nodeCount.data();
  d.n_leaf_tri = (int32_t)n;
  d.leaf_tri = leafTri.data();
  const int rc = upload(&d);
  if (msOut) {
    msOut[0] = msDevice;
    msOut[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return rc;
}

}  // namespace sbr
