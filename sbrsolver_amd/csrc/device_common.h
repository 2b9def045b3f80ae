// Device helpers of the gfx950 engine: BVH closest hit, reception sphere, image correction,
// complex arithmetic and the per-path field.  Compiled with -ffp-contract=off (see sbr_math.h).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "engine.h"
#include "capture_prefilter.h"
#include "correction_prefilter.h"
#include "prefilter.h"

namespace sbr {

__device__ __forceinline__ V3 ld3(const double* p) { return v3(p[0], p[1], p[2]); }

// ---------------------------------------------------------------------------------------------
// Closest hit over the BVH: BVH<Triangle>::RayIntersects (BVH/BVH.h:276-328).
//
// The reference visits, in DFS order, every node whose box the infinite line crosses
// (BVHNode.h:64-86) and keeps the first triangle with strictly smaller distance, i.e. the
// minimum over (distance, DFS position) among the triangles of visited leaves.  Triangles are
// stored in DFS leaf order, so the position is the triangle index and the tie-break is explicit.
// EXACT=true walks every crossed box, as the reference does.  EXACT=false adds two conservative
// cuts that never drop a triangle that could win: boxes wholly behind the origin (a hit needs
// t > EPSILON) and boxes whose entry parameter exceeds the best hit's parameter by a relative
// margin of 1e-7; children are visited nearest entry first so the cut bites early.  A box is
// only entered when the reference's own slab test passes, so the reference's grazing misses are
// reproduced too.
//
// The slab test runs with a per-ray reciprocal direction (one multiply instead of a divide per
// slab).  The reciprocal product differs from the reference's quotient by at most ~2 ulp and has
// the same inf / NaN pattern (both come only from zero direction components), so the pass / fail
// decision can differ only when the entry and exit parameters agree to ~1e-15; inside a band of
// 1e-13 relative the test is redone with the reference's divisions.
// ---------------------------------------------------------------------------------------------
struct Hit {
  int idx;       // DFS triangle index, INT_MAX if none
  double dist;
  double t;
  int nodes;     // BVH nodes entered
  int tests;     // ray-triangle tests
  int exact64;   // child boxes decided by the FP64 slab test (grazing or no single-precision path)
  int redone;    // 1: the wide-tree winner was not confirmed and the ray was redone on the reference tree
};

// Hit point of a closest hit: the same operations tri_intersect applies, so the value is
// bit-identical to the point it computed (kept out of the traversal loop to save registers).
__device__ __forceinline__ V3 hit_point(V3 o, V3 d, const Hit& h) {
  return h.idx == INT_MAX ? v3(0, 0, 0) : vadd(o, vmul(d, h.t));
}

struct RayPre {
  V3 o, d, inv;
  bool safe;     // no subnormal direction component (reciprocal exact in inf / NaN pattern)
};

__device__ __forceinline__ RayPre ray_pre(V3 o, V3 d) {
  RayPre r;
  r.o = o;
  r.d = d;
  r.inv = v3(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
  const double tiny = 1e-290;
  r.safe = !((d.x != 0.0 && fabs(d.x) < tiny) || (d.y != 0.0 && fabs(d.y) < tiny) || (d.z != 0.0 && fabs(d.z) < tiny));
  return r;
}

__device__ __forceinline__ double fin_abs(double x) { return isfinite(x) ? fabs(x) : 0.0; }

__device__ __forceinline__ bool box_test(const double* bmin, const double* bmax, const RayPre& r, double* lo,
                                         double* hi) {
  if (r.safe) {
    double t_max = INFINITY, t_min = -INFINITY;
    const double oo[3] = {r.o.x, r.o.y, r.o.z}, iv[3] = {r.inv.x, r.inv.y, r.inv.z};
#pragma unroll
    for (int i = 0; i < 3; i++) {
      double t1 = (bmin[i] - oo[i]) * iv[i];
      double t2 = (bmax[i] - oo[i]) * iv[i];
      t_max = smin(t_max, smax(t1, t2));
      t_min = smax(t_min, smin(t1, t2));
    }
    double m = 1e-13 * (fin_abs(t_min) + fin_abs(t_max));
    if (t_min > t_max + m) return false;
    if (t_min < t_max - m) {
      *lo = t_min;
      *hi = t_max;
      return true;
    }
  }
  return box_line(bmin, bmax, r.o, r.d, lo, hi);
}

template <bool EXACT>
__device__ __forceinline__ bool keep_box(double lo, double hi, const Hit& h) {
  if (EXACT) return true;
  if (hi < -1e-9) return false;
  if (h.idx != INT_MAX && lo > h.t * (1.0 + 1e-7) + 1e-7) return false;
  return true;
}

// FP64 path (EXACT, or a ray the single-precision tests do not cover).
template <bool EXACT>
__device__ __forceinline__ bool child_exact(const DChildF& c, const DNode* __restrict__ nodes, const RayPre& r,
                                            const Hit& h, float* entry) {
  double lo, hi;
  const DNode& n = nodes[c.item & ~kLeafBit];
  if (!box_test(n.bmin, n.bmax, r, &lo, &hi) || !keep_box<EXACT>(lo, hi, h)) return false;
  *entry = __double2float_rd(lo);
  return true;
}

// Traversal stack of one thread.  Entries are work items -- inner node n, leaf n | kLeafBit,
// leaf-BVH node ~s -- with a lower bound of their entry parameter.  Entries [0, lcap) are in LDS,
// entry e at ln[e * ls] / lt[e * ls] (ls = threads per block, so a wave's entries are contiguous);
// deeper entries, when the tree needs more than lcap, are in device memory, entry e at
// g[(e - lcap) * gs] (gs = threads of the grid).  A short LDS part keeps the blocks per CU that the
// registers allow on scenes with tall trees.
struct TravStack {
  int* ln;
  float* lt;
  int ls;
  int lcap;
  int2* g;
  int64_t gs;
  __device__ __forceinline__ void put(int e, int item, float t) const {
    if (e < lcap) {
      ln[e * ls] = item;
      lt[e * ls] = t;
    } else {
      g[(int64_t)(e - lcap) * gs] = make_int2(item, __float_as_int(t));
    }
  }
  // (item, entry parameter's bits)
  __device__ __forceinline__ int2 get(int e) const {
    if (e < lcap) return make_int2(ln[e * ls], __float_as_int(lt[e * ls]));
    return g[(int64_t)(e - lcap) * gs];
  }
};

// A stack held wholly in LDS (stkN / stkT per thread, stride = threads per block, depth entries).
__device__ __forceinline__ TravStack lds_stack(int* stkN, float* stkT, int stride, int depth) {
  TravStack st;
  st.ln = stkN;
  st.lt = stkT;
  st.ls = stride;
  st.lcap = depth;
  st.g = nullptr;
  st.gs = 0;
  return st;
}

// Exact test of DFS triangle k (Triangle::RayIntersects, Triangle.cpp), keeping the minimum
// (distance, DFS position); clip follows the best hit's parameter for the pre-tests.
template <bool COUNT>
__device__ __forceinline__ void tri_exact(const DTri* __restrict__ tris, int k, V3 o, V3 d, Hit& h, float& clip) {
  if (COUNT) h.tests++;
  const DTri& tr = tris[k];
  V3 p;
  double dist, t;
  if (tri_intersect(ld3(tr.v1), ld3(tr.e1), ld3(tr.e2), o, d, &p, &dist, &t)) {
    if (dist < h.dist || (dist == h.dist && k < h.idx)) {
      h.idx = k;
      h.dist = dist;
      h.t = t;
      clip = __double2float_ru(t * (1.0 + 1e-6));
    }
  }
}

// The triangles [first, first + count) of a leaf.  With the pre-test usable for the ray: the leaf's
// boxes are pre-tested, two loads in flight at a time, into a mask per 64 triangles; then the
// exact tests of the surviving triangles, so that the wave runs as many exact tests as its busiest
// lane needs rather than one wherever any lane's box passed.
template <bool COUNT>
__device__ __forceinline__ void leaf_tests(const DTri* __restrict__ tris, const DBox32* __restrict__ tbox, int first,
                                           int count, const RayF& f, V3 o, V3 d, Hit& h, float& clip) {
  if (f.on) {
    const int end = first + count;
    for (int k0 = first; k0 < end; k0 += 64) {
      const int m = min(64, end - k0);
      unsigned long long pass = 0;
      for (int j = 0; j < m; j += 4) {
        // Four boxes loaded and tested unconditionally (indices clamped), so that their loads
        // are in flight together.
        const int j1 = min(j + 1, m - 1), j2 = min(j + 2, m - 1), j3 = min(j + 3, m - 1);
        const DBox32 b0 = tbox[k0 + j];
        const DBox32 b1 = tbox[k0 + j1];
        const DBox32 b2 = tbox[k0 + j2];
        const DBox32 b3 = tbox[k0 + j3];
        float e;
        pass |= (unsigned long long)(box32_may_hit(b0.lo, b0.hi, f, clip, &e) ? 1u : 0u) << j;
        pass |= (unsigned long long)(box32_may_hit(b1.lo, b1.hi, f, clip, &e) ? 1u : 0u) << j1;
        pass |= (unsigned long long)(box32_may_hit(b2.lo, b2.hi, f, clip, &e) ? 1u : 0u) << j2;
        pass |= (unsigned long long)(box32_may_hit(b3.lo, b3.hi, f, clip, &e) ? 1u : 0u) << j3;
      }
      while (pass) {
        const int j = __builtin_ctzll(pass);
        pass &= pass - 1;
        tri_exact<COUNT>(tris, k0 + j, o, d, h, clip);
      }
    }
  } else {
    for (int k = first; k < first + count; k++) tri_exact<COUNT>(tris, k, o, d, h, clip);
  }
}

// The triangles subTri[first .. first + count - 1] of a leaf-BVH leaf (the ray is covered by the
// pre-test): pre-tests four at a time, then the exact tests of the survivors.
template <bool COUNT>
__device__ __forceinline__ void leaf_tests_listed(const DTri* __restrict__ tris, const DBox32* __restrict__ tbox,
                                                  const int32_t* __restrict__ subTri, int first, int count,
                                                  const RayF& f, V3 o, V3 d, Hit& h, float& clip) {
  for (int j = 0; j < count; j += 4) {
    const int m = min(4, count - j);
    const int k0 = subTri[first + j], k1 = subTri[first + j + min(1, m - 1)], k2 = subTri[first + j + min(2, m - 1)],
              k3 = subTri[first + j + min(3, m - 1)];
    const DBox32 b0 = tbox[k0];
    const DBox32 b1 = tbox[k1];
    const DBox32 b2 = tbox[k2];
    const DBox32 b3 = tbox[k3];
    float e;
    unsigned pass = (box32_may_hit(b0.lo, b0.hi, f, clip, &e) ? 1u : 0u) | (box32_may_hit(b1.lo, b1.hi, f, clip, &e) ? 2u : 0u) |
                    (box32_may_hit(b2.lo, b2.hi, f, clip, &e) ? 4u : 0u) | (box32_may_hit(b3.lo, b3.hi, f, clip, &e) ? 8u : 0u);
    pass &= (1u << m) - 1u;   // the clamped repeats of the last triangle
    while (pass) {
      const int u = __builtin_ctz(pass);
      pass &= pass - 1;
      tri_exact<COUNT>(tris, u == 0 ? k0 : u == 1 ? k1 : u == 2 ? k2 : k3, o, d, h, clip);
    }
  }
}

//
// SUB: the scene has leaf BVHs (otherwise that code is left out of the kernel).
// COUNT: h.nodes / h.tests / h.exact64 are counted (diagnostic builds of the kernels only; the
// counters otherwise stay zero and take no registers in the traversal loop).
// A leaf with a leaf BVH is searched through it (EXACT=false, pre-test usable for the ray): the
// leaf-BVH boxes are unions of the triangles' padded pre-test boxes, so a subtree is skipped only
// when the ray provably misses all its triangles or every hit in it lies beyond the best hit,
// i.e. exactly when the per-triangle pre-test would reject each of its triangles.  The leaf's
// result, the minimum (distance, DFS position) over its hit triangles, is unchanged.
template <bool EXACT, bool SUB, bool COUNT = false>
__device__ __forceinline__ bool closest_hit(const DNode* __restrict__ nodes, const DNodeF* __restrict__ nodesF,
                                            const DTri* __restrict__ tris, const DBox32* __restrict__ tbox,
                                            const DSub* __restrict__ subs, const int32_t* __restrict__ subTri,
                                            float ext, V3 o, V3 d, Hit& h, const TravStack& st) {
  h.idx = INT_MAX;
  h.dist = INFINITY;
  h.t = INFINITY;
  h.nodes = 0;
  h.tests = 0;
  h.exact64 = 0;
  h.redone = 0;
  const RayPre r = ray_pre(o, d);
  const RayF f = rayf_pre(o, d, ext);
  const bool fast = !EXACT && f.on;
  float clip = INFINITY;
  {
    double lo, hi;
    if (!box_test(nodes[0].bmin, nodes[0].bmax, r, &lo, &hi) || !keep_box<EXACT>(lo, hi, h)) return false;
  }
  int sp = 0;
  int cur = nodes[0].count > 0 ? kLeafBit : 0;
  // Exact test of DFS triangle k after its pre-test.
  auto test_tri = [&](int k) {
    float e;
    if (f.on && !box32_may_hit(tbox[k].lo, tbox[k].hi, f, clip, &e)) return;
    tri_exact<COUNT>(tris, k, o, d, h, clip);
  };
  // Next work item from the stack, skipping entries that cannot beat the best hit; false when the
  // stack is empty.
  auto pop = [&]() {
    while (sp > 0) {
      --sp;
      const int2 e = st.get(sp);
      if (!EXACT && h.idx != INT_MAX && (double)__int_as_float(e.y) > h.t * (1.0 + 1e-7) + 1e-7) continue;
      cur = e.x;
      return true;
    }
    return false;
  };
  bool live = true;
  // While-while traversal: each lane descends through inner nodes until it holds a leaf (or a
  // leaf-BVH node) or runs out of work, then the wave tests its lanes' leaves together.  Every lane
  // visits the same items in the same order as a one-item-per-step loop would.
  while (true) {
    while (live && cur >= 0 && !(cur & kLeafBit)) {
      if (COUNT) h.nodes++;
      const DNodeF& nf = nodesF[cur];
      float ta, tb;
      bool ina, inb;
      if (fast) {
        // Whether the traversal enters a child, single-precision path: the grown box decides a
        // miss (or a child wholly behind the origin or beyond the best hit), the shrunk box decides
        // a pass, and a ray between the two takes the reference's FP64 slab test on the exact box.
        // ta / tb: lower bounds of the t of every hit inside the child.
        // Single-precision tests of both children first; then the FP64 slab test of each child
        // whose ray lies between its grown and shrunk boxes, one child per pass, so that a wave
        // runs the FP64 test once for all its lanes' first undecided child and a second time only
        // where a lane has two.
        ina = box32_may_hit(nf.box[0], nf.box[0] + 3, f, clip, &ta);
        inb = box32_may_hit(nf.box[1], nf.box[1] + 3, f, clip, &tb);
        bool und0 = ina && !(nf.ch[0].inner && box32_crossed(nf.box[0] + 6, nf.box[0] + 9, f));
        bool und1 = inb && !(nf.ch[1].inner && box32_crossed(nf.box[1] + 6, nf.box[1] + 9, f));
        while (und0 || und1) {
          const bool first = und0;
          if (COUNT) h.exact64++;
          double lo, hi;
          const DNode& n = nodes[(first ? nf.ch[0].item : nf.ch[1].item) & ~kLeafBit];
          const bool in = box_test(n.bmin, n.bmax, r, &lo, &hi) && keep_box<false>(lo, hi, h);
          if (first) {
            ina = in;
            und0 = false;
          } else {
            inb = in;
            und1 = false;
          }
        }
      } else {
        ina = child_exact<EXACT>(nf.ch[0], nodes, r, h, &ta);
        inb = child_exact<EXACT>(nf.ch[1], nodes, r, h, &tb);
      }
      if (ina && inb) {
        const bool bFirst = !EXACT && tb < ta;   // EXACT: the reference's DFS order
        st.put(sp, bFirst ? nf.ch[0].item : nf.ch[1].item, bFirst ? ta : tb);
        sp++;
        cur = bFirst ? nf.ch[1].item : nf.ch[0].item;
      } else if (ina || inb) {
        cur = ina ? nf.ch[0].item : nf.ch[1].item;
      } else {
        live = pop();
      }
    }
    if (!live) break;
    if (COUNT) h.nodes++;
    if (SUB && cur < 0) {
      const DSub sn = subs[~cur];
      if (sn.count > 0) {
        for (int j = sn.first; j < sn.first + sn.count; j++) test_tri(subTri[j]);
      } else {
        float ta, tb;
        const bool ina = box32_may_hit(subs[sn.first].lo, subs[sn.first].hi, f, clip, &ta);
        const bool inb = box32_may_hit(subs[sn.first + 1].lo, subs[sn.first + 1].hi, f, clip, &tb);
        if (ina && inb) {
          const bool aFirst = ta <= tb;
          st.put(sp, ~(aFirst ? sn.first + 1 : sn.first), aFirst ? tb : ta);
          sp++;
          cur = ~(aFirst ? sn.first : sn.first + 1);
          continue;
        }
        if (ina || inb) {
          cur = ~(ina ? sn.first : sn.first + 1);
          continue;
        }
      }
    } else {
      // a, b, count, sub from the leaf's slot of the single-precision records (LDS when staged)
      const int4 li = *reinterpret_cast<const int4*>(&nodesF[cur & ~kLeafBit].ch[0]);
      if (SUB && fast && li.w >= 0) {
        cur = ~li.w;   // search the leaf through its leaf BVH
        continue;
      }
      leaf_tests<COUNT>(tris, tbox, li.x, li.z, f, o, d, h, clip);
    }
    live = pop();
    if (!live) break;
  }
  return h.idx != INT_MAX;
}

// ---------------------------------------------------------------------------------------------
// Closest hit through the wide tree (DNode4), for rays the single-precision pre-test covers.
//
// The wide tree's boxes are the reference tree's boxes grown by 2 delta, so every triangle that
// tri_intersect hits lies in boxes whose computed intervals are non-empty: the traversal finds the
// minimum (distance, DFS position) over all hit triangles, with the same cuts as the reference-tree
// traversal (boxes wholly behind the origin or beyond the best hit; children visited nearest entry
// first).  The reference only tests the triangles of the leaves whose boxes, and all their
// ancestors' boxes, pass its FP64 slab test (BVHNode.h:64-86).  If the winner's leaf is reached that
// way (wide_reaches), it is also the minimum over the triangles the reference tests, i.e. the
// reference's result; otherwise (a line grazing a box of the winner's path) the caller redoes the
// ray on the reference tree.  No hit at all means the reference finds none either.  Leaves with a
// leaf BVH are wide subtrees of the tree itself (the leaf BVH collapsed four-wide), so the
// traversal has one node kind; SUB: the scene has such leaves (listed-triangle leaves).
// Stack items: a wide node (>= 0) or a leaf (kLeafBit | DLeaf index).
// ---------------------------------------------------------------------------------------------
template <bool SUB, bool COUNT>
__device__ __forceinline__ void wide_hit(const DNode4* __restrict__ n4, const DLeaf* __restrict__ leaves,
                                         const DTri* __restrict__ tris, const DBox32* __restrict__ tbox,
                                         const int32_t* __restrict__ subTri, const RayF& f, V3 o, V3 d, Hit& h,
                                         const TravStack& st) {
  h.idx = INT_MAX;
  h.dist = INFINITY;
  h.t = INFINITY;
  float clip = INFINITY;
  int sp = 0;
  int cur = 0;
  auto pop = [&]() {
    while (sp > 0) {
      --sp;
      const int2 e = st.get(sp);
      if (h.idx != INT_MAX && (double)__int_as_float(e.y) > h.t * (1.0 + 1e-7) + 1e-7) continue;
      cur = e.x;
      return true;
    }
    return false;
  };
  auto push = [&](int item, float t) {
    st.put(sp, item, t);
    sp++;
  };
  bool live = true;
  // While-while traversal as in closest_hit: lanes descend to a leaf, then the wave tests its
  // lanes' leaves together.
  while (true) {
    while (live && cur >= 0 && !(cur & kLeafBit)) {
      if (COUNT) h.nodes++;
      const DNode4& n = n4[cur];
      float t[kWide];
      int it[kWide];
#pragma unroll
      for (int k = 0; k < kWide; k++) {
        float e;
        const bool in = box32_may_hit(n.lo[0][k], n.lo[1][k], n.lo[2][k], n.hi[0][k], n.hi[1][k], n.hi[2][k], f, clip, &e);
        it[k] = n.item[k];
        t[k] = (in && it[k] != -1) ? e : INFINITY;
      }
      // nearest entry first: a sorting network over the (entry, item) pairs
      auto cswap = [&](int a, int b) {
        if (t[b] < t[a]) {
          const float tt = t[a];
          t[a] = t[b];
          t[b] = tt;
          const int ii = it[a];
          it[a] = it[b];
          it[b] = ii;
        }
      };
      if (kWide == 4) {
        cswap(0, 1);
        cswap(2, 3);
        cswap(0, 2);
        cswap(1, 3);
        cswap(1, 2);
      } else {   // eight: Batcher's odd-even merge network
        cswap(0, 1); cswap(2, 3); cswap(4, 5); cswap(6, 7);
        cswap(0, 2); cswap(1, 3); cswap(4, 6); cswap(5, 7);
        cswap(1, 2); cswap(5, 6);
        cswap(0, 4); cswap(1, 5); cswap(2, 6); cswap(3, 7);
        cswap(2, 4); cswap(3, 5);
        cswap(1, 2); cswap(3, 4); cswap(5, 6);
      }
      if (t[0] == INFINITY) {
        live = pop();
      } else {
#pragma unroll
        for (int k = kWide - 1; k >= 1; k--)
          if (t[k] != INFINITY) push(it[k], t[k]);
        cur = it[0];
      }
    }
    if (!live) break;
    if (COUNT) h.nodes++;
    const DLeaf li = leaves[cur & ~kLeafBit];
    if (SUB && li.listed)
      leaf_tests_listed<COUNT>(tris, tbox, subTri, li.first, li.count, f, o, d, h, clip);
    else
      leaf_tests<COUNT>(tris, tbox, li.first, li.count, f, o, d, h, clip);
    live = pop();
    if (!live) break;
  }
}

// Whether the reference's traversal reaches the reference-tree leaf `node`: its box and those of
// all its ancestors pass the FP64 slab test.  Walking up, a node whose shrunk box the line crosses
// (single precision) settles it and all its ancestors; other nodes (flat leaves, mostly) take the
// reference's own slab test with its divisions (box_line), too rare to warrant the reciprocals.
template <bool COUNT>
__device__ __forceinline__ bool wide_reaches(int node, const DVer* __restrict__ ver, const DNode* __restrict__ nodes,
                                             const RayF& f, V3 o, V3 d, Hit& h) {
  while (node >= 0) {
    const DVer v = ver[node];
    if (v.shrunk && box32_crossed(v.lo, v.hi, f)) return true;
    if (COUNT) h.exact64++;
    double lo, hi;
    if (!box_line(nodes[node].bmin, nodes[node].bmax, o, d, &lo, &hi)) return false;
    node = v.parent;
  }
  return true;
}

// Closest hit with the reference's result (BVH<Triangle>::RayIntersects): the wide tree where it
// applies, the reference tree (closest_hit, clipped) for rays the single-precision pre-test does
// not cover and for winners whose leaf the check does not confirm.  nodesF: the reference tree's
// single-precision records (device memory).
template <bool SUB, bool COUNT = false>
__device__ __forceinline__ bool closest_hit_wide(const WideTree& w, const DNode* __restrict__ nodes,
                                                 const DNodeF* __restrict__ nodesF, const DTri* __restrict__ tris,
                                                 const DBox32* __restrict__ tbox, const DSub* __restrict__ subs,
                                                 const int32_t* __restrict__ subTri, float ext, V3 o, V3 d, Hit& h,
                                                 const TravStack& st) {
  h.nodes = 0;
  h.tests = 0;
  h.exact64 = 0;
  h.redone = 0;
  const RayF f = rayf_pre(o, d, ext);
  bool done = false;
  if (f.on) {
    wide_hit<SUB, COUNT>(w.n4, w.leaves, tris, tbox, subTri, f, o, d, h, st);
    done = h.idx == INT_MAX || wide_reaches<COUNT>(tris[h.idx].leaf, w.ver, nodes, f, o, d, h);
  }
  if (!done) {
    const int n0 = h.nodes, t0 = h.tests, x0 = h.exact64;
    closest_hit<false, SUB, COUNT>(nodes, nodesF, tris, tbox, subs, subTri, ext, o, d, h, st);
    if (COUNT) {
      h.nodes += n0;
      h.tests += t0;
      h.exact64 += x0;
      h.redone = 1;
    }
  }
  return h.idx != INT_MAX;
}

// ---------------------------------------------------------------------------------------------
// Reception sphere: HitReceptionSphere (RTSolver.cpp:805-863), and the alpha*L/sqrt(3) model.
// ---------------------------------------------------------------------------------------------
// base: the totalPathLength handed to the node (model 0) or the true length to S (model 1).
// (sphere_hit and its single-precision pre-test: capture_prefilter.h)

// ImagePoint (RTSolver.cpp:910-924)
__device__ __forceinline__ V3 image_point(const DSurf& s, V3 src) {
  double dist = vdot(ld3(s.n), vsub(src, ld3(s.v1)));
  return vsub(src, smul(2 * dist, ld3(s.n)));
}

// ---------------------------------------------------------------------------------------------
// Complex FP64 arithmetic for the field (std::complex<double> semantics up to last-ulp rounding).
// ---------------------------------------------------------------------------------------------
struct cx {
  double r, i;
};
__device__ __forceinline__ cx mkc(double r, double i) { cx z; z.r = r; z.i = i; return z; }
__device__ __forceinline__ cx cadd(cx a, cx b) { return mkc(a.r + b.r, a.i + b.i); }
__device__ __forceinline__ cx csub(cx a, cx b) { return mkc(a.r - b.r, a.i - b.i); }
__device__ __forceinline__ cx cmul(cx a, cx b) { return mkc(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r); }
__device__ __forceinline__ cx cscale(cx a, double s) { return mkc(a.r * s, a.i * s); }
__device__ __forceinline__ cx cneg(cx a) { return mkc(-a.r, -a.i); }
// Smith's division.
__device__ __forceinline__ cx cdiv(cx a, cx b) {
  if (fabs(b.r) >= fabs(b.i)) {
    double q = b.i / b.r, den = b.r + b.i * q;
    return mkc((a.r + a.i * q) / den, (a.i - a.r * q) / den);
  }
  double q = b.r / b.i, den = b.r * q + b.i;
  return mkc((a.r * q + a.i) / den, (a.i * q - a.r) / den);
}
__device__ __forceinline__ double cabs_(cx a) { return hypot(a.r, a.i); }
// Principal square root.
__device__ __forceinline__ cx csqrt_(cx z) {
  if (z.r == 0.0 && z.i == 0.0) return mkc(0.0, z.i);
  double m = cabs_(z);
  if (z.r >= 0.0) {
    double t = sqrt(0.5 * (m + z.r));
    return mkc(t, z.i / (2.0 * t));
  }
  double t = sqrt(0.5 * (m - z.r));
  return mkc(fabs(z.i) / (2.0 * t), copysign(t, z.i));
}
__device__ __forceinline__ cx cexp_(cx z) {
  double e = exp(z.r);
  return mkc(e * cos(z.i), e * sin(z.i));
}
__device__ __forceinline__ cx csin_(cx z) { return mkc(sin(z.r) * cosh(z.i), cos(z.r) * sinh(z.i)); }
__device__ __forceinline__ cx ccos_(cx z) { return mkc(cos(z.r) * cosh(z.i), -sin(z.r) * sinh(z.i)); }
__device__ __forceinline__ cx clog_(cx z) { return mkc(log(cabs_(z)), atan2(z.i, z.r)); }
// asin(z) = -i log(i z + sqrt(1 - z^2)); for real z in [-1, 1] the real routine is used.
__device__ __forceinline__ cx casin_(cx z) {
  if (z.i == 0.0 && fabs(z.r) <= 1.0) return mkc(asin(z.r), z.i);
  cx w = csqrt_(csub(mkc(1.0, 0.0), cmul(z, z)));
  cx u = cadd(mkc(-z.i, z.r), w);
  cx l = clog_(u);
  return mkc(l.i, -l.r);
}

struct V3C {
  cx x, y, z;
};

// Physical constants as the reference's MSVC build evaluates them (Constants.h:6-22).
#define SBR_PI 3.141592653589793
#define SBR_C0 2.99792458e8

// AngleBetween (VecMatDef.h:56-77), result in radians.
__device__ __forceinline__ double angle_between_rad(V3 a, V3 b) {
  a = vnormalized(a);
  b = vnormalized(b);
  double ratio = vdot(a, b);
  if (ratio < 0) return SBR_PI - 2.0 * asin(vnorm(vsub(vneg(a), b)) / 2.0);
  return 2.0 * asin(vnorm(vsub(a, b)) / 2.0);
}

// rows-as-given 3x3 (row-major) times vector, reduction order (m0*v0 + m1*v1) + m2*v2.
__device__ __forceinline__ V3 mat_vec(const double* m, V3 v) {
  return v3((m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[3] * v.x + m[4] * v.y) + m[5] * v.z,
            (m[6] * v.x + m[7] * v.y) + m[8] * v.z);
}
__device__ __forceinline__ cx rdot(double a, double b, double c, cx x, cx y, cx z) {
  return cadd(cadd(cscale(x, a), cscale(y, b)), cscale(z, c));
}
__device__ __forceinline__ V3C mat_vec_c(const double* m, const V3C& v) {
  V3C r;
  r.x = rdot(m[0], m[1], m[2], v.x, v.y, v.z);
  r.y = rdot(m[3], m[4], m[5], v.x, v.y, v.z);
  r.z = rdot(m[6], m[7], m[8], v.x, v.y, v.z);
  return r;
}
__device__ __forceinline__ V3C matT_vec_c(const double* m, const V3C& v) {   // m^T v
  V3C r;
  r.x = rdot(m[0], m[3], m[6], v.x, v.y, v.z);
  r.y = rdot(m[1], m[4], m[7], v.x, v.y, v.z);
  r.z = rdot(m[2], m[5], m[8], v.x, v.y, v.z);
  return r;
}

// TE/TM coefficients (FieldCompute.cpp:237-302) for incidence thetaI from air onto a material
// of relative permittivity relPerm and conductivity sigma; returns the two factors applied to
// (E_theta, E_phi): reflection (refl=true) or transmission.
// infWall (useFresnelCoeff) with slabWidth > 0: a slab of that thickness (ITU-R P.2040: phase thickness
// q = k d sqrt(eps - sin^2 theta), R = G (1 - e^{-2jq}) / (1 - G^2 e^{-2jq}), T = (1 - G^2) e^{-jq} /
// (1 - G^2 e^{-2jq})), an extension; slabWidth = 0 keeps the reference's walls.
__device__ __forceinline__ void fresnel(bool refl, double thetaI, double relPerm, double sigma, double freq,
                                        bool infWall, double slabWidth, cx* fTM, cx* fTE) {
  const double U0 = 4 * SBR_PI * 1e-7;
  const double E0 = 1.0 / (U0 * SBR_C0 * SBR_C0);
  const double ETA0 = U0 * SBR_C0;
  cx et = mkc(relPerm, -1 * sigma / (E0 * 2 * SBR_PI * freq));
  cx ni = csqrt_(mkc(1.0, 0.0));
  cx nt = csqrt_(et);
  cx ti = mkc(thetaI, 0.0);
  cx tt = casin_(cmul(cdiv(ni, nt), csin_(ti)));
  cx etai = cdiv(mkc(ETA0, 0.0), ni), etat = cdiv(mkc(ETA0, 0.0), nt);
  cx ci = ccos_(ti), ct = ccos_(tt), si = csin_(ti);
  cx gtm = cdiv(cadd(cmul(cneg(etat), ct), cmul(etai, ci)), cadd(cmul(etat, ct), cmul(etai, ci)));
  cx a = cdiv(etat, ct), b = cdiv(etai, ci);
  cx gte = cneg(cdiv(csub(a, b), cadd(a, b)));
  if (infWall && slabWidth > 0.0) {
    cx q = cmul(mkc(2 * SBR_PI * slabWidth / (SBR_C0 / freq), 0.0), csqrt_(csub(et, cmul(si, si))));
    cx e1 = cexp_(cmul(mkc(0.0, -1.0), q)), e2 = cmul(e1, e1);
    cx g2[2] = {gtm, gte};
    cx outR[2], outT[2];
    for (int k = 0; k < 2; k++) {
      cx g = g2[k];
      cx den = csub(mkc(1.0, 0.0), cmul(cmul(g, g), e2));
      outR[k] = cdiv(cmul(g, csub(mkc(1.0, 0.0), e2)), den);
      outT[k] = cdiv(cmul(csub(mkc(1.0, 0.0), cmul(g, g)), e1), den);
    }
    if (refl) { *fTM = outR[0]; *fTE = outR[1]; }
    else { *fTM = outT[0]; *fTE = outT[1]; }
    return;
  }
  if (infWall) {
    if (refl) { *fTM = gtm; *fTE = gte; }
    else { *fTM = mkc(0, 0); *fTE = mkc(0, 0); }
    return;
  }
  // Slab of width 0 (the reference passes width = 0): q = 0, phase factor 1.
  const double lamda = SBR_C0 / freq;
  const double width = 0.0;
  cx q = cmul(mkc(2 * SBR_PI * width / lamda, 0.0), csqrt_(csub(nt, cmul(si, si))));
  cx ph = cexp_(cmul(mkc(0.0, -2.0), q));
  cx g2[2] = {gtm, gte};
  cx outR[2], outT[2];
  for (int k = 0; k < 2; k++) {
    cx g = g2[k];
    cx den = csub(mkc(1.0, 0.0), cmul(cmul(g, g), ph));
    outR[k] = cdiv(cmul(g, csub(mkc(1.0, 0.0), ph)), den);
    outT[k] = cdiv(cmul(csub(mkc(1.0, 0.0), cmul(g, g)), ph), den);
  }
  if (refl) { *fTM = outR[0]; *fTE = outR[1]; }
  else { *fTM = outT[0]; *fTE = outT[1]; }
}

}  // namespace sbr
