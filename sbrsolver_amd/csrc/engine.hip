// gfx950 engine of the shoot-bounce-reflect trace-and-field step.
//
// Pipeline of one trace (RTSolver::ExecuteRayTracing, RTSolver.cpp:52-126), for a batch of one or
// more transmitters sharing the scene and the receivers:
//   1. k_ico_flag / k_ico_split / k_ico_dirs: icosahedral launch directions per subdivision level,
//      in the vertex order of Icosahedron.cpp:3-110 (generated once per level and kept).
//   2. k_trace<EXACT, LDS>: the interaction trees of every launch (RayLaunch, RTSolver.cpp:128-257),
//      one persistent launch per tree depth.  Each launch reads its item count from device memory and
//      appends the next depth's items with wave-aggregated atomics, so the bounce loop has no host
//      round trip.  The BVH nodes and pre-test boxes of small scenes are staged in LDS; the traversal stack always lives in LDS.
//   3. k_los: direct-path line of sight per (transmitter, receiver) (DirectPathCorrection,
//      RTSolver.cpp:375-394).
//   4. k_capture: every tree node against the receivers binned in a uniform xy grid
//      (HitReceptionSphere, RTSolver.cpp:805-863); a capture counts when no ancestor captures
//      (RayCapture + Paths, RTSolver.cpp:259-297, Ray.cpp:136-182); image correction in FP64
//      (RTSolver.cpp:299-433); duplicate removal keeps the minimum order key per (transmitter,
//      receiver, coplanar-id sequence) in a lock-free hash table (RemoveDuplicatePath,
//      RTSolver.cpp:435-459).  The order key (root, depth, branch code) is the reference's
//      (launch, BFS position) order.
//   5. k_collect, a radix sort by order key, k_resolve (order key -> tree node), a stable radix sort
//      by (transmitter, receiver), k_field and k_rx_sum: the field of every kept path
//      (FieldForPath, FieldCompute.cpp:71-151) summed per receiver in the reference's order
//      (FieldAtReceiver, FieldCompute.cpp:11-25).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>

#include "dedup_keys.h"
#include "device_common.h"
#include "engine.h"

namespace sbr {

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail(SBR_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
void DBuf<T>::ensure(size_t want) {
  if (want <= n && p) return;
  release();
  size_t cap = want + want / 4 + 64;
  if (hipMalloc((void**)&p, cap * sizeof(T)) != hipSuccess) {
    p = nullptr;
    n = 0;
    (void)hipGetLastError();
    return;
  }
  n = cap;
}
template <typename T>
void DBuf<T>::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  n = 0;
}

bool HostPinned::ensure(size_t bytes) {
  if (bytes <= n && p) return true;
  release();
  const size_t cap = bytes + bytes / 4 + 4096;
  if (hipHostMalloc((void**)&p, cap, hipHostMallocDefault) != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return false;
  }
  n = cap;
  return true;
}
void HostPinned::release() {
  if (p) (void)hipHostFree(p);
  p = nullptr;
  n = 0;
}
static_assert(sizeof(LevelRec) * (kMaxDepth + 2) <= 1024, "level records fit their staging slot");

template struct DBuf<double>;
template struct DBuf<float>;
template struct DBuf<int32_t>;
template struct DBuf<uint8_t>;
template struct DBuf<unsigned long long>;
template struct DBuf<int64_t>;
template struct DBuf<DNode>;
template struct DBuf<DTri>;
template struct DBuf<DBox32>;
template struct DBuf<DSub>;
template struct DBuf<DNode4>;
template struct DBuf<DLeaf>;
template struct DBuf<DVer>;
template struct DBuf<HeavyJob>;
template struct DBuf<DNodeF>;
template struct DBuf<int2>;
template struct DBuf<DSurf>;
template struct DBuf<TNode>;
template struct DBuf<TxDesc>;
template struct DBuf<LevelRec>;
template struct DBuf<RxCell>;
template struct DBuf<unsigned int>;

#define ENSURE(buf, cnt)                                                                         \
  do {                                                                                           \
    (buf).ensure((size_t)(cnt));                                                                 \
    if (!(buf).p) return fail(SBR_E_NOMEM, std::string("device allocation failed: ") + #buf);  \
  } while (0)

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

// =============================================================================================
// 1. Icosahedral launch directions (Icosahedron.cpp:3-110)
// =============================================================================================
// Each edge of the closed mesh belongs to two faces.  Subdivide() appends a new vertex the first
// time an edge is met in (face, edge slot) order, i.e. in the face with the smaller index.
__global__ void k_ico_flag(const int32_t* __restrict__ adj, int32_t nf, int32_t* __restrict__ flag) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= 3 * (int64_t)nf) return;
  int32_t f = (int32_t)(s / 3);
  flag[s] = f < adj[2 * s] ? 1 : 0;
}

// Faces (v0, v1, v2) get midpoints m0 (v0v1), m1 (v1v2), m2 (v2v0) and the children
// (v0,m0,m2), (v1,m1,m0), (v2,m2,m1), (m0,m1,m2) at 4f..4f+3 (Icosahedron.cpp:23-41).
// adj[2*(3f+e)] = neighbour face across edge e, adj[2*(3f+e)+1] = that edge's slot in it.
__global__ void k_ico_split(const int32_t* __restrict__ faces, const int32_t* __restrict__ adj,
                            const int32_t* __restrict__ flag, const int32_t* __restrict__ pos, int32_t nf,
                            int64_t nv, double* __restrict__ verts, int32_t* __restrict__ faces2,
                            int32_t* __restrict__ adj2, int writeChildren) {
  int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int32_t v[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
  int32_t m[3];
#pragma unroll
  for (int e = 0; e < 3; e++) {
    int64_t s = 3 * (int64_t)f + e;
    if (flag[s]) {
      m[e] = (int32_t)(nv + pos[s]);
      int a = v[e], b = v[(e + 1) % 3];
      V3 p = vnormalized(vadd(ld3(verts + 3 * (int64_t)a), ld3(verts + 3 * (int64_t)b)));
      double* o = verts + 3 * (int64_t)m[e];
      o[0] = p.x; o[1] = p.y; o[2] = p.z;
    } else {
      int64_t t = 3 * (int64_t)adj[2 * s] + adj[2 * s + 1];
      m[e] = (int32_t)(nv + pos[t]);
    }
  }
  if (!writeChildren) return;
  int32_t c[4][3] = {{v[0], m[0], m[2]}, {v[1], m[1], m[0]}, {v[2], m[2], m[1]}, {m[0], m[1], m[2]}};
  for (int k = 0; k < 4; k++)
    for (int j = 0; j < 3; j++) faces2[3 * (4 * (int64_t)f + k) + j] = c[k][j];
  // Half edges: child x, slot 0 lies on parent edge x and touches v_x; slot 2 lies on parent
  // edge (x+2)%3 and touches v_x.  Across parent edge E the neighbour g holds the matching half:
  // its child at the slot of the shared vertex.
  auto half = [&](int E, int w, int32_t* outF, int32_t* outS) {
    int64_t s = 3 * (int64_t)f + E;
    int32_t g = adj[2 * s], es = adj[2 * s + 1];
    if (faces[3 * g + es] == w) { *outF = 4 * g + es; *outS = 0; }
    else { *outF = 4 * g + (es + 1) % 3; *outS = 2; }
  };
  for (int x = 0; x < 3; x++) {
    int64_t base = 2 * 3 * (4 * (int64_t)f + x);
    int32_t nf0, ns0, nf2, ns2;
    half(x, v[x], &nf0, &ns0);
    half((x + 2) % 3, v[x], &nf2, &ns2);
    adj2[base + 0] = nf0; adj2[base + 1] = ns0;
    adj2[base + 2] = 4 * f + 3; adj2[base + 3] = (x + 2) % 3;   // interior edge, centre child
    adj2[base + 4] = nf2; adj2[base + 5] = ns2;
  }
  int64_t base = 2 * 3 * (4 * (int64_t)f + 3);
  adj2[base + 0] = 4 * f + 1; adj2[base + 1] = 1;
  adj2[base + 2] = 4 * f + 2; adj2[base + 3] = 1;
  adj2[base + 4] = 4 * f + 0; adj2[base + 5] = 1;
}

// GenerateRaysOnIcosahedron: directions are the normalized vertices, in vertex order.
__global__ void k_ico_dirs(const double* __restrict__ verts, int64_t nv, double* __restrict__ dirs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  V3 d = vnormalized(ld3(verts + 3 * i));
  dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
}

// =============================================================================================
// 2. Interaction trees, one persistent launch per depth (RayLaunch, RTSolver.cpp:128-257)
// =============================================================================================

// Capture region of a segment S->Tg: a capture at C has its closest point P within
// (base + |C - ref|) * rc / (1 - rc) of C, with base/ref per reception model, so only the part of
// the segment inside the receivers' box grown by that bound R can capture.  Returns false when no
// part is; otherwise the clipped segment P0->P1 and R.
struct SegClip {
  V3 P0, P1;
  double R;
  bool all;
};
__device__ __forceinline__ bool seg_clip(const RxGrid& g, V3 S, V3 Tg, double base, int model, double rc,
                                         SegClip& o) {
  const V3 D = vsub(Tg, S);
  double far2 = 0.0;
  for (int a = 0; a < 3; a++) {
    double sa = model == 0 ? 0.0 : vget(S, a);
    double e = fmax(fabs(g.bmin[a] - sa), fabs(g.bmax[a] - sa));
    far2 += e * e;
  }
  double R = rc < 1.0 ? (base + sqrt(far2)) * rc / (1.0 - rc) : INFINITY;
  // A second bound through the segment itself: the unfolded length is base + |cp| <= base +
  // max(|S|, |Tg|) (model 0, cp on the segment) or base + t0 |Tg - S| <= base + |Tg - S| (model 1),
  // so no capture is farther than (base + that) * rc from the segment.
  const double seg = model == 0 ? fmax(vnorm(S), vnorm(Tg)) : vnorm(D);
  const double Rs = (base + seg) * rc;
  if (Rs < R) R = Rs;
  o.R = R * (1.0 + 1e-6) + 1e-6;
  o.all = !(o.R < 1e300);
  double u0 = 0.0, u1 = 1.0;
  if (!o.all) {
    for (int a = 0; a < 3; a++) {
      double lo = g.bmin[a] - o.R, hi = g.bmax[a] + o.R, s = vget(S, a), d = vget(D, a);
      if (fabs(d) < 1e-300) {
        if (s < lo || s > hi) return false;
      } else {
        // the reciprocal's rounding (a few ulps of u) is far inside the 1e-12 widening below
        const double inv = 1.0 / d;
        double ua = (lo - s) * inv, ub = (hi - s) * inv;
        u0 = fmax(u0, fmin(ua, ub) - 1e-12);
        u1 = fmin(u1, fmax(ua, ub) + 1e-12);
        if (u0 > u1) return false;
      }
    }
  }
  o.P0 = vadd(S, vmul(D, u0));
  o.P1 = vadd(S, vmul(D, u1));
  return true;
}

// Necessary condition for seg_clip to pass, cheap enough for every bounce: the segment's bounding
// box meets the receivers' box grown by an upper bound of seg_clip's radius, with margins above
// seg_clip's rounding and its 1e-12 widening of the segment parameter.  capK = rc / (1 - rc) rounded
// up (INFINITY for rc >= 1).
__device__ __forceinline__ bool may_capture(const RxGrid& g, V3 S, V3 Tg, double base, int model, double rc,
                                            double capK) {
  double far2 = 0.0;
  for (int a = 0; a < 3; a++) {
    const double sa = model == 0 ? 0.0 : vget(S, a);
    const double e = fmax(fabs(g.bmin[a] - sa), fabs(g.bmax[a] - sa));
    far2 += e * e;
  }
  double R = (base + sqrt(far2)) * capK;
  const double seg = model == 0 ? fmax(vnorm(S), vnorm(Tg)) : vnorm(vsub(Tg, S));
  const double Rs = (base + seg) * rc;
  if (Rs < R) R = Rs;
  R = R * (1.0 + 2e-6) + 2e-6;
  if (!(R < 1e300)) return true;
  for (int a = 0; a < 3; a++) {
    const double s = vget(S, a), t = vget(Tg, a);
    const double lo = fmin(s, t), hi = fmax(s, t);
    const double m = 1e-9 * (fabs(lo) + fabs(hi) + fabs(g.bmin[a]) + fabs(g.bmax[a]) + R) + 1e-12;
    if (hi < g.bmin[a] - R - m || lo > g.bmax[a] + R + m) return false;
  }
  return true;
}

constexpr int kTraceOversub = 8;   // bounce-kernel blocks per resident block slot
// Each depth's queue is split into kQParts partitions with an append counter each (kQStride words
// apart, one per 128-byte line): a wave appends its children to partition (wave index mod kQParts).
// A single counter for the whole queue serialises every wave's append on one address.
constexpr int kQParts = 64;
constexpr int kQStride = 16;

struct TraceArgs {
  int32_t depth;
  int32_t nTx;
  int32_t stackDepth;
  int32_t nodeBytes;             // single-precision node records staged in LDS (LDS variant)
  int64_t rootBase;              // global root index of the window's first root item
  float ext;                                // scene extent for the single-precision pre-test
  int32_t capture;               // 1: list the nodes whose segment can capture (receivers' box in g)
  unsigned long long nodeCap, queueCap;
  RxGrid g;
  unsigned long long* qCnt;      // append counters: depth d, partition p at (d * kQParts + p) * kQStride
  unsigned long long partCap;    // entries per partition (queueCap / kQParts)
  int32_t stackLds;              // traversal stack entries per thread in LDS (the rest in gStack)
  int2* gStack;                  // deeper entries: (stackDepth - stackLds) per thread of the grid
  WideTree wide;                 // wide traversal tree (kTravWide)
  int32_t wideNodes, wideLeaves; // its sizes (LDS staging)
};

// Flags: [0] bit 0 node overflow, bit 1 queue overflow; [4] rays redone on the reference tree, [5]
// FP64 box tests, [6] node visits, [7] triangle tests (COUNT variant).  [1] receives the number of
// listed nodes after the trace.
// Items of depth d > 0 are the entries of the kQParts partitions of qIn, in partition order; item i
// is node L.base + i, and writes that index into its parent's child link.
// MODE: kTravWide, kTravExact or kTravRef (closest-hit traversal, sbr_trace_params.exactTraversal).
template <int MODE, bool LDS, bool SUB, bool COUNT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) k_trace(TraceArgs A, const TxDesc* __restrict__ txs,
                                                  const DNode* __restrict__ nodes, const DNodeF* __restrict__ gNodesF,
                                                  const DTri* __restrict__ gTris,
                                                  const DBox32* __restrict__ gBox, const DSub* __restrict__ subs,
                                                  const int32_t* __restrict__ subTri,
                                                  const double* __restrict__ triN, const int32_t* __restrict__ qIn,
                                                  int32_t* __restrict__ qOut, TNode* __restrict__ tree,
                                                  LevelRec* __restrict__ lv, unsigned long long* __restrict__ flags,
                                                  uint8_t* __restrict__ candFlag) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ unsigned long long sPre[kQParts + 1];   // items before each partition (depth > 0)
  // The depth's item count is final (the roots, or the previous launch's appends); blocks beyond it
  // leave at once (the grid is sized for the largest depth and oversubscribes the CUs, so that the
  // dispatcher balances the launch's end).
  const LevelRec L = lv[A.depth];
  const int lane = threadIdx.x & 63;
  unsigned long long n = L.count;
  if (A.depth > 0) {
    static_assert(kQParts <= 64, "one partition count per lane of the first wave");
    if (threadIdx.x < 64) {
      unsigned long long c = 0;
      if (lane < kQParts) c = A.qCnt[((size_t)A.depth * kQParts + lane) * kQStride];
      if (c > A.partCap) c = A.partCap;   // an overflowing append was flagged by the appender
      unsigned long long incl = c;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
      }
      if (lane < kQParts) sPre[lane + 1] = incl;
      if (lane == 0) sPre[0] = 0;
    }
    __syncthreads();
    n = sPre[kQParts];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    lv[A.depth + 1].base = L.base + n;
    if (A.depth > 0) lv[A.depth].count = n;
  }
  if ((unsigned long long)blockIdx.x * blockDim.x >= n) return;
  const DNodeF* nodesF = gNodesF;
  const DTri* tris = gTris;
  const DBox32* tbox = gBox;
  WideTree wide = A.wide;
  size_t sceneB = 0;
  if (LDS) {   // single-precision node records in LDS; FP64 boxes, triangles and pre-test boxes via the caches
    int4* d0 = reinterpret_cast<int4*>(smem);
    if (MODE == kTravWide) {   // wide nodes, then their leaves
      const int4* s0 = reinterpret_cast<const int4*>(A.wide.n4);
      const int4* s1 = reinterpret_cast<const int4*>(A.wide.leaves);
      const int n0 = A.wideNodes * (int)(sizeof(DNode4) / 16), n1 = A.wideLeaves * (int)(sizeof(DLeaf) / 16);
      for (int i = threadIdx.x; i < n0 + n1; i += blockDim.x) d0[i] = i < n0 ? s0[i] : s1[i - n0];
      wide.n4 = reinterpret_cast<const DNode4*>(smem);
      wide.leaves = reinterpret_cast<const DLeaf*>(smem + (size_t)n0 * 16);
    } else {
      const int4* s0 = reinterpret_cast<const int4*>(gNodesF);
      for (int i = threadIdx.x; i < A.nodeBytes / 16; i += blockDim.x) d0[i] = s0[i];
      nodesF = reinterpret_cast<const DNodeF*>(smem);
    }
    sceneB = (size_t)A.nodeBytes;
    __syncthreads();
  }
  TravStack st;
  st.ln = reinterpret_cast<int*>(smem + sceneB) + threadIdx.x;
  st.lt = reinterpret_cast<float*>(smem + sceneB + (size_t)A.stackLds * blockDim.x * 4) + threadIdx.x;
  st.ls = blockDim.x;
  st.lcap = A.stackLds;
  st.gs = (int64_t)gridDim.x * blockDim.x;
  st.g = A.gStack ? A.gStack + (int64_t)blockIdx.x * blockDim.x + threadIdx.x : nullptr;

  // An overflow at an earlier depth (flags[0] != 0) means the trace is redone with more room: this
  // depth's queue may hold unwritten entries, so nothing is read from it.
  if (A.depth > 0 && flags[0] != 0ull) return;
  unsigned long long cntNodes = 0, cntTests = 0, cntExact = 0, cntRedo = 0;
  // this wave's partition of the next depth's queue
  const int part = (int)(((unsigned long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kQParts);
  unsigned long long* partCnt = A.qCnt + ((size_t)(A.depth + 1) * kQParts + part) * kQStride;
  int32_t* partQ = qOut + (size_t)part * A.partCap;
  for (unsigned long long b0 = (unsigned long long)blockIdx.x * blockDim.x; b0 < n;
       b0 += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long i = b0 + threadIdx.x;
    const unsigned long long nodeIdx = L.base + i;
    int mask = 0;
    if (i < n && nodeIdx >= A.nodeCap) atomicOr(&flags[0], 1ull);
    // The record is assembled in registers and written with one 128-byte store after the traversal
    // (partial-line stores before and after it would cost a second write-back of the line); the
    // children fill in its child links later.
    TNode nd;
    const bool live = i < n && nodeIdx < A.nodeCap;
    if (live) {
      double totalPL;
      int allow;
      {
        const TxDesc* T;
        if (A.depth == 0) {
          const unsigned long long gi = (unsigned long long)A.rootBase + i;   // global root index
          int lo = 0, hi = A.nTx - 1;   // transmitter owning root item gi
          while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if ((unsigned long long)txs[mid].launchStart <= gi) lo = mid;
            else hi = mid - 1;
          }
          T = &txs[lo];
          const int64_t launch = T->launchOffset + ((int64_t)gi - T->launchStart) * T->launchStride;
          const V3 d = ld3(T->rays + 3 * launch);
          nd.src[0] = T->tx[0]; nd.src[1] = T->tx[1]; nd.src[2] = T->tx[2];
          nd.dir[0] = d.x; nd.dir[1] = d.y; nd.dir[2] = d.z;
          nd.parent = -1;
          nd.gl = (int32_t)gi;
          nd.code = 0;
          nd.type = 0;
          nd.rc = nd.tc = 0;
          nd.txi = (int16_t)lo;
          nd.base = 0.0;
          totalPL = 0.0;
        } else {
          // item i: partition p with sPre[p] <= i < sPre[p + 1]
          int lo = 0, hi = kQParts;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sPre[mid] <= i) lo = mid;
            else hi = mid;
          }
          const int32_t q = qIn[(size_t)lo * A.partCap + (i - sPre[lo])];
          const int32_t pi = q >> 1;
          const bool refl = q & 1;
          // the parent's link to this node (its index was not known when the parent appended it)
          if (refl) tree[pi].childR = (int32_t)nodeIdx;
          else tree[pi].childT = (int32_t)nodeIdx;
          const TNode& P = tree[pi];
          T = &txs[P.txi];
          const V3 pd = ld3(P.dir);
          // Reflect about the hit triangle's stored normal, then normalize (RTSolver.cpp:186-203).
          const V3 d = refl ? vnormalized(vreflect(pd, ld3(triN + 3 * (int64_t)P.hitDfs))) : pd;
          nd.src[0] = P.tgt[0]; nd.src[1] = P.tgt[1]; nd.src[2] = P.tgt[2];
          nd.dir[0] = d.x; nd.dir[1] = d.y; nd.dir[2] = d.z;
          nd.parent = pi;
          nd.gl = P.gl;
          nd.code = (P.code << 1) | (refl ? 1u : 0u);
          nd.type = refl ? 1 : 2;
          nd.rc = P.rc + (refl ? 1 : 0);
          nd.tc = P.tc + (refl ? 0 : 1);
          nd.txi = P.txi;
          // reference sphere: the totalPathLength RayCapture hands on (RTSolver.cpp:288, 293);
          // alpha * L / sqrt(3): the true length travelled
          nd.base = T->receptionModel == 0 ? P.base + P.pl : P.base + vnorm(vsub(ld3(P.tgt), ld3(P.src)));
          totalPL = P.pl;
        }
        const bool depthOk = T->maxTotal <= 0 || (nd.rc + nd.tc) < T->maxTotal;
        allow = ((nd.tc < T->maxT && depthOk) ? 1 : 0) | ((nd.rc < T->maxR && depthOk) ? 2 : 0);
      }
      nd.depth = (uint8_t)A.depth;
      nd.pad0 = 0;
      const V3 src = ld3(nd.src), dir = ld3(nd.dir);
      Hit h;
      bool hit;
      if (MODE == kTravWide)
        hit = closest_hit_wide<SUB, COUNT>(wide, nodes, gNodesF, tris, tbox, subs, subTri, A.ext, src, dir, h, st);
      else
        hit = closest_hit<MODE == kTravExact, SUB, COUNT>(nodes, nodesF, tris, tbox, subs, subTri, A.ext, src, dir, h, st);
      if (hit) {
        const DTri& t = tris[h.idx];
        const V3 hp = hit_point(src, dir, h);
        nd.tgt[0] = hp.x; nd.tgt[1] = hp.y; nd.tgt[2] = hp.z;
        nd.pl = totalPL + h.dist;
        nd.tri = t.id;
        nd.cop = t.cop;
        nd.hitDfs = h.idx;
        mask = allow;   // transmitted child (bit 0, BFS order: first), reflected child (bit 1)
      } else {
        // A miss leaves a 100 m stub below the root, a 1e10 m segment at the root (RTSolver.cpp:224-256);
        // with longMiss (extension) every miss is 1e10 m.
        const double len = A.depth == 0 || txs[nd.txi].longMiss ? 1E10 : 1E2;
        const V3 t = vadd(src, vmul(dir, len));
        nd.tgt[0] = t.x; nd.tgt[1] = t.y; nd.tgt[2] = t.z;
        nd.pl = INFINITY;
        nd.tri = -1;
        nd.cop = -1;
        nd.hitDfs = -1;
      }
      nd.mask = (uint8_t)mask;
      if (COUNT) {
        cntNodes += h.nodes;
        cntTests += h.tests;
        cntExact += h.exact64;
        cntRedo += h.redone;
      }
    }
    // Wave-aggregated append of the children to the wave's partition of the next depth's queue.
    // The wave's transmitted children go first and its reflected ones after them, so that the next
    // depth's waves hold rays of one kind (transmitted rays continue into the wall, reflected ones
    // leave it).  The children's node indices follow from the partition counts once this depth is
    // done; each child then sets its parent's link.
    const int cT = mask & 1, cR = (mask >> 1) & 1;
    int incl = cT | (cR << 16);   // two 16-bit inclusive counts
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      int v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    const int tot = __shfl(incl, 63, 64);
    const int totT = tot & 0xffff, totR = tot >> 16;
    unsigned long long wbase = 0;
    if (lane == 63 && tot > 0) wbase = atomicAdd(partCnt, (unsigned long long)(totT + totR));
    wbase = __shfl(wbase, 63, 64);
    if (cT + cR > 0) {
      if (wbase + (unsigned long long)(totT + totR) > A.partCap) {
        atomicOr(&flags[0], 2ull);
      } else {
        const int32_t nodeI = (int32_t)nodeIdx;
        if (cT) partQ[wbase + (unsigned long long)((incl & 0xffff) - 1)] = nodeI << 1;
        if (cR) partQ[wbase + (unsigned long long)(totT + (incl >> 16) - 1)] = (nodeI << 1) | 1;
      }
    }
    if (live) {
      nd.childT = -1;   // set by the children
      nd.childR = -1;
      const int4* srcw = reinterpret_cast<const int4*>(&nd);
      int4* dstw = reinterpret_cast<int4*>(&tree[nodeIdx]);
#pragma unroll
      for (int w = 0; w < 8; w++) dstw[w] = srcw[w];
    }
    // Whether the node's segment may reach the receivers' capture region: a flag per node (no
    // atomics here), compacted into the capture pass's list after the last depth.
    if (A.capture && live) {
      const TxDesc& T = txs[nd.txi];
      candFlag[nodeIdx] =
          may_capture(A.g, ld3(nd.src), ld3(nd.tgt), nd.base, T.receptionModel, T.rcoef, T.capK) ? 1 : 0;
    }
  }
  if (COUNT) {
    atomicAdd(&flags[6], cntNodes);
    atomicAdd(&flags[7], cntTests);
    atomicAdd(&flags[5], cntExact);
    atomicAdd(&flags[4], cntRedo);
  }
}

// =============================================================================================
// 3. Direct-path line of sight (DirectPathCorrection, RTSolver.cpp:375-394)
// =============================================================================================
// Closest hit in traversal mode MODE (kTravWide, kTravExact, kTravRef).
template <int MODE>
__device__ __forceinline__ bool closest_hit_mode(const WideTree& w, const DNode* __restrict__ nodes,
                                                 const DNodeF* __restrict__ nodesF, const DTri* __restrict__ tris,
                                                 const DBox32* __restrict__ tbox, const DSub* __restrict__ subs,
                                                 const int32_t* __restrict__ subTri, float ext, V3 o, V3 d, Hit& h,
                                                 const TravStack& st) {
  if (MODE == kTravWide) return closest_hit_wide<true>(w, nodes, nodesF, tris, tbox, subs, subTri, ext, o, d, h, st);
  return closest_hit<MODE == kTravExact, true>(nodes, nodesF, tris, tbox, subs, subTri, ext, o, d, h, st);
}

template <int MODE>
__global__ void __launch_bounds__(kBlock) k_los(const TxDesc* __restrict__ txs, int32_t nTx, WideTree w,
                                                const DNode* __restrict__ nodes, const DNodeF* __restrict__ nodesF,
                                                const DTri* __restrict__ tris,
                                                const DBox32* __restrict__ tbox, const DSub* __restrict__ subs,
                                                const int32_t* __restrict__ subTri, float ext,
                                                const double* __restrict__ rx, int32_t m, int32_t stackDepth,
                                                uint8_t* __restrict__ los) {
  extern __shared__ __align__(16) unsigned char smem[];
  int* stkN = reinterpret_cast<int*>(smem) + threadIdx.x;
  float* stkT = reinterpret_cast<float*>(smem + (size_t)stackDepth * blockDim.x * 4) + threadIdx.x;
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)nTx * m) return;
  const TxDesc& T = txs[k / m];
  V3 tx = v3(T.tx[0], T.tx[1], T.tx[2]);
  V3 d = vsub(ld3(rx + 3 * (k % m)), tx);
  Hit h;
  bool hh = closest_hit_mode<MODE>(w, nodes, nodesF, tris, tbox, subs, subTri, ext, tx, d, h,
                                   lds_stack(stkN, stkT, blockDim.x, stackDepth));
  los[k] = (!hh || vnorm(d) < h.dist) ? 1 : 0;
}

template <int MODE>
__global__ void __launch_bounds__(kBlock) k_closest_batch(WideTree w, const DNode* __restrict__ nodes,
                                                          const DNodeF* __restrict__ nodesF,
                                                          const DTri* __restrict__ tris,
                                                          const DBox32* __restrict__ tbox,
                                                          const DSub* __restrict__ subs,
                                                          const int32_t* __restrict__ subTri, float ext,
                                                          const double* __restrict__ org, const double* __restrict__ dirs,
                                                          int32_t n, int32_t stackDepth, int32_t* __restrict__ triOut,
                                                          double* __restrict__ ptOut, double* __restrict__ distOut) {
  extern __shared__ __align__(16) unsigned char smem[];
  int* stkN = reinterpret_cast<int*>(smem) + threadIdx.x;
  float* stkT = reinterpret_cast<float*>(smem + (size_t)stackDepth * blockDim.x * 4) + threadIdx.x;
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Hit h;
  bool hh = closest_hit_mode<MODE>(w, nodes, nodesF, tris, tbox, subs, subTri, ext, ld3(org + 3 * (int64_t)i),
                                   ld3(dirs + 3 * (int64_t)i), h, lds_stack(stkN, stkT, blockDim.x, stackDepth));
  triOut[i] = hh ? tris[h.idx].id : -1;
  const V3 hp = hit_point(ld3(org + 3 * (int64_t)i), ld3(dirs + 3 * (int64_t)i), h);
  ptOut[3 * i] = hp.x; ptOut[3 * i + 1] = hp.y; ptOut[3 * i + 2] = hp.z;
  distOut[i] = h.dist;
}

// =============================================================================================
// 4. Capture, image correction, duplicate removal
// =============================================================================================
struct CaptureArgs {
  int64_t nodes;   // tree nodes (bound of the listed ones)
  int32_t nrx;
  int32_t cellLds;   // > 0: the cell starts (that many, all below 65536) are staged in LDS as uint16
  RxGrid g;
  // 1 / g.cell.  Cell indices of the search region are taken with it; the region's radius carries
  // a margin (1e-6 relative and absolute) far above the rounding that this adds, so a receiver
  // within the capture radius still falls in a cell of the region.
  double invCell;
  unsigned long long tabMask;
  int32_t fuse;   // 1: the wave-per-node kernel deduplicates its nodes' captures itself
};

__device__ __forceinline__ unsigned long long order_key(const TNode& nd) {
  return ((unsigned long long)(uint32_t)nd.gl << 32) | ((unsigned long long)nd.depth << 27) |
         (unsigned long long)(nd.code & ((1u << 27) - 1));
}

// Per-node state of the capture: the segment and the conservative search region.  No per-node
// arrays: ancestors are reached through parent links.  Only the first (capture) half of the records
// is read for a node that captures nothing; the dedup key of the node's coplanar-id sequence and its
// order key are computed at its first capture, and image sources only for records that need a
// correction.
struct NodeCtx {
  V3 S, Tg, P0, P1;
  SphereF sf;      // single-precision pre-test of its reception sphere
  SphereK sk;
  double base, rc, R;
  double invDy;   // 1 / (P1.y - P0.y), 0 when that is below 1e-300 in magnitude
  int model, depth, txi;
  int32_t node, parent;
  bool all, keyed, emitted;
  bool spilled;    // some of its captures became records (k_dedup reads its key)
  bool imgReady;   // its image chain is in LDS (wave-per-node kernel)
  bool chainOk;    // its image chain is held in LDS (wave-per-node kernel, once keyed)
  int64_t seqId;   // id of the path's coplanar-id sequence in the sequence trie (-1 - txi: empty)
  unsigned long long orderKey;
  int64_t losBase;
  int j0, j1;
  ConeF cone;      // the correction's first-step pre-test (state 0: off)
  V3 I;            // the node's image source
};

struct CaptureBufs {
  const TxDesc* txs;
  const TNode* tree;
  const DSurf* surf;
  const double* rx;
  const int32_t* cellStart;
  const RxCell* cellRx;           // receivers in cell order (cellStart: first of each cell)
  const uint8_t* los;
  unsigned long long* tabA;         // group key (sequence id, receiver), 0: empty
  unsigned long long* seqTab;       // sequence trie: exact (parent id, element) keys, kSeqEmpty: empty
  unsigned long long seqMask;
  unsigned long long* tabOrder;
  int32_t* tabRx;
  // [0] records [1] corrected [2] slots used [3] flags (bit 0 table full, bit 2 record buffer full,
  // bit 3 sequence table full) [4] collected [5] heavy nodes [6] candidate sphere tests (count mode)
  // [7] sequence-table slots used [8] records and [9] corrections of the wave-per-node kernel's own
  // duplicate removal (cleared with each candidate pass)
  unsigned long long* counters;
  HeavyJob* heavy;                // nodes left to the wave-per-node kernel
  // (node, receiver) records of the candidate pass, in kRecParts partitions of recPart slots each
  // (appends spread over that many counters, kRecStride words apart, instead of contending on one)
  int2* rec;
  unsigned long long recPart;
  unsigned long long* recCount;
  const unsigned long long* recPrefix;   // dedup pass: records before each partition
  int countTests;
  unsigned long long* keys;       // per node with records: sequence id, order key (3 words, the third unused)
  const int32_t* cand;            // nodes listed by the bounce kernel
  const unsigned long long* nCand;   // their number
  // scene, for the occlusion test of corrected paths (k_dedup<true> only)
  const DNode* nodes;
  const DNodeF* nodesF;
  const DTri* tris;
  const DBox32* tbox;
  const DSub* subs;
  const int32_t* subTri;
  float ext;
  int32_t stackDepth;
  WideTree wide;                  // and the trace's traversal (kTravWide: the wide tree)
  int32_t mode;
  // image source of each tree node (FP64, the operations of GetImageSources), or null: then the
  // correction's first-step pre-test is off
  const double* img;
  unsigned long long heavyCap;    // entries of `heavy`
};

// The wave-per-node kernel removes duplicates of its own nodes' records, against the node's image
// chain (kFuseDepth interactions at most) held in LDS.  The receivers of a chunk of the node's rows
// are listed in LDS first (kList per chunk), so that each step's receivers are known a step ahead;
// those the node's own sphere captures wait in a ring (kPend) for the ancestor rule, 64 at a time.
constexpr int kFuseDepth = 8;
constexpr int kList = 256;
constexpr int kPend = 128;
static_assert(kPend >= 127, "ring bound: under 64 waiting plus a step's 64");

// The correction's first-step pre-test of a node at depth >= 1, from its image source and the
// surface its segment starts on (that of its parent's hit); off at depth 0 (the line of sight decides
// there) or without image sources.  False when no receiver can pass the correction.
__device__ __forceinline__ bool node_cone(const CaptureArgs& A, const CaptureBufs& B, NodeCtx& c) {
  c.cone.state = 0;
  c.I = v3(0.0, 0.0, 0.0);
  if (B.img == nullptr || c.depth == 0) return true;
  c.I = ld3(B.img + 3 * (int64_t)c.node);
  const DSurf& s = B.surf[B.tree[c.parent].tri];
  c.cone = cone_setup(c.I, ld3(s.v1), ld3(s.e1), ld3(s.e2), cone_reach(c.I, A.g.bmin, A.g.bmax));
  return c.cone.state >= 0;
}

// Rows [j0, j1] narrowed to those that can hold a receiver passing the node's correction pre-test.
// Receivers are binned by floor((y - gy0) / cell); the bounds are padded well above the rounding of
// that division and of the reciprocal used here.
__device__ __forceinline__ void cone_rows(const CaptureArgs& A, NodeCtx& c) {
  if (c.cone.state <= 0) return;
  const double band[6] = {A.g.bmin[0], A.g.bmin[1], A.g.bmin[2], A.g.bmax[0], A.g.bmax[1], A.g.bmax[2]};
  double lo, hi;
  if (!cone_interval(c.cone, c.I, band, 1, &lo, &hi)) {
    c.j1 = c.j0 - 1;
    return;
  }
  const double pad = 1e-6 * A.g.cell + 1e-9 * (fabs(A.g.gy0) + fabs(A.g.bmax[1]));
  c.j0 = max(c.j0, (int)fmax(0.0, floor((lo - pad - A.g.gy0) * A.invCell)));
  c.j1 = min(c.j1, (int)fmin((double)(A.g.ny - 1), floor((hi + pad - A.g.gy0) * A.invCell)));
}

// Image source of every tree node of one depth (GetImageSources, RTSolver.cpp:396-433, one
// interaction at a time): the transmitter at depth 0; below it the parent's image source, mirrored
// across the parent's hit surface when the node is a reflection.  These are the operations
// image_source repeats per path, in the same order, so the values are the same.
__global__ void __launch_bounds__(256) k_images(const TNode* __restrict__ tree, const TxDesc* __restrict__ txs,
                                                const DSurf* __restrict__ surf, int64_t base, int64_t count,
                                                int depth, double* __restrict__ img) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = base + i;
    const TNode& nd = tree[n];
    V3 I;
    if (depth == 0) {
      const TxDesc& T = txs[nd.txi];
      I = v3(T.tx[0], T.tx[1], T.tx[2]);
    } else {
      const int32_t p = nd.parent;
      const V3 Ip = ld3(img + 3 * (int64_t)p);
      I = nd.type == 1 ? image_point(surf[tree[p].tri], Ip) : Ip;
    }
    img[3 * n] = I.x;
    img[3 * n + 1] = I.y;
    img[3 * n + 2] = I.z;
  }
}

// Returns false when no receiver can be captured by the node.  CONE: the correction pre-test is
// built for the node (B.img is set); otherwise it is off and its code left out.
template <bool CONE>
__device__ bool node_setup(const CaptureArgs& A, const CaptureBufs& B, int64_t n, NodeCtx& c) {
  const TNode& nd = B.tree[n];   // capture half only
  const TxDesc& T = B.txs[nd.txi];
  c.node = (int32_t)n;
  c.parent = nd.parent;
  c.depth = nd.depth;
  c.txi = nd.txi;
  c.model = T.receptionModel;
  c.rc = T.rcoef;
  c.base = nd.base;
  c.keyed = false;
  c.emitted = false;
  c.spilled = false;
  c.imgReady = false;
  c.S = ld3(nd.src);
  c.Tg = ld3(nd.tgt);
  c.sk = sphere_k(c.rc, c.model);
  c.sf = sphere_f(c.S, c.Tg, c.base, c.sk);
  SegClip sc;
  if (!seg_clip(A.g, c.S, c.Tg, c.base, c.model, c.rc, sc)) return false;
  c.P0 = sc.P0;
  c.P1 = sc.P1;
  c.R = sc.R;
  c.all = sc.all;
  c.j0 = 0;
  c.j1 = A.g.ny - 1;
  const double dy = c.P1.y - c.P0.y;
  c.invDy = fabs(dy) < 1e-300 ? 0.0 : 1.0 / dy;
  if (!c.all) {
    double ylo = fmin(c.P0.y, c.P1.y) - c.R, yhi = fmax(c.P0.y, c.P1.y) + c.R;
    c.j0 = (int)fmax(0.0, floor((ylo - A.g.gy0) * A.invCell));
    c.j1 = (int)fmin((double)(A.g.ny - 1), floor((yhi - A.g.gy0) * A.invCell));
  }
  c.losBase = (int64_t)c.txi * A.nrx;
  if (CONE) {
    if (!node_cone(A, B, c)) return false;
    cone_rows(A, c);
  } else {
    c.cone.state = 0;
  }
  return c.j0 <= c.j1;
}

// node_setup for a node that the thread-per-node pass found heavy: its results from the job.
__device__ __forceinline__ void node_from_job(const CaptureArgs& A, const CaptureBufs& B, const HeavyJob& J,
                                              NodeCtx& c) {
  c.node = J.node;
  const TNode& nd = B.tree[c.node];
  const TxDesc& T = B.txs[nd.txi];
  c.parent = nd.parent;
  c.depth = nd.depth;
  c.txi = nd.txi;
  c.model = T.receptionModel;
  c.rc = T.rcoef;
  c.base = nd.base;
  c.keyed = false;
  c.emitted = false;
  c.spilled = false;
  c.imgReady = false;
  c.S = ld3(nd.src);
  c.Tg = ld3(nd.tgt);
  c.sk = sphere_k(c.rc, c.model);
  c.sf = J.sf;
  c.P0 = ld3(J.P0);
  c.P1 = ld3(J.P1);
  c.R = J.R;
  c.invDy = J.invDy;
  c.all = (J.flags & 1) != 0;
  c.j0 = J.j0;
  c.j1 = J.j1;
  c.losBase = (int64_t)c.txi * A.nrx;
  (void)node_cone(A, B, c);   // the pass listed the node, so some receiver can pass
}

// Id of the sequence (parent sequence id, element) in the trie: the slot holding that exact key,
// claimed by the first thread to insert it.  kSeqNone if the table is full (flagged; the trace is
// redone with a larger table).
__device__ int64_t seq_intern(const CaptureBufs& B, int64_t parent, uint32_t elem) {
  const unsigned long long key = seq_key(parent, elem);
  unsigned long long slot = mix64(key) & B.seqMask;
  for (unsigned long long probe = 0; probe <= B.seqMask && probe < kMaxProbe; probe++) {
    const unsigned long long cur =
        __hip_atomic_load(&B.seqTab[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == key) return (int64_t)slot;
    if (cur == kSeqEmpty) {
      const unsigned long long prev = atomicCAS(&B.seqTab[slot], kSeqEmpty, key);
      if (prev == kSeqEmpty) {
        atomicAdd(&B.counters[7], 1ull);
        return (int64_t)slot;
      }
      if (prev == key) return (int64_t)slot;
    }
    slot = (slot + 1) & B.seqMask;
  }
  atomicOr(&B.counters[3], 8ull);
  return kSeqNone;
}

// Dedup key of the node's path: the id of its ancestors' coplanar-id sequence (root first, without
// the receiver; with typedDedup each element also carries the type of the interaction that follows
// it), and its order key; computed once, at the node's first capture.
__device__ void node_key(const CaptureBufs& B, NodeCtx& c) {
  int32_t anc[kMaxDepth], typ[kMaxDepth];
  int32_t q = c.parent;
  for (int k = c.depth - 1; k >= 0; k--) { anc[k] = B.tree[q].cop; q = B.tree[q].parent; }
  const bool typed = B.txs[c.txi].typedDedup != 0;
  if (typed) {   // the interaction at depth m is node m's type
    int32_t t = c.node;
    for (int m = c.depth; m >= 1; m--) {
      const TNode& nd = B.tree[t];
      typ[m - 1] = nd.type;
      t = nd.parent;
    }
  }
  int64_t id = seq_root(c.txi);
  for (int k = 0; k < c.depth && id != kSeqNone; k++) id = seq_intern(B, id, seq_elem(anc[k], typed ? typ[k] : 0));
  c.seqId = id;
  c.orderKey = order_key(B.tree[c.node]);
  c.keyed = true;
}

// Grid cells [i0, i1] of row j that can hold a receiver within R of the clipped segment (xy).
__device__ __forceinline__ bool row_range_capsule(const CaptureArgs& A, const NodeCtx& c, int j, int* i0, int* i1) {
  if (c.all) {
    *i0 = 0;
    *i1 = A.g.nx - 1;
    return true;
  }
  double xa, xb;
  const double by0 = A.g.gy0 + j * A.g.cell - c.R, by1 = A.g.gy0 + (j + 1) * A.g.cell + c.R;
  if (c.invDy == 0.0) {
    xa = fmin(c.P0.x, c.P1.x);
    xb = fmax(c.P0.x, c.P1.x);
  } else {
    // segment parameters of the row's (grown) y band; the reciprocal's rounding is far inside the
    // 1e-12 widening
    double va = (by0 - c.P0.y) * c.invDy, vb = (by1 - c.P0.y) * c.invDy;
    double w0 = fmax(0.0, fmin(va, vb) - 1e-12), w1 = fmin(1.0, fmax(va, vb) + 1e-12);
    if (w0 > w1) return false;
    double xa0 = c.P0.x + (c.P1.x - c.P0.x) * w0, xa1 = c.P0.x + (c.P1.x - c.P0.x) * w1;
    xa = fmin(xa0, xa1);
    xb = fmax(xa0, xa1);
  }
  *i0 = (int)fmax(0.0, floor((xa - c.R - A.g.gx0) * A.invCell));
  *i1 = (int)fmin((double)(A.g.nx - 1), floor((xb + c.R - A.g.gx0) * A.invCell));
  return *i0 <= *i1;
}

// ... narrowed to the cells that can hold a receiver passing the node's correction pre-test (the
// row's y band padded as in cone_rows).
__device__ __forceinline__ bool row_range(const CaptureArgs& A, const NodeCtx& c, int j, int* i0, int* i1) {
  if (!row_range_capsule(A, c, j, i0, i1)) return false;
  if (c.cone.state <= 0) return true;
  const double padY = 1e-6 * A.g.cell + 1e-9 * (fabs(A.g.gy0) + fabs(A.g.bmax[1]));
  const double band[6] = {A.g.bmin[0], A.g.gy0 + j * A.g.cell - padY, A.g.bmin[2],
                          A.g.bmax[0], A.g.gy0 + (j + 1) * A.g.cell + padY, A.g.bmax[2]};
  double lo, hi;
  if (!cone_interval(c.cone, c.I, band, 0, &lo, &hi)) return false;
  const double pad = 1e-6 * A.g.cell + 1e-9 * (fabs(A.g.gx0) + fabs(A.g.bmax[0]));
  *i0 = max(*i0, (int)fmax(0.0, floor((lo - pad - A.g.gx0) * A.invCell)));
  *i1 = min(*i1, (int)fmin((double)(A.g.nx - 1), floor((hi + pad - A.g.gx0) * A.invCell)));
  return *i0 <= *i1;
}

// Cells of the region's bounding box (rows j0..j1 by the columns of the clipped segment's x extent
// grown by R): an upper bound of the cells the rows hold, up to rounding at a cell edge.
__device__ __forceinline__ int64_t box_cells(const CaptureArgs& A, const NodeCtx& c) {
  if (c.j0 > c.j1) return 0;
  int i0 = 0, i1 = A.g.nx - 1;
  if (!c.all) {
    i0 = (int)fmax(0.0, floor((fmin(c.P0.x, c.P1.x) - c.R - A.g.gx0) * A.invCell));
    i1 = (int)fmin((double)(A.g.nx - 1), floor((fmax(c.P0.x, c.P1.x) + c.R - A.g.gx0) * A.invCell));
  }
  return i0 > i1 ? 0 : (int64_t)(c.j1 - c.j0 + 1) * (i1 - i0 + 1);
}

// Image source k of the path ending at `leaf` (GetImageSources, RTSolver.cpp:396-433): the
// transmitter mirrored across the surfaces of its reflections among the first k interactions.
__device__ V3 image_source(const CaptureBufs& B, const TxDesc& T, int32_t leaf, int depth, int k, int32_t* surfOut) {
  V3 cur = v3(T.tx[0], T.tx[1], T.tx[2]);
  for (int m = 1; m <= k; m++) {
    // node at depth m - 1 (its hit surface) and node at depth m (its interaction type)
    int32_t b = leaf;
    for (int up = depth; up > m; up--) b = B.tree[b].parent;
    const TNode& nb = B.tree[b];
    const TNode& na = B.tree[nb.parent];
    if (nb.type == 1) cur = image_point(B.surf[na.tri], cur);
    if (m == k) *surfOut = na.tri;
  }
  return cur;
}

// Image correction (MultiPathCorrection / DirectPathCorrection): walk back from the receiver,
// re-intersecting each reflection / transmission surface from its image source.  Out of line: it
// runs for few records and would otherwise hold registers across the candidate loop.
__device__ bool corrected(const CaptureBufs& B, const NodeCtx& c, int32_t r, V3 C) {
  if (c.depth == 0) return B.los[c.losBase + r] != 0;
  const TxDesc& T = B.txs[c.txi];
  V3 target = C;
  for (int k = c.depth; k >= 1; k--) {
    int32_t sid = -1;
    const V3 img = image_source(B, T, c.node, c.depth, k, &sid);
    if (sid < 0) return false;
    const DSurf& s = B.surf[sid];
    V3 p;
    double dist;
    if (!tri_intersect(ld3(s.v1), ld3(s.e1), ld3(s.e2), img, vsub(target, img), &p, &dist)) return false;
    target = p;
  }
  return true;
}

// Appends a (node, receiver) record; called by the lanes that have one (wave-aggregated atomic
// over the active lanes, on the partition of the calling wave).
__device__ __forceinline__ void append_record(const CaptureBufs& B, int32_t node, int32_t r) {
  const unsigned long long active = __ballot(1);
  const int leader = __ffsll((unsigned long long)active) - 1;
  const int lane = (int)(threadIdx.x & 63);
  const int rank = __popcll(active & ((1ull << lane) - 1ull));
  const unsigned part = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kRecParts;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(&B.recCount[part * kRecStride], (unsigned long long)__popcll(active));
  base = __shfl(base, leader, 64);
  const unsigned long long pos = base + (unsigned long long)rank;
  if (pos < B.recPart) B.rec[part * B.recPart + pos] = make_int2(node, r);
  else atomicOr(&B.counters[3], 4ull);
}

// Candidate pass, one receiver against one node: reception sphere and the ancestor rule (a node
// yields a path iff it captures the receiver and no proper ancestor does).  Only the capture halves
// of the records are read.
__device__ __forceinline__ void capture_one(const CaptureBufs& B, NodeCtx& c, const RxCell& x,
                                            unsigned long long& myTests) {
  myTests++;
  const V3 C = ld3(x.p);
  const int32_t r = x.r;
  if (!cone_may_pass(c.cone, c.I, C)) return;   // the correction would fail
  const int pv = sphere_pre(c.sf, c.sk, c.S, C);
  if (pv < 0 || (pv == 0 && !sphere_hit(c.S, c.Tg, C, c.base, c.model, c.rc))) return;
  int32_t q = c.parent;
  while (q >= 0) {
    const TNode& a = B.tree[q];
    if (sphere_hit(ld3(a.src), ld3(a.tgt), C, a.base, c.model, c.rc)) return;
    q = a.parent;
  }
  append_record(B, c.node, r);
  c.emitted = true;
  c.spilled = true;
}

// Candidate tests of the wave-per-node kernel.  The node's own reception sphere first; the receivers
// it captures are then tested against the ancestors (the ancestor rule), read from the wave's copy
// of their AncRecs in LDS (the box first: a receiver outside it is not captured by that ancestor).
// qRest continues the walk in the tree past the copied ones.
__device__ __forceinline__ bool own_captures(const NodeCtx& c, V3 C) {
  const int pv = sphere_pre(c.sf, c.sk, c.S, C);
  return pv > 0 || (pv == 0 && sphere_hit(c.S, c.Tg, C, c.base, c.model, c.rc));
}

__device__ __forceinline__ bool no_ancestor_captures(const CaptureBufs& B, const NodeCtx& c, V3 C, const AncRec* anc,
                                                     int nAnc, int32_t qRest) {
  for (int j = 0; j < nAnc; j++) {
    const AncRec& a = anc[j];
    if (C.x < a.box[0] || C.x > a.box[3] || C.y < a.box[1] || C.y > a.box[4] || C.z < a.box[2] || C.z > a.box[5])
      continue;
    const V3 Sa = v3(a.cap[0], a.cap[1], a.cap[2]);
    const int pa = sphere_pre(a.f, c.sk, Sa, C);
    if (pa > 0 || (pa == 0 && sphere_hit(Sa, v3(a.cap[3], a.cap[4], a.cap[5]), C, a.cap[6], c.model, c.rc)))
      return false;
  }
  for (int32_t q = qRest; q >= 0;) {
    const TNode& a = B.tree[q];
    if (sphere_hit(ld3(a.src), ld3(a.tgt), C, a.base, c.model, c.rc)) return false;
    q = a.parent;
  }
  return true;
}

// Box of the points a segment's reception sphere can reach: the segment's box grown by
// (base + max(|S|, |T|)) rc (model 0: |cp| <= max(|S|, |T|) on the segment) or (base + |T - S|) rc,
// widened by 1e-6 of itself and 1e-9 of the magnitudes involved (far above the FP64 test's rounding).
__device__ __forceinline__ void capture_box(V3 S, V3 T, double base, int model, double rc, double* bx) {
  const double L = vnorm(vsub(T, S));
  const double m = fmax(vnorm(S), vnorm(T));
  const double R = (base + (model == 0 ? m : L)) * rc;
  const double g = R * (1.0 + 1e-6) + 1e-9 * (m + L + base) + 1e-30;
  bx[0] = fmin(S.x, T.x) - g;
  bx[1] = fmin(S.y, T.y) - g;
  bx[2] = fmin(S.z, T.z) - g;
  bx[3] = fmax(S.x, T.x) + g;
  bx[4] = fmax(S.y, T.y) + g;
  bx[5] = fmax(S.z, T.z) + g;
  if (!(g < 1e300)) {   // no bound (infinite or NaN radius): every receiver is in the box
    bx[0] = bx[1] = bx[2] = -INFINITY;
    bx[3] = bx[4] = bx[5] = INFINITY;
  }
}

// After a node's candidates: the first lane of the wave that emitted records for it stores its key
// (all lanes of the wave-per-node kernel share the node; in the thread-per-node kernel each lane
// has its own).
__device__ __forceinline__ void store_key(const CaptureBufs& B, NodeCtx& c) {
  if (!c.keyed) node_key(B, c);
  unsigned long long* k = B.keys + 3 * (int64_t)c.node;
  k[0] = (unsigned long long)c.seqId;
  k[1] = c.orderKey;
}

__device__ bool corrected_points(V3 tx, const TNode* __restrict__ tree, const DSurf* __restrict__ surf,
                                 const int32_t* chain, int depth, V3 C, V3* pts);

// Occlusion test of a corrected multi-bounce path (extension, checkOcclusion): no scene triangle
// nearer than the end of any segment (a hit at the end itself is the next surface or, for the last
// segment, lies beyond the receiver).  Closest hits through the wide tree (traversal mode 0) or the
// clipped reference tree (the other modes; all give the reference's closest hit).
__device__ bool unobstructed(const CaptureBufs& B, const NodeCtx& c, V3 C, int* stkN, float* stkT) {
  int32_t chain[kMaxDepth + 1];
  int32_t q = c.node;
  for (int k = c.depth; k >= 0; k--) { chain[k] = q; q = B.tree[q].parent; }
  const TxDesc& T = B.txs[c.txi];
  V3 pts[kMaxDepth + 2];
  if (!corrected_points(v3(T.tx[0], T.tx[1], T.tx[2]), B.tree, B.surf, chain, c.depth, C, pts)) return false;
  for (int k = 0; k <= c.depth; k++) {
    const V3 d = vsub(pts[k + 1], pts[k]);
    Hit h;
    const TravStack st = lds_stack(stkN, stkT, blockDim.x, B.stackDepth);
    const bool hit =
        B.mode == kTravWide
            ? closest_hit_wide<true>(B.wide, B.nodes, B.nodesF, B.tris, B.tbox, B.subs, B.subTri, B.ext, pts[k], d, h, st)
            : closest_hit<false, true>(B.nodes, B.nodesF, B.tris, B.tbox, B.subs, B.subTri, B.ext, pts[k], d, h, st);
    if (hit && h.dist < vnorm(d) * (1.0 - 1e-7)) return false;
  }
  return true;
}

// The prune: false when the group already holds a corrected path with a smaller order key (keys only
// decrease, so such a record cannot win and needs no correction).  A stale read is only less
// selective.
__device__ __forceinline__ bool group_may_win(const CaptureArgs& A, const CaptureBufs& B, unsigned long long gk,
                                              unsigned long long orderKey) {
  unsigned long long slot = mix64(gk) & A.tabMask;
  for (unsigned long long probe = 0; probe <= A.tabMask && probe < kMaxProbe; probe++) {
    const unsigned long long a = B.tabA[slot];
    if (a == 0ull) return true;
    if (a == gk) return B.tabOrder[slot] >= orderKey;
    slot = (slot + 1) & A.tabMask;
  }
  return true;
}

// Inserts a corrected path: the group's slot (claimed by the first insertion, which also records the
// receiver) keeps the smallest order key.  Returns 1 if this call claimed the slot; flags the table
// full when no slot is left within kMaxProbe probes (the trace is then redone with a larger table:
// a nearly full table would otherwise make the probes of one pass run over most of it).
__device__ __forceinline__ unsigned long long group_insert(const CaptureArgs& A, const CaptureBufs& B,
                                                           unsigned long long gk, unsigned long long orderKey,
                                                           int64_t rxOut) {
  unsigned long long slot = mix64(gk) & A.tabMask;
  for (unsigned long long probe = 0; probe <= A.tabMask && probe < kMaxProbe; probe++) {
    const unsigned long long prev = atomicCAS(&B.tabA[slot], 0ull, gk);
    if (prev == 0ull || prev == gk) {
      if (prev == 0ull) B.tabRx[slot] = (int32_t)rxOut;
      atomicMin(&B.tabOrder[slot], orderKey);
      return prev == 0ull ? 1ull : 0ull;
    }
    slot = (slot + 1) & A.tabMask;
  }
  atomicOr(&B.counters[3], 1ull);
  return 0ull;
}

// Records before each partition of the candidate pass's record buffer (one block of kRecParts
// threads): prefix[p] for p < kRecParts, the total at prefix[kRecParts] and the largest partition's
// count, overflowing or not, at prefix[kRecParts + 1] (for resizing).
__global__ void __launch_bounds__(kRecParts) k_rec_prefix(const unsigned long long* __restrict__ count,
                                                          unsigned long long recPart,
                                                          unsigned long long* __restrict__ prefix) {
  __shared__ unsigned long long sw[kRecParts / 64], sm[kRecParts / 64];
  const int p = threadIdx.x, lane = p & 63, w = p >> 6;
  const unsigned long long c = count[(size_t)p * kRecStride];
  const unsigned long long v = c < recPart ? c : recPart;
  unsigned long long incl = v, mx = c;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long u = __shfl_up(incl, off, 64);
    if (lane >= off) incl += u;
    const unsigned long long m2 = __shfl_xor(mx, off, 64);
    mx = m2 > mx ? m2 : mx;
  }
  if (lane == 63) {
    sw[w] = incl;
    sm[w] = mx;
  }
  __syncthreads();
  unsigned long long before = 0, total = 0, big = 0;
  for (int k = 0; k < kRecParts / 64; k++) {
    if (k < w) before += sw[k];
    total += sw[k];
    big = sm[k] > big ? sm[k] : big;
  }
  prefix[p] = before + incl - v;
  if (p == 0) {
    prefix[kRecParts] = total;
    prefix[kRecParts + 1] = big;
  }
}

// Dedup pass, one thread per record: the group key of the node's coplanar-id sequence, the prune
// against the key its group already holds, image correction (and, with OCCL, the occlusion test of
// the transmitters that ask for it), insertion.
template <bool OCCL>
__global__ void __launch_bounds__(256) k_dedup(CaptureArgs A, CaptureBufs B) {
  extern __shared__ __align__(16) unsigned char smemD[];
  const int64_t nRec = (int64_t)B.recPrefix[kRecParts];   // records of the candidate pass
  unsigned long long myCor = 0, myUsed = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nRec; k += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = kRecParts - 1;   // partition holding record k
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (B.recPrefix[mid] <= (unsigned long long)k) lo = mid;
      else hi = mid - 1;
    }
    const int2 rr = B.rec[lo * B.recPart + ((unsigned long long)k - B.recPrefix[lo])];
    const int32_t r = rr.y;
    NodeCtx c;
    const TNode& nd = B.tree[rr.x];
    c.node = rr.x;
    c.parent = nd.parent;
    c.depth = nd.depth;
    c.txi = nd.txi;
    c.losBase = (int64_t)c.txi * A.nrx;
    const unsigned long long* key = B.keys + 3 * (int64_t)c.node;
    c.seqId = (int64_t)key[0];
    c.orderKey = key[1];
    c.keyed = true;
    // the sequence table filled up (flagged, the trace is redone): no key
    const bool keyed = c.seqId != kSeqNone;
    const unsigned long long gk = group_key(c.seqId, r);
    const bool keep = keyed && group_may_win(A, B, gk, c.orderKey);
    bool ok = keep && corrected(B, c, r, ld3(B.rx + 3 * (int64_t)r));
    if (OCCL && ok && c.depth > 0 && B.txs[c.txi].checkOcclusion) {
      int* stkN = reinterpret_cast<int*>(smemD) + threadIdx.x;
      float* stkT = reinterpret_cast<float*>(smemD + (size_t)B.stackDepth * blockDim.x * 4) + threadIdx.x;
      ok = unobstructed(B, c, ld3(B.rx + 3 * (int64_t)r), stkN, stkT);
    }
    if (ok) {
      myCor++;
      myUsed += group_insert(A, B, gk, c.orderKey, c.losBase + r);
    }
  }
  // one atomic per block and counter (appends to one address serialise)
  for (int off = 32; off > 0; off >>= 1) {
    myCor += __shfl_xor(myCor, off, 64);
    myUsed += __shfl_xor(myUsed, off, 64);
  }
  __shared__ unsigned long long sCor[4], sUsed[4];
  if ((threadIdx.x & 63) == 0) {
    sCor[threadIdx.x >> 6] = myCor;
    sUsed[threadIdx.x >> 6] = myUsed;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long c = sCor[0] + sCor[1] + sCor[2] + sCor[3], u = sUsed[0] + sUsed[1] + sUsed[2] + sUsed[3];
    if (c) atomicAdd(&B.counters[1], c);
    if (u) atomicAdd(&B.counters[2], u);
  }
}

__device__ __forceinline__ void flush_tests(const CaptureBufs& B, unsigned long long myTests) {
  if (!B.countTests) return;
  for (int off = 32; off > 0; off >>= 1) myTests += __shfl_xor(myTests, off, 64);
  if ((threadIdx.x & 63) == 0 && myTests) atomicAdd(&B.counters[6], myTests);
}

// Thread per listed node (those whose segment reaches the receivers' capture region).  Nodes whose
// candidate region spans more than heavy_cells(model) grid cells are left to k_capture_heavy, which
// spreads one node over a wave.  The split was measured on the urban workload: 12 cells for
// alpha * L / sqrt(3) (6-24 within noise), 48 for the reference sphere, whose nodes' cells mostly
// fail the correction pre-test per receiver (capture 23.0 ms at 12, 17.3 at 48-64, 20.3 with no
// wave-per-node kernel at all).
__device__ __forceinline__ int heavy_cells(int model) { return model == 0 ? 48 : 12; }

template <bool CONE>
__device__ __forceinline__ void capture_pass(const CaptureArgs& A, const CaptureBufs& B) {
  extern __shared__ uint16_t sCell[];
  if (A.cellLds > 0) {   // cell starts in LDS: one global round trip less per row of cells
    for (int k = threadIdx.x; k < A.cellLds; k += blockDim.x) sCell[k] = (uint16_t)B.cellStart[k];
    __syncthreads();
  }
  auto cellStart = [&](int k) { return A.cellLds > 0 ? (int)sCell[k] : B.cellStart[k]; };
  const int64_t nc = (int64_t)B.nCand[0];   // listed nodes (written after the bounce loop)
  unsigned long long myTests = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += (int64_t)gridDim.x * blockDim.x) {
    const int32_t n = B.cand[k];
    NodeCtx c;
    if (node_setup<CONE>(A, B, n, c)) {
      // The decision is the loop's own exit flag: a `cells > limit` test after the loop was merged by
      // the compiler with the identical in-loop test, whose lane mask only holds the lanes still
      // looping in the last iteration (lanes that left earlier were then taken for light ones).
      // Rows are counted only when the bounding box of the region could hold more than limit cells.
      int cells = 0;
      bool heavy = false;
      const int limit = heavy_cells(c.model);
      const int jEnd = box_cells(A, c) > limit ? c.j1 : c.j0 - 1;
      for (int j = c.j0; j <= jEnd; j++) {
        int i0, i1;
        if (row_range(A, c, j, &i0, &i1)) cells += i1 - i0 + 1;
        if (cells > limit) {
          heavy = true;
          break;
        }
      }
      const unsigned long long h = heavy ? atomicAdd(&B.counters[5], 1ull) : 0ull;
      if (heavy && h >= B.heavyCap) {
        atomicOr(&B.counters[3], 16ull);   // job list full: the capture is redone with a larger one
      } else if (heavy) {
        HeavyJob& J = B.heavy[h];
        for (int a = 0; a < 3; a++) {
          J.P0[a] = vget(c.P0, a);
          J.P1[a] = vget(c.P1, a);
        }
        J.R = c.R;
        J.invDy = c.invDy;
        J.sf = c.sf;
        J.node = (int32_t)n;
        J.j0 = c.j0;
        J.j1 = c.j1;
        int nA = 0;
        int32_t q = c.parent;
        while (q >= 0 && nA < kJobAnc) {
          J.anc[nA++] = q;
          q = B.tree[q].parent;
        }
        J.flags = (c.all ? 1 : 0) | (nA << 1);
        J.qRest = q;
      } else {
        for (int j = c.j0; j <= c.j1; j++) {   // receivers of a row's cells i0..i1 are contiguous
          int i0, i1;
          if (!row_range(A, c, j, &i0, &i1)) continue;
          const int q1 = cellStart(j * A.g.nx + i1 + 1);
          for (int q = cellStart(j * A.g.nx + i0); q < q1; q++) capture_one(B, c, B.cellRx[q], myTests);
        }
        if (c.spilled) store_key(B, c);
      }
    }
  }
  flush_tests(B, myTests);
}

// Without the correction pre-test the pass fits 3 waves per SIMD with no spills (166 VGPRs, 172
// unbounded); with it, the bound would spill, so that variant keeps 2.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_capture(CaptureArgs A,
                                                                                   CaptureBufs B) {
  capture_pass<false>(A, B);
}
__global__ void __launch_bounds__(256) k_capture_cone(CaptureArgs A, CaptureBufs B) { capture_pass<true>(A, B); }

// Hit triangle (low 32 bits), coplanar id (bits 32-61) and interaction type (bits 62-63) of a tree
// node, packed in one word of its AncRec.
__device__ __forceinline__ double anc_ids(const TNode& a) {
  const unsigned long long w = (unsigned long long)(uint32_t)a.tri |
                               ((unsigned long long)(((uint32_t)a.cop & ((1u << 30) - 1)) | ((uint32_t)a.type << 30)) << 32);
  return __longlong_as_double((long long)w);
}
__device__ __forceinline__ int32_t anc_tri(const AncRec& a) { return (int32_t)(uint32_t)__double_as_longlong(a.cap[7]); }
__device__ __forceinline__ int32_t anc_cop(const AncRec& a) {
  return (int32_t)(((unsigned long long)__double_as_longlong(a.cap[7]) >> 32) & ((1u << 30) - 1));
}
__device__ __forceinline__ int anc_type(const AncRec& a) {
  return (int)((unsigned long long)__double_as_longlong(a.cap[7]) >> 62);
}

// node_key's sequence id from the ancestors in LDS (every ancestor has a hit, so its coplanar id
// fits the 30 bits kept): the same elements, interned in the same order.
__device__ int64_t node_seq_lds(const CaptureBufs& B, const NodeCtx& c, const AncRec* anc, int nodeType) {
  const bool typed = B.txs[c.txi].typedDedup != 0;
  int64_t id = seq_root(c.txi);
  for (int k = 0; k < c.depth && id != kSeqNone; k++) {
    // element k: the coplanar id of the ancestor at depth k, typed by the interaction at depth k + 1
    const int t = k + 1 == c.depth ? nodeType : anc_type(anc[c.depth - 2 - k]);
    id = seq_intern(B, id, seq_elem(anc_cop(anc[c.depth - 1 - k]), typed ? t : 0));
  }
  return id;
}

// Image chain of the node for the correction, into the wave's LDS (12 doubles per interaction m =
// 1..depth: image source m, then v1, e1, e2 of the surface of interaction m).  The image sources are
// the prefixes of one pass over the path, the same operations image_source repeats for every k.
// Returns false if an interaction has no surface (no corrected path then).
__device__ bool image_chain_lds(const CaptureBufs& B, const NodeCtx& c, const AncRec* anc, int nodeType, double* img) {
  const TxDesc& T = B.txs[c.txi];
  V3 cur = v3(T.tx[0], T.tx[1], T.tx[2]);
  for (int m = 1; m <= c.depth; m++) {
    const int32_t sid = anc_tri(anc[c.depth - m]);   // hit surface of the node at depth m - 1
    if (sid < 0) return false;
    const DSurf& sf = B.surf[sid];
    const int t = m == c.depth ? nodeType : anc_type(anc[c.depth - 1 - m]);
    if (t == 1) cur = image_point(sf, cur);
    double* o = img + 12 * (m - 1);
    o[0] = cur.x; o[1] = cur.y; o[2] = cur.z;
    for (int a = 0; a < 3; a++) {
      o[3 + a] = sf.v1[a];
      o[6 + a] = sf.e1[a];
      o[9 + a] = sf.e2[a];
    }
  }
  return true;
}

// corrected() over an image chain in LDS: the same re-intersections, from the receiver back.
__device__ __forceinline__ bool corrected_chain(const double* img, int depth, V3 C) {
  V3 target = C;
  for (int k = depth; k >= 1; k--) {
    const double* e = img + 12 * (k - 1);
    const V3 I = v3(e[0], e[1], e[2]);
    V3 p;
    double dist;
    if (!tri_intersect(v3(e[3], e[4], e[5]), v3(e[6], e[7], e[8]), v3(e[9], e[10], e[11]), I, vsub(target, I), &p,
                       &dist))
      return false;
    target = p;
  }
  return true;
}

// The node's image chain, and its dedup key, each taken once, by the first lane, when the wave first
// needs it (called by the whole wave, where its control flow is uniform).
__device__ void node_chain_wave(const CaptureBufs& B, NodeCtx& c, const AncRec* anc, double* img) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int ok = 0;
  if ((threadIdx.x & 63) == 0) ok = c.depth == 0 || image_chain_lds(B, c, anc, B.tree[c.node].type, img) ? 1 : 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  c.chainOk = __shfl(ok, 0, 64) != 0;
  c.imgReady = true;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ void node_seq_wave(const CaptureBufs& B, NodeCtx& c, const AncRec* anc) {
  int64_t seq = 0;
  if ((threadIdx.x & 63) == 0) seq = node_seq_lds(B, c, anc, B.tree[c.node].type);
  c.seqId = __shfl(seq, 0, 64);
  c.orderKey = order_key(B.tree[c.node]);
  c.keyed = true;
}

// The image correction of a candidate (node c, receiver r at C) in the wave-per-node kernel, against
// the image chain in LDS (the line of sight for a node at depth 0).
__device__ __forceinline__ bool corrected_fast(const CaptureBufs& B, const NodeCtx& c, V3 C, int32_t r,
                                               const double* img) {
  return c.depth == 0 ? B.los[c.losBase + r] != 0 : (c.chainOk && corrected_chain(img, c.depth, C));
}

// Insertion of a corrected record that passed the ancestor rule: the prune, then the group table.
__device__ __forceinline__ void dedup_insert(const CaptureArgs& A, const CaptureBufs& B, const NodeCtx& c, int32_t r,
                                             unsigned long long& myCor, unsigned long long& myUsed) {
  if (c.seqId == kSeqNone) return;   // a full sequence table is flagged and the trace redone
  const unsigned long long gk = group_key(c.seqId, r);
  if (!group_may_win(A, B, gk, c.orderKey)) return;
  myCor++;
  myUsed += group_insert(A, B, gk, c.orderKey, c.losBase + r);
}

// AncRec of tree node a (transmitter T).
__device__ __forceinline__ void anc_rec(const TNode& a, const TxDesc& T, AncRec& r) {
  r.cap[0] = a.src[0]; r.cap[1] = a.src[1]; r.cap[2] = a.src[2];
  r.cap[3] = a.tgt[0]; r.cap[4] = a.tgt[1]; r.cap[5] = a.tgt[2];
  r.cap[6] = a.base;
  r.cap[7] = anc_ids(a);
  r.f = sphere_f(ld3(a.src), ld3(a.tgt), a.base, sphere_k(T.rcoef, T.receptionModel));
  r.pad = 0.0f;
  capture_box(ld3(a.src), ld3(a.tgt), a.base, T.receptionModel, T.rcoef, r.box);
}

// The same value on every lane, as the compiler can see (kept in scalar registers).
__device__ __forceinline__ float wave_uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double wave_uniform(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ V3 wave_uniform(V3 v) { return v3(wave_uniform(v.x), wave_uniform(v.y), wave_uniform(v.z)); }

// A node's state is the same on every lane of the wave-per-node kernel; saying so keeps it out of
// the vector registers.
__device__ __forceinline__ void ctx_uniform(NodeCtx& c) {
  c.S = wave_uniform(c.S);
  c.Tg = wave_uniform(c.Tg);
  c.P0 = wave_uniform(c.P0);
  c.P1 = wave_uniform(c.P1);
  c.base = wave_uniform(c.base);
  c.rc = wave_uniform(c.rc);
  c.R = wave_uniform(c.R);
  c.invDy = wave_uniform(c.invDy);
  c.sf.dx = wave_uniform(c.sf.dx);
  c.sf.dy = wave_uniform(c.sf.dy);
  c.sf.dz = wave_uniform(c.sf.dz);
  c.sf.l2inv = wave_uniform(c.sf.l2inv);
  c.sf.linv = wave_uniform(c.sf.linv);
  c.sf.sx = wave_uniform(c.sf.sx);
  c.sf.sy = wave_uniform(c.sf.sy);
  c.sf.sz = wave_uniform(c.sf.sz);
  c.sf.base = wave_uniform(c.sf.base);
  c.sf.mT = wave_uniform(c.sf.mT);
  c.sf.m0 = wave_uniform(c.sf.m0);
  c.sk.rc = wave_uniform(c.sk.rc);
  c.sk.mX = wave_uniform(c.sk.mX);
  c.sk.model = wave_uniform(c.sk.model);
  c.model = wave_uniform(c.model);
  c.depth = wave_uniform(c.depth);
  c.j0 = wave_uniform(c.j0);
  c.j1 = wave_uniform(c.j1);
  for (int k = 0; k < 4; k++) {
    for (int a = 0; a < 3; a++) c.cone.n[k][a] = wave_uniform(c.cone.n[k][a]);
    c.cone.m[k] = wave_uniform(c.cone.m[k]);
  }
  c.cone.state = wave_uniform(c.cone.state);
  c.I = wave_uniform(c.I);
}

// Wave per heavy node.  The node's rows are taken 64 at a time, one row's cell range per lane; the
// receivers of a row's range are contiguous in cell order.  A wave scan of their counts places them
// in a list in LDS (kList at a time), from which they are dealt to the lanes one each, a step's
// receivers loaded while the previous step's are tested (neighbouring lanes read neighbouring
// receivers), so a node costs about receivers / 64 steps whatever the shape of its region.  The
// rows and cells are those that can hold a receiver passing the correction's first-step pre-test
// (node_cone), and each receiver takes that pre-test before its sphere test.  The receivers the
// node's own sphere captures gather in a ring in LDS and are taken 64 at a time.  A
// path is kept for a receiver iff the node captures it, no ancestor does, its image correction
// succeeds and it has the smallest order key of its group; the first three are independent tests of
// the pair, so here the correction (against the image chain in LDS, no memory traffic) comes before
// the ancestor rule, which then runs only for the few pairs it passes.  With records for k_dedup
// (occlusion tests, deep nodes) the ancestor rule comes first and the records are appended.
// 3 waves per SIMD (168 VGPRs, a few spilled): reference-sphere capture 17.3 -> 16.4 ms per step and the
// 256-transmitter batch 86.8 -> 85.1 ms against 4 waves (100 VGPRs spilled); 2 and 5 no better.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_capture_heavy(CaptureArgs A,
                                                                                   CaptureBufs B) {
  // appended by k_capture (beyond the list's capacity only counted: the capture is then redone)
  const int64_t nHeavy = (int64_t)std::min<unsigned long long>(B.counters[5], B.heavyCap);
  const int lane = threadIdx.x & 63;
  const unsigned long long lanesBelow = (1ull << lane) - 1ull;
  const int64_t wave = wave_uniform((int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int64_t nWaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long myTests = 0;
  // per wave: the current node's nearest ancestors (kFuseDepth at most) and their node indices, its
  // image chain, the ring (positions and receiver indices) and the list
  __shared__ AncRec sAnc[4][kFuseDepth];
  __shared__ double sImg[4][12 * kFuseDepth];
  __shared__ double sRingC[4][3 * kPend];
  __shared__ int32_t sRingR[4][kPend];
  __shared__ int32_t sList[4][kList];
  AncRec* anc = sAnc[threadIdx.x >> 6];
  double* img = sImg[threadIdx.x >> 6];
  double* ringC = sRingC[threadIdx.x >> 6];
  int32_t* ringR = sRingR[threadIdx.x >> 6];
  int32_t* list = sList[threadIdx.x >> 6];
  unsigned long long myRec = 0, myCor = 0, myUsed = 0;
  for (int64_t h = wave; h < nHeavy; h += nWaves) {   // wave-uniform
    NodeCtx c;
    node_from_job(A, B, B.heavy[h], c);
    ctx_uniform(c);
    // The nearest ancestors (kFuseDepth at most, listed in the job): lane k writes the AncRec of
    // ancestor k to the wave's LDS.  A wave's LDS accesses complete in order, and the fences and wave
    // barriers keep the compiler from moving the other lanes' reads above the writes they follow, or
    // the next node's writes above them.
    static_assert(kFuseDepth == kJobAnc, "the job lists the ancestors held in LDS");
    const bool fuse = A.fuse != 0 && c.depth <= kFuseDepth;   // the same on every lane
    const HeavyJob& J = B.heavy[h];
    const int nAnc = wave_uniform(J.flags >> 1);
    int32_t qRest = J.qRest;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < nAnc) anc_rec(B.tree[J.anc[lane]], B.txs[c.txi], anc[lane]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    qRest = wave_uniform(qRest);
    int pHead = 0, pTail = 0;  // ring: receivers the node captures, waiting (the same on every lane)
    // the n (<= 64) receivers at the head of the ring
    auto take = [&](int n) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      V3 C = v3(0.0, 0.0, 0.0);
      int32_t r = 0;
      if (lane < n) {
        const int e = (pHead + lane) & (kPend - 1);
        C = v3(ringC[3 * e], ringC[3 * e + 1], ringC[3 * e + 2]);
        r = ringR[e];
      }
      pHead += n;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (fuse) {
        if (!c.imgReady) node_chain_wave(B, c, anc, img);
        bool ok = lane < n && corrected_fast(B, c, C, r, img);
        if (__ballot(ok) != 0ull) {
          ok = ok && no_ancestor_captures(B, c, C, anc, nAnc, qRest);
          const unsigned long long m = __ballot(ok);
          if (m != 0ull) {
            if (!c.keyed) node_seq_wave(B, c, anc);
            if (lane == 0) myRec += (unsigned long long)__popcll(m);
            if (ok) dedup_insert(A, B, c, r, myCor, myUsed);
          }
        }
      } else if (lane < n && no_ancestor_captures(B, c, C, anc, nAnc, qRest)) {
        append_record(B, c.node, r);
        c.spilled = true;
      }
    };
    for (int jb = c.j0; jb <= c.j1; jb += 64) {
      // lane: row jb + lane, whose cells i0..i1 hold the receivers qa..qa+cnt-1 (contiguous)
      const int j = jb + lane;
      int i0 = 0, i1 = -1, qa = 0, cnt = 0;
      if (j <= c.j1 && row_range(A, c, j, &i0, &i1)) {
        qa = B.cellStart[j * A.g.nx + i0];
        cnt = B.cellStart[j * A.g.nx + i1 + 1] - qa;
      }
      int incl = cnt;   // inclusive scan of the rows' receiver counts over the lanes (all lanes active)
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
      }
      const int excl = incl - cnt;
      const int total = wave_uniform(__shfl(incl, 63, 64));
      for (int cb = 0; cb < total; cb += kList) {   // wave-uniform
        const int nCh = total - cb < kList ? total - cb : kList;
        // the chunk's list: row j's receivers k with excl + k in [cb, cb + nCh)
        {
          const int k0 = cb - excl > 0 ? cb - excl : 0;
          const int k1 = cb + nCh - excl < cnt ? cb + nCh - excl : cnt;
          for (int k = k0; k < k1; k++) list[excl + k - cb] = qa + k;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        V3 C = v3(0.0, 0.0, 0.0);
        int32_t r = 0;
        if (lane < nCh) {
          const RxCell& x = B.cellRx[list[lane]];
          C = ld3(x.p);
          r = x.r;
        }
        for (int base = 0; base < nCh; base += 64) {   // wave-uniform trip count
          // next step's receiver, loaded while this step's is tested
          const int fN = base + 64 + lane;
          V3 CN = v3(0.0, 0.0, 0.0);
          int32_t rN = 0;
          if (fN < nCh) {
            const RxCell& x = B.cellRx[list[fN]];
            CN = ld3(x.p);
            rN = x.r;
          }
          const bool have = base + lane < nCh;
          myTests += have ? 1 : 0;
          const bool hit = have && cone_may_pass(c.cone, c.I, C) && own_captures(c, C);
          const unsigned long long m = __ballot(hit);
          if (hit) {
            const int e = (pTail + __popcll(m & lanesBelow)) & (kPend - 1);
            ringC[3 * e] = C.x;
            ringC[3 * e + 1] = C.y;
            ringC[3 * e + 2] = C.z;
            ringR[e] = r;
          }
          pTail += __popcll(m);
          if (pTail - pHead >= 64) take(64);
          C = CN;
          r = rN;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    if (pTail > pHead) take(pTail - pHead);
    const unsigned long long sp = __ballot(c.spilled);
    if (sp && lane == __ffsll(sp) - 1) store_key(B, c);
  }
  flush_tests(B, myTests);
  for (int off = 32; off > 0; off >>= 1) {
    myRec += __shfl_xor(myRec, off, 64);
    myCor += __shfl_xor(myCor, off, 64);
    myUsed += __shfl_xor(myUsed, off, 64);
  }
  if (lane == 0) {
    if (myRec) atomicAdd(&B.counters[8], myRec);
    if (myCor) atomicAdd(&B.counters[9], myCor);
    if (myUsed) atomicAdd(&B.counters[2], myUsed);
  }
}

// Groups whose smallest order key has its root in [glLo, glHi) (those first seen in the window).
// Indices of the flagged nodes in node order, in two passes over tiles of kFlagTile flags (16 per
// thread): per-tile counts, then (after a scan of the counts) each tile writes its indices.
constexpr int kFlagTile = 4096;

__device__ __forceinline__ int flags16(const uint8_t* __restrict__ flags, int64_t n, int64_t at, uint4* v) {
  if (at + 16 <= n) {
    *v = *reinterpret_cast<const uint4*>(flags + at);
  } else {
    unsigned w[4] = {0u, 0u, 0u, 0u};
    for (int k = 0; k < 16; k++)
      if (at + k < n && flags[at + k]) w[k >> 2] |= 1u << (8 * (k & 3));
    *v = make_uint4(w[0], w[1], w[2], w[3]);
  }
  return __popc(v->x) + __popc(v->y) + __popc(v->z) + __popc(v->w);   // flags are 0 or 1
}

// Sum over the block (256 threads); every thread gets the total and its exclusive prefix.
__device__ __forceinline__ int block_scan256(int c, int* total) {
  __shared__ int waveSum[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = c;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  if (lane == 63) waveSum[wave] = incl;
  __syncthreads();
  int before = 0, all = 0;
  for (int w = 0; w < 4; w++) {
    if (w < wave) before += waveSum[w];
    all += waveSum[w];
  }
  *total = all;
  return before + incl - c;
}

__global__ void __launch_bounds__(256) k_flag_count(const uint8_t* __restrict__ flags, int64_t n,
                                                    int32_t* __restrict__ tileCount) {
  uint4 v;
  const int c = flags16(flags, n, (int64_t)blockIdx.x * kFlagTile + 16 * threadIdx.x, &v);
  int total;
  (void)block_scan256(c, &total);
  if (threadIdx.x == 0) tileCount[blockIdx.x] = total;
}

__global__ void __launch_bounds__(256) k_flag_scatter(const uint8_t* __restrict__ flags, int64_t n,
                                                      const int32_t* __restrict__ tileCount,
                                                      const int32_t* __restrict__ tileOffset, int32_t nTiles,
                                                      int32_t* __restrict__ out, unsigned long long* __restrict__ nOut) {
  const int64_t at = (int64_t)blockIdx.x * kFlagTile + 16 * threadIdx.x;
  uint4 v;
  const int c = flags16(flags, n, at, &v);
  int total;
  int pos = tileOffset[blockIdx.x] + block_scan256(c, &total);
  const unsigned w[4] = {v.x, v.y, v.z, v.w};
  for (int k = 0; k < 16; k++)
    if ((w[k >> 2] >> (8 * (k & 3))) & 1u) out[pos++] = (int32_t)(at + k);
  if (blockIdx.x == nTiles - 1 && threadIdx.x == 0)
    *nOut = (unsigned long long)tileOffset[nTiles - 1] + (unsigned long long)tileCount[nTiles - 1];
}

// Table slots per thread: 16 (a block takes 4096 contiguous slots) on large tables; one on tables
// up to 4 M slots, where 16 would leave most CUs idle and the loop's dependent loads would set the time.
constexpr int kCollectPerThread = 16;
constexpr unsigned long long kCollectSmall = 1ull << 22;

template <int PER>
__global__ void __launch_bounds__(256) k_collect(const unsigned long long* __restrict__ tabA,
                                                 const unsigned long long* __restrict__ tabOrder,
                                                 const int32_t* __restrict__ tabRx, unsigned long long cap,
                                                 unsigned long long glLo, unsigned long long glHi,
                                                 unsigned long long* __restrict__ winKey, int32_t* __restrict__ winRx,
                                                 unsigned long long* __restrict__ counter) {
  // Two passes over the block's tile: count, a block scan and one atomic per block for the base,
  // then write (the slots are read again from the cache).  One append per block rather than per wave:
  // appends to a single counter serialise.  The order of the winners does not matter (they are
  // sorted by key next).
  const unsigned long long tile = (unsigned long long)blockIdx.x * (256ull * PER);
  auto takes = [&](unsigned long long s, unsigned long long* key) {
    if (s >= cap || tabA[s] == 0ull) return false;
    *key = tabOrder[s];
    const unsigned long long gl = *key >> 32;
    return gl >= glLo && gl < glHi;
  };
  int cnt = 0;
  for (int i = 0; i < PER; i++) {
    unsigned long long key;
    if (takes(tile + threadIdx.x + 256ull * i, &key)) cnt++;
  }
  int total = 0;
  const int excl = block_scan256(cnt, &total);
  __shared__ unsigned long long sBase;
  if (threadIdx.x == 0) sBase = total > 0 ? atomicAdd(counter, (unsigned long long)total) : 0ull;
  __syncthreads();
  if (cnt == 0) return;
  unsigned long long w = sBase + (unsigned long long)excl;
  for (int i = 0; i < PER; i++) {
    const unsigned long long sl = tile + threadIdx.x + 256ull * i;
    unsigned long long key;
    if (takes(sl, &key)) {
      winKey[w] = key;
      winRx[w] = tabRx[sl];
      w++;
    }
  }
}

// Tree node of every kept path (order keys, window roots from rootBase): from the root, follow the
// branch code's bits through the child links.
__global__ void k_resolve(const TNode* __restrict__ tree, const unsigned long long* __restrict__ keys, int64_t W,
                          int64_t rootBase, int32_t* __restrict__ winNode) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const unsigned long long key = keys[w];
  const int depth = (int)((key >> 27) & 31ull);
  const uint32_t code = (uint32_t)(key & ((1ull << 27) - 1));
  int32_t n = (int32_t)((int64_t)(key >> 32) - rootBase);
  for (int l = depth - 1; l >= 0; l--) {
    const TNode& t = tree[n];
    n = ((code >> l) & 1u) ? t.childR : t.childT;
  }
  winNode[w] = n;
}

__global__ void k_iota(int32_t* __restrict__ a, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (int32_t)i;
}

// =============================================================================================
// 5. Field of every kept path (FieldForPath, FieldCompute.cpp:71-151) and per-receiver sums
// =============================================================================================
// Corrected vertices of the path of node n to receiver C: pts[0] = tx, pts[1..depth] the
// re-intersections, pts[depth+1] = C.  Returns false if the correction fails.
__device__ bool corrected_points(V3 tx, const TNode* __restrict__ tree, const DSurf* __restrict__ surf,
                                 const int32_t* chain, int depth, V3 C, V3* pts) {
  V3 img[kMaxDepth + 1];
  int32_t isurf[kMaxDepth + 1];
  V3 cur = tx;
  img[0] = cur;
  for (int k = 0; k < depth; k++) {
    const TNode& a = tree[chain[k]];
    const TNode& b = tree[chain[k + 1]];
    if (a.tri < 0) return false;
    if (b.type == 1) cur = image_point(surf[a.tri], cur);
    else if (b.type != 2) return false;
    img[k + 1] = cur;
    isurf[k + 1] = a.tri;
  }
  pts[0] = img[0];
  pts[depth + 1] = C;
  V3 target = C;
  for (int k = depth; k >= 1; k--) {
    const DSurf& s = surf[isurf[k]];
    V3 p;
    double dist;
    if (!tri_intersect(ld3(s.v1), ld3(s.e1), ld3(s.e2), img[k], vsub(target, img[k]), &p, &dist)) return false;
    pts[k] = p;
    target = p;
  }
  return true;
}

// Spherical angles of a local vector (CartesianToSpherical, VecMatDef.h:132-155).
__device__ __forceinline__ void cart_to_sph(V3 v, double* r, double* th, double* ph) {
  *r = vnorm(v);
  if (v.x == 0.0 && v.y == 0.0 && v.z == 0.0) *th = SBR_PI / 2;
  else *th = atan2(sqrt(v.x * v.x + v.y * v.y), v.z);
  double p = atan2(v.y, v.x);
  if (p < 0) p += 2 * SBR_PI;
  *ph = p;
}

struct FieldArgs {
  int64_t winners;
  int32_t nrx;
  // Tabulated transmitter pattern (antennaType 3): patNT x patNP grid points, theta-major, theta
  // from 0 to 180 degrees and phi from 0 to 360 - 360 / patNP degrees; 4 doubles per point, the
  // theta and phi field coefficients j sqrt(60 G) e^{j psi} as (re, im) pairs.
  const double* pat;
  int32_t patNT, patNP;
};

// Field of the tabulated pattern in the transmitter's spherical frame (per unit power): the
// coefficients bilinearly interpolated in (theta, phi) degrees, periodic in phi.
__device__ __forceinline__ void pattern_lookup(const FieldArgs& F, double th, double ph, cx* eth, cx* eph) {
  const double u = th * (180.0 / SBR_PI) * (double)(F.patNT - 1) / 180.0;
  int i0 = (int)floor(u);
  i0 = i0 < 0 ? 0 : (i0 > F.patNT - 2 ? F.patNT - 2 : i0);
  const double a = u - (double)i0;
  const double v = ph * (180.0 / SBR_PI) * (double)F.patNP / 360.0;
  const double fv = floor(v);
  const double b = v - fv;
  const int j0 = (((int)fv % F.patNP) + F.patNP) % F.patNP, j1 = (j0 + 1) % F.patNP;
  const double w00 = (1.0 - a) * (1.0 - b), w01 = (1.0 - a) * b, w10 = a * (1.0 - b), w11 = a * b;
  const double* c00 = F.pat + 4 * ((int64_t)i0 * F.patNP + j0);
  const double* c01 = F.pat + 4 * ((int64_t)i0 * F.patNP + j1);
  const double* c10 = F.pat + 4 * ((int64_t)(i0 + 1) * F.patNP + j0);
  const double* c11 = F.pat + 4 * ((int64_t)(i0 + 1) * F.patNP + j1);
  double o[4];
  for (int q = 0; q < 4; q++) o[q] = w00 * c00[q] + w01 * c01[q] + w10 * c10[q] + w11 * c11[q];
  *eth = mkc(o[0], o[1]);
  *eph = mkc(o[2], o[3]);
}

// One segment of a path as the field transport reads it: its end points, the surface it ends on
// (whose frame is the next one; -1: the receiver), how it left its start surface (0: from the
// transmitter or no interaction, 1 reflected, 2 transmitted) and that surface's material.
struct FieldSeg {
  V3 a, b;
  int32_t surf;
  int32_t type;
  int32_t mat;
};

// FieldForPath (FieldCompute.cpp:71-151) over nseg segments given by seg(k); frames(id) is the
// row-major coordSys of surface id.  Writes the path's field (3 complex) to out.
template <typename SegF, typename FrameF>
__device__ __forceinline__ void path_field(const FieldArgs& F, const TxDesc& K, const double* __restrict__ mats,
                                           int nseg, SegF seg, FrameF frames, double* __restrict__ out) {
  const double lamda = SBR_C0 / K.freq;
  const double kw = (2 * SBR_PI) / lamda;
  const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double* next = K.txcs;
  const double* cur = I3;
  V3C tot = {mkc(0, 0), mkc(0, 0), mkc(0, 0)};
  double L = 0.0;
  V3 prevA = v3(0, 0, 0), prevB = v3(0, 0, 0);
  for (int k = 0; k < nseg; k++) {
    const FieldSeg sg = seg(k);
    V3 vg = vsub(sg.b, sg.a);
    V3 vl = mat_vec(next, vg);
    double r, th, ph;
    cart_to_sph(vl, &r, &th, &ph);
    L += r;
    cur = next;
    next = sg.surf == -1 ? I3 : frames(sg.surf);
    if (K.physicalField && k > 0) {
      // Extension: the global field split at the interaction into the component along s (normal to
      // the plane of incidence) and along p (in it, s x direction before and after), with the
      // Fresnel coefficients of the incidence angle in [0, 90] degrees; the reference's TE
      // coefficient is minus the usual s one (reflection), its TM coefficient the usual p one.
      const V3 din = vnormalized(vsub(prevB, prevA)), dout = vnormalized(vg);
      prevA = sg.a;
      prevB = sg.b;
      const V3 nrm = vnormalized(v3(cur[6], cur[7], cur[8]));
      double cth = fabs(vdot(din, nrm));
      cth = cth > 1.0 ? 1.0 : cth;
      V3 s = vcross(din, nrm);
      if (vnorm(s) < 1e-12) {   // normal incidence: any direction normal to din
        const V3 ax = fabs(din.x) <= fabs(din.y) && fabs(din.x) <= fabs(din.z) ? v3(1, 0, 0)
                      : (fabs(din.y) <= fabs(din.z) ? v3(0, 1, 0) : v3(0, 0, 1));
        s = vcross(din, ax);
      }
      s = vnormalized(s);
      const V3 pin = vcross(s, din), pout = vcross(s, dout);
      cx fTM, fTE;
      fresnel(sg.type == 1, acos(cth), mats[6 * sg.mat + 3], mats[6 * sg.mat + 5], K.freq, K.useFresnel != 0,
              K.wallWidth, &fTM, &fTE);
      const cx cs = sg.type == 1 ? cneg(fTE) : fTE, cp = fTM;
      const cx es = cadd(cadd(cscale(tot.x, s.x), cscale(tot.y, s.y)), cscale(tot.z, s.z));
      const cx ep = cadd(cadd(cscale(tot.x, pin.x), cscale(tot.y, pin.y)), cscale(tot.z, pin.z));
      const cx as = cmul(cs, es), ap = cmul(cp, ep);
      tot.x = cadd(cscale(as, s.x), cscale(ap, pout.x));
      tot.y = cadd(cscale(as, s.y), cscale(ap, pout.y));
      tot.z = cadd(cscale(as, s.z), cscale(ap, pout.z));
      continue;
    }
    prevA = sg.a;
    prevB = sg.b;
    V3C inc;
    if (k == 0) {
      inc.x = mkc(0, 0);
      inc.y = mkc(0, 0);
      inc.z = mkc(0, 0);
      if (K.antennaType == 0) inc.x = mkc(0, sqrt(60 * K.pt));
      else if (K.antennaType == 1) inc.y = mkc(0, sqrt(60 * K.pt * 1.5) * sin(th));
      else if (K.antennaType == 2) inc.y = mkc(0, sqrt(60 * K.pt) * (cos(SBR_PI * cos(th) / 2.0) / sin(th)));
      else if (K.antennaType == 3) {
        cx eth, eph;
        pattern_lookup(F, th, ph, &eth, &eph);
        inc.y = cscale(eth, sqrt(K.pt));
        inc.z = cscale(eph, sqrt(K.pt));
      }
    } else {
      inc = matT_vec_c(cur, tot);
      double st = sin(th), sp = sin(ph), ct = cos(th), cp = cos(ph);
      const double Tm[9] = {st * cp, ct * cp, -1 * sp, st * sp, ct * sp, cp, ct, -1 * st, 0.0};
      inc = matT_vec_c(Tm, inc);
    }
    if (sg.type == 1 || sg.type == 2) {
      double ang = angle_between_rad(vg, v3(cur[6], cur[7], cur[8]));
      cx fTM, fTE;
      fresnel(sg.type == 1, ang, mats[6 * sg.mat + 3], mats[6 * sg.mat + 5], K.freq, K.useFresnel != 0, K.wallWidth,
              &fTM, &fTE);
      tot.x = inc.x;
      tot.y = cmul(inc.y, fTM);
      tot.z = cmul(inc.z, fTE);
    } else {
      tot = inc;
    }
    double st = sin(th), sp = sin(ph), ct = cos(th), cp = cos(ph);
    const double Tm[9] = {st * cp, ct * cp, -1 * sp, st * sp, ct * sp, cp, ct, -1 * st, 0.0};
    tot = mat_vec_c(Tm, tot);
    tot = matT_vec_c(cur, tot);
  }
  cx e = cexp_(mkc(-0.0 * kw * L, -1.0 * kw * L));
  cx f[3] = {cdiv(cmul(tot.x, e), mkc(L, 0)), cdiv(cmul(tot.y, e), mkc(L, 0)), cdiv(cmul(tot.z, e), mkc(L, 0))};
  for (int c = 0; c < 3; c++) { out[2 * c] = f[c].r; out[2 * c + 1] = f[c].i; }
}

// winTxRx / winPerm: kept paths sorted by (transmitter * nrx + receiver), each pointing into the
// order-key-sorted arrays (winNode).
__global__ void __launch_bounds__(256) k_field(FieldArgs F, const TxDesc* __restrict__ txs, const TNode* __restrict__ tree,
                                               const DTri* __restrict__ tris, const DSurf* __restrict__ surf,
                                               const double* __restrict__ mats, const double* __restrict__ rx,
                                               const int32_t* __restrict__ winTxRx, const int32_t* __restrict__ winPerm,
                                               const int32_t* __restrict__ winNode, double* __restrict__ winField) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= F.winners) return;
  const int64_t n = winNode[winPerm[w]];
  const TNode& leaf = tree[n];
  const int depth = leaf.depth;
  const TxDesc& K = txs[leaf.txi];
  int32_t chain[kMaxDepth + 1];
  {
    int64_t c = n;
    for (int k = depth; k >= 0; k--) { chain[k] = (int32_t)c; c = tree[c].parent; }
  }
  const V3 C = ld3(rx + 3 * (int64_t)(winTxRx[w] % F.nrx));
  const V3 txp = v3(K.tx[0], K.tx[1], K.tx[2]);
  V3 pts[kMaxDepth + 2];
  if (depth == 0) {
    pts[0] = txp;
    pts[1] = C;
  } else {
    (void)corrected_points(txp, tree, surf, chain, depth, C, pts);
  }
  // segment k: corrected points k -> k + 1; it ends on the surface its tree node hit (the last one
  // on the receiver); its interaction is its node's type, at the surface the previous node hit
  auto seg = [&](int k) {
    const TNode& sg = tree[chain[k]];
    FieldSeg f;
    f.a = pts[k];
    f.b = pts[k + 1];
    f.surf = k == depth ? -1 : sg.tri;
    f.type = sg.type;
    f.mat = k > 0 ? tris[tree[chain[k - 1]].hitDfs].mat : 0;
    return f;
  };
  auto frames = [&](int32_t id) { return surf[id].cs; };
  path_field(F, K, mats, depth + 1, seg, frames, winField + 6 * w);
}

// Fields of caller-given paths (sbr_path_fields): path p has count[p] segments from off[p]; one
// thread per path.  Segment s: src / tgt (3 doubles each), the surface it ends on, its interaction
// type and material; frames[id * 9 ...] the row-major coordSys of surface id.
__global__ void __launch_bounds__(256) k_field_paths(FieldArgs F, const TxDesc* __restrict__ K,
                                                     const double* __restrict__ mats, int64_t nPaths,
                                                     const int64_t* __restrict__ off, const int32_t* __restrict__ count,
                                                     const double* __restrict__ src, const double* __restrict__ tgt,
                                                     const int32_t* __restrict__ surfId, const int32_t* __restrict__ type,
                                                     const int32_t* __restrict__ mat, const double* __restrict__ frames,
                                                     double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nPaths) return;
  const int64_t o = off[p];
  auto seg = [&](int k) {
    const int64_t q = o + k;
    FieldSeg f;
    f.a = ld3(src + 3 * q);
    f.b = ld3(tgt + 3 * q);
    f.surf = surfId[q];
    f.type = type[q];
    f.mat = mat[q];
    return f;
  };
  auto fr = [&](int32_t id) { return frames + 9 * (int64_t)id; };
  path_field(F, *K, mats, count[p], seg, fr, out + 6 * p);
}

// Reflection and transmission coefficients between two media of complex relative permittivity
// epsI (incident side) and epsT, for the sweeps of RefCoeffTest / TransCoeffTest (FieldCompute.cpp:
// 27-69, 236-302): Snell's angle theta_t = asin(n_i / n_t sin theta_i), Gamma_TM =
// (-eta_t cos t + eta_i cos i) / (eta_t cos t + eta_i cos i), Gamma_TE = -(eta_t / cos t - eta_i /
// cos i) / (eta_t / cos t + eta_i / cos i); an infinite wall transmits nothing, otherwise the
// reference's slab of the given width (phase q = 2 pi w / lambda sqrt(n_t - sin^2 theta_i), factor
// e^{-2jq}).  out: rTM, tTM, rTE, tTE.
__device__ void coeffs_general(cx epsI, cx epsT, double thetaI, double lamda, double width, bool infWall, cx* out) {
  const double ETA0 = 4 * SBR_PI * 1e-7 * SBR_C0;
  const cx ti = mkc(thetaI, 0.0);
  const cx ni = csqrt_(epsI), nt = csqrt_(epsT);
  const cx tt = casin_(cmul(cdiv(ni, nt), csin_(ti)));
  const cx etai = cdiv(mkc(ETA0, 0.0), ni), etat = cdiv(mkc(ETA0, 0.0), nt);
  const cx ci = ccos_(ti), ct = ccos_(tt), si = csin_(ti);
  const cx gtm = cdiv(cadd(cmul(cneg(etat), ct), cmul(etai, ci)), cadd(cmul(etat, ct), cmul(etai, ci)));
  const cx a = cdiv(etat, ct), b = cdiv(etai, ci);
  const cx gte = cneg(cdiv(csub(a, b), cadd(a, b)));
  const cx g[2] = {gtm, gte};
  for (int k = 0; k < 2; k++) {
    if (infWall) {
      out[2 * k] = g[k];
      out[2 * k + 1] = mkc(0, 0);
      continue;
    }
    const cx q = cmul(mkc(2 * SBR_PI * width / lamda, 0.0), csqrt_(csub(nt, cmul(si, si))));
    const cx ph = cexp_(cmul(mkc(0.0, -2.0), q));
    const cx den = csub(mkc(1.0, 0.0), cmul(cmul(g[k], g[k]), ph));
    out[2 * k] = cdiv(cmul(g[k], csub(mkc(1.0, 0.0), ph)), den);
    out[2 * k + 1] = cdiv(cmul(csub(mkc(1.0, 0.0), cmul(g[k], g[k])), ph), den);
  }
}

// Angles 0 .. pi/2 in n even steps (Eigen's LinSpaced(n, 0, pi/2): low + i * step, the last one
// exactly pi/2) and their coefficients: theta[n], out[n][4][2] (rTM, tTM, rTE, tTE).
__global__ void k_coeff_sweep(int32_t n, double lamda, cx epsI, cx epsT, double width, int32_t infWall,
                              double* __restrict__ theta, double* __restrict__ out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hi = SBR_PI / 2;
  const double step = n > 1 ? hi / (double)(n - 1) : 0.0;
  const double th = (n == 1 || i == n - 1) ? hi : (double)i * step;
  cx c[4];
  coeffs_general(epsI, epsT, th, lamda, width, infWall != 0, c);
  theta[i] = th;
  for (int k = 0; k < 4; k++) { out[8 * i + 2 * k] = c[k].r; out[8 * i + 2 * k + 1] = c[k].i; }
}

// Per (transmitter, receiver): running sum of its kept paths' fields in (launch, BFS) order,
// continued from the value the previous launch windows left.
__global__ void k_rx_sum(const int32_t* __restrict__ winTxRx, int64_t W, const double* __restrict__ winField,
                         int64_t nOut, double* __restrict__ field) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nOut) return;
  int64_t lo = 0, hi = W;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (winTxRx[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  if (lo >= W || winTxRx[lo] != r) return;
  double acc[6];
  for (int c = 0; c < 6; c++) acc[c] = field[6 * r + c];
  for (int64_t w = lo; w < W && winTxRx[w] == r; w++)
    for (int c = 0; c < 6; c++) acc[c] += winField[6 * w + c];
  for (int c = 0; c < 6; c++) field[6 * r + c] = acc[c];
}

// Path export for the kept paths (one row per segment).
__global__ void k_paths(int64_t W, int32_t nrx, const TxDesc* __restrict__ txs, const TNode* __restrict__ tree,
                        const DSurf* __restrict__ surf, const double* __restrict__ rx,
                        const int32_t* __restrict__ winTxRx, const int32_t* __restrict__ winPerm,
                        const int32_t* __restrict__ winNode, const int64_t* __restrict__ off, int32_t* __restrict__ stri,
                        int32_t* __restrict__ scop, int32_t* __restrict__ stype, double* __restrict__ ssrc,
                        double* __restrict__ stgt) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const int64_t n = winNode[winPerm[w]];
  const int depth = tree[n].depth;
  const TxDesc& K = txs[tree[n].txi];
  int32_t chain[kMaxDepth + 1];
  int64_t c = n;
  for (int k = depth; k >= 0; k--) { chain[k] = (int32_t)c; c = tree[c].parent; }
  const V3 C = ld3(rx + 3 * (int64_t)(winTxRx[w] % nrx));
  const V3 txp = v3(K.tx[0], K.tx[1], K.tx[2]);
  V3 pts[kMaxDepth + 2];
  if (depth == 0) {
    pts[0] = txp;
    pts[1] = C;
  } else {
    (void)corrected_points(txp, tree, surf, chain, depth, C, pts);
  }
  const int64_t o = off[w];
  for (int k = 0; k <= depth; k++) {
    const TNode& sg = tree[chain[k]];
    stri[o + k] = k == depth ? -1 : sg.tri;
    scop[o + k] = k == depth ? -1 : sg.cop;
    stype[o + k] = sg.type;
    ssrc[3 * (o + k)] = pts[k].x; ssrc[3 * (o + k) + 1] = pts[k].y; ssrc[3 * (o + k) + 2] = pts[k].z;
    stgt[3 * (o + k)] = pts[k + 1].x; stgt[3 * (o + k) + 1] = pts[k + 1].y; stgt[3 * (o + k) + 2] = pts[k + 1].z;
  }
}

// Per kept path of a window (sorted position w): the row of the path store (tx * nrx + rx, launch,
// segment count, field) and the segment count for the offsets' scan.
__global__ void k_path_meta(int64_t W, const TxDesc* __restrict__ txs, int32_t nrx,
                            const unsigned long long* __restrict__ winKey, const int32_t* __restrict__ winPerm,
                            const int32_t* __restrict__ winTxRx, const double* __restrict__ winField,
                            int64_t* __restrict__ segCnt, int32_t* __restrict__ oTxRx, int32_t* __restrict__ oLaunch,
                            int32_t* __restrict__ oNseg, double* __restrict__ oField) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const unsigned long long k = winKey[winPerm[w]];
  const int32_t nseg = (int32_t)((k >> 27) & 31ull) + 1;
  const int64_t gl = (int64_t)(k >> 32);
  const int32_t tr = winTxRx[w];
  const TxDesc& T = txs[tr / nrx];
  oTxRx[w] = tr;
  oLaunch[w] = (int32_t)(T.launchOffset + (gl - T.launchStart) * T.launchStride);
  oNseg[w] = nseg;
  segCnt[w] = nseg;
  for (int c = 0; c < 6; c++) oField[6 * w + c] = winField[6 * w + c];
}

// Store offsets of a window's paths (segBase + the window's exclusive scan) and, in segOff[W], the
// window's segment total.
__global__ void k_path_off(int64_t W, int64_t segBase, const int64_t* __restrict__ segCnt, int64_t* __restrict__ segOff,
                           int64_t* __restrict__ oOff) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  oOff[w] = segBase + segOff[w];
  if (w == W - 1) segOff[W] = segOff[w] + segCnt[w];
}

// =============================================================================================
// Host orchestration
// =============================================================================================
Engine::~Engine() {
  for (auto& kv : rayCache) delete kv.second;
  for (auto& e : evSide)
    if (e) (void)hipEventDestroy(e);
  if (sideStream) (void)hipStreamDestroy(sideStream);
  if (stream) (void)hipStreamDestroy(stream);
}

static int scratch_address(int dev, unsigned char** out);   // staging array of closest_hit_noalloc

int Engine::init(int dev) {
  if (dev == -1) {   // host-only context: scene preparation and queries, no device work
    device = -1;
    return SBR_OK;
  }
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (dev < 0 || dev >= count) return fail(SBR_E_ARG, "no such HIP device");
  device = dev;
  HIPCHK(hipSetDevice(dev));
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&sideStream, hipStreamNonBlocking));
  for (auto& e : evSide) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  numCU = prop.multiProcessorCount;
  const int rc = scratch_address(dev, &scratchPtr);   // this device is current
  if (rc != SBR_OK) return fail(rc, "hipGetSymbolAddress of the staging array failed");
  return SBR_OK;
}

int Engine::set_materials(const sbr_material* m, int n) {
  if (!m || n <= 0) return fail(SBR_E_ARG, "materials: empty list");
  mats.assign((const double*)m, (const double*)m + 6 * (size_t)n);
  if (device < 0) return SBR_OK;
  HIPCHK(hipSetDevice(device));
  ENSURE(dMats, mats.size());
  HIPCHK(hipMemcpyAsync(dMats.p, mats.data(), mats.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SBR_OK;
}

int Engine::set_pattern(const float* gi, int nt, int np) {
  if (!gi || nt < 2 || np < 1 || (int64_t)nt * np > (int64_t(1) << 24))
    return fail(SBR_E_ARG, "pattern: need n_theta >= 2, n_phi >= 1 and a table of n_theta * n_phi rows");
  std::vector<double> c(4 * (size_t)nt * np);
  for (size_t k = 0; k < (size_t)nt * np; k++) {
    for (int h = 0; h < 2; h++) {   // theta, phi: gain dB, phase degrees -> j sqrt(60 G) e^{j psi}
      const double g = gi[4 * k + 2 * h], psi = gi[4 * k + 2 * h + 1];
      if (!std::isfinite(g) || !std::isfinite(psi)) return fail(SBR_E_ARG, "pattern: values must be finite");
      const double amp = std::sqrt(60.0 * std::pow(10.0, g / 10.0)), ang = psi * (3.141592653589793 / 180.0);
      c[4 * k + 2 * h] = -amp * std::sin(ang);
      c[4 * k + 2 * h + 1] = amp * std::cos(ang);
    }
  }
  hPattern.assign(gi, gi + 4 * (size_t)nt * np);
  patNT = nt;
  patNP = np;
  if (device < 0) return SBR_OK;
  HIPCHK(hipSetDevice(device));
  ENSURE(dPattern, c.size());
  HIPCHK(hipMemcpyAsync(dPattern.p, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SBR_OK;
}

// Base icosahedron (Icosahedron.cpp:50-67) and its edge adjacency.
static void ico_base(std::vector<double>& v, std::vector<int32_t>& f, std::vector<int32_t>& adj) {
  const double X = .525731112119133606, Z = .850650808352039932, N = 0.0;
  const double b[12][3] = {{-X, N, Z}, {X, N, Z}, {-X, N, -Z}, {X, N, -Z}, {N, Z, X}, {N, Z, -X},
                           {N, -Z, X}, {N, -Z, -X}, {Z, X, N}, {-Z, X, N}, {Z, -X, N}, {-Z, -X, N}};
  const int t[20][3] = {{0, 4, 1}, {0, 9, 4}, {9, 5, 4}, {4, 5, 8}, {4, 8, 1}, {8, 10, 1}, {8, 3, 10},
                        {5, 3, 8}, {5, 2, 3}, {2, 7, 3}, {7, 10, 3}, {7, 6, 10}, {7, 11, 6}, {11, 0, 6},
                        {0, 1, 6}, {6, 1, 10}, {9, 0, 11}, {9, 11, 2}, {9, 2, 5}, {7, 2, 11}};
  v.assign(&b[0][0], &b[0][0] + 36);
  f.assign(&t[0][0], &t[0][0] + 60);
  adj.assign(120, -1);
  for (int a = 0; a < 20; a++)
    for (int e = 0; e < 3; e++) {
      int p = t[a][e], q = t[a][(e + 1) % 3];
      for (int g = 0; g < 20; g++) {
        if (g == a) continue;
        for (int e2 = 0; e2 < 3; e2++) {
          int p2 = t[g][e2], q2 = t[g][(e2 + 1) % 3];
          if ((p2 == p && q2 == q) || (p2 == q && q2 == p)) { adj[2 * (3 * a + e)] = g; adj[2 * (3 * a + e) + 1] = e2; }
        }
      }
    }
}

int Engine::generate_rays(int tess, DBuf<double>& out, int64_t* nOut) {
  if (device < 0) return fail(SBR_E_STATE, "host-only context has no device");
  if (tess < 0 || tess > kMaxTess) return fail(SBR_E_ARG, "tessellation must be in [0, " + std::to_string(kMaxTess) + "]");
  const int64_t nvFinal = 10 * (int64_t(1) << (2 * tess)) + 2;
  std::vector<double> v0;
  std::vector<int32_t> f0, a0;
  ico_base(v0, f0, a0);
  ENSURE(dVerts, 3 * (size_t)nvFinal);
  const int64_t nfMax = 20 * (int64_t(1) << (2 * std::max(0, tess - 1)));
  ENSURE(dFaces[0], 3 * nfMax);
  ENSURE(dAdj[0], 6 * nfMax);
  if (tess > 1) {
    ENSURE(dFaces[1], 3 * nfMax);
    ENSURE(dAdj[1], 6 * nfMax);
  }
  ENSURE(dFlag, 3 * nfMax);
  ENSURE(dScan, 3 * nfMax);
  HIPCHK(hipMemcpyAsync(dVerts.p, v0.data(), v0.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dFaces[0].p, f0.data(), f0.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dAdj[0].p, a0.data(), a0.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  int cur = 0;
  int64_t nv = 12, nf = 20;
  for (int l = 0; l < tess; l++) {
    const int64_t ns = 3 * nf;
    k_ico_flag<<<blocks_for(ns, 256), 256, 0, stream>>>(dAdj[cur].p, (int32_t)nf, dFlag.p);
    size_t tmpBytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmpBytes, dFlag.p, dScan.p, (int)ns, stream));
    ENSURE(dScratch, tmpBytes);
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(dScratch.p, tmpBytes, dFlag.p, dScan.p, (int)ns, stream));
    const bool last = l == tess - 1;
    k_ico_split<<<blocks_for(nf, 256), 256, 0, stream>>>(dFaces[cur].p, dAdj[cur].p, dFlag.p, dScan.p,
                                                         (int32_t)nf, nv, dVerts.p, last ? nullptr : dFaces[1 - cur].p,
                                                         last ? nullptr : dAdj[1 - cur].p, last ? 0 : 1);
    HIPCHK(hipGetLastError());
    nv = nv + 3 * nf / 2;   // each of the 3F/2 edges adds one vertex
    nf *= 4;
    cur = 1 - cur;
  }
  if (nv != nvFinal) return fail(SBR_E_INTERNAL, "icosahedron vertex count mismatch");
  ENSURE(out, 3 * nv);
  k_ico_dirs<<<blocks_for(nv, 256), 256, 0, stream>>>(dVerts.p, nv, out.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream));
  *nOut = nv;
  return SBR_OK;
}

int Engine::rays_for(int tess, const double** out, int64_t* nOut) {
  auto it = rayCache.find(tess);
  if (it == rayCache.end()) {
    DBuf<double>* b = new DBuf<double>();
    int64_t n = 0;
    int rc = generate_rays(tess, *b, &n);
    if (rc) {
      delete b;
      return rc;
    }
    it = rayCache.emplace(tess, b).first;
  }
  *out = it->second->p;
  *nOut = 10 * (int64_t(1) << (2 * tess)) + 2;
  return SBR_OK;
}

static size_t stack_bytes(int height, int block) { return (size_t)height * block * 8; }

// Staging area of sbr_closest_hit_noalloc: a fixed array in device memory, one per GPU, held from the
// library's first use on that GPU until the process ends.  Host code never passes the array's name
// to a kernel or a copy (on the host it is a stand-in, not a device address): it uses the address
// hipGetSymbolAddress returns on each device (Engine::scratchPtr, looked up once per device).
// Layout for n rays, each part 256-byte aligned: origins (3 n doubles), directions (3 n doubles),
// hit points (3 n doubles), distances (n doubles), triangle ids (n int32).
static __device__ __align__(256) unsigned char scratch[SBR_NOALLOC_SCRATCH_BYTES];
static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
static size_t noalloc_bytes(size_t n) {
  return 3 * align256(3 * n * sizeof(double)) + align256(n * sizeof(double)) + align256(n * sizeof(int32_t));
}
static_assert(SBR_NOALLOC_MAX_RAYS * 84 + 5 * 256 <= SBR_NOALLOC_SCRATCH_BYTES, "noalloc ray limit fits the array");

static std::mutex gScratchMu;
static unsigned char* gScratchAddr[64];   // per device ordinal; nullptr until looked up

static int scratch_address(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> g(gScratchMu);
  if (dev < 0 || dev >= 64) return SBR_E_ARG;
  if (!gScratchAddr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return SBR_E_HIP;
    }
    gScratchAddr[dev] = static_cast<unsigned char*>(p);
  }
  *out = gScratchAddr[dev];
  return SBR_OK;
}

WideTree Engine::wide_tree() const { return WideTree{dNodes4.p, dLeaves.p, dVer.p}; }

// One closest-hit query per ray on the device (mode: kTravWide, kTravExact, kTravRef).
hipError_t Engine::launch_closest(int mode, const double* o, const double* d, int n, int32_t* tri, double* pt,
                                  double* dist) {
  const size_t sm = stack_bytes(stackDepth(), kBlock);
  auto k = mode == kTravExact ? k_closest_batch<kTravExact>
                              : mode == kTravRef ? k_closest_batch<kTravRef> : k_closest_batch<kTravWide>;
  k<<<blocks_for(n, kBlock), kBlock, sm, stream>>>(wide_tree(), dNodes.p, dNodesF.p, dTris.p, dTriBox32.p, dSub.p,
                                                   dSubTri.p, sceneExtent, o, d, n, stackDepth(), tri, pt, dist);
  return hipGetLastError();
}

int Engine::closest_hit_noalloc(const double* o, const double* d, int n, int exact, int32_t* tri, double* pt,
                                double* dist) {
  if (device < 0) return fail(SBR_E_STATE, "host-only context has no device");
  if (!haveScene) return fail(SBR_E_STATE, "no scene uploaded");
  if (n > SBR_NOALLOC_MAX_RAYS)
    return fail(SBR_E_LIMIT, "closest_hit_noalloc: " + std::to_string(n) + " rays exceed the staging array's " +
                                 std::to_string(SBR_NOALLOC_MAX_RAYS) + " (use sbr_closest_hit or split the call)");
  if (n <= 0) return SBR_OK;
  if (!scratchPtr) return fail(SBR_E_STATE, "closest_hit_noalloc: staging array address not known");
  if (noalloc_bytes((size_t)n) > SBR_NOALLOC_SCRATCH_BYTES) return fail(SBR_E_INTERNAL, "noalloc layout");
  HIPCHK(hipSetDevice(device));
  unsigned char* base = scratchPtr;
  double* sO = reinterpret_cast<double*>(base);
  double* sD = reinterpret_cast<double*>(base + align256(3 * (size_t)n * sizeof(double)));
  double* sP = reinterpret_cast<double*>(base + 2 * align256(3 * (size_t)n * sizeof(double)));
  double* sDist = reinterpret_cast<double*>(base + 3 * align256(3 * (size_t)n * sizeof(double)));
  int32_t* sT = reinterpret_cast<int32_t*>(base + 3 * align256(3 * (size_t)n * sizeof(double)) +
                                           align256((size_t)n * sizeof(double)));
  // Every output entry of the n rays is written by the kernel, so nothing needs clearing: what the
  // previous call left in the array is never read.
  HIPCHK(hipMemcpyAsync(sO, o, 3 * sizeof(double) * n, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(sD, d, 3 * sizeof(double) * n, hipMemcpyHostToDevice, stream));
  HIPCHK(launch_closest(exact, sO, sD, n, sT, sP, sDist));
  HIPCHK(hipMemcpyAsync(tri, sT, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(pt, sP, 3 * sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(dist, sDist, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SBR_OK;
}

int Engine::closest_hit(const double* o, const double* d, int n, int exact, int32_t* tri, double* pt, double* dist) {
  if (device < 0) return fail(SBR_E_STATE, "host-only context has no device");
  if (!haveScene) return fail(SBR_E_STATE, "no scene uploaded");
  if (n <= 0) return SBR_OK;
  HIPCHK(hipSetDevice(device));
  DBuf<double> dO, dD, dP, dDist;
  DBuf<int32_t> dT;
  dO.ensure(3 * (size_t)n); dD.ensure(3 * (size_t)n); dP.ensure(3 * (size_t)n); dDist.ensure(n); dT.ensure(n);
  if (!dO.p || !dD.p || !dP.p || !dDist.p || !dT.p) return fail(SBR_E_NOMEM, "closest_hit buffers");
  HIPCHK(hipMemcpyAsync(dO.p, o, 3 * sizeof(double) * n, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dD.p, d, 3 * sizeof(double) * n, hipMemcpyHostToDevice, stream));
  HIPCHK(launch_closest(exact, dO.p, dD.p, n, dT.p, dP.p, dDist.p));
  HIPCHK(hipMemcpyAsync(tri, dT.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(pt, dP.p, 3 * sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(dist, dDist.p, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SBR_OK;
}

static double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

// Bounce loop of the root items [rootBase, rootBase + nRoot): one persistent k_trace launch per
// depth, no host synchronisation in between.  Returns SBR_E_LIMIT when the node or queue buffers
// overflowed, with *need a lower bound of the nodes the window needs; *ms is the kernels' time.
int Engine::trace_stage(const std::vector<TxDesc>& txs, int64_t rootBase, int64_t nRoot, int maxDepth, int mode,
                        bool count, const RxGrid* capG, int64_t* need, float* ms) {
  *need = 0;
  *ms = 0;
  if (nRoot == 0) {
    nodes = capNodes = 0;
    return SBR_OK;
  }
  // Traversal stack: the entries that leave LDS room for kTraceBlocksPerCU blocks per CU (the bound
  // the kernel's registers set) next to the scene, if staged, and the kernel's own LDS; deeper entries
  // of tall trees go to device memory.
  const int depthAll = stackDepth();
  auto ldsEntries = [&](size_t scene) {
    const size_t perBlock = kLdsPerCU / kTraceBlocksPerCU - 1024;
    const int64_t e = perBlock > scene ? (int64_t)((perBlock - scene) / ((size_t)kBlock * 8)) : 0;
    return (int)std::max<int64_t>(4, std::min<int64_t>(depthAll, e));
  };
  // LDS-staged part of the scene: the wide nodes and leaves, or the reference tree's records.
  const size_t stageBytes = mode == kTravWide ? wideBytes : sceneBytes;
  int kGlob = ldsEntries(0), kLds = ldsEntries(stageBytes);
  if (tune.stackLds > 0) kGlob = kLds = std::min(tune.stackLds, depthAll);
  const size_t stk = stack_bytes(kGlob, kBlock), stkL = stack_bytes(kLds, kBlock);
  // Scene in LDS or read through the caches: automatic choice takes LDS staging unless it lowers the
  // number of resident blocks per CU.
  const bool ldsFits = stageBytes + stkL <= kSceneLdsMax;
  const bool sub = !hSub.empty();
  using KT = decltype(&k_trace<kTravWide, false, false, false>);
  auto pick = [](bool l, bool sb, auto k00, auto k01, auto k10, auto k11) -> KT {
    return l ? (sb ? k11 : k10) : (sb ? k01 : k00);
  };
  auto variant = [&](bool l) -> KT {
    if (count) {
      if (mode == kTravExact) return l ? k_trace<kTravExact, true, false, true> : k_trace<kTravExact, false, false, true>;
      if (mode == kTravRef)
        return pick(l, sub, k_trace<kTravRef, false, false, true>, k_trace<kTravRef, false, true, true>,
                    k_trace<kTravRef, true, false, true>, k_trace<kTravRef, true, true, true>);
      return pick(l, sub, k_trace<kTravWide, false, false, true>, k_trace<kTravWide, false, true, true>,
                  k_trace<kTravWide, true, false, true>, k_trace<kTravWide, true, true, true>);
    }
    if (mode == kTravExact) return l ? k_trace<kTravExact, true, false, false> : k_trace<kTravExact, false, false, false>;
    if (mode == kTravRef)
      return pick(l, sub, k_trace<kTravRef, false, false, false>, k_trace<kTravRef, false, true, false>,
                  k_trace<kTravRef, true, false, false>, k_trace<kTravRef, true, true, false>);
    return pick(l, sub, k_trace<kTravWide, false, false, false>, k_trace<kTravWide, false, true, false>,
                k_trace<kTravWide, true, false, false>, k_trace<kTravWide, true, true, false>);
  };
  int occGlobal = 1, occLds = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occGlobal, variant(false), kBlock, stk));
  if (ldsFits) HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occLds, variant(true), kBlock, stageBytes + stkL));
  bool lds = ldsFits && (tune.ldsScene == 1 || (tune.ldsScene < 0 && occLds >= occGlobal));
  const int kStack = lds ? kLds : kGlob;
  const size_t smem = (lds ? stageBytes : 0) + stack_bytes(kStack, kBlock);
  TraceArgs A{};
  A.nTx = (int32_t)txs.size();
  A.stackDepth = stackDepth();
  A.nodeBytes = (int32_t)stageBytes;
  A.wide = wide_tree();
  A.wideNodes = (int32_t)hNodes4.size();
  A.wideLeaves = (int32_t)hLeaves.size();
  A.ext = sceneExtent;
  A.rootBase = rootBase;
  A.nodeCap = nodeCap;
  A.queueCap = nodeCap;
  ENSURE(dQCnt, (size_t)(kMaxDepth + 2) * kQParts * kQStride);
  A.qCnt = dQCnt.p;
  A.partCap = nodeCap / kQParts;
  A.capture = capG ? 1 : 0;
  if (capG) A.g = *capG;
  auto kern = variant(lds);
  int perCU = std::max(1, lds ? occLds : occGlobal);
  if (tune.blocksPerCU > 0) perCU = std::min(perCU, tune.blocksPerCU);
  // Resident blocks times kTraceOversub (capped by twice the roots' blocks): blocks that finish early
  // are replaced by waiting ones, so the launch ends evenly.
  const int oversub = tune.oversub > 0 ? tune.oversub : kTraceOversub;
  const unsigned grid = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>((int64_t)numCU * perCU * oversub, blocks_for(std::max<int64_t>(2 * nRoot, 1), kBlock)));
  stats.trace_lds = lds ? 1 : 0;
  stats.trace_blocks_per_cu = perCU;
  A.stackLds = kStack;
  A.gStack = nullptr;
  if (kStack < depthAll) {
    ENSURE(dStackG, (size_t)(depthAll - kStack) * grid * kBlock);
    A.gStack = dStackG.p;
  }
  LevelRec* lv0 = pinSmall.at<LevelRec>(kPinLevels0);
  for (int d = 0; d < kMaxDepth + 2; d++) lv0[d] = LevelRec{0, 0};
  lv0[0].count = (unsigned long long)nRoot;
  HIPCHK(hipMemcpyAsync(dLevels.p, lv0, sizeof(LevelRec) * (kMaxDepth + 2), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemsetAsync(dTraceCounters.p, 0, 8 * sizeof(unsigned long long), stream));
  HIPCHK(hipMemsetAsync(dQCnt.p, 0, sizeof(unsigned long long) * (kMaxDepth + 2) * kQParts * kQStride, stream));
  hipEvent_t kev[2];
  HIPCHK(hipEventCreate(&kev[0]));
  HIPCHK(hipEventCreate(&kev[1]));
  struct KevGuard { hipEvent_t* e; ~KevGuard() { (void)hipEventDestroy(e[0]); (void)hipEventDestroy(e[1]); } } kguard{kev};
  HIPCHK(hipEventRecord(kev[0], stream));
  for (int d = 0; d <= maxDepth; d++) {
    A.depth = d;
    const int32_t* qin = dQueue[d & 1].p;
    int32_t* qout = dQueue[(d + 1) & 1].p;
    kern<<<grid, kBlock, smem, stream>>>(A, dTx.p, dNodes.p, dNodesF.p, dTris.p, dTriBox32.p, dSub.p, dSubTri.p, dTriN.p, qin,
                                         qout, dTree.p, dLevels.p, dTraceCounters.p, dCandFlag.p);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(kev[1], stream));
  const LevelRec* lv = pinSmall.at<LevelRec>(kPinLevels);
  const unsigned long long* ctr = pinSmall.at<unsigned long long>(kPinTraceCtr);
  HIPCHK(hipMemcpyAsync(pinSmall.at<LevelRec>(kPinLevels), dLevels.p, sizeof(LevelRec) * (kMaxDepth + 2),
                        hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(pinSmall.at<unsigned long long>(kPinTraceCtr), dTraceCounters.p, 8 * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  float kms = 0;
  HIPCHK(hipEventElapsedTime(&kms, kev[0], kev[1]));
  unsigned long long total = 0;
  int deepest = 0;
  for (int d = 0; d <= maxDepth; d++) {
    total += lv[d].count;
    if (lv[d].count) deepest = d;
  }
  if ((ctr[0] & 3ull) || total > nodeCap) {
    *need = (int64_t)std::max<unsigned long long>(total, nodeCap + 1);
    return SBR_E_LIMIT;
  }
  *ms = kms;
  nodes = (int64_t)total;
  capNodes = 0;
  if (capG && nodes > 0) {   // listed nodes, in node order; the count stays on the device
    const int32_t nTiles = (int32_t)((nodes + kFlagTile - 1) / kFlagTile);
    ENSURE(dTileCount, nTiles);
    ENSURE(dTileOffset, nTiles);
    k_flag_count<<<nTiles, 256, 0, stream>>>(dCandFlag.p, nodes, dTileCount.p);
    HIPCHK(hipGetLastError());
    size_t tb = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, dTileCount.p, dTileOffset.p, nTiles, stream));
    ENSURE(dScratch, tb);
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(dScratch.p, tb, dTileCount.p, dTileOffset.p, nTiles, stream));
    k_flag_scatter<<<nTiles, 256, 0, stream>>>(dCandFlag.p, nodes, dTileCount.p, dTileOffset.p, nTiles, dCand.p,
                                               dTraceCounters.p + 1);
    HIPCHK(hipGetLastError());
  }
  stats.ms_k_trace += kms;
  stats.n_k_trace += maxDepth + 1;
  stats.bounces += nodes;
  stats.max_depth = std::max<int64_t>(stats.max_depth, deepest);
  stats.node_visits += count ? (int64_t)ctr[6] : 0;
  stats.tri_tests += count ? (int64_t)ctr[7] : 0;
  stats.box_tests_fp64 += count ? (int64_t)ctr[5] : 0;
  stats.wide_redone += count ? (int64_t)ctr[4] : 0;
  return SBR_OK;
}

int Engine::trace(const sbr_trace_params* P, int nt, const double* rx, int m) {
  if (device < 0) return fail(SBR_E_STATE, "host-only context has no device");
  if (!haveScene) return fail(SBR_E_STATE, "no scene uploaded");
  if (mats.empty()) return fail(SBR_E_STATE, "no materials set");
  if (!P || nt <= 0 || nt > 32767 || m < 0 || (m > 0 && !rx)) return fail(SBR_E_ARG, "trace: bad arguments");
  if ((int64_t)nt * m >= (int64_t(1) << 31)) return fail(SBR_E_LIMIT, "transmitters x receivers above 2^31");
  for (int32_t t = 0; t < ntri; t++)
    if (hMat[t] < 0 || hMat[t] * 6 >= (int)mats.size()) return fail(SBR_E_ARG, "triangle material id without a material");
  for (int t = 1; t < nt; t++)
    if ((P[t].q1MeshIndexing != 0) != (P[0].q1MeshIndexing != 0))
      return fail(SBR_E_ARG, "q1MeshIndexing must be the same for every transmitter of a batch");
  HIPCHK(hipSetDevice(device));
  if (!pinSmall.ensure(kPinSmall)) return fail(SBR_E_NOMEM, "pinned staging");
  haveResult = false;
  lastParams.assign(P, P + nt);   // read by the path export of every window (multi-window traces)
  stats = sbr_stats{};
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t ev[5];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  struct EvGuard { hipEvent_t* e; ~EvGuard() { for (int i = 0; i < 5; i++) (void)hipEventDestroy(e[i]); } } guard{ev};
  HIPCHK(hipEventRecord(ev[0], stream));

  if (ncop >= (1 << 30)) return fail(SBR_E_LIMIT, "coplanar ids above 2^30 do not fit the duplicate-removal keys");
  // ---- transmitters and their launch directions
  std::vector<TxDesc> txs(nt);
  int maxDepth = 0;
  bool anyExact = false, anyRef = false, count = false, anyOcclusion = false, anyModel0 = false;
  int mode = kTravWide;   // one traversal per call: the reference order if any transmitter asks for it
  int64_t nRoot = 0;
  for (int t = 0; t < nt; t++) {
    const sbr_trace_params& p = P[t];
    if (p.maxR < 0 || p.maxT < 0) return fail(SBR_E_ARG, "maxR / maxT must be >= 0");
    int md = p.maxR + p.maxT;
    if (p.maxTotal > 0) md = std::min(md, p.maxTotal);
    if (md > kMaxDepth) return fail(SBR_E_LIMIT, "tree depth above the device limit");
    maxDepth = std::max(maxDepth, md);
    const int stride = p.launchStride > 0 ? p.launchStride : 1;
    if (p.launchOffset < 0 || p.launchOffset >= stride) return fail(SBR_E_ARG, "launchOffset must be in [0, stride)");
    if (p.antennaType < 0 || p.antennaType > 3) return fail(SBR_E_ARG, "antennaType must be 0, 1, 2 or 3");
    if (p.antennaType == 3 && patNT == 0) return fail(SBR_E_STATE, "antennaType 3 needs a pattern (sbr_set_tx_pattern)");
    if (p.receptionModel < 0 || p.receptionModel > 1) return fail(SBR_E_ARG, "receptionModel must be 0 or 1");
    if (!(p.frequency > 0)) return fail(SBR_E_ARG, "frequency must be > 0");
    if (p.exactTraversal < 0 || p.exactTraversal > 2) return fail(SBR_E_ARG, "exactTraversal must be 0, 1 or 2");
    anyExact = anyExact || p.exactTraversal == kTravExact;
    anyRef = anyRef || p.exactTraversal == kTravRef;
    count = count || p.countTraversal;
    mode = anyExact ? kTravExact : anyRef ? kTravRef : kTravWide;
    TxDesc& T = txs[t];
    memset(&T, 0, sizeof T);
    memcpy(T.tx, p.tx, sizeof T.tx);
    memcpy(T.txcs, p.txCoordSys, sizeof T.txcs);
    T.freq = p.frequency;
    T.pt = p.pt;
    if (p.receptionModel == 0) {
      // RTSolver.cpp:839, 841 with txTesslation = tess + 1 (RTSolver.cpp:70)
      double beta = 1.0 / (3.0 * (p.tess + 1)) * std::acos(-1.0 / std::sqrt(5.0));
      T.rcoef = std::tan(beta);
    } else {
      double alpha = (69.0 * 3.141592653589793 / 180.0) / std::ldexp(1.0, p.tess);
      T.rcoef = alpha / std::sqrt(3.0);
    }
    T.capK = T.rcoef < 1.0 ? T.rcoef / (1.0 - T.rcoef) * (1.0 + 1e-12) : INFINITY;
    int64_t nl = 0;
    int rc = rays_for(p.tess, &T.rays, &nl);
    if (rc) return rc;
    T.launchStride = stride;
    T.launchOffset = p.launchOffset;
    T.nLaunch = nl > p.launchOffset ? (nl - p.launchOffset + stride - 1) / stride : 0;
    T.launchStart = nRoot;
    nRoot += T.nLaunch;
    T.maxR = p.maxR;
    T.maxT = p.maxT;
    T.maxTotal = p.maxTotal;
    T.useFresnel = p.useFresnel;
    T.antennaType = p.antennaType;
    if (!(p.wallWidth >= 0.0) || !std::isfinite(p.wallWidth)) return fail(SBR_E_ARG, "wallWidth must be finite and >= 0");
    T.wallWidth = p.wallWidth;
    if (p.typedDedup != 0 && p.typedDedup != 1) return fail(SBR_E_ARG, "typedDedup must be 0 or 1");
    T.typedDedup = p.typedDedup;
    if (p.checkOcclusion != 0 && p.checkOcclusion != 1) return fail(SBR_E_ARG, "checkOcclusion must be 0 or 1");
    T.checkOcclusion = p.checkOcclusion;
    if (p.longMiss != 0 && p.longMiss != 1) return fail(SBR_E_ARG, "longMiss must be 0 or 1");
    T.longMiss = p.longMiss;
    if (p.physicalField != 0 && p.physicalField != 1) return fail(SBR_E_ARG, "physicalField must be 0 or 1");
    T.physicalField = p.physicalField;
    anyOcclusion = anyOcclusion || p.checkOcclusion != 0;
    anyModel0 = anyModel0 || p.receptionModel == 0;
    T.receptionModel = p.receptionModel;
  }
  if (nRoot >= (int64_t(1) << 31)) return fail(SBR_E_LIMIT, "more than 2^31 launches in one trace");
  ENSURE(dTx, nt);
  HIPCHK(hipMemcpyAsync(dTx.p, txs.data(), sizeof(TxDesc) * nt, hipMemcpyHostToDevice, stream));
  HIPCHK(hipEventRecord(ev[1], stream));
  stats.launches = nRoot;

  // ---- receivers: grid (host binning) and direct-path line of sight for every pair
  winners = 0;
  nrx = m;
  ntx = nt;
  const int64_t nOut = (int64_t)nt * m;
  bool losWait = false;   // the capture waits for the side stream's line of sight once
  RxGrid g{};
  if (m > 0) {
    // Uniform xy grid with about one receiver per cell.
    for (int a = 0; a < 3; a++) { g.bmin[a] = INFINITY; g.bmax[a] = -INFINITY; }
    for (int r = 0; r < m; r++)
      for (int a = 0; a < 3; a++) {
        if (!std::isfinite(rx[3 * r + a])) return fail(SBR_E_ARG, "receiver coordinates must be finite");
        g.bmin[a] = std::min(g.bmin[a], rx[3 * r + a]);
        g.bmax[a] = std::max(g.bmax[a], rx[3 * r + a]);
      }
    double ex = g.bmax[0] - g.bmin[0], ey = g.bmax[1] - g.bmin[1];
    double cell = std::sqrt(std::max(ex * ey, 1e-12) / std::max(1, m));
    cell = std::max(cell, std::max(ex, ey) / 4096.0);
    if (!(cell > 0)) cell = 1.0;
    g.cell = cell;
    g.gx0 = g.bmin[0];
    g.gy0 = g.bmin[1];
    g.nx = std::max(1, (int)std::floor(ex / cell) + 1);
    g.ny = std::max(1, (int)std::floor(ey / cell) + 1);
    grid = g;
    // Staged in page-locked memory: receivers in cell order, cell starts, receivers as given.
    const size_t nCell = (size_t)g.nx * g.ny + 1;
    const size_t offStart = sizeof(RxCell) * (size_t)m, offRx = offStart + ((sizeof(int32_t) * nCell + 63) & ~size_t(63));
    if (!pinIn.ensure(offRx + 3 * sizeof(double) * (size_t)m)) return fail(SBR_E_NOMEM, "pinned receiver staging");
    RxCell* cellRx = pinIn.at<RxCell>(0);
    int32_t* cellCount = pinIn.at<int32_t>(offStart);
    std::fill(cellCount, cellCount + nCell, 0);
    std::vector<int32_t> cellOf(m);
    for (int r = 0; r < m; r++) {
      int ix = std::min(g.nx - 1, std::max(0, (int)std::floor((rx[3 * r] - g.gx0) / cell)));
      int iy = std::min(g.ny - 1, std::max(0, (int)std::floor((rx[3 * r + 1] - g.gy0) / cell)));
      cellOf[r] = iy * g.nx + ix;
      cellCount[cellOf[r] + 1]++;
    }
    for (size_t c = 1; c < nCell; c++) cellCount[c] += cellCount[c - 1];
    std::vector<int32_t> fillPos(cellCount, cellCount + nCell - 1);
    for (int r = 0; r < m; r++) {
      RxCell& x = cellRx[fillPos[cellOf[r]]++];
      for (int a = 0; a < 3; a++) x.p[a] = rx[3 * r + a];
      x.r = r;
      x.pad = 0;
    }
    ENSURE(dRx, 3 * (size_t)m);
    ENSURE(dCellStart, nCell);
    ENSURE(dCellRx, m);
    ENSURE(dLos, nOut);
    memcpy(pinIn.at<double>(offRx), rx, 3 * sizeof(double) * m);
    HIPCHK(hipMemcpyAsync(dRx.p, pinIn.at<double>(offRx), 3 * sizeof(double) * m, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dCellStart.p, cellCount, sizeof(int32_t) * nCell, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dCellRx.p, cellRx, sizeof(RxCell) * m, hipMemcpyHostToDevice, stream));
    // Line of sight for every pair.  A small one (a few blocks per CU at most) runs on the side
    // stream beside the bounce loop, whose launch tails it fills, and the capture waits for it; a
    // large one would hold CUs that the persistent bounce kernel needs, so it runs in line.
    const size_t sm = stack_bytes(stackDepth(), kBlock);
    const bool side = blocks_for(nOut, kBlock) <= (int64_t)numCU / 2;
    hipStream_t ls = side ? sideStream : stream;
    if (side) {
      HIPCHK(hipEventRecord(evSide[0], stream));
      HIPCHK(hipStreamWaitEvent(sideStream, evSide[0], 0));
    }
    auto kl = mode == kTravExact ? k_los<kTravExact> : mode == kTravRef ? k_los<kTravRef> : k_los<kTravWide>;
    kl<<<blocks_for(nOut, kBlock), kBlock, sm, ls>>>(dTx.p, nt, wide_tree(), dNodes.p, dNodesF.p, dTris.p, dTriBox32.p,
                                                     dSub.p, dSubTri.p, sceneExtent, dRx.p, m, stackDepth(), dLos.p);
    HIPCHK(hipGetLastError());
    if (side) {
      HIPCHK(hipEventRecord(evSide[1], sideStream));
      losWait = true;
    }
  }
  HIPCHK(hipEventRecord(ev[2], stream));

  // ---- launch windows: trees, capture, kept paths and field sums per window of root items.
  // Order keys grow with the global root index, so a duplicate-removal group first seen in window w
  // is final when w ends (later windows only hold larger keys): its path is resolved and its field
  // added while w's trees are still resident.  Receiver sums continue in key order across windows.
  const DSurf* surf = P[0].q1MeshIndexing ? dSurfQ1.p : dSurfId.p;
  // Tables start at the last trace's final size, so a repeated trace does not restart; a trace far
  // smaller than the last (a single transmitter after a batch, a small query after a large one)
  // scales it down by the larger of its output and launch ratios, since clearing and scanning the
  // tables is a cost per entry.  A table that fills up restarts the trace with a larger one.
  const int64_t nLaunches = nRoot;
  int64_t want = std::max<int64_t>(1 << 16, 4 * nOut);
  if (tune.tableEntries > 0) {
    want = std::max<int64_t>(1024, tune.tableEntries);
  } else if (tabCapHint > 0) {
    double scale = 1.0;
    if (16 * nOut < tabHintOut && 16 * nLaunches < tabHintLaunches)
      scale = std::max((double)nOut / (double)tabHintOut, (double)nLaunches / (double)tabHintLaunches);
    want = std::max<int64_t>(want, (int64_t)std::ceil((double)tabCapHint * scale));
  }
  size_t cap = 1;
  while (cap < (size_t)want) cap <<= 1;
  int tabRestarts = 0;
  ENSURE(dLevels, kMaxDepth + 2);
  ENSURE(dCounters, 10);
  ENSURE(dTraceCounters, 8);
  // at least 256 entries per queue partition
  const size_t budget =
      std::max<size_t>(tune.windowNodes > 0 ? (size_t)tune.windowNodes : kWindowNodes, (size_t)kQParts * 256);
  // Bound of the tree size per root item: one node per interaction sequence, i.e. sum over the
  // allowed (reflections r, transmissions t) of C(r + t, r), for the largest transmitter.
  double perRootBound = 1.0;
  for (int t = 0; t < nt; t++) {
    double b = 0.0;
    for (int r = 0; r <= P[t].maxR; r++)
      for (int k = 0; k <= P[t].maxT; k++) {
        if (P[t].maxTotal > 0 && r + k > P[t].maxTotal) continue;
        double c = 1.0;   // C(r + k, r)
        for (int j = 1; j <= k; j++) c = c * (double)(r + j) / (double)j;
        b += c;
      }
    perRootBound = std::max(perRootBound, b);
  }
  double perRootMax = 0;   // densest window of this trace (nodes per root item)
  bool done = false;
  for (int tabAttempt = 0; tabAttempt < 12 && !done; tabAttempt++) {
    stats.bounces = stats.node_visits = stats.tri_tests = stats.capture_records = stats.corrected = 0;
    stats.box_tests_fp64 = 0;
    stats.ms_k_trace = stats.ms_k_capture = stats.ms_trace = stats.ms_capture = stats.ms_field = 0;
    stats.n_k_trace = 0;
    stats.max_depth = 0;
    stats.heavy_nodes = 0;
    stats.capture_nodes = 0;
    stats.capture_tests = 0;
    stats.windows = 0;
    winners = 0;
    multiWindow = false;
    cache = PathCache{};
    cacheReady = false;
    store.np = store.ns = 0;
    if (m > 0) {
      ENSURE(dTabA, cap);
      ENSURE(dSeqTab, cap);
      ENSURE(dTabOrder, cap);
      ENSURE(dTabRx, cap);
      ENSURE(dField, 6 * (size_t)nOut);
      HIPCHK(hipMemsetAsync(dTabA.p, 0, cap * sizeof(unsigned long long), stream));
      HIPCHK(hipMemsetAsync(dSeqTab.p, 0xFF, cap * sizeof(unsigned long long), stream));
      HIPCHK(hipMemsetAsync(dTabOrder.p, 0xFF, cap * sizeof(unsigned long long), stream));
      HIPCHK(hipMemsetAsync(dField.p, 0, 6 * sizeof(double) * nOut, stream));
      HIPCHK(hipMemsetAsync(dCounters.p, 0, 10 * sizeof(unsigned long long), stream));
    }
    unsigned long long usedBefore = 0, corBefore = 0;
    // Nodes per root item for sizing windows: the last trace's densest window, then the densest
    // window of this trace so far, never above the tree-size bound (with which no window overflows).
    double perRoot = perRootHint > 0 ? std::min(perRootHint, perRootBound) : perRootBound;
    bool tableFull = false;
    int64_t ws = 0;
    while (ws < nRoot || (nRoot == 0 && ws == 0)) {
      // trees of roots [ws, ws + W), sized for 1.25x the expected nodes so that a window denser than
      // the estimate seldom overflows (an overflow reruns the window)
      int64_t W = std::max<int64_t>(1, std::min<int64_t>(nRoot - ws, (int64_t)((double)budget / (1.25 * perRoot))));
      if (nRoot == 0) W = 0;
      float msT = 0;
      for (int attempt = 0;; attempt++) {
        // at least 256 queue entries per partition (a wave appends up to 128 children)
        const size_t want = std::min(budget, std::max((size_t)((double)W * perRoot * 1.25) + 1024,
                                                      (size_t)kQParts * 256));
        if (nodeCap < want) nodeCap = want;
        if (nodeCap > budget) nodeCap = budget;
        ENSURE(dTree, nodeCap);
        ENSURE(dQueue[0], nodeCap);
        ENSURE(dQueue[1], nodeCap);
        if (m > 0) {
          ENSURE(dCand, nodeCap);
          ENSURE(dCandFlag, nodeCap);
        }
        nodeCap = std::min(dTree.n, std::min(dQueue[0].n, dQueue[1].n));
        if (m > 0) nodeCap = std::min(nodeCap, std::min(dCand.n, dCandFlag.n));
        int64_t need = 0;
        int rc = trace_stage(txs, ws, W, maxDepth, mode, count, m > 0 ? &g : nullptr, &need, &msT);
        if (rc == SBR_OK) break;
        if (rc != SBR_E_LIMIT || attempt > 40) return rc == SBR_E_LIMIT ? fail(rc, "tree buffers could not be sized") : rc;
        // Overflowed (need is only a lower bound): first use the whole window budget, then halve
        // the window until its trees fit.
        perRoot = std::min(perRootBound, std::max(perRoot, (double)need / (double)std::max<int64_t>(W, 1)));
        if (nodeCap < budget) {
          nodeCap = std::max(nodeCap, std::min(budget, (size_t)need * 2 + 1024));
        } else {
          if (W == 1) return fail(SBR_E_LIMIT, "one launch needs more tree nodes than the window budget");
          W = std::max<int64_t>(1, W / 2);
        }
      }
      stats.ms_trace += msT;
      stats.windows++;
      if (W > 0) {
        perRootMax = std::max(perRootMax, (double)nodes / (double)W);
        perRoot = std::min(perRootBound, std::max(1.0, perRootMax));
      }
      const bool last = ws + W >= nRoot;
      if (ws == 0 && !last) multiWindow = true;
      if (m == 0) {
        ws += std::max<int64_t>(W, 1);
        continue;
      }

      // capture + correction + duplicate removal for this window
      // Reference-sphere transmitters capture over large regions, where the correction's first-step
      // pre-test (from each node's image source) rejects most candidates before their sphere tests.
      const bool cone = anyModel0 && nodes > 0;
      if (cone) {
        ENSURE(dImg, 3 * (size_t)nodes);
        const LevelRec* lvw = pinSmall.at<LevelRec>(kPinLevels);
        for (int d = 0; d <= maxDepth; d++) {
          if (lvw[d].count == 0) break;
          k_images<<<(unsigned)std::min<int64_t>(blocks_for((int64_t)lvw[d].count, 256), (int64_t)numCU * 8), 256, 0,
                     stream>>>(dTree.p, dTx.p, surf, (int64_t)lvw[d].base, (int64_t)lvw[d].count, d, dImg.p);
          HIPCHK(hipGetLastError());
        }
      }
      hipEvent_t kev[4];
      for (auto& e : kev) HIPCHK(hipEventCreate(&e));
      struct KevGuard { hipEvent_t* e; ~KevGuard() { for (int i = 0; i < 4; i++) (void)hipEventDestroy(e[i]); } } kguard{kev};
      // heavy jobs: a list sized from the last trace's (a full list flags the pass, which is redone
      // with a larger one), at most one per node
      if (heavyCap == 0) heavyCap = 1 << 16;
      size_t hcap = std::min<size_t>(heavyCap, (size_t)std::max<int64_t>(nodes, 1));
      ENSURE(dHeavy, hcap);
      ENSURE(dNodeKey, 3 * (size_t)std::max<int64_t>(nodes, 1));
      CaptureArgs CA{};
      CA.nodes = nodes;   // bound of the listed nodes; their count is read on the device
      {
        const int64_t nCell = (int64_t)g.nx * g.ny + 1;
        CA.cellLds = (m < 65536 && nCell <= 16384) ? (int32_t)nCell : 0;
      }
      CA.nrx = m;
      CA.g = g;
      CA.invCell = 1.0 / g.cell;
      CA.tabMask = cap - 1;
      CA.fuse = anyOcclusion ? 0 : 1;   // the occlusion test stays in k_dedup
      if (losWait) {
        HIPCHK(hipStreamWaitEvent(stream, evSide[1], 0));
        losWait = false;
      }
      HIPCHK(hipEventRecord(kev[0], stream));
      // candidate pass: (node, receiver) records, then the dedup pass over them, with no host round
      // trip in between (the records' partition prefix is taken on the device); redone with a larger
      // record buffer if it filled (the duplicate-removal tables only ever take smaller keys, so the
      // redo over all records gives what one pass would)
      unsigned long long nRec = 0, nHeavy = 0;
      unsigned long long* ctr = pinSmall.at<unsigned long long>(kPinCapCtr);
      ENSURE(dRecCount, kRecParts * kRecStride);
      ENSURE(dRecPrefix, kRecParts + 2);
      for (int recAttempt = 0;; recAttempt++) {
        if (recCap < (size_t)kRecParts * 1024) recCap = std::max<size_t>((size_t)kRecParts * 2048, (size_t)nodes);
        ENSURE(dRec, recCap);
        recCap = dRec.n;
        const unsigned long long part = recCap / kRecParts;
        HIPCHK(hipMemsetAsync(dRecCount.p, 0, sizeof(unsigned long long) * kRecParts * kRecStride, stream));
        HIPCHK(hipMemsetAsync(dCounters.p + 3, 0, sizeof(unsigned long long), stream));
        HIPCHK(hipMemsetAsync(dCounters.p + 5, 0, 2 * sizeof(unsigned long long), stream));
        HIPCHK(hipMemsetAsync(dCounters.p + 8, 0, 2 * sizeof(unsigned long long), stream));
        CaptureBufs CB{dTx.p, dTree.p, surf, dRx.p, dCellStart.p, dCellRx.p, dLos.p, dTabA.p, dSeqTab.p,
                       (unsigned long long)(cap - 1), dTabOrder.p, dTabRx.p, dCounters.p, dHeavy.p, dRec.p, part, dRecCount.p,
                       dRecPrefix.p, count ? 1 : 0, dNodeKey.p, dCand.p, dTraceCounters.p + 1};
        CB.img = cone ? dImg.p : nullptr;
        CB.heavyCap = hcap;
        CB.nodes = dNodes.p;
        CB.nodesF = dNodesF.p;
        CB.tris = dTris.p;
        CB.tbox = dTriBox32.p;
        CB.subs = dSub.p;
        CB.subTri = dSubTri.p;
        CB.ext = sceneExtent;
        CB.stackDepth = stackDepth();
        CB.wide = wide_tree();
        CB.mode = mode;
        if (nodes > 0) {
          auto kc = cone ? k_capture_cone : k_capture;
          kc<<<(unsigned)std::min<int64_t>(blocks_for(nodes, 256), (int64_t)numCU * 16), 256,
               sizeof(uint16_t) * (size_t)CA.cellLds, stream>>>(CA, CB);
          HIPCHK(hipGetLastError());
          // wave per heavy node; the count is read on the device
          k_capture_heavy<<<blocks_for((int64_t)numCU * 32 * 64, 256), 256, 0, stream>>>(CA, CB);
          HIPCHK(hipGetLastError());
          k_rec_prefix<<<1, kRecParts, 0, stream>>>(dRecCount.p, part, dRecPrefix.p);
          HIPCHK(hipGetLastError());
          // dedup pass: a grid-stride loop over the records, their count read on the device
          const unsigned dg = (unsigned)std::min<int64_t>(blocks_for((int64_t)recCap, 256), (int64_t)numCU * 16);
          if (anyOcclusion)
            k_dedup<true><<<dg, 256, stack_bytes(stackDepth(), 256), stream>>>(CA, CB);
          else
            k_dedup<false><<<dg, 256, 0, stream>>>(CA, CB);
          HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(kev[1], stream));
        HIPCHK(hipMemcpyAsync(ctr, dCounters.p, 10 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(ctr + 10, dRecPrefix.p + kRecParts, 2 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(ctr + 12, dTraceCounters.p + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              stream));
        HIPCHK(hipStreamSynchronize(stream));
        nHeavy = ctr[5];
        capNodes = (int64_t)ctr[12];
        nRec = nodes > 0 ? ctr[10] : 0;
        if (!(ctr[3] & 20ull)) {
          stats.capture_tests += (int64_t)ctr[6];
          break;
        }
        if (recAttempt > 4) return fail(SBR_E_LIMIT, "capture record buffer could not be sized");
        if (ctr[3] & 4ull) {
          const unsigned long long maxPart = ctr[11];
          recCap = std::max<size_t>(recCap * 2, (size_t)(maxPart + maxPart / 4 + 1024) * kRecParts);
        }
        if (ctr[3] & 16ull) {
          heavyCap = std::max<size_t>(heavyCap * 2, (size_t)(ctr[5] + ctr[5] / 4 + 1024));
          hcap = std::min<size_t>(heavyCap, (size_t)std::max<int64_t>(nodes, 1));
          ENSURE(dHeavy, hcap);
        }
      }
      stats.heavy_nodes += (int64_t)nHeavy;
      stats.capture_nodes += capNodes;
      float kms = 0;
      HIPCHK(hipEventElapsedTime(&kms, kev[0], kev[1]));
      stats.ms_k_capture += kms;
      stats.ms_capture += kms;
      if ((ctr[3] & 9ull) || ctr[2] * 10 > cap * 7 || ctr[7] * 10 > cap * 7) {
        tableFull = true;
        break;
      }
      stats.capture_records += (int64_t)(nRec + ctr[8]);
      stats.corrected += (int64_t)(ctr[1] - corBefore + ctr[9]);
      corBefore = ctr[1];
      const int64_t Wn = (int64_t)(ctr[2] - usedBefore);   // groups first seen in this window
      usedBefore = ctr[2];

      // kept paths of this window: collect, sort by order key, resolve nodes, sort by (tx, rx),
      // fields, receiver sums
      HIPCHK(hipEventRecord(kev[2], stream));
      if (Wn > 0) {
        const int64_t Wc = Wn;
        ENSURE(dWinKey, Wc);
        ENSURE(dWinKey2, Wc);
        ENSURE(dWinRx, Wc);
        ENSURE(dWinRx2, Wc);
        ENSURE(dWinIdx, Wc);
        ENSURE(dWinIdx2, Wc);
        ENSURE(dWinNode, Wc);
        ENSURE(dWinField, 6 * Wc);
        HIPCHK(hipMemsetAsync(dCounters.p + 4, 0, sizeof(unsigned long long), stream));
        const bool smallTab = cap <= kCollectSmall;
        auto kcol = smallTab ? k_collect<1> : k_collect<kCollectPerThread>;
        kcol<<<blocks_for((int64_t)cap, 256 * (smallTab ? 1 : kCollectPerThread)), 256, 0, stream>>>(
            dTabA.p, dTabOrder.p, dTabRx.p, cap, (unsigned long long)ws, (unsigned long long)(ws + W), dWinKey.p,
            dWinRx.p, dCounters.p + 4);
        HIPCHK(hipGetLastError());
        int txrxBits = 1;
        while ((int64_t(1) << txrxBits) < nOut) txrxBits++;
        size_t tb1 = 0, tb2 = 0;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, dWinKey.p, dWinKey2.p, dWinRx.p, dWinRx2.p, (int)Wn, 0,
                                                  64, stream));
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, dWinRx2.p, dWinRx.p, dWinIdx.p, dWinIdx2.p, (int)Wn, 0,
                                                  txrxBits, stream));
        ENSURE(dScratch, std::max(tb1, tb2));
        // by order key -> dWinKey2 (keys), dWinRx2 (tx * nrx + rx)
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(dScratch.p, tb1, dWinKey.p, dWinKey2.p, dWinRx.p, dWinRx2.p, (int)Wn,
                                                  0, 64, stream));
        k_resolve<<<blocks_for(Wn, 256), 256, 0, stream>>>(dTree.p, dWinKey2.p, Wn, ws, dWinNode.p);
        k_iota<<<blocks_for(Wn, 256), 256, 0, stream>>>(dWinIdx.p, Wn);
        HIPCHK(hipGetLastError());
        // stable by tx * nrx + rx -> dWinRx (keys), dWinIdx2 (index into the key-sorted arrays)
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(dScratch.p, tb2, dWinRx2.p, dWinRx.p, dWinIdx.p, dWinIdx2.p, (int)Wn,
                                                  0, txrxBits, stream));
        FieldArgs FA{Wn, m, dPattern.p, patNT, patNP};
        k_field<<<blocks_for(Wn, 256), 256, 0, stream>>>(FA, dTx.p, dTree.p, dTris.p, surf, dMats.p, dRx.p, dWinRx.p,
                                                          dWinIdx2.p, dWinNode.p, dWinField.p);
        HIPCHK(hipGetLastError());
        k_rx_sum<<<blocks_for(nOut, 256), 256, 0, stream>>>(dWinRx.p, Wn, dWinField.p, nOut, dField.p);
        HIPCHK(hipGetLastError());
        if (multiWindow) {   // this window's paths to the device store while its trees are resident
          int rc = store_window_paths(Wn, maxDepth);
          if (rc) return rc;
        }
      }
      HIPCHK(hipEventRecord(kev[3], stream));
      HIPCHK(hipStreamSynchronize(stream));
      HIPCHK(hipEventElapsedTime(&kms, kev[2], kev[3]));
      stats.ms_field += kms;
      winners = Wn;
      if (multiWindow && Wn > 0) {
        store.np += Wn;
        store.ns += *pinSmall.at<int64_t>(kPinSegTotal);
      }
      ws += std::max<int64_t>(W, 1);
    }
    if (tableFull) {
      cap <<= 2;
      tabRestarts++;
      continue;
    }
    done = true;
  }
  if (!done) return fail(SBR_E_LIMIT, "dedup table could not be sized");
  if (nRoot > 0 && perRootMax > 0) perRootHint = perRootMax;
  if (m > 0) {
    tabCapHint = cap;
    tabHintOut = nOut;
    tabHintLaunches = nLaunches;
  }
  stats.table_entries = m > 0 ? (int64_t)cap : 0;
  stats.table_restarts = tabRestarts;
  if (multiWindow) winners = store.np;
  HIPCHK(hipEventRecord(ev[4], stream));
  HIPCHK(hipStreamSynchronize(stream));
  stats.winners = winners;
  float a = 0;
  HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
  stats.ms_raygen = a;
  HIPCHK(hipEventElapsedTime(&a, ev[1], ev[2]));
  stats.ms_capture += a;   // receiver binning upload and line of sight
  stats.launches = nRoot;
  // Receiver fields read back into page-locked memory, where the result accessors find them.
  hFieldN = 0;
  if (nOut > 0) {
    if (!pinOut.ensure(6 * sizeof(double) * nOut)) return fail(SBR_E_NOMEM, "pinned field staging");
    HIPCHK(hipMemcpyAsync(pinOut.p, dField.p, 6 * sizeof(double) * nOut, hipMemcpyDeviceToHost, stream));
  }
  HIPCHK(hipStreamSynchronize(stream));
  hFieldN = 6 * (size_t)nOut;
  stats.ms_total = ms_since(t0);
  haveResult = true;
  return SBR_OK;
}

// Kept paths of the window whose winners are in dWin* (winners of them).
template <typename F>
static void for_each_trace_buffer(Engine& e, F&& f) {
  f(e.dTx); f(e.dTree); f(e.dQueue[0]); f(e.dQueue[1]); f(e.dLevels); f(e.dScratch); f(e.dRx);
  f(e.dCellStart); f(e.dCellRx); f(e.dLos); f(e.dTabA); f(e.dSeqTab); f(e.dTabOrder); f(e.dTabRx);
  f(e.dHeavy); f(e.dCand); f(e.dCandFlag); f(e.dTileCount); f(e.dTileOffset); f(e.dRec); f(e.dRecCount);
  f(e.dRecPrefix); f(e.dNodeKey); f(e.dWinKey); f(e.dWinKey2); f(e.dWinRx); f(e.dWinRx2); f(e.dWinIdx);
  f(e.dWinIdx2); f(e.dWinNode); f(e.dWinField); f(e.dField); f(e.dImg); f(e.dQCnt); f(e.dStackG);
}

size_t Engine::held_bytes() const {
  size_t b = 0;
  for_each_trace_buffer(const_cast<Engine&>(*this), [&](auto& buf) { b += buf.n * sizeof(*buf.p); });
  for (const auto& kv : rayCache) b += kv.second->n * sizeof(double);
  return b;
}

void Engine::trim() {
  if (device >= 0) (void)hipSetDevice(device);
  for_each_trace_buffer(*this, [](auto& buf) { buf.release(); });
  nodeCap = 0;
  recCap = 0;
  tabCapHint = 0;
  perRootHint = 0;
  haveResult = false;
  multiWindow = false;
  cache = PathCache{};
}

int Engine::paths_window(int32_t* ptx, int32_t* prx, int32_t* plaunch, int32_t* pnseg, int64_t* poff, int32_t* stri,
                         int32_t* scop, int32_t* stype, double* ssrc, double* stgt, int64_t* nPaths, int64_t* nSegs,
                         bool countOnly) {
  HIPCHK(hipSetDevice(device));
  std::vector<int32_t> txrx(winners), perm(winners);
  std::vector<unsigned long long> key(winners);
  if (winners > 0) {
    HIPCHK(hipMemcpyAsync(txrx.data(), dWinRx.p, sizeof(int32_t) * winners, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(perm.data(), dWinIdx2.p, sizeof(int32_t) * winners, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(key.data(), dWinKey2.p, sizeof(unsigned long long) * winners, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  std::vector<int64_t> off(winners + 1, 0);
  for (int64_t w = 0; w < winners; w++) off[w + 1] = off[w] + (int64_t)((key[perm[w]] >> 27) & 31ull) + 1;
  *nPaths = winners;
  *nSegs = off[winners];
  if (countOnly) return SBR_OK;
  std::vector<TxDesc> txs(ntx);
  HIPCHK(hipMemcpyAsync(txs.data(), dTx.p, sizeof(TxDesc) * ntx, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  for (int64_t w = 0; w < winners; w++) {
    const int t = txrx[w] / nrx;
    const int64_t gl = (int64_t)(key[perm[w]] >> 32);
    if (ptx) ptx[w] = t;
    prx[w] = txrx[w] % nrx;
    plaunch[w] = (int32_t)(txs[t].launchOffset + (gl - txs[t].launchStart) * txs[t].launchStride);
    pnseg[w] = (int32_t)(off[w + 1] - off[w]);
    poff[w] = off[w];
  }
  if (winners == 0) return SBR_OK;
  const int64_t ns = off[winners];
  DBuf<int64_t> dOff;
  DBuf<int32_t> dI;
  DBuf<double> dS;
  dOff.ensure(winners);
  dI.ensure(3 * ns);
  dS.ensure(6 * ns);
  if (!dOff.p || !dI.p || !dS.p) return fail(SBR_E_NOMEM, "path export buffers");
  HIPCHK(hipMemcpyAsync(dOff.p, off.data(), sizeof(int64_t) * winners, hipMemcpyHostToDevice, stream));
  const DSurf* surf = lastParams[0].q1MeshIndexing ? dSurfQ1.p : dSurfId.p;
  k_paths<<<blocks_for(winners, 128), 128, 0, stream>>>(winners, nrx, dTx.p, dTree.p, surf, dRx.p, dWinRx.p,
                                                         dWinIdx2.p, dWinNode.p, dOff.p, dI.p, dI.p + ns, dI.p + 2 * ns,
                                                         dS.p, dS.p + 3 * ns);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(stri, dI.p, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(scop, dI.p + ns, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(stype, dI.p + 2 * ns, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(ssrc, dS.p, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(stgt, dS.p + 3 * ns, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SBR_OK;
}

int Engine::paths(int32_t* ptx, int32_t* prx, int32_t* plaunch, int32_t* pnseg, int64_t* poff, int32_t* stri,
                  int32_t* scop, int32_t* stype, double* ssrc, double* stgt, int64_t* nPaths, int64_t* nSegs,
                  bool countOnly) {
  if (!haveResult) return fail(SBR_E_STATE, "no trace result");
  if (!multiWindow)
    return paths_window(ptx, prx, plaunch, pnseg, poff, stri, scop, stype, ssrc, stgt, nPaths, nSegs, countOnly);
  if (countOnly) {
    *nPaths = store.np;
    *nSegs = store.ns;
    return SBR_OK;
  }
  int rc = ensure_path_cache();
  if (rc) return rc;
  const int64_t np = (int64_t)cache.prx.size(), ns = (int64_t)cache.stri.size();
  *nPaths = np;
  *nSegs = ns;
  if (countOnly) return SBR_OK;
  if (ptx) std::copy(cache.ptx.begin(), cache.ptx.end(), ptx);
  std::copy(cache.prx.begin(), cache.prx.end(), prx);
  std::copy(cache.plaunch.begin(), cache.plaunch.end(), plaunch);
  std::copy(cache.pnseg.begin(), cache.pnseg.end(), pnseg);
  std::copy(cache.poff.begin(), cache.poff.end(), poff);
  std::copy(cache.stri.begin(), cache.stri.end(), stri);
  std::copy(cache.scop.begin(), cache.scop.end(), scop);
  std::copy(cache.stype.begin(), cache.stype.end(), stype);
  std::copy(cache.ssrc.begin(), cache.ssrc.end(), ssrc);
  std::copy(cache.stgt.begin(), cache.stgt.end(), stgt);
  return SBR_OK;
}

int Engine::path_tx(int32_t* out, int64_t n) {
  if (!haveResult) return fail(SBR_E_STATE, "no trace result");
  if (n != winners || (!out && n > 0)) return fail(SBR_E_ARG, "path tx: count mismatch");
  if (n == 0) return SBR_OK;
  if (multiWindow) {
    int rc = ensure_path_cache();
    if (rc) return rc;
    std::copy(cache.ptx.begin(), cache.ptx.end(), out);
    return SBR_OK;
  }
  std::vector<int32_t> txrx(n);
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipMemcpy(txrx.data(), dWinRx.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  for (int64_t w = 0; w < n; w++) out[w] = txrx[w] / nrx;
  return SBR_OK;
}

int Engine::path_fields(double* out, int64_t n) {
  if (!haveResult) return fail(SBR_E_STATE, "no trace result");
  if (n != winners || (!out && n > 0)) return fail(SBR_E_ARG, "path fields: count mismatch");
  if (n == 0) return SBR_OK;
  if (multiWindow) {
    int rc = ensure_path_cache();
    if (rc) return rc;
    std::copy(cache.pfield.begin(), cache.pfield.end(), out);
    return SBR_OK;
  }
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipMemcpy(out, dWinField.p, 6 * sizeof(double) * n, hipMemcpyDeviceToHost));
  return SBR_OK;
}

// Grows a device buffer to hold `want` elements, keeping the first `used` (the stream waits for the
// copy before the old buffer is freed).
template <typename T>
static bool grow_keep(DBuf<T>& b, size_t used, size_t want, hipStream_t s) {
  if (want <= b.n && b.p) return true;
  DBuf<T> nb;
  nb.ensure(std::max(want, 2 * b.n));
  if (!nb.p) return false;
  if (used > 0 && (hipMemcpyAsync(nb.p, b.p, used * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess ||
                   hipStreamSynchronize(s) != hipSuccess))
    return false;
  std::swap(b.p, nb.p);
  std::swap(b.n, nb.n);
  return true;   // nb now holds the old buffer and frees it
}

// Appends the current window's kept paths (winners w in dWin*, Wn of them) to the device store:
// rows, segment offsets (scan of the counts) and segments (k_paths).  Enqueued only; the window's
// segment total lands in pinned memory for the host to add after the window's synchronisation.
int Engine::store_window_paths(int64_t Wn, int maxDepth) {
  const int64_t np = store.np, ns = store.ns;
  const size_t segMax = (size_t)(ns + Wn * (int64_t)(maxDepth + 1));
  if (!grow_keep(store.txrx, np, np + Wn, stream) || !grow_keep(store.launch, np, np + Wn, stream) ||
      !grow_keep(store.nseg, np, np + Wn, stream) || !grow_keep(store.off, np, np + Wn, stream) ||
      !grow_keep(store.field, 6 * np, 6 * (np + Wn), stream) || !grow_keep(store.tri, ns, segMax, stream) ||
      !grow_keep(store.cop, ns, segMax, stream) || !grow_keep(store.type, ns, segMax, stream) ||
      !grow_keep(store.src, 3 * ns, 3 * segMax, stream) || !grow_keep(store.tgt, 3 * ns, 3 * segMax, stream))
    return fail(SBR_E_NOMEM, "path store");
  ENSURE(dSegCnt, Wn);
  ENSURE(dSegOff, Wn + 1);
  k_path_meta<<<blocks_for(Wn, 256), 256, 0, stream>>>(Wn, dTx.p, nrx, dWinKey2.p, dWinIdx2.p, dWinRx.p, dWinField.p,
                                                        dSegCnt.p, store.txrx.p + np, store.launch.p + np,
                                                        store.nseg.p + np, store.field.p + 6 * np);
  HIPCHK(hipGetLastError());
  size_t tb = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, dSegCnt.p, dSegOff.p, (int)Wn, stream));
  ENSURE(dScratch, tb);
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(dScratch.p, tb, dSegCnt.p, dSegOff.p, (int)Wn, stream));
  k_path_off<<<blocks_for(Wn, 256), 256, 0, stream>>>(Wn, ns, dSegCnt.p, dSegOff.p, store.off.p + np);
  HIPCHK(hipGetLastError());
  const DSurf* surf = lastParams[0].q1MeshIndexing ? dSurfQ1.p : dSurfId.p;
  k_paths<<<blocks_for(Wn, 128), 128, 0, stream>>>(Wn, nrx, dTx.p, dTree.p, surf, dRx.p, dWinRx.p, dWinIdx2.p, dWinNode.p,
                                                    dSegOff.p, store.tri.p + ns, store.cop.p + ns, store.type.p + ns,
                                                    store.src.p + 3 * ns, store.tgt.p + 3 * ns);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(pinSmall.at<int64_t>(kPinSegTotal), dSegOff.p + Wn, sizeof(int64_t), hipMemcpyDeviceToHost,
                        stream));
  return SBR_OK;
}

// The host copy of the path store, in the order of a single-window trace.
int Engine::ensure_path_cache() {
  if (!multiWindow || cacheReady) return SBR_OK;
  HIPCHK(hipSetDevice(device));
  const int64_t np = store.np, ns = store.ns;
  PathCache c;
  std::vector<int32_t> txrx(np);
  c.ptx.resize(np);
  c.prx.resize(np);
  c.plaunch.resize(np);
  c.pnseg.resize(np);
  c.poff.resize(np);
  c.pfield.resize(6 * np);
  c.stri.resize(ns);
  c.scop.resize(ns);
  c.stype.resize(ns);
  c.ssrc.resize(3 * ns);
  c.stgt.resize(3 * ns);
  if (np > 0) {
    HIPCHK(hipMemcpyAsync(txrx.data(), store.txrx.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(c.plaunch.data(), store.launch.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(c.pnseg.data(), store.nseg.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(c.poff.data(), store.off.p, sizeof(int64_t) * np, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(c.pfield.data(), store.field.p, sizeof(double) * 6 * np, hipMemcpyDeviceToHost, stream));
  }
  if (ns > 0) {
    HIPCHK(hipMemcpyAsync(c.stri.data(), store.tri.p, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(c.scop.data(), store.cop.p, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(c.stype.data(), store.type.p, sizeof(int32_t) * ns, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(c.ssrc.data(), store.src.p, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(c.stgt.data(), store.tgt.p, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, stream));
  }
  HIPCHK(hipStreamSynchronize(stream));
  for (int64_t k = 0; k < np; k++) {
    c.ptx[k] = txrx[k] / nrx;
    c.prx[k] = txrx[k] % nrx;
  }
  cache = std::move(c);
  int rc = finish_path_cache();
  if (rc) return rc;
  cacheReady = true;
  return SBR_OK;
}

// Windows come in launch order and each window's paths are (tx, rx) then launch ordered: a stable
// sort by (tx, rx) gives the single-window order.
int Engine::finish_path_cache() {
  const int64_t np = (int64_t)cache.prx.size();
  std::vector<int64_t> perm(np);
  for (int64_t k = 0; k < np; k++) perm[k] = k;
  std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) {
    return cache.ptx[a] != cache.ptx[b] ? cache.ptx[a] < cache.ptx[b] : cache.prx[a] < cache.prx[b];
  });
  PathCache c;
  c.ptx.resize(np);
  c.prx.resize(np);
  c.plaunch.resize(np);
  c.pnseg.resize(np);
  c.poff.resize(np);
  c.pfield.resize(6 * np);
  int64_t seg = 0;
  for (int64_t k = 0; k < np; k++) {
    const int64_t q = perm[k];
    c.ptx[k] = cache.ptx[q];
    c.prx[k] = cache.prx[q];
    c.plaunch[k] = cache.plaunch[q];
    c.pnseg[k] = cache.pnseg[q];
    c.poff[k] = seg;
    for (int j = 0; j < 6; j++) c.pfield[6 * k + j] = cache.pfield[6 * q + j];
    for (int32_t j = 0; j < cache.pnseg[q]; j++) {
      const int64_t from = cache.poff[q] + j;
      c.stri.push_back(cache.stri[from]);
      c.scop.push_back(cache.scop[from]);
      c.stype.push_back(cache.stype[from]);
      for (int a = 0; a < 3; a++) {
        c.ssrc.push_back(cache.ssrc[3 * from + a]);
        c.stgt.push_back(cache.stgt[3 * from + a]);
      }
    }
    seg += cache.pnseg[q];
  }
  cache = std::move(c);
  return SBR_OK;
}

int Engine::trees(int32_t* launch, int32_t* depth, int64_t* code, int32_t* tri, int32_t* cop, int32_t* type,
                  double* src, double* tgt, double* pl) {
  if (!haveResult) return fail(SBR_E_STATE, "no trace result");
  if (multiWindow) return fail(SBR_E_STATE, "interaction trees are kept for single-window traces only");
  HIPCHK(hipSetDevice(device));
  std::vector<TNode> h(nodes);
  std::vector<TxDesc> txs(ntx);
  if (nodes > 0) {
    HIPCHK(hipMemcpyAsync(h.data(), dTree.p, sizeof(TNode) * nodes, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(txs.data(), dTx.p, sizeof(TxDesc) * ntx, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  std::vector<int64_t> order(nodes);
  for (int64_t i = 0; i < nodes; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    if (h[a].gl != h[b].gl) return h[a].gl < h[b].gl;
    if (h[a].depth != h[b].depth) return h[a].depth < h[b].depth;
    return h[a].code < h[b].code;
  });
  for (int64_t k = 0; k < nodes; k++) {
    const TNode& t = h[order[k]];
    const TxDesc& T = txs[t.txi];
    launch[k] = (int32_t)(T.launchOffset + ((int64_t)t.gl - T.launchStart) * T.launchStride);
    depth[k] = t.depth;
    code[k] = t.code;
    tri[k] = t.tri;
    cop[k] = t.cop;
    type[k] = t.type;
    for (int a = 0; a < 3; a++) { src[3 * k + a] = t.src[a]; tgt[3 * k + a] = t.tgt[a]; }
    pl[k] = t.pl;
  }
  return SBR_OK;
}

int Engine::path_fields_given(const sbr_field_params* P, const sbr_material* m, int nMats, int64_t nPaths,
                              const int64_t* off, const int32_t* cnt, const double* src, const double* tgt,
                              const int32_t* surfId, const int32_t* type, const int32_t* mat, int nSurf,
                              const double* frames, double* out) {
  if (device < 0) return fail(SBR_E_STATE, "host-only context has no device");
  if (!P || nPaths < 0 || nMats < 0 || nSurf < 0) return fail(SBR_E_ARG, "path_fields: bad arguments");
  if (nPaths == 0) return SBR_OK;
  if (!off || !cnt || !src || !tgt || !surfId || !type || !mat || !out || (nSurf > 0 && !frames) || (nMats > 0 && !m))
    return fail(SBR_E_ARG, "path_fields: missing arrays");
  if (P->antennaType < 0 || P->antennaType > 3) return fail(SBR_E_ARG, "path_fields: antennaType must be 0..3");
  if (P->antennaType == 3 && patNT == 0) return fail(SBR_E_STATE, "path_fields: antennaType 3 without a pattern");
  if (!(P->frequency > 0)) return fail(SBR_E_ARG, "path_fields: frequency must be positive");
  // segments are checked on the host: every index the kernel follows stays in range
  int64_t nSeg = 0;
  for (int64_t q = 0; q < nPaths; q++) {
    if (cnt[q] < 1 || cnt[q] > kMaxDepth + 1) return fail(SBR_E_LIMIT, "path_fields: a path has 0 or too many segments");
    if (off[q] < 0) return fail(SBR_E_ARG, "path_fields: negative segment offset");
    nSeg = std::max(nSeg, off[q] + cnt[q]);
  }
  for (int64_t q = 0; q < nPaths; q++)
    for (int k = 0; k < cnt[q]; k++) {
      const int64_t g = off[q] + k;
      if (surfId[g] < -1 || surfId[g] >= nSurf) return fail(SBR_E_ARG, "path_fields: surface id out of range");
      if (type[g] < 0 || type[g] > 2) return fail(SBR_E_ARG, "path_fields: interaction type must be 0, 1 or 2");
      if (type[g] != 0 && (mat[g] < 0 || mat[g] >= nMats))
        return fail(SBR_E_ARG, "path_fields: material id out of range");
    }
  HIPCHK(hipSetDevice(device));
  TxDesc K{};
  for (int k = 0; k < 9; k++) K.txcs[k] = P->txCoordSys[k];
  K.freq = P->frequency;
  K.pt = P->pt;
  K.useFresnel = P->useFresnel ? 1 : 0;
  K.antennaType = P->antennaType;
  K.wallWidth = P->wallWidth;
  K.physicalField = P->physicalField ? 1 : 0;
  std::vector<int32_t> matc(mat, mat + nSeg);
  for (auto& v : matc) v = std::max(v, 0);   // unused for type 0; kept a valid index
  DBuf<TxDesc> dK;
  DBuf<double> dM, dSrc, dTgt, dFr, dOut;
  DBuf<int64_t> dOff;
  DBuf<int32_t> dCnt, dSurfI, dType, dMat;
  ENSURE(dK, 1);
  ENSURE(dM, std::max<size_t>(6 * (size_t)nMats, 6));
  ENSURE(dSrc, 3 * (size_t)nSeg);
  ENSURE(dTgt, 3 * (size_t)nSeg);
  ENSURE(dFr, std::max<size_t>(9 * (size_t)nSurf, 9));
  ENSURE(dOut, 6 * (size_t)nPaths);
  ENSURE(dOff, nPaths);
  ENSURE(dCnt, nPaths);
  ENSURE(dSurfI, nSeg);
  ENSURE(dType, nSeg);
  ENSURE(dMat, nSeg);
  HIPCHK(hipMemcpyAsync(dK.p, &K, sizeof K, hipMemcpyHostToDevice, stream));
  if (nMats > 0) HIPCHK(hipMemcpyAsync(dM.p, m, 6 * sizeof(double) * nMats, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dSrc.p, src, 3 * sizeof(double) * nSeg, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dTgt.p, tgt, 3 * sizeof(double) * nSeg, hipMemcpyHostToDevice, stream));
  if (nSurf > 0) HIPCHK(hipMemcpyAsync(dFr.p, frames, 9 * sizeof(double) * nSurf, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dOff.p, off, sizeof(int64_t) * nPaths, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dCnt.p, cnt, sizeof(int32_t) * nPaths, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dSurfI.p, surfId, sizeof(int32_t) * nSeg, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dType.p, type, sizeof(int32_t) * nSeg, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dMat.p, matc.data(), sizeof(int32_t) * nSeg, hipMemcpyHostToDevice, stream));
  FieldArgs F{};
  F.winners = nPaths;
  F.pat = dPattern.p;
  F.patNT = patNT;
  F.patNP = patNP;
  k_field_paths<<<blocks_for(nPaths, 256), 256, 0, stream>>>(F, dK.p, dM.p, nPaths, dOff.p, dCnt.p, dSrc.p, dTgt.p,
                                                             dSurfI.p, dType.p, dMat.p, dFr.p, dOut.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dOut.p, 6 * sizeof(double) * nPaths, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SBR_OK;
}

int Engine::fresnel_sweep(int n, double freq, const double* epsI, const double* epsT, double width, int infWall,
                          double* theta, double* out) {
  if (device < 0) return fail(SBR_E_STATE, "host-only context has no device");
  if (n <= 0 || !(freq > 0) || !epsI || !epsT || !theta || !out) return fail(SBR_E_ARG, "fresnel_sweep: bad arguments");
  HIPCHK(hipSetDevice(device));
  DBuf<double> dT, dO;
  ENSURE(dT, n);
  ENSURE(dO, 8 * (size_t)n);
  cx ei, et;
  ei.r = epsI[0];
  ei.i = epsI[1];
  et.r = epsT[0];
  et.i = epsT[1];
  k_coeff_sweep<<<blocks_for(n, 256), 256, 0, stream>>>(n, SBR_C0 / freq, ei, et, width, infWall, dT.p, dO.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(theta, dT.p, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(out, dO.p, 8 * sizeof(double) * n, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return SBR_OK;
}

}  // namespace sbr
