// Internal interface of the gfx950 engine behind the C-ABI (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "capture_prefilter.h"
#include "sbr_capi.h"
#include "sbrsolver/sbr_math.h"

namespace sbr {

constexpr int kMaxTess = 11;        // icosahedron subdivision levels accepted (level 11: 41.9 M launches)
constexpr int kMaxDepth = 24;       // reflections + transmissions per path (tree depth)
constexpr int kMaxStack = 64;       // BVH traversal stack (tree height bound, checked on upload)
constexpr int kBlock = 256;         // threads per block of the traversal kernels
constexpr size_t kSceneLdsMax = 96 * 1024;   // scenes up to this size are staged in LDS
constexpr size_t kLdsPerCU = 160 * 1024;     // LDS of a gfx950 compute unit
constexpr int kTraceBlocksPerCU = 4;         // bounce-kernel blocks per CU its registers allow (127 VGPRs)
// Tree nodes per launch window: 2^28 (34 GB of records, about 44 GB with the queues, candidate list
// and keys; 2^27 measured 12 % slower on the dense indoor trace, 28 windows against 12, and 3 % on
// the 256-transmitter batch).
constexpr size_t kWindowNodes = size_t(1) << 28;

// ---- scene, device layout ---------------------------------------------------------------------
// BVH node in depth-first order (children[0] subtree first), 64 B.
struct alignas(64) DNode {
  double bmin[3];
  double bmax[3];
  int32_t a;      // inner: first child (DFS index); leaf: first triangle (DFS index)
  int32_t b;      // inner: second child; leaf: unused
  int32_t count;  // leaf: triangle count (>0); inner: 0
  int32_t sub;    // leaf with more than kSubLeaf triangles: root of its leaf BVH (DSub index), else -1
};
// Node of a leaf BVH, 32 B.  Large leaves of the reference tree (a flat, tessellated floor ends as
// one leaf of thousands of triangles: its box has no surface area, so the SAH never splits it) get
// a BVH of their own over single-precision pre-test boxes.  It only prunes what provably cannot hit
// or cannot beat the best hit, so the leaf's result is unchanged.
struct alignas(32) DSub {
  float lo[3];
  float hi[3];
  int32_t first;  // inner: first child (the second follows); leaf: first entry of subTri
  int32_t count;  // leaf: triangle count (>0); inner: 0
};
// Reference-tree node as the single-precision traversal reads it, 128 B: the boxes of both children
// stored in the parent, each as the exact box grown by 2 delta (a ray whose computed interval of it
// is empty misses the exact box by more than delta, where the reference's FP64 slab test cannot
// pass) and shrunk by delta (a non-empty computed interval puts the line at depth delta / 2 inside
// the exact box, where the FP64 test cannot fail).  Only rays in between take the FP64 test.
constexpr int32_t kLeafBit = 1 << 30;   // work item of a leaf: DFS node index | kLeafBit
struct DChildF {
  int32_t item;    // child's work item (node index, | kLeafBit for a leaf)
  int32_t inner;   // 1: the shrunk box is non-empty
};
struct alignas(128) DNodeF {
  float box[2][12];   // per child: grown lo[3], hi[3], shrunk lo[3], hi[3]
  DChildF ch[2];      // a leaf's slot: {a, b}, {count, sub} of its DNode (no boxes)
  int32_t pad[4];
};
static_assert(sizeof(DNodeF) == 128, "DNodeF layout");
// Closest-hit traversal (sbr_trace_params.exactTraversal, sbr_closest_hit's `exact`): the wide tree
// with the winner confirmed on the reference tree, the reference tree in the reference's order
// (every crossed box), or the reference tree with the clipping cuts.  All give the reference's result.
constexpr int kTravWide = 0;
constexpr int kTravExact = 1;
constexpr int kTravRef = 2;

// Node of the wide traversal tree, 128 B.  It holds up to four nodes of the reference tree from at
// most two levels below its own: starting from the node's two children, the inner child with the
// largest box surface is replaced by its two children until four are held.  Each child's box is the
// exact box grown by 2 delta and rounded outward (a ray whose computed interval of it is empty
// misses the exact box by more than delta), as single-precision bounds per axis; an empty slot
// has NaN bounds, which no interval test passes (and item -1, which the traversal checks too).  A
// reference leaf with a leaf BVH is a subtree of such nodes over the leaf BVH's nodes (their boxes:
// unions of the triangles' padded pre-test boxes).  The wide tree only finds the triangles a ray may
// hit; whether the reference's traversal reaches the winner's leaf is checked on the reference
// tree itself (DVer).
constexpr int kWide = 4;   // children per wide node (4: 128 B nodes; 8, 256 B, measured slower)
struct alignas(128) DNode4 {
  float lo[3][kWide];   // lo[axis][child]
  float hi[3][kWide];
  int32_t item[kWide];  // inner: DNode4 index; leaf: kLeafBit | DLeaf index; empty slot: -1
  int32_t pad[kWide == 4 ? 4 : 8];
};
static_assert(sizeof(DNode4) == (kWide == 4 ? 128 : 256), "DNode4 layout");
// Leaf of the wide tree: a reference-tree leaf (its triangles first .. first + count - 1 in DFS
// order), or a leaf of a leaf BVH (the wide tree holds leaf BVHs as subtrees of its own; its
// triangles are subTri[first .. first + count - 1]).
struct alignas(16) DLeaf {
  int32_t first;
  int32_t count;
  int32_t listed;  // 1: a leaf-BVH leaf (triangles listed in subTri)
  int32_t pad;
};
// Reference-tree node for the check of a winner's leaf, 32 B: the exact box shrunk by delta and
// rounded inward (a non-empty computed interval puts the line at depth delta / 2 inside the exact
// box and so inside the box of every ancestor, whose FP64 slab tests then cannot fail), and the
// parent.
struct alignas(32) DVer {
  float lo[3];
  float hi[3];
  int32_t parent;   // -1 at the root
  int32_t shrunk;   // 1: the shrunk box is non-empty and implies the ancestors' tests
};
// The wide tree as the kernels take it.
struct WideTree {
  const DNode4* n4;
  const DLeaf* leaves;
  const DVer* ver;
};
constexpr int kSubLeaf = 16;      // leaves above this size get a leaf BVH (16 against 64: dense indoor
                                  // bounce kernel -11 %, urban and city unchanged; 8: urban +8 %)
// Probes of the open-addressing tables (sequence trie, groups) before a table counts as full: at the
// 70 % load that also restarts a trace the expected run is under ten slots.
constexpr unsigned long long kMaxProbe = 4096;
constexpr int kRecParts = 1024;   // partitions of the capture record buffer (one append counter each)
constexpr int kRecStride = 16;    // counters 128 B apart
constexpr int kSubLeafSize = 4;   // triangles per leaf-BVH leaf
// Triangle as the traversal reads it, in leaf (DFS) order, 96 B.
struct alignas(32) DTri {
  double v1[3];
  double e1[3];   // v2 - v1
  double e2[3];   // v3 - v1
  int32_t id;     // triangleId
  int32_t cop;    // coplanarId
  int32_t mat;    // materialId
  int32_t leaf;   // the reference-tree leaf holding the triangle (DFS node index)
  double pad2;
};
// Conservative single-precision box (rounded outward and padded), 32 B: a ray that provably misses
// it misses the exact FP64 box, so it serves as a cheap pre-test in front of the exact tests.
struct alignas(32) DBox32 {
  float lo[3];
  float hi[3];
  float pad[2];
};
// Surface data addressed by a path's hitSurfaceID (image points, re-intersection, field frame).
struct alignas(64) DSurf {
  double v1[3];
  double e1[3];
  double e2[3];
  double n[3];
  double cs[9];   // coordSys, row-major
  double pad[3];
};

// ---- transmitters --------------------------------------------------------------------------------
struct TxDesc {
  double tx[3];
  double txcs[9];
  double freq, pt, rcoef;
  const double* rays;        // launch directions of this transmitter's icosahedron level
  int64_t launchStart;       // global index of its first root item
  int64_t nLaunch;           // root items (selected launches)
  int32_t launchStride, launchOffset;
  int32_t maxR, maxT, maxTotal;
  int32_t useFresnel, antennaType, receptionModel;
  double capK;               // rcoef / (1 - rcoef) rounded up (INFINITY for rcoef >= 1): capture prefilter
  double wallWidth;          // > 0: slab walls of this thickness (physical mode), 0: reference walls
  int32_t typedDedup;        // 1: duplicate-removal groups also keyed by the interaction types
  int32_t checkOcclusion;    // 1: corrected multi-bounce paths must be unobstructed
  int32_t longMiss;          // 1: a miss after a bounce extends 1e10 m (0: a 100 m stub)
  int32_t physicalField;     // 1: s/p transport in the plane of incidence (0: reference rotations)
};

// ---- interaction tree ---------------------------------------------------------------------------
// One RayLaunch call (one bounce), 128 B in two 64 B halves.  The first half is all that the
// receiver capture reads for a node and its ancestors; the second serves the bounce loop, the dedup
// key, the path export and the field.  The hit triangle's material is tris[hitDfs].mat; the launch
// index within the transmitter follows from gl (see TxDesc).
struct alignas(128) TNode {
  // capture half
  double src[3];
  double tgt[3];
  double base;       // length the reception radius grows with: the totalPathLength RayCapture hands
                     // this node (reference sphere) or the true length from the transmitter to src
  int32_t parent;    // node index of the parent, -1 for a root
  uint8_t depth;
  uint8_t pad0;
  int16_t txi;       // transmitter index in the batch
  // bounce half
  double dir[3];     // direction the segment was traced with
  double pl;         // pathLength as stored by the reference (cumulative; INF for a miss)
  int32_t gl;        // global root index of the tree (transmitters' root items concatenated)
  int32_t tri;       // hitSurfaceID, -1 on a miss
  int32_t cop;       // hitCoplanarId, -1 on a miss
  int32_t hitDfs;    // DFS index of the hit triangle, -1 on a miss
  uint32_t code;     // branch code: bit per level, transmit = 0, reflect = 1, first level most significant
  uint8_t type;      // 0 from the transmitter, 1 reflected, 2 transmitted
  uint8_t rc, tc;    // reflections / transmissions so far
  uint8_t mask;      // children created: bit 0 transmitted, bit 1 reflected
  int32_t childT;    // node index of the transmitted child, -1 if none
  int32_t childR;    // node index of the reflected child, -1 if none
};
static_assert(sizeof(TNode) == 128, "TNode layout");

// What the capture's ancestor rule reads of a tree node, 160 B, held in LDS by the wave-per-node
// kernel for the ancestors of its node: the capture half (src, tgt, base, and the hit ids packed in
// one word), the single-precision pre-test of its reception sphere, and the box its sphere can
// reach.
struct alignas(16) AncRec {
  double cap[8];
  SphereF f;
  float pad;
  double box[6];
};
static_assert(sizeof(AncRec) == 160, "AncRec layout");

// A node left to the wave-per-node capture, 160 B: what the thread-per-node pass computed for it
// (the segment clipped to the receivers' region, the search radius and rows, the sphere pre-test)
// and its nearest ancestors (node indices, nearest first; nAnc of them, the walk continuing at qRest).
constexpr int kJobAnc = 8;
struct alignas(16) HeavyJob {
  double P0[3], P1[3];
  double R, invDy;
  SphereF sf;
  int32_t node;
  int32_t j0, j1;
  int32_t flags;   // bit 0: every cell is searched (no radius bound); bits 1..: nAnc
  int32_t qRest;
  int32_t anc[kJobAnc];
};
static_assert(sizeof(HeavyJob) == 160, "HeavyJob layout");

// Per-depth work bookkeeping on the device.
struct LevelRec {
  unsigned long long base;   // node index of the level's first node
  unsigned long long count;  // items (nodes) of the level
};

// Receiver in cell order of the receivers' grid, 32 B: position and index.
struct alignas(32) RxCell {
  double p[3];
  int32_t r;
  int32_t pad;
};

struct RxGrid {
  double bmin[3], bmax[3];
  double gx0, gy0, cell;
  int32_t nx, ny;
};

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;  // capacity in elements
  void ensure(size_t want);
  void release();
  ~DBuf() { release(); }
};

// Page-locked host memory (hipHostMalloc) for the per-trace copies, so that they go by DMA without a
// pageable bounce.  Grown only while the stream is idle (at the start of a trace).
struct HostPinned {
  unsigned char* p = nullptr;
  size_t n = 0;  // bytes
  HostPinned() = default;
  HostPinned(const HostPinned&) = delete;
  HostPinned& operator=(const HostPinned&) = delete;
  bool ensure(size_t bytes);
  void release();
  ~HostPinned() { release(); }
  template <typename T>
  T* at(size_t off) { return reinterpret_cast<T*>(p + off); }
};
// Offsets in the small staging area: bookkeeping read back or sent per launch window.
constexpr size_t kPinLevels = 0;         // LevelRec[kMaxDepth + 2] read back
constexpr size_t kPinLevels0 = 1024;     // LevelRec[kMaxDepth + 2] sent
constexpr size_t kPinTraceCtr = 2048;    // 8 counters of the bounce kernel
constexpr size_t kPinCapCtr = 2176;      // capture counters (10), record total and largest partition, listed nodes
constexpr size_t kPinSegTotal = 2304;    // segments of a window's kept paths (multi-window traces)
constexpr size_t kPinSmall = 4096;

// Launch configuration of the traversal kernel (sbr_set_tuning, or chosen by sbr_autotune).
struct Tuning {
  int ldsScene = -1;      // -1: LDS staging when it fits and keeps occupancy, 1: whenever it fits, 0: never
  int blocksPerCU = 0;    // persistent blocks per CU; 0: occupancy maximum
  int64_t windowNodes = 0;   // tree nodes per launch window; 0: kWindowNodes
  int stackLds = 0;       // traversal stack entries per ray in LDS; 0: what leaves room for the blocks per CU
  int oversub = 0;        // bounce-kernel blocks per resident block; 0: kTraceOversub
  int64_t tableEntries = 0;   // starting size of the duplicate-removal tables; 0: sized from the last trace
};

// Kept paths gathered window by window when a trace needs several launch windows.
struct PathCache {
  std::vector<int32_t> ptx, prx, plaunch, pnseg, stri, scop, stype;
  std::vector<int64_t> poff;
  std::vector<double> ssrc, stgt, pfield;
};
// Kept paths of a multi-window trace, appended window by window on the device (a window's paths in
// (tx, rx), then launch order; segments at off).  Copied to the host cache, in the order of a
// single-window trace, when first asked for.
struct PathStore {
  DBuf<int32_t> txrx, launch, nseg, tri, cop, type;
  DBuf<int64_t> off;
  DBuf<double> field, src, tgt;
  int64_t np = 0, ns = 0;
};

struct Engine {
  int device = 0;
  Tuning tune;
  hipStream_t stream = nullptr;
  hipStream_t sideStream = nullptr;   // line of sight, concurrent with the bounce loop
  hipEvent_t evSide[2] = {nullptr, nullptr};   // receivers uploaded; line of sight done
  std::string err;
  int numCU = 256;

  // scene (host copies for queries)
  int ntri = 0, nnode = 0, ncop = 0, treeHeight = 0;
  // Traversal stack entries per ray: one deferred sibling per internal node of the current path.
  // A leaf BVH continues the path below its leaf, adding its own internal levels.
  // The wide traversal defers up to three siblings per wide node of its path (wideStack).
  int stackDepth() const { return std::max(1, std::max(treeHeight - 1, wideStack) + subHeight); }
  int wideStack = 0;
  size_t sceneBytes = 0, wideBytes = 0;
  std::vector<DNode> hNodes;
  std::vector<DTri> hTris;
  std::vector<double> hTriN;                // normals in DFS order
  std::vector<DSurf> hSurfQ1, hSurfId;      // by hitSurfaceID: mesh-vector mapping / identity
  std::vector<int32_t> hCop, hMat, hMeshTriId;
  std::vector<double> hVerts, hNorms;       // by triangleId
  DBuf<DNode> dNodes;
  std::vector<DNodeF> hNodesF;
  DBuf<DNodeF> dNodesF;
  DBuf<DTri> dTris;
  std::vector<DNode4> hNodes4;     // wide traversal tree (root at 0)
  std::vector<DLeaf> hLeaves;      // its leaves
  std::vector<DVer> hVer;          // per reference-tree node: the winner check
  DBuf<DNode4> dNodes4;
  DBuf<DLeaf> dLeaves;
  DBuf<DVer> dVer;
  std::vector<DBox32> hTriBox32;   // single-precision pre-test box per DFS triangle
  DBuf<DBox32> dTriBox32;
  std::vector<DSub> hSub;          // leaf BVHs of large leaves
  std::vector<int32_t> hSubTri;    // their triangles (DFS indices)
  DBuf<DSub> dSub;
  DBuf<int32_t> dSubTri;
  int subHeight = 0;               // height of the tallest leaf BVH
  float sceneExtent = 1.0f, boxDelta = 1e-3f;
  DBuf<double> dTriN;
  DBuf<DSurf> dSurfQ1, dSurfId;
  std::vector<double> mats;                 // 6 per material
  DBuf<double> dMats;
  // tabulated transmitter pattern (antennaType 3): the rows as given (4 floats per grid point) and
  // the device table of field coefficients (4 doubles per grid point)
  std::vector<float> hPattern;
  DBuf<double> dPattern;
  int32_t patNT = 0, patNP = 0;
  bool haveScene = false;

  // launch directions per icosahedron level (generated once, kept)
  std::map<int, DBuf<double>*> rayCache;
  DBuf<int32_t> dFaces[2], dAdj[2];
  DBuf<int32_t> dFlag, dScan;
  DBuf<double> dVerts;

  // per-trace buffers
  DBuf<TxDesc> dTx;
  DBuf<TNode> dTree;
  DBuf<int32_t> dQueue[2];       // (parent << 1) | reflect
  DBuf<unsigned long long> dQCnt;   // per depth: the queue partitions' append counters
  DBuf<int2> dStackG;            // traversal stack entries beyond the LDS part (tall trees)
  DBuf<LevelRec> dLevels;
  DBuf<uint8_t> dScratch;        // library temp storage (scan / sort)
  DBuf<double> dRx;
  HostPinned pinSmall, pinIn, pinOut;   // bookkeeping; receivers in; fields out
  DBuf<int32_t> dCellStart;
  DBuf<RxCell> dCellRx;
  DBuf<uint8_t> dLos;
  DBuf<unsigned long long> dTabA, dSeqTab, dTabOrder;   // group table keys, sequence trie, order keys
  DBuf<int32_t> dTabRx;
  DBuf<HeavyJob> dHeavy;         // nodes left to the wave-per-node capture
  DBuf<int32_t> dCand;           // nodes whose segment may reach the receivers' capture region
  DBuf<uint8_t> dCandFlag;       // per node: 1 if listed (written by the bounce kernel)
  DBuf<int32_t> dTileCount, dTileOffset;   // compaction of the flags: per-tile counts and offsets
  DBuf<int2> dRec;               // (node, receiver) records of the capture's candidate pass
  DBuf<unsigned long long> dRecCount, dRecPrefix;   // per-partition record counts / prefix
  DBuf<unsigned long long> dNodeKey;   // dedup and order keys of the nodes with records
  DBuf<double> dImg;             // image source of each tree node (reference-sphere traces)
  size_t recCap = 0;
  size_t heavyCap = 0;   // entries of the heavy-job list (grown when a capture pass fills it)
  DBuf<unsigned long long> dCounters;        // capture: records, corrected, groups, flags, collected, heavy
  DBuf<unsigned long long> dTraceCounters;   // bounce loop: overflow flags, node visits, triangle tests
  DBuf<unsigned long long> dWinKey, dWinKey2;
  DBuf<int32_t> dWinRx, dWinRx2, dWinIdx, dWinIdx2, dWinNode;
  DBuf<double> dWinField;
  DBuf<double> dField;

  // last result
  bool haveResult = false;
  std::vector<sbr_trace_params> lastParams;
  sbr_stats stats{};
  int nrx = 0, ntx = 0;
  int64_t nodes = 0;
  int64_t capNodes = 0;          // of them listed for the capture (known after the capture pass)
  int64_t winners = 0;
  size_t nodeCap = 0;
  size_t tabCapHint = 0;   // dedup table size of the last trace
  int64_t tabHintOut = 0, tabHintLaunches = 0;   // its (transmitter, receiver) outputs and launches
  double perRootHint = 0;  // tree nodes per root item of the last trace's densest window (first window size)
  size_t hFieldN = 0;   // doubles of the last trace's receiver fields, held in pinOut
  const double* field_host() const { return reinterpret_cast<const double*>(pinOut.p); }
  RxGrid grid{};

  ~Engine();
  int fail(int code, const std::string& msg) { err = msg; return code; }
  int init(int dev);
  int upload(const sbr_scene_desc* d);
  // Triangles, coplanar ids and a linear BVH built on the device from STL-style facets, then uploaded
  // as upload() does (ingest.hip).  msOut (may be null): [0] until the description is on the host,
  // [1] including the upload.
  int ingest_soup(const float* coords9, const float* normals3, const int32_t* matIds, int nf, double* msOut);
  int clone_scene(const Engine& src);   // scene and materials of another engine (device copies made here)
  int set_materials(const sbr_material* m, int n);
  int set_pattern(const float* gainInfo, int nTheta, int nPhi);
  int rays_for(int tess, const double** out, int64_t* nOut);   // cached launch directions
  int trace(const sbr_trace_params* p, int ntx, const double* rx, int m);
  int closest_hit(const double* o, const double* d, int n, int exact, int32_t* tri, double* pt, double* dist);
  // closest_hit with the rays staged in the library's fixed device array (no allocation).
  int closest_hit_noalloc(const double* o, const double* d, int n, int exact, int32_t* tri, double* pt, double* dist);
  unsigned char* scratchPtr = nullptr;   // device address of that array on this engine's device
  int paths(int32_t* ptx, int32_t* prx, int32_t* plaunch, int32_t* pnseg, int64_t* poff, int32_t* stri,
            int32_t* scop, int32_t* stype, double* ssrc, double* stgt, int64_t* nPaths, int64_t* nSegs,
            bool countOnly);
  int trees(int32_t* launch, int32_t* depth, int64_t* code, int32_t* tri, int32_t* cop, int32_t* type, double* src,
            double* tgt, double* pl);
  int path_tx(int32_t* out, int64_t n);
  int path_fields(double* out, int64_t n);
  // FieldForPath on the device for caller-given paths (sbr_path_fields).
  int path_fields_given(const sbr_field_params* p, const sbr_material* m, int nMats, int64_t nPaths, const int64_t* off,
                        const int32_t* cnt, const double* src, const double* tgt, const int32_t* surfId,
                        const int32_t* type, const int32_t* mat, int nSurf, const double* frames, double* out);
  // Reflection / transmission coefficient sweep over incidence angles (sbr_fresnel_sweep).
  int fresnel_sweep(int n, double freq, const double* epsI, const double* epsT, double width, int infWall,
                    double* theta, double* out);
  // Device bytes held between traces: the per-trace buffers (kept at their largest size so that the
  // next trace need not reallocate) and the cached launch directions.  The scene is not counted.
  size_t held_bytes() const;
  // Releases the per-trace buffers and the sizing hints that would grow them back to the last
  // trace's size (the result of the last trace is dropped; the scene and directions stay).
  void trim();

  bool multiWindow = false;   // last trace ran in several launch windows (paths kept in `store`)
  PathStore store;
  PathCache cache;            // host copy of `store`, made on the first request
  bool cacheReady = false;
  DBuf<int64_t> dSegCnt, dSegOff;

 private:
  int generate_rays(int tess, DBuf<double>& out, int64_t* nOut);
  int trace_stage(const std::vector<TxDesc>& txs, int64_t rootBase, int64_t nRoot, int maxDepth, int mode,
                  bool count, const RxGrid* capG, int64_t* need, float* ms);
  WideTree wide_tree() const;
  hipError_t launch_closest(int mode, const double* o, const double* d, int n, int32_t* tri, double* pt, double* dist);
  int paths_window(int32_t* ptx, int32_t* prx, int32_t* plaunch, int32_t* pnseg, int64_t* poff, int32_t* stri,
                   int32_t* scop, int32_t* stype, double* ssrc, double* stgt, int64_t* nPaths, int64_t* nSegs,
                   bool countOnly);
  int store_window_paths(int64_t w, int maxDepth);
  int ensure_path_cache();
  int finish_path_cache();
};

}  // namespace sbr

// Handle behind the C-ABI's sbr_ctx.
struct sbr_ctx {
  sbr::Engine eng;
};
