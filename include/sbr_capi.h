/* C-ABI of the MI355X (gfx950) shoot-bounce-reflect library, libsbr_amd.so.
 *
 * This is the thin layer under the C++ RTSolver / Preprocessor / FieldCompute call surface
 * (include/sbrsolver/).  Plain pointers and sizes only; no C++ exceptions cross it.
 *
 * Conventions
 *   - Status: every entry point returns int, 0 = OK, negative = error class (SBR_E_*); the
 *     message of the last error on a context is returned by sbr_last_error().
 *   - Ownership: the caller owns every host buffer it passes; the context owns every device
 *     buffer and frees it in sbr_destroy().  Host buffers are read during the call only.
 *   - Threading: one context per host thread; a context is not thread-safe.
 *
 * Reference interfaces each entry point replaces (paths relative to SBRSolver_rev2/):
 *   sbr_scene_from_stl     Preprocessor::StlToGeometry + BVH<Triangle>::ConstructBVH
 *                          (Preprocessor.cpp:69-114, BVH/BVH.h:51-55, SBRSolver_rev2.cpp:67-68)
 *   sbr_scene_from_soup    same, from the facet arrays stl_reader delivers (StlReader.h:637-700)
 *   sbr_scene_from_soup_device  the same triangle set (Preprocessor.cpp:69-246, StlReader.h:448-508)
 *                          built on the device, with a linear BVH in place of BVH/BVH.h:135-274
 *   sbr_scene_upload       the BVH<Triangle>& / std::vector<Triangle*>& pair that
 *                          RTSolver::Init (RTSolver.cpp:13-22) and ExecuteRayTracing take
 *   sbr_materials          the MaterialProperties[] of RTSolver::Init (RTSolver.cpp:13-22)
 *   sbr_trace              RTSolver::ExecuteRayTracing (RTSolver.cpp:52-126) through
 *                          FieldCompute::FieldAtReceiver (FieldCompute.cpp:11-25)
 *   sbr_result_fields      RTSolver::_efield (RTSolver.h:49) as SaveFieldAsCsv writes it
 *                          (RTSolver.cpp:692-717)
 *   sbr_result_paths       RTSolver::_rayPaths after ImagePathCorrection + RemoveDuplicatePath
 *                          (RTSolver.h:45, RTSolver.cpp:116-121) as SavePathsAsVtk walks it
 *                          (RTSolver.cpp:532-538)
 *   sbr_generate_rays      GenerateRaysOnIcosahedron (Icosahedron.cpp:80-110)
 *   sbr_closest_hit        BVH<Triangle>::RayIntersects (BVH/BVH.h:276-328)
 *   sbr_path_fields        FieldCompute::FieldForPath (FieldCompute.cpp:71-151), the per-path step
 *                          of FieldCompute::FieldAtReceiver (FieldCompute.cpp:11-25) as the
 *                          FieldCompute class (include/sbrsolver/FieldCompute.h) calls it
 *   sbr_fresnel_sweep      FieldCompute::RefCoeffTest / TransCoeffTest (FieldCompute.cpp:27-69)
 *                          with GetTransAngle / GetTMCoeff / GetTECoeff (FieldCompute.cpp:236-302)
 */
#ifndef SBR_CAPI_H
#define SBR_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBR_OK 0
#define SBR_E_ARG -1       /* invalid argument */
#define SBR_E_STATE -2     /* call out of order (no scene, no result, ...) */
#define SBR_E_HIP -3       /* HIP runtime error */
#define SBR_E_NOMEM -4     /* device memory exhausted or over the budget */
#define SBR_E_IO -5        /* file could not be read / parsed */
#define SBR_E_LIMIT -6     /* a fixed limit was exceeded (tree depth, hash table, stack) */
#define SBR_E_INTERNAL -7  /* internal consistency check failed */

typedef struct sbr_ctx sbr_ctx;

/* MaterialProperties (common/DataStructures.h:20-28), same field order and layout. */
typedef struct {
  double frequency;
  double transmissionLoss;
  double reflectionLoss;
  double relPermittivityRe;
  double relPermittivityIm;
  double relConductivity;
} sbr_material;

/* Flattened scene as sbr_scene_upload takes it.  Triangles are addressed by triangleId
 * (0..n_tri-1); the BVH lists nodes in the reference's creation order with their triangle lists.
 *   tri_v        [n_tri][9]  v1, v2, v3 (FP64)
 *   tri_n        [n_tri][3]  normal stored in the triangle (the STL file normal)
 *   tri_cs       [n_tri][9]  surface frame coordSys, row-major (Triangle.cpp:37-49)
 *   tri_cop      [n_tri]     coplanar id
 *   tri_mat      [n_tri]     material id
 *   mesh_tri_id  [n_tri]     triangleId of the triangle at position k of the caller's mesh
 *                            vector (the reference indexes that vector by triangleId, Q1)
 *   node_box     [n_node][6] min xyz, max xyz
 *   node_child   [n_node][2] children (-1 for a leaf)
 *   node_first, node_count   leaf triangle list as a slice of leaf_tri
 *   leaf_tri     [n_leaf_tri] triangle ids */
typedef struct {
  int32_t n_tri;
  const double* tri_v;
  const double* tri_n;
  const double* tri_cs;
  const int32_t* tri_cop;
  const int32_t* tri_mat;
  const int32_t* mesh_tri_id;
  int32_t n_node;
  const double* node_box;
  const int32_t* node_child;
  const int32_t* node_first;
  const int32_t* node_count;
  int32_t n_leaf_tri;
  const int32_t* leaf_tri;
} sbr_scene_desc;

/* One transmitter and the trace options (RTSolver::ExecuteRayTracing arguments + extensions). */
typedef struct {
  double tx[3];
  double txCoordSys[9];  /* row-major Mat3 */
  int32_t maxR;          /* maxReflectionCount */
  int32_t maxT;          /* maxTransmissionCount */
  int32_t maxTotal;      /* extension: cap on reflections + transmissions; <= 0: none (reference) */
  double frequency;
  double pt;
  int32_t useFresnel;    /* useFresnelCoeff */
  int32_t tess;          /* txTesslation: icosahedron subdivision level */
  int32_t antennaType;   /* 0 isotropic, 1 Hertzian (reference default), 2 half-wave dipole,
                            3 the context's tabulated pattern (sbr_set_tx_pattern) */
  int32_t receptionModel;/* 0 reference sphere (RTSolver.cpp:838-841); 1 alpha*L/sqrt(3) */
  int32_t launchStride;  /* trace launches i with i % stride == offset (1 = all) */
  int32_t launchOffset;
  int32_t q1MeshIndexing;/* 1 (reference): surface data of mesh[hitSurfaceID]; 0: by id */
  int32_t exactTraversal;/* closest-hit traversal, all with the reference's result: 0 the wide tree
                            (four children per node, single-precision boxes), its winner's leaf
                            confirmed by the reference tree's own FP64 box tests (the ray is redone on
                            the reference tree otherwise); 1 the reference tree, every box the line
                            touches (reference order); 2 the reference tree with the clipping cuts.
                            One traversal per call: 1 if any transmitter asks for it, else 2 if any
                            does, else 0 */
  int32_t countTraversal;/* 1: count BVH node visits and triangle tests (sbr_stats) */
  double wallWidth;      /* extension: > 0 with useFresnel = 1: every wall is a slab of this thickness
                            (m) with the ITU-R P.2040 slab coefficients, so transmissions carry field;
                            0: the reference's infinite wall (reflection Gamma, transmission 0) */
  int32_t typedDedup;    /* extension: 1 keys duplicate removal by the coplanar ids and the interaction
                            types (reflection / transmission) of a path; 0: coplanar ids only, as the
                            reference (RTSolver.cpp:435-459, 875-883) */
  int32_t checkOcclusion;/* extension: 1 drops a corrected multi-bounce path when a scene triangle lies
                            on one of its segments (closest hit nearer than the segment's end); 0:
                            no such test, as the reference (RTSolver.cpp:330-373) */
  int32_t longMiss;      /* extension: 1 extends a ray that misses the scene after a bounce to 1e10 m,
                            as a missing launch ray; 0: a 100 m stub (RTSolver.cpp:224-256), so no
                            receiver beyond it captures the ray */
  int32_t physicalField; /* extension: 1 carries the field through each interaction in the global frame,
                            split into the components perpendicular (s) and parallel (p) to the plane
                            of incidence with the incidence angle in [0, 90] degrees; 0: the
                            reference's rotations through the surface frames (FieldCompute.cpp:71-151) */
} sbr_trace_params;

/* Counters and stage times of the last trace. */
typedef struct {
  int64_t launches;
  int64_t bounces;          /* RayLaunch calls = interaction-tree nodes */
  int64_t capture_records;  /* (node, receiver) candidates handed to duplicate removal: captured by the
                               node and by no ancestor; in the wave-per-node capture, which corrects
                               before it applies the ancestor rule, only those whose image correction
                               also succeeds */
  int64_t corrected;        /* records that survived image correction */
  int64_t winners;          /* paths kept after duplicate removal */
  int64_t max_depth;
  double ms_raygen, ms_trace, ms_capture, ms_field, ms_total;  /* stage times (HIP events) */
  double ms_k_trace;        /* summed duration of the closest-hit kernel launches (one per depth) */
  int64_t n_k_trace;        /* number of those launches */
  double ms_k_capture;      /* duration of the capture kernel */
  int64_t node_visits;      /* BVH nodes entered (countTraversal = 1) */
  int64_t tri_tests;        /* ray-triangle tests (countTraversal = 1) */
  int32_t trace_lds;        /* 1 if the traversal staged the scene in LDS */
  int32_t trace_blocks_per_cu;
  int64_t heavy_nodes;      /* tree nodes captured wave-cooperatively (large candidate regions) */
  int64_t windows;          /* launch windows the trace ran in (tree buffers bounded by window_nodes) */
  int64_t capture_tests;    /* reception-sphere tests of the capture's candidate pass (countTraversal) */
  int64_t box_tests_fp64;   /* child boxes the traversal decided with the FP64 slab test (countTraversal) */
  int64_t capture_nodes;    /* tree nodes whose segment reaches the receivers' capture region */
  int64_t wide_redone;      /* bounces whose wide-tree winner the reference tree's box tests did not
                               confirm, redone on the reference tree (countTraversal) */
  int64_t table_entries;    /* size of the duplicate-removal tables the trace finished with */
  int64_t table_restarts;   /* restarts of the trace with larger tables (a table filled up) */
} sbr_stats;

/* Launch configuration of the traversal kernel (no reference counterpart).  lds_scene: -1 stage
 * the scene in LDS when it fits and does not lower the resident blocks per CU, 1 whenever it
 * fits, 0 never; blocks_per_cu: persistent blocks per compute unit, 0 = the occupancy maximum;
 * window_nodes: interaction-tree nodes held on the device at once (128 B each), 0 = 2^28; smaller
 * values than 16384 are raised to 16384.  A trace whose trees need more runs in several launch
 * windows.  stack_lds: traversal stack entries per ray held in LDS (1..64), the deeper ones in
 * device memory; 0 = as many as leave LDS room for the blocks per CU the kernel's registers allow.
 * oversub: blocks of the bounce kernel's grid per resident block (1..64; the chunks of 256 items
 * that the hardware dispatcher hands to freed block slots), 0 = 8.  table_entries: starting size of
 * the duplicate-removal hash tables (rounded up to a power of two, at least 1024; a table that
 * fills up restarts the trace with one four times larger), 0 = sized from the last trace: its final
 * size, scaled down with the trace when both the (transmitter, receiver) outputs and the launches
 * are under a sixteenth of the last trace's, and at least 4 entries per output and 65,536.  Results
 * do not depend on any of them. */
typedef struct {
  int32_t lds_scene;
  int32_t blocks_per_cu;
  int64_t window_nodes;
  int32_t stack_lds;
  int32_t oversub;
  int64_t table_entries;
} sbr_tuning;

/* Autotuner (no reference counterpart): times the traversal kernel on a sample of the given trace
 * (every launch_stride-th launch, chosen so that about 2^18 launches run) for each launch
 * configuration -- scene staged in LDS or not and persistent blocks per CU from the occupancy
 * maximum down, then for the best of those the LDS stack entries and the grid oversubscription --
 * keeps the fastest, sets it on the context and returns it in *chosen.  The last trace result of
 * the context is replaced by the last trial. */
int sbr_autotune(sbr_ctx* ctx, const sbr_trace_params* params, const double* rx, int32_t n_rx, sbr_tuning* chosen);

/* device >= 0: HIP device ordinal.  device == -1: host-only context (scene preparation and
 * scene queries work; device entry points return SBR_E_STATE). */
int sbr_create(int device, sbr_ctx** out);
int sbr_destroy(sbr_ctx* ctx);
const char* sbr_last_error(const sbr_ctx* ctx);
int sbr_version(void);

int sbr_scene_from_stl(sbr_ctx* ctx, const char* path);
int sbr_scene_from_soup(sbr_ctx* ctx, const float* coords9, const float* normals3, const int32_t* material_ids,
                        int32_t n_facets);
int sbr_scene_upload(sbr_ctx* ctx, const sbr_scene_desc* desc);
/* Scene ingestion on the device for large scenes (no reference counterpart as one call): the
 * triangle set of sbr_scene_from_soup (vertex dedup, degenerate facets dropped, file normals,
 * coplanar ids; identical to the host's) built with sorts and a union-find on the device, and a
 * linear BVH (Morton order of the triangle centres, ties by triangle id, leaves of at most 4
 * triangles) instead of the reference's SAH.  The mesh vector keeps triangle-id order.  Traces run
 * the reference's algorithm over this tree (sbr_scene_bvh_dfs exports it).  ms_out[2] (may be
 * NULL): milliseconds until the scene description was built, and including the upload. */
int sbr_scene_from_soup_device(sbr_ctx* ctx, const float* coords9, const float* normals3, const int32_t* material_ids,
                               int32_t n_facets, double* ms_out);
/* Scene as stored: per triangleId coplanar id, material id, vertices, normal; mesh order. */
int sbr_scene_info(sbr_ctx* ctx, int32_t* n_tri, int32_t* n_node, int32_t* n_coplanar);
int sbr_scene_triangles(sbr_ctx* ctx, int32_t* coplanar, int32_t* material, double* verts9, double* norm3,
                        int32_t* mesh_tri_id);
/* BVH in DFS order as the device holds it: node_box [6], node_a/node_b (children, or
 * first/count for a leaf marked by node_leaf=1), leaf triangle ids in DFS order. */
int sbr_scene_bvh_dfs(sbr_ctx* ctx, double* node_box, int32_t* node_a, int32_t* node_b, int32_t* node_leaf,
                      int32_t* tri_dfs_id);

int sbr_materials(sbr_ctx* ctx, const sbr_material* mats, int32_t count);
/* Tabulated transmitter pattern for antennaType 3 (extension: the reference reads TXPattern.dat
 * into a GainMap, Preprocessor.cpp:3-35, but FieldCompute.cpp:117 always uses the analytic
 * pattern).  gain_info holds n_theta * n_phi rows of GainInfo (DataStructures.h:32-37): theta gain
 * dB, theta phase deg, phi gain dB, phi phase deg; row i * n_phi + j is theta = i * 180 / (n_theta - 1)
 * and phi = j * 360 / n_phi degrees (TXPattern.dat: 181 x 360 at 1 degree).  A row gives the field
 * coefficients j sqrt(60 Pt G) e^{j psi} (FieldCompute.cpp:114-116: sqrt(eta Pt G / 2 pi)); between
 * rows they are interpolated bilinearly in (theta, phi), periodic in phi.  Copied to the device;
 * replaces any earlier table; servers created from the context take it over. */
int sbr_set_tx_pattern(sbr_ctx* ctx, const float* gain_info, int32_t n_theta, int32_t n_phi);
int sbr_set_tuning(sbr_ctx* ctx, const sbr_tuning* tuning);
int sbr_get_tuning(sbr_ctx* ctx, sbr_tuning* tuning);

int sbr_generate_rays(sbr_ctx* ctx, int32_t tess, double* dirs_out, int64_t capacity, int64_t* n_out);
/* exact: the traversal, as sbr_trace_params.exactTraversal (0, 1 or 2; SBR_E_ARG otherwise). */
int sbr_closest_hit(sbr_ctx* ctx, const double* orig, const double* dirs, int32_t n, int32_t exact,
                    int32_t* tri_out, double* point_out, double* dist_out);
/* Allocation-free variant of sbr_closest_hit (no counterpart in the reference; for callers that
 * cannot allocate device memory: code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as sbr_closest_hit, with n at most
 * SBR_NOALLOC_MAX_RAYS (SBR_E_LIMIT beyond, nothing run).  The rays are staged in a fixed array of
 * SBR_NOALLOC_SCRATCH_BYTES in device memory, one per GPU, held from the library's first use on that
 * GPU until the process ends; no device memory is allocated or freed by the call.  All calls on one
 * GPU share that array: calls that overlap on one GPU (from several threads or contexts) race.  The
 * call returns after its results are on the host. */
#define SBR_NOALLOC_SCRATCH_BYTES (32u << 20)
#define SBR_NOALLOC_MAX_RAYS 399442
int sbr_closest_hit_noalloc(sbr_ctx* ctx, const double* orig, const double* dirs, int32_t n, int32_t exact,
                            int32_t* tri_out, double* point_out, double* dist_out);

int sbr_trace(sbr_ctx* ctx, const sbr_trace_params* params, const double* rx, int32_t n_rx);

/* Field of caller-given paths: FieldCompute::FieldForPath (FieldCompute.cpp:71-151) on the device,
 * the same transport the trace applies to its kept paths.  Path p has seg_count[p] segments (1 ..
 * 25) starting at seg_off[p]; segment s: seg_src[s][3] -> seg_tgt[s][3], seg_surf[s] the surface it
 * ends on (its hitSurfaceID; -1 at the receiver), seg_type[s] how it left its start surface (0 from
 * the transmitter, 1 reflected, 2 transmitted: reflectionMaterialId / penetrationMaterialId >= 0)
 * and seg_mat[s] that surface's material (an index into mats, unused for type 0).  surf_cs[id][9]
 * is the row-major coordSys of surface id (mesh[id]->coordSys).  fields_out[n_paths][3][2].  The
 * scene of the context is not used; antennaType 3 uses its pattern table. */
typedef struct {
  double txCoordSys[9];  /* row-major Mat3 */
  double frequency;
  double pt;
  int32_t useFresnel;
  int32_t antennaType;   /* 0 isotropic, 1 Hertzian (FieldCompute.cpp:117), 2 half-wave, 3 table */
  double wallWidth;      /* extension, as sbr_trace_params */
  int32_t physicalField; /* extension, as sbr_trace_params */
  int32_t reserved;
} sbr_field_params;
int sbr_path_fields(sbr_ctx* ctx, const sbr_field_params* params, const sbr_material* mats, int32_t n_mats,
                    int64_t n_paths, const int64_t* seg_off, const int32_t* seg_count, const double* seg_src,
                    const double* seg_tgt, const int32_t* seg_surf, const int32_t* seg_type, const int32_t* seg_mat,
                    int32_t n_surf, const double* surf_cs, double* fields_out);
/* Coefficient sweep of FieldCompute::RefCoeffTest / TransCoeffTest (FieldCompute.cpp:27-69): n
 * incidence angles 0 .. pi/2 evenly spaced (theta_out[n], radians) between media of complex relative
 * permittivity eps_i = {re, im} and eps_t, at `frequency`; inf_wall = 1 is the reference's infinite
 * wall (no transmission), 0 its slab of `width` m.  coeff_out[n][4][2]: rTM, tTM, rTE, tTE. */
int sbr_fresnel_sweep(sbr_ctx* ctx, int32_t n, double frequency, const double* eps_i, const double* eps_t,
                      double width, int32_t inf_wall, double* theta_out, double* coeff_out);
/* Several transmitters over the same scene and receivers in one pass (no reference counterpart:
 * the reference runs one transmitter per ExecuteRayTracing call).  params[n_tx]; every
 * transmitter's result equals what sbr_trace gives for it alone. */
int sbr_trace_batch(sbr_ctx* ctx, const sbr_trace_params* params, int32_t n_tx, const double* rx, int32_t n_rx);
int sbr_result_stats(sbr_ctx* ctx, sbr_stats* out);
/* Complex field per (transmitter, receiver): out[n_tx][n_rx][3][2] (Ex re, im, Ey re, im, Ez re,
 * im); n = n_tx * n_rx. */
int sbr_result_fields(sbr_ctx* ctx, double* out, int32_t n);
/* Kept paths, (transmitter, receiver) ascending, then launch ascending, then the path order of a
 * launch (breadth first, transmitted child before reflected child). */
int sbr_result_path_counts(sbr_ctx* ctx, int64_t* n_paths, int64_t* n_segs);
int sbr_result_paths(sbr_ctx* ctx, int32_t* path_rx, int32_t* path_launch, int32_t* path_nseg, int64_t* path_off,
                     int32_t* seg_tri, int32_t* seg_cop, int32_t* seg_type, double* seg_src, double* seg_tgt);
/* Transmitter index of every kept path (batched traces), in the order of sbr_result_paths. */
int sbr_result_path_tx(sbr_ctx* ctx, int32_t* out, int64_t n_paths);
/* Field of every kept path, in the order of sbr_result_paths: out[n_paths][3][2]. */
int sbr_result_path_fields(sbr_ctx* ctx, double* out, int64_t n_paths);
/* Interaction trees of the last trace, one row per node, ordered by (launch, depth, branch). */
int sbr_result_tree_count(sbr_ctx* ctx, int64_t* n_nodes);
int sbr_result_trees(sbr_ctx* ctx, int32_t* launch, int32_t* depth, int64_t* code, int32_t* tri, int32_t* cop,
                     int32_t* type, double* src, double* tgt, double* pl);

/* ---- Serving loop (no reference counterpart: the reference has no serving mode) -------------
 * A server keeps the scene and material tables of a context resident on its device and runs
 * coverage queries (one transmitter, its own receivers) submitted from any host thread.  It owns
 * `slots` engines, each with its own stream and worker thread: a query waiting in the queue is
 * admitted as soon as a slot is free and the device memory of all slots with it stays within the
 * HBM budget, so queries that arrive while others run overlap with them on the device (continuous
 * batching).  The budget covers what each slot holds between queries (per-trace buffers kept at
 * their largest size, cached launch directions) and, while a query runs, the larger of that and the
 * query's estimate; a slot holding more than budget / slots after a query releases its per-trace
 * buffers.  The scene copies are outside the budget.  A query's result does not depend on what
 * runs beside it.  Threading: every sbr_server call is thread-safe; the strings returned by
 * sbr_server_last_error and sbr_request_error are copies owned by the calling thread, valid until
 * its next call of either. */
#define SBR_PENDING 1      /* sbr_poll: the query has not finished yet */

typedef struct sbr_server sbr_server;

typedef struct {
  int32_t slots;             /* concurrent queries on the device; 0 = 3 */
  int32_t max_queued;        /* queries waiting for admission before sbr_submit refuses; 0 = 4096 */
  int64_t hbm_budget_bytes;  /* admission budget for the slots' device memory; 0 = 3/4 of the free memory */
} sbr_server_config;

typedef struct {
  int32_t state;             /* 0 queued, 1 running, 2 done, 3 failed */
  int32_t status;            /* SBR_OK or the SBR_E_* of a failed query */
  int32_t n_rx;
  int32_t slot;              /* engine that ran it, -1 before admission */
  int64_t launches, bounces, paths;
  int64_t est_bytes;         /* admission estimate */
  double ms_queued;          /* submit -> admission */
  double ms_run;             /* admission -> result on the host */
  double ms_total;           /* submit -> result on the host */
} sbr_request_info;

typedef struct {
  int64_t submitted, completed, failed, refused, queued, running;
  int64_t bounces;           /* of completed queries */
  int64_t budget_bytes, peak_inflight_bytes;   /* peak of the slots' counted device memory */
  double p50_ms, p99_ms, mean_ms, max_ms;   /* submit -> result latency of completed queries */
  double window_s;           /* first submission -> last completion */
  double req_per_s;          /* completed / window_s */
  double mray_bounces_per_s; /* bounces / window_s / 1e6 */
} sbr_server_stats;

/* Scene and materials are copied from `src` (which must hold both); the server is independent of
 * `src` afterwards.  A host-only `src` (device -1) gives a server whose queries fail with
 * SBR_E_STATE (queue, admission and error paths without a device). */
int sbr_server_create(sbr_ctx* src, const sbr_server_config* cfg, sbr_server** out);
/* Drains: waits for running queries, fails the queued ones with SBR_E_STATE, frees everything. */
int sbr_server_destroy(sbr_server* s);
const char* sbr_server_last_error(const sbr_server* s);
/* Queues a query; rx[n_rx][3] is copied.  SBR_E_ARG for tess outside [0, 11], SBR_E_LIMIT when
 * the queue is full, SBR_E_NOMEM when the query alone exceeds the HBM budget.  *id identifies it for the calls below. */
int sbr_submit(sbr_server* s, const sbr_trace_params* params, const double* rx, int32_t n_rx, int64_t* id);
/* Waits up to wait_ms (< 0: no limit) for the query.  Returns SBR_OK (done), SBR_PENDING, the
 * query's SBR_E_* status if it failed, or SBR_E_ARG for an unknown id.  info may be NULL. */
int sbr_poll(sbr_server* s, int64_t id, int32_t wait_ms, sbr_request_info* info);
/* Error message of a failed query. */
const char* sbr_request_error(sbr_server* s, int64_t id);
/* Fields of a finished query: out[n_rx][3][2], as sbr_result_fields. */
int sbr_request_fields(sbr_server* s, int64_t id, double* out, int32_t n_rx);
/* Forgets a finished query (its result is freed; its latency stays in the statistics). */
int sbr_request_release(sbr_server* s, int64_t id);
int sbr_server_get_stats(sbr_server* s, sbr_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* SBR_CAPI_H */
