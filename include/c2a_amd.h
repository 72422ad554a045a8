/*
 * c2a_amd.h -- C ABI of the MI355X conservative-advancement CCD engine.
 *
 * This is the drop-in boundary for the reference's CCD path
 * (EwhaGlab/C2A):
 *   C2A_Solve               C2A/C2A.h:23-35,   C2A/src/C2A.cpp:2315-2444
 *   C2A_QueryTimeOfContact  C2A/C2A.h:274-281, C2A/src/C2A.cpp:1987-2146
 *   C2A_TimeOfContactStep   C2A/src/C2A.cpp:1778-1931 (not declared in C2A.h)
 *   C2A_QueryContact        C2A/C2A.h:284-289, C2A/src/C2A.cpp:1937-1966
 *   C2A_Model Begin/AddTri/EndModel  C2A/C2A_Internal.h:57-64,
 *                                    C2A/src/C2A_PQP.cpp:82-417
 * The C++ layer in include/c2a/ (C2A.h, C2A_Internal.h, InterpMotion.h,
 * LinearMath.h) rebuilds the reference's source interface on top of these
 * entry points.
 *
 * Rules: plain C types only; every entry point returns a c2a_status; no
 * exception crosses the boundary; the caller owns host buffers, the library
 * owns device buffers; handles are opaque.  Results match the reference
 * semantics bit for bit (FP64 throughout, same operation order).
 */
#ifndef C2A_AMD_H
#define C2A_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  C2A_OK = 0,
  C2A_ERR_BAD_ARG = 1,
  C2A_ERR_OOM = 2,
  C2A_ERR_HIP = 3,
  C2A_ERR_NO_DEVICE = 4,
  C2A_ERR_UNSUPPORTED = 5,
  C2A_ERR_EMPTY_MODEL = 6,  /* PQP_ERR_BUILD_EMPTY_MODEL (C2A_PQP.cpp:352-357) */
  C2A_ERR_CACHE = 7         /* c2a_mesh_load: not a hierarchy file, damaged, or
                               built from other triangles */
} c2a_status;

typedef struct c2a_mesh c2a_mesh;
typedef struct c2a_batch c2a_batch;
typedef struct c2a_scorer c2a_scorer;   /* learned front ordering, see below */

/* One CCD query: start/end poses of object 1 then object 2, each a
 * row-major 3x3 rotation followed by the translation (12 doubles).  The
 * meshes are indices into the mesh table passed with the batch.  seed_tri is
 * the initial triangle pair that bounds the first distance (the reference's
 * dres.last_triA/B, C2A.cpp:1812-1813), as indices in the mesh's build
 * order; -1 means triangle 0 (the model's last_tri after EndModel). */
typedef struct {
  double pose[4][12];
  int32_t mesh[2];
  int32_t seed_tri[2];
} c2a_query;

/* Query parameters.  Defaults (c2a_params_default) are C2A_Solve's:
 * tol_d = tol_t = 1e-4, 150 iterations, contacts counted within
 * 2*distance + 0.001 of the TOC pose. */
typedef struct {
  double tol_d;
  double tol_t;
  int32_t max_iter;
  int32_t with_contacts;
  double contact_scale;
  double contact_offset;
  int32_t spec_depth;       /* speculative subtree depth per round, 0 = auto */
  int32_t waves_per_query;  /* wavefronts cooperating on one query; this
                               build provides 1 (0 = 1) */
  int32_t replay_log;       /* expansions logged per CA step for the next step's
                               replay; -1 (default) = from the largest mesh
                               (8192, or ntri / 8 rounded up to a power of
                               two); 0 disables replay */
  int32_t order;            /* C2A_ORDER_PLAIN (the reference's closer-child-first
                               order, bit-exact) or C2A_ORDER_LEARNED */
  const c2a_scorer *scorer; /* C2A_ORDER_LEARNED: visits the sibling with the
                               lower score first; read at each run; never
                               prunes a pair */
  double sample_rate;       /* fraction of replayed pairs logged as training
                               samples (selected by hashing their key) */
  int64_t sample_cap;       /* rows of the sample buffer; 0 = no logging */
  int32_t miss_depth;       /* speculative depth of the first round below a
                               replayed pair the step must expand further
                               than the last one (default 2; 0 = spec_depth) */
  int32_t ext_levels;       /* levels computed ahead, during the replay, below
                               replayed pairs that newly pass the descend test
                               at the step's initial bound; -1 (default) = the
                               autotuned value of each query's size bucket
                               when spec_depth is 0 (0 under the learned
                               order or with sample logging), else 0;
                               0 = off.  A query's step skips it after a
                               step with fewer misses than one in 64
                               expansions. */
  int32_t translation;      /* 1 (default): queries whose two rotations are
                               below 1e-8 take C2A_Solve's translation-only
                               branch (C2A.cpp:1362-1521, 2032-2049: one
                               traversal, toc = mint, object 2's velocity
                               ignored); 0: the rotation path, as a direct
                               C2A_QueryTimeOfContact call does */
  int32_t strict;           /* 0 (default): the reference's semantics, bit for
                               bit.  1: conservative mode -- the motion bounds
                               cross the world-frame rotation axis (the
                               reference uses the body-frame axis,
                               InterpMotion.cpp:784-791, 838-845, with world
                               directions, C2A.cpp:1154-1164, 1240-1242),
                               every tested leaf pair bounds mint (the
                               reference: only improving ones, C2A.cpp:1150),
                               a touching leaf pair bounds it by 0, and
                               translation-only queries take the rotation
                               path (the translation branch ignores object
                               2's velocity, C2A.cpp:1472-1474) */
  int32_t tail_waves;       /* wavefronts per query in the tail pass; this
                               build provides 4 (0 = 4) */
  int64_t tail_budget;      /* tail pass: a query that has made this many BV
                               tests when a CA step ends continues, from
                               that step on, in a second launch with
                               tail_waves wavefronts per query.  Results are
                               unchanged.  0 (default) = one launch.  Plain
                               order without sample logging only (otherwise
                               one launch).  It pays only when a few queries
                               run far longer than all others: the second
                               launch starts when the first one ends. */
} c2a_params;

enum { C2A_ORDER_PLAIN = 0, C2A_ORDER_LEARNED = 1 };

/* Per-query result (C2A_TimeOfContactResult, C2A.h:161-224, plus the
 * reference's per-model last_tri after the query). */
typedef struct {
  double toc;
  double distance;
  double mint;
  double R_toc[9];
  double T_toc[3];
  double p1[3];
  double p2[3];
  int64_t bv_tests;
  int64_t tri_tests;
  int32_t numCA;
  int32_t collisionfree;
  int32_t num_contact;
  int32_t last_tri[2];      /* the models' last_tri after the query; for a
                               query the translation-only branch ran, the
                               result's last_triA/B (the last improving pair,
                               C2A.cpp:1413-1414) -- that branch leaves the
                               models' last_tri alone */
  int32_t flags;            /* C2A_FLAG_* */
  /* res->R, res->T: model 2's pose in model 1's frame at the last CA step
   * (C2A.cpp:1793-1796; the translation-only branch: at the advanced pose,
   * C2A.cpp:1900-1904), and the step's error schedule and UpboundTOC */
  double R[9];
  double T[3];
  double abs_err, rel_err, UpboundTOC;
} c2a_result;

enum {
  C2A_FLAG_ITER_CAP = 1,          /* the 150-iteration cap was hit */
  C2A_FLAG_TRANSLATION_ONLY = 2,  /* both rotations below 1e-8 (the
                                     translation-only branch ran when
                                     params->translation) */
  C2A_FLAG_STACK_OVERFLOW = 4,    /* traversal deeper than the stack (never
                                     with capacities chosen from mesh depth) */
  C2A_FLAG_BAD_INPUT = 8          /* c2a_ccd_batch_noalloc: the query's mesh
                                     ids or seed triangles are out of range;
                                     it was skipped (other fields zero) */
};

void c2a_params_default(c2a_params *p);

/* Builds the RSS hierarchy on the host (PQP-style, C2A_Build.cpp); it is
 * uploaded to the current device the first time a launch uses it.  tri_xyz9: ntri x 9 doubles (p1, p2, p3).  tri_ids / vert_idx3 may be
 * NULL (ids default to the input order). */
int c2a_mesh_create(const double *tri_xyz9, const int32_t *tri_ids,
                    const int32_t *vert_idx3, int64_t ntri, c2a_mesh **out);
int c2a_mesh_destroy(c2a_mesh *m);
/* Number of triangles, bounding volumes and maximum depth. */
int c2a_mesh_info(const c2a_mesh *m, int64_t *ntri, int64_t *nbv, int32_t *max_depth);
/* c2a_mesh_create with the hierarchy built on the device (C2A_Build.cpp:
 * 393-574 in top-down levels, one workgroup per large node): the same nodes
 * and triangle order, bit for bit, as the host build; the device copy stays
 * resident for the queries.  stream may be NULL. */
int c2a_mesh_create_device(const double *tri_xyz9, const int32_t *tri_ids,
                           const int32_t *vert_idx3, int64_t ntri, void *stream,
                           c2a_mesh **out);

/* Refit (no counterpart in the reference, which rebuilds: C2A_PQP.cpp:331-417):
 * keeps the tree, the triangle order and the node numbers, and fits every
 * node's RSS volume again -- with the build's fit: covariance orientation,
 * rectangle and radius, parent-relative frames -- to moved triangles
 * tri_xyz9 (ntri x 9 doubles, the same triangles in the order given at
 * creation).  Each node is fitted over its triangles in their final order,
 * so a refit at unchanged triangles equals the build up to the order of the
 * fit's sums and corner pass.  on_device = 1 runs it on the GPU (and keeps
 * the device copy), 0 on the host.  For deforming meshes: the volumes stay
 * conservative, the tree's quality degrades with the motion.  A refitted
 * mesh's cache file matches no build input. */
int c2a_mesh_refit(c2a_mesh *m, const double *tri_xyz9, int32_t on_device, void *stream);

/* On-disk cache of a built hierarchy (no counterpart in the reference, which
 * rebuilds in every EndModel, C2A_PQP.cpp:331-417).  The file holds the
 * nodes, the triangles in build order with their ids and vertex indices, a
 * digest of the triangles, ids and vertex indices the mesh was created from,
 * and a digest of the payload.  c2a_mesh_load with tri_xyz9 != NULL checks
 * that those inputs (as passed to c2a_mesh_create) match the file's digest;
 * a mismatch, a damaged or foreign file returns C2A_ERR_CACHE.  Little-endian
 * hosts only. */
int c2a_mesh_save(const c2a_mesh *m, const char *path);
int c2a_mesh_load(const char *path, const double *tri_xyz9, const int32_t *tri_ids,
                  const int32_t *vert_idx3, int64_t ntri, c2a_mesh **out);

/* Caller ids (AddTri "id") of the triangles in build order. */
int c2a_mesh_tri_ids(const c2a_mesh *m, int32_t *ids_out);
/* Copies the built hierarchy back to the host: nodes as nbv x 32 doubles
 * (R[9], Tr[3], l[2], r, angR, Rloc[9], size, then first_child and depth
 * stored as doubles), triangles as ntri x 9 doubles in build order. */
int c2a_mesh_export(const c2a_mesh *m, double *nodes32, double *tris9);

/* Runs nq independent queries on the GPU (C2A_QueryTimeOfContact semantics,
 * plus C2A_Solve's contact count when params->with_contacts).  Host in/out;
 * stream may be NULL (the null stream). */
int c2a_ccd_batch(c2a_mesh *const *meshes, int32_t nmesh, const c2a_query *q, int64_t nq,
                  const c2a_params *params, c2a_result *out, void *stream);

/* Allocation-free batch.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  The c2a_ccd_batch
 * computation (plain order, params->strict honoured) with device-resident
 * inputs and outputs; it calls no hipMalloc, hipFree or other allocating
 * routine and returns once the work is queued on `stream` (the caller
 * synchronises).
 *   motions     nq x 2 x 24 doubles in device memory: per query, object 1's
 *               then object 2's record as c2a_motion_prepare writes it
 *   mesh_pairs  nq x 2 int32 in device memory (indices into `meshes`)
 *   seed_tris   nq x 2 int32 in device memory, or NULL (triangle 0)
 *   out         nq results in device memory; a query whose mesh ids or seed
 *               triangles are out of range gets C2A_FLAG_BAD_INPUT and no
 *               other work
 * Every mesh must already be on the device (uploaded by an earlier
 * c2a_ccd_batch or c2a_batch_create with it); at most C2A_NOALLOC_MAX_MESHES
 * meshes.  The library's scratch is a fixed C2A_NOALLOC_SCRATCH_MIB MiB per
 * device, held from the first call until the process ends: the work-queue
 * counter, the mesh table and one replay arena per resident query
 * (params->replay_log, -1 = 1024 expansions; the grid shrinks to the
 * arenas that fit; a log whose single arena does not fit is rejected with
 * C2A_ERR_BAD_ARG).  Calls that overlap on one device race on that
 * scratch: serialise them (one stream, or events). */
#define C2A_NOALLOC_SCRATCH_MIB 128
#define C2A_NOALLOC_MAX_MESHES 32
int c2a_ccd_batch_noalloc(c2a_mesh *const *meshes, int32_t nmesh, const double *motions,
                          const int32_t *mesh_pairs, const int32_t *seed_tris, int64_t nq,
                          const c2a_params *params, c2a_result *out, void *stream);

/* Device-resident batches: inputs uploaded once, run repeatedly, results
 * fetched on demand (used to time the kernel with inputs already in HBM).
 * A batch keeps the device addresses of its meshes and, with the learned
 * order, reads its scorer at every run: the meshes and the scorer must
 * outlive the batch (destroy the batch first). */
int c2a_batch_create(c2a_mesh *const *meshes, int32_t nmesh, const c2a_query *q, int64_t nq,
                     const c2a_params *params, c2a_batch **out);
int c2a_batch_run(c2a_batch *b, void *stream);
/* c2a_batch_run through a HIP graph: the first call captures the batch's
 * launch sequence (the work-queue reset, a kernel, then the persistent CCD
 * kernel, which runs every query's whole CA loop) and later calls replay it.
 * Results are those of c2a_batch_run.  Not for batches with a scorer or
 * sample logging. */
int c2a_batch_run_graph(c2a_batch *b, void *stream);
int c2a_batch_fetch(c2a_batch *b, c2a_result *out, void *stream);
int c2a_batch_destroy(c2a_batch *b);
/* Kernel-only device time of the last c2a_batch_run, in milliseconds,
 * measured with HIP events on the stream the kernel was launched on. */
int c2a_batch_last_ms(c2a_batch *b, float *ms);
/* Queries the last run handed to the tail pass (c2a_params.tail_budget);
 * waits for the batch's stream.  0 when the batch runs in one launch.  No
 * counterpart in the reference (diagnostics of this build's launches). */
int c2a_batch_tail_count(c2a_batch *b, int64_t *n);
/* Training samples logged by the last c2a_batch_run (params->sample_cap > 0;
 * replay_log > 0): every CA step after the first replays the previous
 * step's traversal, and each replayed pair whose replayed subtree holds a
 * leaf pair is a sample -- its 32 scorer features (BF16 values as FP32), the
 * label log(closest replayed triangle distance in that subtree + 1e-4) and
 * a key (query << 34 | CA step << 26 | record).  Up to `cap` rows sorted by
 * key are written; *n_out receives the number of samples selected, which
 * may exceed sample_cap (the rows kept past the cap depend on timing). */
int c2a_batch_samples(c2a_batch *b, float *feats, float *labels, int64_t *keys, int64_t cap,
                      int64_t *n_out, int64_t *n_written);

/* One conservative-advancement step at explicit poses (backs
 * C2A_TimeOfContactStep).  state_in: numCA, previous mint and UpboundTOC as
 * the CA loop leaves them; seed: the initial triangle pair.  Writes
 * distance, mint, abs/rel error, counters and the updated last_tri. */
typedef struct {
  double R1[9], T1[3], R2[9], T2[3];   /* current absolute poses */
  double pose_s[2][12], pose_e[2][12]; /* start/end poses defining the motions */
  int32_t mesh[2];
  int32_t seed_tri[2];
  int32_t last_tri[2];  /* the models' last_tri before the step (kept when no
                           leaf improves the bound); -1 = seed_tri */
  int32_t numCA;
  int32_t pad;
  double mint_prev;
  double UpboundTOC;
} c2a_step_query;
typedef struct {
  double distance, mint, abs_err, rel_err;
  double p1[3], p2[3];
  int64_t bv_tests, tri_tests;
  int32_t last_tri[2];
  int32_t flags, pad;
  double R[9], T[3];    /* res->R, res->T of the step (C2A.cpp:1793-1796) */
} c2a_step_result;
int c2a_ccd_step(c2a_mesh *const *meshes, int32_t nmesh, const c2a_step_query *q, int64_t nq,
                 c2a_step_result *out, void *stream);

/* Contact enumeration at explicit poses (C2A_QueryContact): counts the
 * triangle pairs within `threshold` (distance <= threshold). */
int c2a_contact_count(c2a_mesh *const *meshes, int32_t nmesh, const double *poses /* nq x 2 x 12 */,
                      const int32_t *mesh_pairs /* nq x 2 */, const double *thresholds, int64_t nq,
                      int32_t *counts_out, void *stream);

/* Contact list at explicit poses (C2A_QueryContact's cont_l,
 * C2A.cpp:1525-1631, ContactF of C2A/C2A.h:116-153): every triangle pair
 * within `threshold` whose closest features are classified, in the
 * reference's visiting order (its std::list is built with push_front, so
 * cont_l lists them in the reverse order).  out holds nq x cap records (the
 * first min(count, cap) of query i at out + i * cap); counts_out (nq) gets
 * every query's full count, as c2a_contact_count. */
typedef struct {
  int32_t tri[2];       /* TriangleID_A/B: triangle indices in build order */
  int32_t type[2];      /* FeatureType_A/B: 1 vertex, 2 edge, 3 face */
  int32_t fid[2][3];    /* FeatureID_A/B: vertex ids (AddTri i1, i2, i3) of the
                           feature: the vertex; the edge's two vertices; the
                           face's three; -1 where unused */
  double pa[3];         /* P_A: closest point on triangle A, model 1's frame */
  double pb[3];         /* P_B: closest point on triangle B, model 2's frame */
  double distance;
} c2a_contact;
int c2a_contact_list(c2a_mesh *const *meshes, int32_t nmesh, const double *poses /* nq x 2 x 12 */,
                     const int32_t *mesh_pairs /* nq x 2 */, const double *thresholds,
                     int64_t nq, int32_t cap, c2a_contact *out, int32_t *counts_out,
                     void *stream);

/* Static distance query at explicit poses (C2A_Distance with qsize <= 2,
 * C2A/src/C2A_PQP.cpp:970-1057 -> C2ADistanceRecurse :481-620): the closest
 * triangle pair within the relative / absolute error bounds, searched depth
 * first, closer pair first, from the bound of the seed pair.  seed_tris (nq x
 * 2, or NULL for triangle 0) plays the models' last_tri; the result's
 * last_tri is the models' last_tri after the query (build order; caller ids
 * through c2a_mesh_tri_ids).  p1 is in model 1's frame, p2 in model 2's. */
typedef struct {
  double distance;
  double p1[3], p2[3];
  int64_t bv_tests, tri_tests;
  int32_t last_tri[2];
  int32_t flags;        /* C2A_FLAG_STACK_OVERFLOW: the search was cut short */
  int32_t pad;
} c2a_distance_result;
int c2a_distance(c2a_mesh *const *meshes, int32_t nmesh, const double *poses /* nq x 2 x 12 */,
                 const int32_t *mesh_pairs /* nq x 2 */, const int32_t *seed_tris /* nq x 2 */,
                 double rel_err, double abs_err, int64_t nq, c2a_distance_result *out,
                 void *stream);

/* The same search entering only pairs whose RSS volumes overlap, the root
 * pair included: C2A_Collide's C2A_DistanceResult form
 * (C2A/src/C2A_PQP.cpp:1199-1279 -> C2ACollideRecurse :1059-1193).  Pairs
 * farther apart are never visited, so the result is the seed pair's distance
 * unless an overlapping pair is closer. */
int c2a_collide_distance(c2a_mesh *const *meshes, int32_t nmesh, const double *poses,
                         const int32_t *mesh_pairs, const int32_t *seed_tris, double rel_err,
                         double abs_err, int64_t nq, c2a_distance_result *out, void *stream);

/* C2A_Distance with qsize > 2 (C2A/src/C2A_PQP.cpp:1031-1037): the
 * best-first search C2ADistanceQueueRecurse (:624-786) over a bounded
 * priority queue of qsize tests; a full queue runs a nested search.  Same
 * arguments and result as c2a_distance, which qsize <= 2 calls. */
int c2a_distance_queue(c2a_mesh *const *meshes, int32_t nmesh, const double *poses,
                       const int32_t *mesh_pairs, const int32_t *seed_tris, double rel_err,
                       double abs_err, int32_t qsize, int64_t nq, c2a_distance_result *out,
                       void *stream);

/* Result of the static queries below.  Triangles are build-order indices
 * (caller ids through c2a_mesh_tri_ids); p1 is in model 1's frame, p2 in
 * model 2's. */
typedef struct {
  double distance;
  double p1[3], p2[3];
  int64_t bv_tests, tri_tests;
  int32_t tri[2];       /* the pair found (contact space: the models' last_tri
                           after the query, the seed pair when none) */
  int32_t found;        /* contact space: a pair closer than the bound;
                           tolerance: closer_than_tolerance; collide: the
                           number of contacting pairs */
  int32_t flags;        /* C2A_FLAG_STACK_OVERFLOW: the search was cut short */
} c2a_static_result;

/* C2A_Distance_At_ContactSpace (C2A/src/C2A_PQP.cpp:1418-1496 ->
 * :1282-1415): the depth-first search from distance = upper_bounds[i] with
 * zero error bounds and no seed pair, which stops at the first triangle pair
 * closer than the bound.  found = 0: distance = the bound, p1/p2 unset. */
int c2a_distance_contact_space(c2a_mesh *const *meshes, int32_t nmesh, const double *poses,
                               const int32_t *mesh_pairs, const int32_t *seed_tris,
                               const double *upper_bounds, int64_t nq, c2a_static_result *out,
                               void *stream);

/* C2A_Tolerance (C2A/src/C2A_PQP.cpp:1762-1845): is some triangle pair
 * within tolerances[i] (negative = 0)?  Depth first, closer pair first
 * (:1498-1604), or best first over a queue of qsize tests when qsize > 2
 * (:1606-1759); stops at the first pair found.  found = closer_than_tolerance;
 * distance, p1, p2 and tri are set only when found. */
int c2a_tolerance(c2a_mesh *const *meshes, int32_t nmesh, const double *poses,
                  const int32_t *mesh_pairs, const double *tolerances, int32_t qsize,
                  int64_t nq, c2a_static_result *out, void *stream);

/* C2A_Collide, PQP_CollideResult form (C2A/src/C2A_PQP.cpp:911-968 ->
 * CollideRecurse :798-909): the intersecting triangle pairs under pairs of
 * overlapping RSS volumes, in visiting order.  flag = 1 (C2A_ALL_CONTACTS)
 * lists all, 2 (C2A_FIRST_CONTACT) stops at the first.  pairs_out holds nq x
 * pair_cap x 2 build-order triangles (the first min(found, pair_cap) of
 * query i at pairs_out + 2 * i * pair_cap); found counts them all. */
int c2a_collide_pairs(c2a_mesh *const *meshes, int32_t nmesh, const double *poses,
                      const int32_t *mesh_pairs, int32_t flag, int64_t nq, int32_t pair_cap,
                      int32_t *pairs_out, c2a_static_result *out, void *stream);

/* Linear motion between two poses, as the kernel integrates it
 * (CInterpMotion_Linear ctor + velocity(), InterpMotion.cpp:148-270,
 * 486-513): motion24 = R0[9], T0[3], q0[4] (x,y,z,w), cv[3], axis[3],
 * angVel, 0.  Host only. */
int c2a_motion_prepare(const double *pose_start, const double *pose_end, double *motion24);
/* Pose at t (CInterpMotion_Linear::integrate, InterpMotion.cpp:516-568):
 * pose12 = R[9] then T[3]; quat4 (optional) = (x,y,z,w). */
int c2a_motion_pose(const double *motion24, double t, double *pose12, double *quat4);

/* Learned front ordering (no counterpart in the reference: a pair scorer
 * that orders the traversal and never prunes).  A scorer is a
 * 32 -> 64 -> 64 -> 1 ReLU network with BF16 weights on the matrix cores and
 * FP32 accumulation and master parameters (6337 FP32 values: W1[32][64],
 * b1[64], W2[64][64], b2[64], w3[64], b3).  The feature list is frozen in
 * c2a_amd/csrc/scorer.h; the target is log(closest triangle distance in the
 * pair's subtree + 1e-4). */
int c2a_scorer_create(uint64_t seed, c2a_scorer **out);   /* He-uniform weights */
int c2a_scorer_destroy(c2a_scorer *s);
int c2a_scorer_set_params(c2a_scorer *s, const float *params6337);
int c2a_scorer_get_params(const c2a_scorer *s, float *params6337);
/* Parameter file: "C2ASCR01", uint32 count (6337), FP32 parameters, all
 * little-endian.  Loading checks the header, the count and finiteness. */
int c2a_scorer_save(const c2a_scorer *s, const char *path);
int c2a_scorer_load(const char *path, c2a_scorer **out);
/* Scores of n feature rows (n x 32 FP32). */
int c2a_scorer_infer(const c2a_scorer *s, const float *feats, int64_t n, float *out);
/* Gradient (6337) and value of the mean squared error over n rows. */
int c2a_scorer_grad(const c2a_scorer *s, const float *feats, const float *labels, int64_t n,
                    float *grad_out, float *loss_out);
/* Mini-batch Adam for `epochs` passes in a seeded shuffled order; loss_out
 * (optional) receives the mean training loss of each epoch. */
int c2a_scorer_train(c2a_scorer *s, const float *feats, const float *labels, int64_t n,
                     int32_t epochs, int32_t batch, float lr, uint64_t seed, float *loss_out);

const char *c2a_status_string(int status);
/* Thread-local text of the last error, "" if none. */
const char *c2a_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
