/*
 * c2a_oracle.h -- FP64 CPU oracle for the conservative-advancement CCD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (EwhaGlab/C2A + the PQP routines it calls).  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it, and
 * only as the checker.  The product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Where the
 * arithmetic lives in PQP (GammaUNC/PQP, not vendored, version unpinned) the
 * published algorithm is restated and the citation names the C2A call site.
 *
 * Pinning status of each part (what checks it besides self-consistency;
 * the tests in tests/test_oracle_pinning.py hold the checks):
 *   rect_dist          pinned (brute force over edge pairs + corner/face
 *                      projections, dense sampling, symmetry)
 *   tri_dist           pinned (brute force 9 seg-seg + 6 vertex-face, sampling)
 *   motion / integrate pinned (closed forms t=0, t=1, constant-axis slerp)
 *   motion bounds      pinned (closed form + sampling with R0 = I)
 *   bvh build          pinned by invariants (containment, 2n-1 nodes,
 *                      permutation, angular radius closed form); the exact
 *                      tree shape depends on PQP's Meigen -> parity unpinned
 *   toc step / CA loop pinned by brute-force distance in exact mode,
 *                      scripted state machine, dense-sampling conservativeness
 *   contact count      pinned (brute-force count of pairs within threshold)
 *   static distance    pinned (brute-force minimum over all triangle pairs at
 *                      zero error bounds; error-bound envelope; closest
 *                      points at the reported distance in both frames)
 *   distance queue     pinned (brute-force minimum at zero error bounds,
 *                      error-bound envelope, the depth-first form's
 *                      minimum at zero bounds); the order of equal-d
 *                      tests in PQP's BVTQ is parity unpinned
 *   contact-space dist pinned (brute force: found iff some pair is closer
 *                      than the bound, the pair found is closer, the first
 *                      improving leaf of the depth-first order)
 *   tolerance          pinned (brute force: closer iff the minimum pair
 *                      distance <= tolerance; reported pair within it)
 *   TriContact         pinned (agrees with tri_dist == 0 away from touching;
 *                      constructed piercing / coplanar-overlap / separated)
 *   collide pairs      pinned (brute force: every contacting pair under
 *                      overlapping volumes = every pair with tri_dist == 0)
 */
#ifndef C2A_ORACLE_H
#define C2A_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Triangle record (PQP Tri + C2A_Tri index triple, C2A_Tri.h:46-52). */
typedef struct {
  double p[3][3];   /* vertices p1,p2,p3 */
  int id;           /* caller id (AddTri "i") */
  int idx[3];       /* vertex indices (AddTri i1,i2,i3) */
} c2o_tri;

/* RSS bounding volume (PQP BV + C2A_BV extras used on the CCD path,
 * C2A_BV.h:33-77). R/Tr are parent-relative after the build. */
typedef struct {
  double R[3][3];
  double Tr[3];
  double l[2];
  double r;
  double R_loc[3][3];   /* model-frame orientation snapshot (C2A_Build.cpp:439) */
  double angR;          /* max |vertex| from the model origin (C2A_BV.cpp:71-111) */
  double size;          /* PQP BV::GetSize for RSS: sqrt(l0^2+l1^2) + 2r */
  int first_child;      /* >=0: child bv pair index; <0: leaf, tri = -fc-1 */
  int depth;
} c2o_bv;

typedef struct {
  int ntri;
  c2o_tri *tris;        /* reordered by the build (split_tris swaps) */
  int nbv;
  c2o_bv *bvs;
  int last_tri;         /* PQP_Model::last_tri as an index into tris */
} c2o_model;

/* Linear motion (CInterpMotion_Linear, InterpMotion.cpp:486-568). */
typedef struct {
  double R0[3][3], T0[3];
  double R1[3][3], T1[3];
  double q0[4];         /* quaternion of R0, (x,y,z,w) */
  double cv[3];         /* T1 - T0 */
  double axis[3];       /* body-frame unit axis */
  double angVel;        /* rotation angle over t in [0,1] */
  double axw[3];        /* world-frame axis R0 * axis (strict bounds) */
} c2o_motion;

/* Per-query result, mirroring C2A_TimeOfContactResult (C2A.h:161-224). */
typedef struct {
  double toc;
  double distance;
  double mint;
  double p1[3], p2[3];
  double R_toc[3][3], T_toc[3];
  double abs_err, rel_err, UpboundTOC;
  int numCA;
  int collisionfree;
  int num_contact;
  long long num_bv_tests;
  long long num_tri_tests;
  int last_triA, last_triB;   /* seed pair (caller-owned in the reference) */
  int out_last_triA, out_last_triB; /* model last_tri after the query
                                       (translation-only branch: the
                                       result's last_triA/B, the models'
                                       last_tri being left alone) */
  int flags;                  /* C2O_FLAG_* */
  double R[3][3], T[3];       /* res->R, res->T: model 2 in model 1's frame at
                                 the last step (C2A.cpp:1793-1796) */
} c2o_result;

enum {
  C2O_FLAG_ITER_CAP = 1,        /* hit the 150-iteration cap (C2A.cpp:2079) */
  C2O_FLAG_TRANSLATION_ONLY = 2 /* both angVel < 1e-8: the translation-only
                                   branch ran (C2A.cpp:1362-1521, 1818-1852) */
};

/* Contact record produced by the contact enumeration (ContactF, C2A.h:116-153). */
typedef struct {
  int triA, triB;        /* indices into the reordered tris arrays */
  int typeA, typeB;      /* 1 vertex, 2 edge, 3 face */
  int fidA[3], fidB[3];
  double PA[3], PB[3];
  double distance;
} c2o_contact;

/* sin/cos used by integrate: 0 = the hand-written reduction + polynomial
 * shared with the GPU path (default), 1 = libm. */
void c2o_set_libm_trig(int use_libm);
/* Strict (conservative) mode, off by default (= the reference): world-frame
 * rotation axis in both motion bounds (the reference crosses the body-frame
 * axis with world-frame directions, InterpMotion.cpp:784-791, 838-845 vs
 * C2A.cpp:1154-1164, 1240-1242), every tested leaf pair bounds mint (the
 * reference takes only pairs with dTri <= distance, C2A.cpp:1150), and a
 * touching leaf pair bounds it by 0 (the reference's zero-length normalise
 * gives NaN and no update). */
void c2o_set_strict(int on);
int c2o_get_strict(void);
void c2o_sincos(double x, double *s, double *c);

/* ---- primitives (exported for the pinning tests) ---- */
double c2o_rect_dist(const double Rab[3][3], const double Tab[3],
                     const double a[2], const double b[2],
                     double P[3], double Q[3], double S[3]);
double c2o_tri_dist(double P[3], double Q[3], const double S[3][3],
                    const double T[3][3], int *typeA, int *fidA,
                    int *typeB, int *fidB);
void c2o_quat_from_matrix(const double R[3][3], double q[4]);
void c2o_matrix_from_quat(const double q[4], double R[3][3]);
void c2o_motion_init(c2o_motion *m, const double R0[3][3], const double T0[3],
                     const double R1[3][3], const double T1[3]);
void c2o_motion_integrate(const c2o_motion *m, double t, double R[3][3],
                          double T[3], double qua7[7]);
double c2o_motion_bound(const c2o_motion *m, double angR, const double N[3]);
double c2o_motion_bound_leaf(const c2o_motion *m, double angR, const double S[3]);

/* ---- model build (C2A_Model::BeginModel/AddTri/EndModel) ---- */
int c2o_model_build(c2o_model *m, int n, const double *xyz9, const int *ids,
                    const int *idx3);
void c2o_model_free(c2o_model *m);
void c2o_jacobi_eigen(const double A[3][3], double V[3][3], double d[3]);

/* ---- the CCD path ---- */
typedef struct {
  double tol_d, tol_t;
  int max_iter;          /* 150 in the reference */
} c2o_params;

/* One C2A_TimeOfContactStep at the given absolute poses.  res must hold
 * numCA, mint (previous step) and UpboundTOC as the CA loop leaves them. */
void c2o_toc_step(const c2o_model *m1, const c2o_model *m2,
                  const c2o_motion *mo1, const c2o_motion *mo2,
                  const double R1[3][3], const double T1[3],
                  const double R2[3][3], const double T2[3],
                  c2o_model *upd1, c2o_model *upd2, c2o_result *res);

/* C2A_QueryTimeOfContact (C2A.cpp:1987-2146).  Seeds the first-step bound
 * with (res->last_triA, res->last_triB).  Updates m*->last_tri like the
 * reference.  Returns toc.  With both rotations below 1e-8 it takes
 * C2A_Solve's translation-only branch (toc = mint after one traversal)
 * unless c2o_set_solve_translation(0). */
double c2o_query_toc(c2o_model *m1, c2o_model *m2, c2o_motion *mo1,
                     c2o_motion *mo2, c2o_result *res, double tol_d,
                     double tol_t, int max_iter);
void c2o_set_solve_translation(int on);   /* default 1 (C2A_Solve semantics);
                                            strict mode always takes the
                                            rotation path */

/* The CA loop (C2A.cpp:2029-2143) driven by a scripted step: the k-th step
 * call returns distance dist[k] and mint mint[k] (the last entry repeats).
 * Writes toc, collisionfree, numCA and flags to res, and per call (up to
 * cap) the lamda, UpboundTOC and numCA it was made with.  Returns the
 * number of step calls. */
int c2o_ca_loop_scripted(const double *dist, const double *mint, int n, double tol_d,
                         double tol_t, int max_iter, c2o_result *res, double *lamda_out,
                         double *up_out, int *numca_out, int cap);

/* Contact enumeration at absolute poses (C2A_QueryContact,
 * C2A.cpp:1937-1966 -> 1727-1775 -> 1525-1725).  Returns the count; fills
 * up to cap records in reverse DFS order (the reference push_fronts). */
int c2o_query_contact(const c2o_model *m1, const c2o_model *m2,
                      const double R1[3][3], const double T1[3],
                      const double R2[3][3], const double T2[3],
                      double threshold, c2o_contact *out, int cap,
                      long long *bv_tests);

/* Static distance query (C2A_Distance with qsize <= 2, C2A_PQP.cpp:970-1057).
 * The first bound is the pair (m1->last_tri, m2->last_tri); both are updated
 * like the reference's.  p1 in model 1's frame, p2 in model 2's. */
typedef struct {
  double distance;
  double p1[3], p2[3];
  long long bv_tests, tri_tests;
  int last_tri[2];
} c2o_distance_result;
void c2o_distance(c2o_model *m1, c2o_model *m2, const double R1[3][3], const double T1[3],
                  const double R2[3][3], const double T2[3], double rel_err, double abs_err,
                  c2o_distance_result *r);
/* C2A_Collide, C2A_DistanceResult form (C2A_PQP.cpp:1199-1279): the same
 * search entering only pairs whose RSS volumes overlap, the root included. */
void c2o_collide_distance(c2o_model *m1, c2o_model *m2, const double R1[3][3],
                          const double T1[3], const double R2[3][3], const double T2[3],
                          double rel_err, double abs_err, c2o_distance_result *r);
/* nq independent queries (poses24: model 1's pose then model 2's, row-major
 * R then T), query i seeded with seeds[2i], seeds[2i+1] (NULL: triangle 0). */
void c2o_distance_batch(const c2o_model *m1, const c2o_model *m2, int nq, const double *poses24,
                        const int *seeds, double rel_err, double abs_err, int nthreads,
                        c2o_distance_result *out);

/* C2A_Distance with qsize > 2 (C2A_PQP.cpp:1031-1037 -> the best-first
 * search C2ADistanceQueueRecurse, :624-786) over a restatement of PQP's
 * BVTQ heap; qsize <= 2 is c2o_distance. */
void c2o_distance_queue(c2o_model *m1, c2o_model *m2, const double R1[3][3], const double T1[3],
                        const double R2[3][3], const double T2[3], double rel_err,
                        double abs_err, int qsize, c2o_distance_result *r);
/* C2A_Distance_At_ContactSpace (C2A_PQP.cpp:1282-1496): returns 1 when a
 * triangle pair closer than upper_bound was found (distance, p1, p2 and
 * last_tri then describe it), else 0 with distance = upper_bound. */
int c2o_distance_cs(c2o_model *m1, c2o_model *m2, const double R1[3][3], const double T1[3],
                    const double R2[3][3], const double T2[3], double upper_bound,
                    c2o_distance_result *r);

/* C2A_Tolerance (C2A_PQP.cpp:1762-1845, PQP_ToleranceResult fields). */
typedef struct {
  int closer;           /* closer_than_tolerance */
  int tri[2];           /* build-order triangles of the pair found, else -1 */
  double tolerance;     /* max(tolerance, 0) */
  double distance;      /* the pair's distance when closer, else 0 */
  double p1[3], p2[3];  /* closest points (model 1's / model 2's frame) when closer */
  long long bv_tests, tri_tests;
} c2o_tolerance_result;
void c2o_tolerance(const c2o_model *m1, const c2o_model *m2, const double R1[3][3],
                   const double T1[3], const double R2[3][3], const double T2[3],
                   double tolerance, int qsize, c2o_tolerance_result *r);

/* C2A_Collide, PQP_CollideResult form (C2A_PQP.cpp:911-968): the first cap
 * contacting pairs (build-order triangle indices) go to pairs[2k], pairs[2k+1]. */
typedef struct {
  long long bv_tests, tri_tests;
  int num_pairs;
  int pad;
} c2o_collide_result;
void c2o_collide_pairs(const c2o_model *m1, const c2o_model *m2, const double R1[3][3],
                       const double T1[3], const double R2[3][3], const double T2[3], int flag,
                       int cap, int *pairs, c2o_collide_result *r);
/* PQP's TriContact: 1 when the triangles intersect or touch. */
int c2o_tri_contact(const double P1[3], const double P2[3], const double P3[3],
                    const double Q1[3], const double Q2[3], const double Q3[3]);

void c2o_distance_queue_batch(const c2o_model *m1, const c2o_model *m2, int nq,
                              const double *poses24, const int *seeds, double rel_err,
                              double abs_err, int qsize, c2o_distance_result *out);
void c2o_distance_cs_batch(const c2o_model *m1, const c2o_model *m2, int nq,
                           const double *poses24, const int *seeds, const double *upper,
                           c2o_distance_result *out, int *found);
void c2o_tolerance_batch(const c2o_model *m1, const c2o_model *m2, int nq, const double *poses24,
                         const double *tolerance, int qsize, c2o_tolerance_result *out);
void c2o_collide_pairs_batch(const c2o_model *m1, const c2o_model *m2, int nq,
                             const double *poses24, int flag, int cap, int *pairs,
                             c2o_collide_result *out);

/* C2A_Solve (C2A.cpp:2315-2444): tol_d = tol_t = 1e-4, contact threshold
 * 2d+0.001.  Seeds from m*->last_tri like the demo (mainTorusknot.cpp:314). */
int c2o_solve(const double R00[3][3], const double T00[3],
              const double R01[3][3], const double T01[3], c2o_model *m1,
              const double R10[3][3], const double T10[3],
              const double R11[3][3], const double T11[3], c2o_model *m2,
              double R0out[3][3], double T0out[3],
              double R1out[3][3], double T1out[3], c2o_result *res);

/* Batch helper for tests/baseline: runs nq independent queries with
 * C2A_QueryTimeOfContact semantics; pose arrays are nq x 12 doubles
 * (row-major R then T); seeds nq x 2 ints (-1 = use model last_tri).
 * chain != 0 carries last_tri across queries like the demo.  nthreads > 1
 * runs queries in parallel (chain must be 0). */
int c2o_batch_toc(const c2o_model *m1, const c2o_model *m2, int nq,
                  const double *pose00, const double *pose01,
                  const double *pose10, const double *pose11,
                  const int *seeds, double tol_d, double tol_t, int max_iter,
                  int with_contacts, int chain, int nthreads, c2o_result *out);

#ifdef __cplusplus
}
#endif
#endif
