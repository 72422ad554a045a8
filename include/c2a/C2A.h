// C2A.h -- source-compatible CCD interface of the reference (C2A/C2A.h:9-35,
// 161-289; C2A_TimeOfContactStep from C2A.cpp:46-49, 1778-1789), implemented
// on the MI355X engine through the C ABI in c2a_amd.h.  An existing caller
// recompiles against these headers and links libc2a_amd.so.
//
// Semantics kept from the reference: C2A_Solve runs the CA query with
// tolerance_d = tolerance_t = 1e-4 (th_ca is ignored, C2A.cpp:2384-2385),
// seeds the first distance from dres.last_triA/B (the caller sets them, as
// the demo does from the models' last_tri), updates each model's last_tri,
// returns TOCFound, and for a collision writes the TOC poses and counts the
// triangle pairs within 2 * distance + 0.001 (C2A.cpp:2411-2435).
// After a collision, as in the reference, the contact query at the TOC pose
// leaves mint = 1, abs_err = rel_err = 0, UpboundTOC = 1, res->R/T at the
// TOC pose and the contact pairs in cont_l (C2A.cpp:1727-1775, push_front
// order).  Differences: no console output per call; query_time_secs is the
// wall time of the engine call.

#ifndef C2A_COMPAT_C2A_H
#define C2A_COMPAT_C2A_H

#include <list>

#include "C2A_Internal.h"
#include "InterpMotion.h"
#include "LinearMath.h"

enum C2A_Result { OK, TOCFound, CollisionFound, CollisionFree, CollisionNotFound };

// Contact pair record (C2A.h:116-153, ListTri.h:43-61).  Feature types:
// 1 vertex, 2 edge, 3 face; feature ids are AddTri's vertex indices;
// triangle ids are build-order indices; P_A is in model 1's frame, P_B in
// model 2's.  Ordered by distance.
struct ContactF {
  ContactF(int FeatureType_A, int FeatureType_B, int FeatureID_A[3], int FeatureID_B[3],
           int TriangleID_A, int TriangleID_B, PQP_REAL P_A[3], PQP_REAL P_B[3],
           PQP_REAL Distance)
      : FeatureType_A(FeatureType_A), FeatureType_B(FeatureType_B),
        TriangleID_A(TriangleID_A), TriangleID_B(TriangleID_B), Distance(Distance) {
    IcI(this->FeatureID_A, FeatureID_A);
    IcI(this->FeatureID_B, FeatureID_B);
    VcV_L(this->P_A, P_A);
    VcV_L(this->P_B, P_B);
  }
  void IcI(int a[3], int b[3]) { a[0] = b[0]; a[1] = b[1]; a[2] = b[2]; }
  void VcV_L(PQP_REAL a[3], PQP_REAL b[3]) { a[0] = b[0]; a[1] = b[1]; a[2] = b[2]; }

  int FeatureType_A, FeatureType_B;
  int FeatureID_A[3], FeatureID_B[3];
  int TriangleID_A, TriangleID_B;
  PQP_REAL P_A[3], P_B[3];
  PQP_REAL Distance;

  friend bool operator>(const ContactF& a, const ContactF& b) { return a.Distance > b.Distance; }
  friend bool operator<(const ContactF& a, const ContactF& b) { return a.Distance < b.Distance; }
};
typedef std::list<ContactF> ContactFList;
typedef std::list<ContactF>::iterator ContactFListIterator;

struct C2A_TimeOfContactResult {
  int num_bv_tests;
  int num_tri_tests;
  double query_time_secs;
  ContactFList cont_l;
  PQP_REAL R[3][3];
  PQP_REAL T[3];
  PQP_REAL rel_err;
  PQP_REAL abs_err;
  PQP_REAL UpboundTOC;
  PQP_REAL distance;
  PQP_REAL p1[3];
  PQP_REAL p2[3];
  int qsize;
  PQP_REAL mint;
  PQP_REAL toc;          // 0 if collision free (C2A.cpp:2098, 2121)
  bool collisionfree;
  int numCA;
  PQP_REAL R_toc[3][3], T_toc[3];
  int num_contact;
  Tri* last_triA;        // seed pair of the first distance (caller-owned)
  Tri* last_triB;

  C2A_TimeOfContactResult();
  int NumBVTests() { return num_bv_tests; }
  int NumTriTests() { return num_tri_tests; }
  double QueryTimeSecs() { return query_time_secs; }
  PQP_REAL Distance() { return distance; }
  const PQP_REAL* P1() { return p1; }
  const PQP_REAL* P2() { return p2; }
};

C2A_Result C2A_Solve(Transform* trans00, Transform* trans01, C2A_Model* obj1_tested,
                     Transform* trans10, Transform* trans11, C2A_Model* obj2_tested,
                     Transform& trans0, Transform& trans1, PQP_REAL& time_of_contact,
                     int& number_of_iteration, int& number_of_contact, PQP_REAL th_ca,
                     C2A_TimeOfContactResult& dres);

PQP_REAL C2A_QueryTimeOfContact(CInterpMotion* objmotion1, CInterpMotion* objmotion2,
                                C2A_TimeOfContactResult* res, C2A_Model* o1, C2A_Model* o2,
                                PQP_REAL tolerance_d, PQP_REAL tolerance_t, int qsize = 2);

// One CA step at the given poses (the defined 11-argument form, C2A.cpp:46-49).
int C2A_TimeOfContactStep(CInterpMotion* objmotion1, CInterpMotion* objmotion2,
                          C2A_TimeOfContactResult* res, PQP_REAL R1[3][3], PQP_REAL T1[3],
                          C2A_Model* o1, PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2,
                          PQP_REAL tolerance_t, PQP_REAL tolerance_d);

// Result of the static distance query (C2A_Internal.h:89-93 over PQP's
// PQP_DistanceResult): R, T take model 2 into model 1's frame; p1 is in
// model 1's frame, p2 in model 2's; t1, t2 are the caller ids of the
// closest pair.
struct C2A_DistanceResult {
  PQP_REAL R[3][3];
  PQP_REAL T[3];
  PQP_REAL rel_err;
  PQP_REAL abs_err;
  PQP_REAL distance;
  PQP_REAL p1[3];
  PQP_REAL p2[3];
  int qsize;
  int num_bv_tests;
  int num_tri_tests;
  double query_time_secs;
  int t1;
  int t2;

  int NumBVTests() { return num_bv_tests; }
  int NumTriTests() { return num_tri_tests; }
  double QueryTimeSecs() { return query_time_secs; }
  PQP_REAL Distance() { return distance; }
  const PQP_REAL* P1() { return p1; }
  const PQP_REAL* P2() { return p2; }
};

// Closest triangle pair of two models at the given poses, within the
// relative and absolute error bounds (C2A.h:255-259, C2A_PQP.cpp:970-1057).
// The first bound is the models' last_tri pair; both are updated.  qsize > 2
// runs the best-first search over a queue of qsize tests
// (C2ADistanceQueueRecurse, C2A_PQP.cpp:624-786).  Returns PQP_OK,
// PQP_ERR_UNPROCESSED_MODEL, or -1 when the engine reports an error.
int C2A_Distance(C2A_DistanceResult* result, PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1,
                 PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2, PQP_REAL rel_err,
                 PQP_REAL abs_err, int qsize = 2);

// C2A_Distance restricted to pairs whose RSS volumes overlap, the root pair
// included (C2A.h:263-267, C2A_PQP.cpp:1199-1279).  Returns PQP_OK,
// PQP_ERR_UNPROCESSED_MODEL, or -1 when the engine reports an error.
int C2A_Collide(C2A_DistanceResult* result, PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1,
                PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2, PQP_REAL rel_err,
                PQP_REAL abs_err, int qsize = 2);

// C2A_Distance from distance = upper_bound, zero error bounds and no seed
// pair, stopping at the first triangle pair closer than the bound
// (C2A.h:241-244, C2A_PQP.cpp:1282-1496).  When one is found, distance, p1,
// p2, t1, t2 and the models' last_tri describe it; otherwise distance is the
// bound and, as in the reference, only p2 is mapped into model 2's frame.
int C2A_Distance_At_ContactSpace(C2A_DistanceResult* result, PQP_REAL R1[3][3],
                                 PQP_REAL T1[3], C2A_Model* o1, PQP_REAL R2[3][3],
                                 PQP_REAL T2[3], C2A_Model* o2, PQP_REAL upper_bound);

const int C2A_ALL_CONTACTS = 1;   // find all pairwise intersecting triangles
const int C2A_FIRST_CONTACT = 2;  // report the first intersecting pair found

// PQP's collision result (PQP 1.3 PQP_CollideResult; the reference passes it
// to C2A_Collide, C2A.h:250-253): the intersecting pairs as caller ids.
struct CollisionPair {
  int id1;
  int id2;
};
struct PQP_CollideResult {
  int num_bv_tests;
  int num_tri_tests;
  double query_time_secs;
  PQP_REAL R[3][3];
  PQP_REAL T[3];
  int num_pairs_alloced;
  int num_pairs;
  CollisionPair* pairs;

  PQP_CollideResult() : num_bv_tests(0), num_tri_tests(0), query_time_secs(0.0),
                        num_pairs_alloced(0), num_pairs(0), pairs(nullptr) {}
  ~PQP_CollideResult() { delete[] pairs; }
  PQP_CollideResult(const PQP_CollideResult&) = delete;
  PQP_CollideResult& operator=(const PQP_CollideResult&) = delete;
  void FreePairsList() {
    num_pairs = num_pairs_alloced = 0;
    delete[] pairs;
    pairs = nullptr;
  }
  void SizeTo(int n) {   // grows or shrinks the list, keeping the pairs
    if (n < num_pairs) return;
    CollisionPair* p = new CollisionPair[n > 0 ? n : 1];
    for (int i = 0; i < num_pairs; i++) p[i] = pairs[i];
    delete[] pairs;
    pairs = p;
    num_pairs_alloced = n;
  }
  void Add(int a, int b) {
    if (num_pairs >= num_pairs_alloced) SizeTo(num_pairs_alloced * 2 + 8);
    pairs[num_pairs].id1 = a;
    pairs[num_pairs].id2 = b;
    num_pairs++;
  }
  int NumBVTests() { return num_bv_tests; }
  int NumTriTests() { return num_tri_tests; }
  double QueryTimeSecs() { return query_time_secs; }
  int Colliding() { return num_pairs > 0; }
  int NumPairs() { return num_pairs; }
  int Id1(int k) { return pairs[k].id1; }
  int Id2(int k) { return pairs[k].id2; }
};

// Every intersecting triangle pair under pairs of overlapping RSS volumes, in
// visiting order; C2A_FIRST_CONTACT stops at the first (C2A.h:250-253,
// C2A_PQP.cpp:798-968).  Returns PQP_OK, PQP_ERR_UNPROCESSED_MODEL, or -1
// when the engine reports an error.
int C2A_Collide(PQP_CollideResult* result, PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1,
                PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2, int flag = C2A_ALL_CONTACTS);

// PQP's tolerance result (PQP 1.3 PQP_ToleranceResult, filled by
// C2A_Tolerance, C2A_PQP.cpp:1762-1845).
struct PQP_ToleranceResult {
  int num_bv_tests;
  int num_tri_tests;
  double query_time_secs;
  PQP_REAL R[3][3];
  PQP_REAL T[3];
  int closer_than_tolerance;
  int qsize;
  PQP_REAL tolerance;
  PQP_REAL distance;
  PQP_REAL p1[3];
  PQP_REAL p2[3];

  int NumBVTests() { return num_bv_tests; }
  int NumTriTests() { return num_tri_tests; }
  double QueryTimeSecs() { return query_time_secs; }
  int CloserThanTolerance() { return closer_than_tolerance; }
  PQP_REAL Distance() { return distance; }
  const PQP_REAL* P1() { return p1; }
  const PQP_REAL* P2() { return p2; }
};

// Whether some triangle pair is within `tolerance` (negative = 0), depth
// first (qsize <= 2) or best first over a queue of qsize tests, stopping at
// the first pair found (C2A_PQP.cpp:1498-1845).  distance, p1 and p2 are set
// when closer; as in the reference, p2 is mapped into model 2's frame either
// way.  The reference defines C2A_Tolerance without declaring it in C2A.h,
// so this declaration has no default qsize.  Returns PQP_OK,
// PQP_ERR_UNPROCESSED_MODEL, or -1 when the engine reports an error.
int C2A_Tolerance(PQP_ToleranceResult* result, PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1,
                  PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2, PQP_REAL tolerance,
                  int qsize);

// Counts and lists (res->cont_l) the triangle pairs within `threshold` at
// the motions' current poses (C2A.cpp:1937-1966).
PQP_REAL C2A_QueryContact(CInterpMotion* objmotion1, CInterpMotion* objmotion2,
                          C2A_TimeOfContactResult* res, C2A_Model* o1, C2A_Model* o2,
                          double threshold);

#endif
