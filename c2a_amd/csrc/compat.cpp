// compat.cpp -- the reference's C++ CCD interface (include/c2a/C2A.h) over the
// C ABI (include/c2a_amd.h): models become engine meshes, motions become
// engine motion records, and each query is a one-query batch on the GPU.
#include <string.h>

#include <chrono>
#include <vector>

#include "../../include/c2a/C2A.h"
#include "../../include/c2a_amd.h"
#include "c2a_math.h"

namespace {

void pose12_from(const Transform& t, double* p) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) p[3 * i + j] = t.Rotation()[i][j];
  for (int i = 0; i < 3; i++) p[9 + i] = t.Translation()[i];
}

void transform_from(const double* p, Transform& t) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t.Rotation()[i][j] = p[3 * i + j];
  for (int i = 0; i < 3; i++) t.Translation()[i] = p[9 + i];
}

void pose12_from(const PQP_REAL R[3][3], const PQP_REAL T[3], double* p) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) p[3 * i + j] = R[i][j];
  for (int i = 0; i < 3; i++) p[9 + i] = T[i];
}

int seed_of(const C2A_Model* m, const Tri* t) {
  const int i = m->tri_index(t);
  return (i >= 0 && i < m->num_tris) ? i : 0;
}

void set_last(C2A_Model* m, int i) {
  if (m->tris && i >= 0 && i < m->num_tris) m->last_tri = m->tris + i;
}

// Query of two motions as the engine takes it (start and end poses).
c2a_query make_query(const CInterpMotion* m1, const CInterpMotion* m2,
                     const C2A_TimeOfContactResult* res, const C2A_Model* o1,
                     const C2A_Model* o2) {
  c2a_query q;
  memset(&q, 0, sizeof(q));
  pose12_from(m1->transform_s, q.pose[0]);
  pose12_from(m1->transform_t, q.pose[1]);
  pose12_from(m2->transform_s, q.pose[2]);
  pose12_from(m2->transform_t, q.pose[3]);
  q.mesh[0] = 0;
  q.mesh[1] = 1;
  q.seed_tri[0] = seed_of(o1, res->last_triA);
  q.seed_tri[1] = seed_of(o2, res->last_triB);
  return q;
}

void copy_result(const c2a_result& r, C2A_TimeOfContactResult* res) {
  res->toc = r.toc;
  res->distance = r.distance;
  res->mint = r.mint;
  for (int i = 0; i < 3; i++) {
    res->p1[i] = r.p1[i];
    res->p2[i] = r.p2[i];
    res->T_toc[i] = r.T_toc[i];
    for (int j = 0; j < 3; j++) res->R_toc[i][j] = r.R_toc[3 * i + j];
  }
  res->num_bv_tests = (int)r.bv_tests;
  res->num_tri_tests = (int)r.tri_tests;
  res->numCA = r.numCA;
  res->collisionfree = r.collisionfree != 0;
  res->num_contact = r.num_contact;
  for (int i = 0; i < 3; i++) {
    res->T[i] = r.T[i];
    for (int j = 0; j < 3; j++) res->R[i][j] = r.R[3 * i + j];
  }
  res->abs_err = r.abs_err;
  res->rel_err = r.rel_err;
  res->UpboundTOC = r.UpboundTOC;
}

// C2A_QueryContact at absolute poses (C2A.cpp:1937-1966 ->
// C2A_TimeOfContactStep_Contact, 1727-1775): res->R/T at these poses, the
// error schedule zeroed, mint 1, UpboundTOC 1, and the contact pairs within
// `threshold` in cont_l (the reference's push_front order).  Returns the
// engine status.
int contact_query(c2a_mesh* const* meshes, const double* poses24, double threshold,
                  C2A_TimeOfContactResult* res) {
  const int32_t pairs[2] = {0, 1};
  int32_t count = 0;
  res->num_contact = 0;
  res->UpboundTOC = 1;
  res->cont_l.clear();
  if (c2a_contact_count(meshes, 2, poses24, pairs, &threshold, 1, &count, nullptr) != C2A_OK)
    return -1;
  std::vector<c2a_contact> recs((size_t)count);
  if (count > 0 && c2a_contact_list(meshes, 2, poses24, pairs, &threshold, 1, count, recs.data(),
                                    &count, nullptr) != C2A_OK)
    return -1;
  for (int k = 0; k < count; k++) {
    c2a_contact& c = recs[(size_t)k];
    res->cont_l.push_front(ContactF(c.type[0], c.type[1], c.fid[0], c.fid[1], c.tri[0], c.tri[1],
                                    c.pa, c.pb, c.distance));
  }
  res->num_contact = count;
  double R1[9], T1[3], R2[9], T2[3], Rr[9], Tr[3], u[3];
  for (int i = 0; i < 9; i++) { R1[i] = poses24[i]; R2[i] = poses24[12 + i]; }
  for (int i = 0; i < 3; i++) { T1[i] = poses24[9 + i]; T2[i] = poses24[21 + i]; }
  c2a::mtxm(Rr, R1, R2);
  c2a::sub3(u, T2, T1);
  c2a::mtxv(Tr, R1, u);
  for (int i = 0; i < 3; i++) {
    res->T[i] = Tr[i];
    for (int j = 0; j < 3; j++) res->R[i][j] = Rr[3 * i + j];
  }
  res->rel_err = 0;
  res->abs_err = 0;
  res->mint = 1;
  return C2A_OK;
}

}  // namespace

// ---------------------------------------------------------------- LinearMath
void Matrix3x3::Set_Value(const Quaternion& q) {
  const double qq[4] = {q.X(), q.Y(), q.Z(), q.W()};
  c2a::matrix_from_quat(qq, m_);
}

// the engine's matrix-to-quaternion conversion (the motion constructor's)
Quaternion Matrix3x3::Quaternion_() {
  double q[4];
  c2a::quat_from_matrix(m_, q);
  return Quaternion(q[0], q[1], q[2], q[3]);
}

// ---------------------------------------------------------------- C2A_Model
C2A_Model::C2A_Model()
    : tris(nullptr), num_tris(0), last_tri(nullptr), build_state(0), mesh_(nullptr) {}

C2A_Model::~C2A_Model() {
  if (mesh_) c2a_mesh_destroy(mesh_);
}

int C2A_Model::BeginModel(int n) {
  if (mesh_) c2a_mesh_destroy(mesh_);
  mesh_ = nullptr;
  xyz_.clear();
  ids_.clear();
  vidx_.clear();
  built_.clear();
  if (n > 0) xyz_.reserve((size_t)n * 9);
  tris = nullptr;
  num_tris = 0;
  last_tri = nullptr;
  build_state = 1;
  return PQP_OK;
}

int C2A_Model::AddTri(const PQP_REAL* p1, const PQP_REAL* p2, const PQP_REAL* p3, int id) {
  return AddTri(p1, p2, p3, id, -1, -1, -1);
}

int C2A_Model::AddTri(const PQP_REAL* p1, const PQP_REAL* p2, const PQP_REAL* p3, int i, int i1,
                      int i2, int i3) {
  if (build_state == 0) BeginModel();
  else if (build_state == 2) return PQP_ERR_BUILD_OUT_OF_SEQUENCE;
  for (int k = 0; k < 3; k++) xyz_.push_back(p1[k]);
  for (int k = 0; k < 3; k++) xyz_.push_back(p2[k]);
  for (int k = 0; k < 3; k++) xyz_.push_back(p3[k]);
  ids_.push_back(i);
  vidx_.push_back(i1);
  vidx_.push_back(i2);
  vidx_.push_back(i3);
  num_tris = (int)ids_.size();
  return PQP_OK;
}

int C2A_Model::EndModel() {
  if (build_state == 2) return PQP_ERR_BUILD_OUT_OF_SEQUENCE;
  const int64_t n = (int64_t)ids_.size();
  if (n == 0) return PQP_ERR_BUILD_EMPTY_MODEL;
  if (c2a_mesh_create(xyz_.data(), ids_.data(), vidx_.data(), n, &mesh_) != C2A_OK)
    return PQP_ERR_MODEL_OUT_OF_MEMORY;
  std::vector<double> t9((size_t)n * 9);
  std::vector<int32_t> ids((size_t)n);
  c2a_mesh_export(mesh_, nullptr, t9.data());
  c2a_mesh_tri_ids(mesh_, ids.data());
  built_.resize((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    Tri& t = built_[(size_t)i];
    for (int k = 0; k < 3; k++) {
      t.p1[k] = t9[(size_t)i * 9 + k];
      t.p2[k] = t9[(size_t)i * 9 + 3 + k];
      t.p3[k] = t9[(size_t)i * 9 + 6 + k];
    }
    t.id = ids[(size_t)i];
  }
  tris = built_.data();
  num_tris = (int)n;
  last_tri = tris;   // C2A_PQP.cpp:401
  build_state = 2;
  return PQP_OK;
}

// ------------------------------------------------------------------- motions
bool CInterpMotion::integrate(const double dt, PQP_REAL R[3][3], PQP_REAL T[3]) {
  double p[12];
  if (c2a_motion_pose(m_, dt, p, nullptr) != C2A_OK) return false;
  transform_from(p, transform);
  for (int i = 0; i < 3; i++) {
    T[i] = p[9 + i];
    for (int j = 0; j < 3; j++) R[i][j] = p[3 * i + j];
  }
  return true;
}

void CInterpMotion_Linear::init(const double* ps, const double* pe) {
  c2a_motion_prepare(ps, pe, m_);
  transform_from(ps, transform_s);
  transform_from(pe, transform_t);
  transform = transform_s;
  m_axis.Set_Value(m_[19], m_[20], m_[21]);
  m_angVel = m_[22];
  m_toc_delta = 0.0;
}

CInterpMotion_Linear::CInterpMotion_Linear(const PQP_REAL R0[3][3], const PQP_REAL T0[3],
                                           const PQP_REAL R1[3][3], const PQP_REAL T1[3]) {
  double ps[12], pe[12];
  pose12_from(R0, T0, ps);
  pose12_from(R1, T1, pe);
  init(ps, pe);
}

CInterpMotion_Linear::CInterpMotion_Linear(const PQP_REAL q0[7], const PQP_REAL q1[7]) {
  double ps[12], pe[12];
  const double a[4] = {q0[1], q0[2], q0[3], q0[0]}, b[4] = {q1[1], q1[2], q1[3], q1[0]};
  c2a::matrix_from_quat(a, ps);
  c2a::matrix_from_quat(b, pe);
  for (int i = 0; i < 3; i++) {
    ps[9 + i] = q0[4 + i];
    pe[9 + i] = q1[4 + i];
  }
  init(ps, pe);
}

bool CInterpMotion_Linear::integrate(const double dt, PQP_REAL qua[7]) {
  double p[12], q[4];
  if (c2a_motion_pose(m_, dt, p, q) != C2A_OK) return false;
  transform_from(p, transform);
  qua[0] = q[3];
  qua[1] = q[0];
  qua[2] = q[1];
  qua[3] = q[2];
  for (int i = 0; i < 3; i++) qua[4 + i] = p[9 + i];
  return true;
}

// ------------------------------------------------------------------- results
C2A_TimeOfContactResult::C2A_TimeOfContactResult()
    : num_bv_tests(0), num_tri_tests(0), query_time_secs(0.0), cont_l(), rel_err(0.0),
      abs_err(0.0), UpboundTOC(1.0), distance(0.0), qsize(2), mint(0.0), toc(0.0),
      collisionfree(false), numCA(0), num_contact(0), last_triA(nullptr), last_triB(nullptr) {
  memset(R, 0, sizeof(R));
  memset(T, 0, sizeof(T));
  memset(p1, 0, sizeof(p1));
  memset(p2, 0, sizeof(p2));
  memset(R_toc, 0, sizeof(R_toc));
  memset(T_toc, 0, sizeof(T_toc));
}

// ------------------------------------------------------------------- queries
PQP_REAL C2A_QueryTimeOfContact(CInterpMotion* m1, CInterpMotion* m2, C2A_TimeOfContactResult* res,
                                C2A_Model* o1, C2A_Model* o2, PQP_REAL tolerance_d,
                                PQP_REAL tolerance_t, int qsize) {
  (void)qsize;   // unused by the reference as well
  c2a_mesh* meshes[2] = {o1->engine_mesh(), o2->engine_mesh()};
  c2a_query q = make_query(m1, m2, res, o1, o2);
  c2a_params p;
  c2a_params_default(&p);
  p.tol_d = tolerance_d;
  p.tol_t = tolerance_t;
  p.with_contacts = 0;
  // The translation-only branch follows the global b_TanslationCCD, which
  // only C2A_Solve sets (C2A.cpp:2390-2395); a direct call with that flag
  // never set takes the rotation path.  (The reference's flag stays set
  // after a translation-only C2A_Solve; that leftover state is not kept.)
  p.translation = 0;
  c2a_result r;
  const auto t0 = std::chrono::steady_clock::now();
  if (c2a_ccd_batch(meshes, 2, &q, 1, &p, &r, nullptr) != C2A_OK) return -1.0;
  res->query_time_secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  copy_result(r, res);
  res->num_contact = 0;
  set_last(o1, r.last_tri[0]);
  set_last(o2, r.last_tri[1]);
  if (!res->collisionfree) {
    PQP_REAL Rt[3][3], Tt[3];
    m1->integrate(res->toc, Rt, Tt);
    m2->integrate(res->toc, Rt, Tt);
  }
  return res->toc;
}

C2A_Result C2A_Solve(Transform* trans00, Transform* trans01, C2A_Model* obj1, Transform* trans10,
                     Transform* trans11, C2A_Model* obj2, Transform& trans0, Transform& trans1,
                     PQP_REAL& time_of_contact, int& number_of_iteration,
                     int& number_of_contact, PQP_REAL th_ca, C2A_TimeOfContactResult& dres) {
  (void)th_ca;   // ignored by the reference (C2A.cpp:2384-2385)
  double a0[12], a1[12], b0[12], b1[12];
  pose12_from(*trans00, a0);
  pose12_from(*trans01, a1);
  pose12_from(*trans10, b0);
  pose12_from(*trans11, b1);
  PQP_REAL R0[3][3], T0[3], R1[3][3], T1[3], R2[3][3], T2[3], R3[3][3], T3[3];
  for (int i = 0; i < 3; i++) {
    T0[i] = a0[9 + i]; T1[i] = a1[9 + i]; T2[i] = b0[9 + i]; T3[i] = b1[9 + i];
    for (int j = 0; j < 3; j++) {
      R0[i][j] = a0[3 * i + j]; R1[i][j] = a1[3 * i + j];
      R2[i][j] = b0[3 * i + j]; R3[i][j] = b1[3 * i + j];
    }
  }
  CInterpMotion_Linear motion1(R0, T0, R1, T1);
  CInterpMotion_Linear motion2(R2, T2, R3, T3);
  c2a_mesh* meshes[2] = {obj1->engine_mesh(), obj2->engine_mesh()};
  c2a_query q = make_query(&motion1, &motion2, &dres, obj1, obj2);
  c2a_params p;
  c2a_params_default(&p);   // 1e-4 / 1e-4
  p.with_contacts = 0;      // the contact query below counts and lists them
  c2a_result r;
  number_of_iteration = 0;
  const auto t0 = std::chrono::steady_clock::now();
  if (c2a_ccd_batch(meshes, 2, &q, 1, &p, &r, nullptr) != C2A_OK) return CollisionNotFound;
  copy_result(r, &dres);
  dres.cont_l.clear();
  if (r.flags & C2A_FLAG_TRANSLATION_ONLY) {
    // the translation-only branch updates dres.last_triA/B to its last
    // improving pair and leaves the models' last_tri (C2A.cpp:1413-1414)
    if (r.last_tri[0] >= 0 && r.last_tri[0] < obj1->num_tris) dres.last_triA = obj1->tris + r.last_tri[0];
    if (r.last_tri[1] >= 0 && r.last_tri[1] < obj2->num_tris) dres.last_triB = obj2->tris + r.last_tri[1];
  } else {
    set_last(obj1, r.last_tri[0]);
    set_last(obj2, r.last_tri[1]);
  }
  dres.num_contact = 0;
  if (!dres.collisionfree) {
    double pb[12], poses[24];
    for (int i = 0; i < 3; i++) {
      trans0.Translation()[i] = r.T_toc[i];
      for (int j = 0; j < 3; j++) trans0.Rotation()[i][j] = r.R_toc[3 * i + j];
    }
    c2a_motion_pose(motion2.engine_motion(), dres.toc, pb, nullptr);
    transform_from(pb, trans1);
    // C2A_QueryContact at the TOC pose within 2 d + 0.001 (C2A.cpp:2433-2435)
    for (int i = 0; i < 9; i++) poses[i] = r.R_toc[i];
    for (int i = 0; i < 3; i++) poses[9 + i] = r.T_toc[i];
    for (int i = 0; i < 12; i++) poses[12 + i] = pb[i];
    if (contact_query(meshes, poses, 2 * dres.distance + 0.001, &dres) != C2A_OK)
      return CollisionNotFound;
  }
  dres.query_time_secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  time_of_contact = dres.toc;
  number_of_contact = dres.num_contact;
  number_of_iteration = dres.numCA;
  return TOCFound;
}

int C2A_TimeOfContactStep(CInterpMotion* m1, CInterpMotion* m2, C2A_TimeOfContactResult* res,
                          PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1, PQP_REAL R2[3][3],
                          PQP_REAL T2[3], C2A_Model* o2, PQP_REAL tolerance_t,
                          PQP_REAL tolerance_d) {
  (void)tolerance_t;   // the step itself does not use them (C2A.cpp:1778-1931)
  (void)tolerance_d;
  c2a_mesh* meshes[2] = {o1->engine_mesh(), o2->engine_mesh()};
  c2a_step_query q;
  memset(&q, 0, sizeof(q));
  double p1[12], p2[12];
  pose12_from(R1, T1, p1);
  pose12_from(R2, T2, p2);
  memcpy(q.R1, p1, 9 * sizeof(double));
  memcpy(q.T1, p1 + 9, 3 * sizeof(double));
  memcpy(q.R2, p2, 9 * sizeof(double));
  memcpy(q.T2, p2 + 9, 3 * sizeof(double));
  pose12_from(m1->transform_s, q.pose_s[0]);
  pose12_from(m1->transform_t, q.pose_e[0]);
  pose12_from(m2->transform_s, q.pose_s[1]);
  pose12_from(m2->transform_t, q.pose_e[1]);
  q.mesh[0] = 0;
  q.mesh[1] = 1;
  q.seed_tri[0] = seed_of(o1, res->last_triA);
  q.seed_tri[1] = seed_of(o2, res->last_triB);
  q.last_tri[0] = seed_of(o1, o1->last_tri);
  q.last_tri[1] = seed_of(o2, o2->last_tri);
  q.numCA = res->numCA;
  q.mint_prev = res->mint;
  q.UpboundTOC = res->UpboundTOC;
  c2a_step_result r;
  if (c2a_ccd_step(meshes, 2, &q, 1, &r, nullptr) != C2A_OK) return -1;
  res->distance = r.distance;
  res->mint = r.mint;
  res->abs_err = r.abs_err;
  res->rel_err = r.rel_err;
  for (int i = 0; i < 3; i++) {
    res->p1[i] = r.p1[i];
    res->p2[i] = r.p2[i];
    res->T[i] = r.T[i];
    for (int j = 0; j < 3; j++) res->R[i][j] = r.R[3 * i + j];
  }
  res->num_bv_tests += (int)r.bv_tests;
  res->num_tri_tests += (int)r.tri_tests;
  set_last(o1, r.last_tri[0]);
  set_last(o2, r.last_tri[1]);
  return PQP_OK;
}

namespace {

int distance_query(C2A_DistanceResult* res, PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1,
                   PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2, PQP_REAL rel_err,
                   PQP_REAL abs_err, int qsize, bool overlap_only) {
  if (o1->build_state != 2 || o2->build_state != 2) return PQP_ERR_UNPROCESSED_MODEL;
  const auto t0 = std::chrono::steady_clock::now();
  c2a_mesh* meshes[2] = {o1->engine_mesh(), o2->engine_mesh()};
  double poses[24];
  pose12_from(R1, T1, poses);
  pose12_from(R2, T2, poses + 12);
  const int32_t pairs[2] = {0, 1};
  const int32_t seeds[2] = {o1->tri_index(o1->last_tri), o2->tri_index(o2->last_tri)};
  c2a_distance_result r;
  // C2A_Collide's distance form ignores qsize (C2A_PQP.cpp:1200-1279)
  const int rc = overlap_only
                     ? c2a_collide_distance(meshes, 2, poses, pairs, seeds, rel_err, abs_err, 1,
                                            &r, nullptr)
                     : c2a_distance_queue(meshes, 2, poses, pairs, seeds, rel_err, abs_err,
                                          qsize, 1, &r, nullptr);
  if (rc != C2A_OK) return -1;
  // [R, T] = [R1, T1]^T [R2, T2] (C2A_PQP.cpp:990-993)
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++)
      res->R[i][j] = R1[0][i] * R2[0][j] + R1[1][i] * R2[1][j] + R1[2][i] * R2[2][j];
    res->T[i] = R1[0][i] * (T2[0] - T1[0]) + R1[1][i] * (T2[1] - T1[1]) +
                R1[2][i] * (T2[2] - T1[2]);
  }
  res->rel_err = rel_err;
  res->abs_err = abs_err;
  res->distance = r.distance;
  for (int i = 0; i < 3; i++) {
    res->p1[i] = r.p1[i];
    res->p2[i] = r.p2[i];
  }
  if (!overlap_only && qsize > 2) res->qsize = qsize;   // C2A_PQP.cpp:1034
  res->num_bv_tests = (int)r.bv_tests;
  res->num_tri_tests = (int)r.tri_tests;
  o1->last_tri = &o1->tris[r.last_tri[0]];
  o2->last_tri = &o2->tris[r.last_tri[1]];
  res->t1 = o1->last_tri->id;
  res->t2 = o2->last_tri->id;
  res->query_time_secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PQP_OK;
}

// [R, T] = [R1, T1]^T [R2, T2]: model 2 in model 1's frame (C2A_PQP.cpp:990-993)
void relative_pose(PQP_REAL R1[3][3], PQP_REAL T1[3], PQP_REAL R2[3][3], PQP_REAL T2[3],
                   PQP_REAL R[3][3], PQP_REAL T[3]) {
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++)
      R[i][j] = R1[0][i] * R2[0][j] + R1[1][i] * R2[1][j] + R1[2][i] * R2[2][j];
    T[i] = R1[0][i] * (T2[0] - T1[0]) + R1[1][i] * (T2[1] - T1[1]) +
           R1[2][i] * (T2[2] - T1[2]);
  }
}

// p2 <- R^T (p2 - T): the reference's closing step of a static query, run
// on a point it did not compute too (C2A_PQP.cpp:1484-1488, 1834-1838).
void p2_to_frame2(PQP_REAL R[3][3], PQP_REAL T[3], PQP_REAL p2[3]) {
  PQP_REAL u[3] = {p2[0] - T[0], p2[1] - T[1], p2[2] - T[2]};
  for (int i = 0; i < 3; i++) p2[i] = R[0][i] * u[0] + R[1][i] * u[1] + R[2][i] * u[2];
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int C2A_Distance_At_ContactSpace(C2A_DistanceResult* res, PQP_REAL R1[3][3], PQP_REAL T1[3],
                                 C2A_Model* o1, PQP_REAL R2[3][3], PQP_REAL T2[3],
                                 C2A_Model* o2, PQP_REAL upper_bound) {
  if (o1->build_state != 2 || o2->build_state != 2) return PQP_ERR_UNPROCESSED_MODEL;
  const auto t0 = std::chrono::steady_clock::now();
  c2a_mesh* meshes[2] = {o1->engine_mesh(), o2->engine_mesh()};
  double poses[24];
  pose12_from(R1, T1, poses);
  pose12_from(R2, T2, poses + 12);
  const int32_t pairs[2] = {0, 1};
  const int32_t seeds[2] = {o1->tri_index(o1->last_tri), o2->tri_index(o2->last_tri)};
  c2a_static_result r;
  if (c2a_distance_contact_space(meshes, 2, poses, pairs, seeds, &upper_bound, 1, &r, nullptr) !=
      C2A_OK)
    return -1;
  relative_pose(R1, T1, R2, T2, res->R, res->T);
  res->distance = r.distance;
  res->abs_err = 0.0;
  res->rel_err = 0.0;
  res->num_bv_tests = (int)r.bv_tests;
  res->num_tri_tests = (int)r.tri_tests;
  if (r.found) {   // C2A_PQP.cpp:1311-1322
    for (int i = 0; i < 3; i++) {
      res->p1[i] = r.p1[i];
      res->p2[i] = r.p2[i];
    }
    o1->last_tri = &o1->tris[r.tri[0]];
    o2->last_tri = &o2->tris[r.tri[1]];
    res->t1 = o1->last_tri->id;
    res->t2 = o2->last_tri->id;
  } else {
    p2_to_frame2(res->R, res->T, res->p2);
  }
  res->query_time_secs = seconds_since(t0);
  return PQP_OK;
}

int C2A_Collide(PQP_CollideResult* res, PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1,
                PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2, int flag) {
  if (o1->build_state != 2 || o2->build_state != 2) return PQP_ERR_UNPROCESSED_MODEL;
  const auto t0 = std::chrono::steady_clock::now();
  c2a_mesh* meshes[2] = {o1->engine_mesh(), o2->engine_mesh()};
  double poses[24];
  pose12_from(R1, T1, poses);
  pose12_from(R2, T2, poses + 12);
  const int32_t mp[2] = {0, 1};
  const int32_t f = flag == C2A_FIRST_CONTACT ? 2 : 1;   // C2A_PQP.cpp:878, 898
  c2a_static_result r;
  std::vector<int32_t> list(2 * 256);
  if (c2a_collide_pairs(meshes, 2, poses, mp, f, 1, 256, list.data(), &r, nullptr) != C2A_OK)
    return -1;
  if (r.found > 256) {   // the list was cut: run again with room for every pair
    list.assign(2 * (size_t)r.found, 0);
    if (c2a_collide_pairs(meshes, 2, poses, mp, f, 1, r.found, list.data(), &r, nullptr) !=
        C2A_OK)
      return -1;
  }
  relative_pose(R1, T1, R2, T2, res->R, res->T);
  res->num_bv_tests = (int)r.bv_tests;
  res->num_tri_tests = (int)r.tri_tests;
  res->num_pairs = 0;   // keeps the list's memory (C2A_PQP.cpp:930-932)
  for (int k = 0; k < r.found; k++)
    res->Add(o1->tris[list[2 * k]].id, o2->tris[list[2 * k + 1]].id);
  res->query_time_secs = seconds_since(t0);
  return PQP_OK;
}

int C2A_Tolerance(PQP_ToleranceResult* res, PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1,
                  PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2, PQP_REAL tolerance,
                  int qsize) {
  if (o1->build_state != 2 || o2->build_state != 2) return PQP_ERR_UNPROCESSED_MODEL;
  const auto t0 = std::chrono::steady_clock::now();
  c2a_mesh* meshes[2] = {o1->engine_mesh(), o2->engine_mesh()};
  double poses[24];
  pose12_from(R1, T1, poses);
  pose12_from(R2, T2, poses + 12);
  const int32_t mp[2] = {0, 1};
  c2a_static_result r;
  if (c2a_tolerance(meshes, 2, poses, mp, &tolerance, qsize, 1, &r, nullptr) != C2A_OK)
    return -1;
  relative_pose(R1, T1, R2, T2, res->R, res->T);
  res->tolerance = tolerance < 0.0 ? 0.0 : tolerance;   // C2A_PQP.cpp:1786-1787
  res->num_bv_tests = (int)r.bv_tests;
  res->num_tri_tests = (int)r.tri_tests;
  res->closer_than_tolerance = r.found;
  if (qsize > 2) res->qsize = qsize;
  if (r.found) {
    res->distance = r.distance;
    for (int i = 0; i < 3; i++) {
      res->p1[i] = r.p1[i];
      res->p2[i] = r.p2[i];
    }
  } else {
    p2_to_frame2(res->R, res->T, res->p2);
  }
  res->query_time_secs = seconds_since(t0);
  return PQP_OK;
}

int C2A_Distance(C2A_DistanceResult* res, PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1,
                 PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2, PQP_REAL rel_err,
                 PQP_REAL abs_err, int qsize) {
  return distance_query(res, R1, T1, o1, R2, T2, o2, rel_err, abs_err, qsize, false);
}

int C2A_Collide(C2A_DistanceResult* res, PQP_REAL R1[3][3], PQP_REAL T1[3], C2A_Model* o1,
                PQP_REAL R2[3][3], PQP_REAL T2[3], C2A_Model* o2, PQP_REAL rel_err,
                PQP_REAL abs_err, int qsize) {
  return distance_query(res, R1, T1, o1, R2, T2, o2, rel_err, abs_err, qsize, true);
}

PQP_REAL C2A_QueryContact(CInterpMotion* m1, CInterpMotion* m2, C2A_TimeOfContactResult* res,
                          C2A_Model* o1, C2A_Model* o2, double threshold) {
  c2a_mesh* meshes[2] = {o1->engine_mesh(), o2->engine_mesh()};
  double poses[24];
  pose12_from(m1->transform, poses);
  pose12_from(m2->transform, poses + 12);
  if (contact_query(meshes, poses, threshold, res) != C2A_OK) return -1.0;
  return PQP_OK;
}
