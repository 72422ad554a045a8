// ccd_core.h -- the per-group CA loop and exact-order BVTT walk shared by the
// gfx950 kernel (ccd_kernel.hip) and the host-side test harness that runs the
// same code lane by lane.
//
// One lane group per query.  The group leader walks the reference's DFS
// (TOCStepRecurse_Dis, C2A.cpp:1113-1354) over a stack in LDS and makes every
// prune/descend/leaf decision in reference order; when it needs pair values
// it does not have, the whole group computes a speculative depth-K subtree
// below that pair in one round (round_lane).  Pair values never depend on the
// traversal state, so speculation cannot change a result.
#pragma once

#include <stdint.h>

#include "bvh.h"
#include "c2a_math.h"
#include "ccd_types.h"

namespace c2a {

C2A_HD int clz32(unsigned x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __clz((int)x);
#else
  return __builtin_clz(x);
#endif
}

C2A_HD int64_t next_ticket(int64_t* counter) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (int64_t)atomicAdd((unsigned long long*)counter, 1ull);
#else
  return (*counter)++;
#endif
}

struct Entry {           // 128 B stack entry: a considered-but-undecided pair
  double R[9], T[3];     // pair transform (model-2 BV in model-1 BV frame)
  double d, mt;          // RSS distance and admissible time of the pair
  int32_t b1, b2;        // node indices
  int32_t sg;            // (generation << 8) | (slot + 1) of cached values, 0 = none
  int32_t flags;         // bit0 leaf1, bit1 leaf2, bit2 root (always visited)
};

struct Slot {            // speculative pair values, heap-indexed below X
  double R[9], T[3];
  double d, mt;
  double dTri, mleaf;    // leaf pairs: triangle distance and leaf admissible time
  double p[3], q[3];     // leaf pairs: closest points in model-1 frame
  int32_t b1, b2;
  int32_t flags;         // bit0 leaf1, bit1 leaf2, bit2 triangle features set
  int32_t valid;
  int32_t t1, t2;
  int32_t pad[2];
};

template <int K>
struct Shape {
  static constexpr int NSLOT = (1 << (K + 1)) - 1;
  static constexpr int WORK = NSLOT - 1 > 1 ? NSLOT - 1 : 1;
  static constexpr int G = WORK <= 1 ? 1 : WORK <= 2 ? 2 : WORK <= 4 ? 4 : WORK <= 8 ? 8
                         : WORK <= 16 ? 16 : WORK <= 32 ? 32 : 64;
};

template <int K, int CAP>
struct Group {
  Entry stack[CAP];
  Slot slot[Shape<K>::NSLOT];
  Entry X;               // root of the pending round
  double Rrel[9], Trel[3];   // model 2 in model 1's frame (leaf tests)
  double r1[9], tt1[3];      // current pose of object 1 (world directions)
  Motion mo[2];
  int32_t mesh[2];
  int32_t req;           // 1: a round is requested for X
  int32_t contact;       // 1: contact-enumeration traversal (no motion bounds)
  int32_t gen;
  int32_t pad;
};

enum : int {
  ST_FETCH = 0, ST_STEP = 1, ST_WALK = 2, ST_CONTACT_WALK = 3, ST_DONE = 4
};

// Leader-lane registers: the CA loop and step state of one query.
struct Leader {
  int state;
  int64_t qi;
  double lamda, lastLamda;
  int nItrs;
  double D, M, abs_err, rel_err, Up;   // res->distance, mint, errors, UpboundTOC
  int numCA;
  int64_t bv, tri;
  int last1, last2, seed1, seed2;
  double p1[3], p2[3];
  int sp;
  int collisionfree;
  double toc;
  int flags;
  int num_contact;
  double thr;
  int mode;              // MODE_QUERY / MODE_STEP / MODE_CONTACT
};

C2A_HD int node_leaf(const GpuNode* n) { return n->first_child < 0; }

C2A_HD void copy_entry(Entry& dst, const double* R, const double* T, double d,
                                           double mt, int b1, int b2, int sg, int flags) {
  for (int i = 0; i < 9; i++) dst.R[i] = R[i];
  for (int i = 0; i < 3; i++) dst.T[i] = T[i];
  dst.d = d;
  dst.mt = mt;
  dst.b1 = b1;
  dst.b2 = b2;
  dst.sg = sg;
  dst.flags = flags;
}

// Leaf-pair values: C2A.cpp:1145-1166 up to the D-dependent update.
C2A_HD void leaf_values(Slot& s, const MeshDev& M1, const MeshDev& M2, const double* Rrel,
                            const double* Trel, const double* r1, const double* tt1,
                            const Motion& mo1, const Motion& mo2, int b1, int b2, bool contact) {
  const GpuNode* A1 = M1.nodes + b1;
  const GpuNode* A2 = M2.nodes + b2;
  int t1 = -A1->first_child - 1, t2 = -A2->first_child - 1;
  double p[3], q[3];
  int feat[4];
  double dTri = tri_distance(Rrel, Trel, M1.tris + 9 * (int64_t)t1, M2.tris + 9 * (int64_t)t2, p, q,
                             feat);
  double mleaf = 0.0;
  if (!contact) {
    double tmp1[3], tmp2[3], w1[3], w2[3], S1[3], S2[3];
    mxv(tmp1, r1, p);
    add3(w1, tmp1, tt1);
    mxv(tmp2, r1, q);
    add3(w2, tmp2, tt1);
    sub3(S1, w2, w1);
    neg3(S2, S1);
    double mb1 = motion_bound_leaf(mo1, A1->angR, S1);
    double mb2 = motion_bound_leaf(mo2, A2->angR, S2);
    mleaf = dTri / (mb1 + mb2);
    if (mleaf < 0.0) mleaf = 0.0;
  }
  s.dTri = dTri;
  s.mleaf = mleaf;
  for (int i = 0; i < 3; i++) { s.p[i] = p[i]; s.q[i] = q[i]; }
  s.t1 = t1;
  s.t2 = t2;
  s.flags = (s.flags & 3) | ((feat[0] != -1 && feat[2] != -1) ? 4 : 0);
}

// One lane of a round: slot h = gl + 1 of the depth-K subtree below X
// (children chosen by the descend rule C2A.cpp:1195, transforms
// C2A.cpp:1199-1226, values C2A.cpp:1237-1276).
template <int K, int CAP>
C2A_HD void round_lane(Group<K, CAP>& g, int gl, const KArgs& a) {
  const MeshDev M1 = a.meshes[g.mesh[0]];
  const MeshDev M2 = a.meshes[g.mesh[1]];
  const bool contact = g.contact != 0;
  if ((g.X.flags & 3) == 3) {
    if (gl == 0) {
      Slot& s = g.slot[0];
      s.b1 = g.X.b1;
      s.b2 = g.X.b2;
      s.flags = 3;
      s.valid = 1;
      leaf_values(s, M1, M2, g.Rrel, g.Trel, g.r1, g.tt1, g.mo[0], g.mo[1], g.X.b1, g.X.b2, contact);
    }
    return;
  }
  const int h = gl + 1;
  if (h >= Shape<K>::NSLOT) return;
  const int lvl = 31 - clz32(h + 1);
  double R[9], T[3];
  for (int i = 0; i < 9; i++) R[i] = g.X.R[i];
  for (int i = 0; i < 3; i++) T[i] = g.X.T[i];
  int b1 = g.X.b1, b2 = g.X.b2;
  Slot& s = g.slot[h];
  for (int lv = 0; lv < lvl; lv++) {
    const GpuNode* n1 = M1.nodes + b1;
    const GpuNode* n2 = M2.nodes + b2;
    const int fc1 = n1->first_child, fc2 = n2->first_child;
    const bool l1 = fc1 < 0, l2 = fc2 < 0;
    if (l1 && l2) {
      s.valid = 0;
      return;
    }
    const int bit = ((h + 1) >> (lvl - 1 - lv)) & 1;
    double Rn[9], Tn[3], Tt[3];
    if (l2 || (!l1 && (n1->size > n2->size))) {
      const int c = fc1 + bit;
      const GpuNode* cn = M1.nodes + c;
      mtxm(Rn, cn->R, R);
      sub3(Tt, T, cn->Tr);
      mtxv(Tn, cn->R, Tt);
      b1 = c;
    } else {
      const int c = fc2 + bit;
      const GpuNode* cn = M2.nodes + c;
      mxm(Rn, R, cn->R);
      mxvpv(Tn, R, cn->Tr, T);
      b2 = c;
    }
    for (int i = 0; i < 9; i++) R[i] = Rn[i];
    for (int i = 0; i < 3; i++) T[i] = Tn[i];
  }
  const GpuNode* A1 = M1.nodes + b1;
  const GpuNode* A2 = M2.nodes + b2;
  double S[3];
  double d = rect_dist(R, T, A1->l, A2->l, S);
  d -= (A1->r + A2->r);
  d = (d < 0.0) ? 0.0 : d;
  double mt = 0.0;
  if (!contact) {
    if (d != 0.0) {
      double tmp[3], Sw[3], Sn[3];
      mxv(tmp, A1->Rloc, S);
      mxv(Sw, g.r1, tmp);
      neg3(Sn, Sw);
      double mb1 = motion_bound(g.mo[0], A1->angR, Sw);
      double mb2 = motion_bound(g.mo[1], A2->angR, Sn);
      mt = d / (mb1 + mb2);
      if (mt <= 0) mt = 0.0;
    }
  }
  for (int i = 0; i < 9; i++) s.R[i] = R[i];
  for (int i = 0; i < 3; i++) s.T[i] = T[i];
  s.d = d;
  s.mt = mt;
  s.b1 = b1;
  s.b2 = b2;
  s.flags = node_leaf(A1) | (node_leaf(A2) << 1);
  s.valid = 1;
  if ((s.flags & 3) == 3)
    leaf_values(s, M1, M2, g.Rrel, g.Trel, g.r1, g.tt1, g.mo[0], g.mo[1], b1, b2, contact);
}

template <int K, int CAP>
C2A_HD void request_round(Group<K, CAP>& g, Leader& L, Entry& e) {
  int gen = g.gen + 1;
  if (gen >= 0x7fffff) {            // generation wrap: forget every cached link
    for (int i = 0; i < L.sp; i++) g.stack[i].sg = 0;
    gen = 1;
  }
  g.gen = gen;
  e.sg = (gen << 8) | 1;
  g.X = e;
  Slot& s0 = g.slot[0];
  s0.b1 = e.b1;
  s0.b2 = e.b2;
  s0.d = e.d;
  s0.mt = e.mt;
  s0.flags = e.flags & 3;
  s0.valid = 1;
  g.req = 1;
}

// Root pair of a traversal at absolute poses (C2A.cpp:1793-1806).
template <int K, int CAP>
C2A_HD void start_traversal(Group<K, CAP>& g, Leader& L, const MeshDev& M1, const MeshDev& M2,
                                const double* R1, const double* T1, const double* R2,
                                const double* T2) {
  double Tt[3], Rt[9], R[9], T[3];
  mtxm(g.Rrel, R1, R2);
  sub3(Tt, T2, T1);
  mtxv(g.Trel, R1, Tt);
  for (int i = 0; i < 9; i++) g.r1[i] = R1[i];
  for (int i = 0; i < 3; i++) g.tt1[i] = T1[i];
  const GpuNode* r1n = M1.nodes;
  const GpuNode* r2n = M2.nodes;
  mxm(Rt, g.Rrel, r2n->R);
  mtxm(R, r1n->R, Rt);
  mxvpv(Tt, g.Rrel, r2n->Tr, g.Trel);
  sub3(Tt, Tt, r1n->Tr);
  mtxv(T, r1n->R, Tt);
  g.gen = 0;
  L.sp = 1;
  copy_entry(g.stack[0], R, T, 0.0, 0.0, 0, 0, 0,
             node_leaf(r1n) | (node_leaf(r2n) << 1) | 4);
}

// The reference DFS as a stack machine (C2A.cpp:1113-1354 and, in contact
// mode, 1525-1725).  Returns true when a round has been requested.
template <int K, int CAP>
C2A_HD bool walk(Group<K, CAP>& g, Leader& L) {
  const bool contact = g.contact != 0;
  while (L.sp > 0) {
    Entry& e = g.stack[L.sp - 1];
    const int flags = e.flags;
    if (!(flags & 4)) {
      bool desc;
      if (contact) desc = e.d < L.thr;   // abs = rel = 0 (C2A.cpp:1760-1761)
      else desc = e.mt < L.Up && ((e.d < (L.D - L.abs_err)) || (e.d * (1 + L.rel_err) < L.D));
      if (!desc) {
        if (!contact && e.mt < L.M) L.M = e.mt;
        L.sp--;
        continue;
      }
    }
    const int sg = e.sg;
    const bool cached = (sg >> 8) == g.gen && (sg & 0xff) != 0;
    const int s = (sg & 0xff) - 1;
    if ((flags & 3) == 3) {
      if (!cached) {
        request_round(g, L, e);
        return true;
      }
      const Slot& sl = g.slot[s];
      if (!contact) {
        if (sl.dTri <= L.D) {
          L.D = sl.dTri;
          for (int i = 0; i < 3; i++) { L.p1[i] = sl.p[i]; L.p2[i] = sl.q[i]; }
          if (sl.mleaf <= L.M) L.M = sl.mleaf;
          L.last1 = sl.t1;
          L.last2 = sl.t2;
        }
        L.tri++;
      } else {
        if ((sl.flags & 4) && sl.dTri <= L.thr) L.num_contact++;
      }
      L.sp--;
      continue;
    }
    const int lvl = cached ? 31 - clz32(s + 1) : K;
    if (!cached || lvl >= K) {
      request_round(g, L, e);
      return true;
    }
    const int sa = 2 * s + 1, sc = 2 * s + 2;
    if (!contact) L.bv += 2;
    L.sp--;
    if (L.sp + 2 > CAP) {
      L.flags |= C2A_FLAG_STACK_OVERFLOW;
      L.sp = 0;
      return false;
    }
    const Slot& A = g.slot[sa];
    const Slot& C = g.slot[sc];
    const int sgA = (g.gen << 8) | (sa + 1), sgC = (g.gen << 8) | (sc + 1);
    if (C.d < A.d) {   // closer child visited first (C2A.cpp:1281)
      copy_entry(g.stack[L.sp++], A.R, A.T, A.d, A.mt, A.b1, A.b2, sgA, A.flags & 3);
      copy_entry(g.stack[L.sp++], C.R, C.T, C.d, C.mt, C.b1, C.b2, sgC, C.flags & 3);
    } else {
      copy_entry(g.stack[L.sp++], C.R, C.T, C.d, C.mt, C.b1, C.b2, sgC, C.flags & 3);
      copy_entry(g.stack[L.sp++], A.R, A.T, A.d, A.mt, A.b1, A.b2, sgA, A.flags & 3);
    }
  }
  return false;
}

template <int K, int CAP>
C2A_HD void begin_step(Group<K, CAP>& g, Leader& L, const KArgs& a) {
  const MeshDev M1 = a.meshes[g.mesh[0]];
  const MeshDev M2 = a.meshes[g.mesh[1]];
  double R1[9], T1[3], R2[9], T2[3];
  if (L.numCA == 0) {
    for (int i = 0; i < 9; i++) { R1[i] = g.mo[0].R0[i]; R2[i] = g.mo[1].R0[i]; }
    for (int i = 0; i < 3; i++) { T1[i] = g.mo[0].T0[i]; T2[i] = g.mo[1].T0[i]; }
  } else {
    motion_integrate(g.mo[0], L.lamda, R1, T1, nullptr);
    motion_integrate(g.mo[1], L.lamda, R2, T2, nullptr);
  }
  start_traversal(g, L, M1, M2, R1, T1, R2, T2);
  // initial bound from the caller-seeded pair (C2A.cpp:1812-1858)
  double p[3], q[3];
  L.D = tri_distance(g.Rrel, g.Trel, M1.tris + 9 * (int64_t)L.seed1, M2.tris + 9 * (int64_t)L.seed2,
                     p, q, nullptr);
  if (L.numCA == 0) L.M = 1;
  // controlled error schedule (C2A.cpp:1869-1891)
  if (L.M <= 0.005 || L.D <= 0.5 || L.numCA > 5) {
    L.abs_err = 0;
    L.rel_err = 0;
  } else {
    L.abs_err = 1e+30;
    L.rel_err = (L.numCA <= 2) ? 3 : 0.5;
  }
  L.M = 1;
  g.contact = 0;
}

// Finishes the CA loop (C2A.cpp:2125-2143) and optionally starts the contact
// traversal at the TOC pose (C2A_Solve, C2A.cpp:2411-2435).
template <int K, int CAP>
C2A_HD void finish_query(Group<K, CAP>& g, Leader& L, const KArgs& a, bool collisionfree) {
  c2a_result* r = a.out + L.qi;
  r->toc = L.toc;
  r->distance = L.D;
  r->mint = L.M;
  for (int i = 0; i < 3; i++) { r->p1[i] = L.p1[i]; r->p2[i] = L.p2[i]; }
  r->bv_tests = L.bv;
  r->tri_tests = L.tri;
  r->numCA = L.numCA;
  r->collisionfree = collisionfree ? 1 : 0;
  r->last_tri[0] = L.last1;
  r->last_tri[1] = L.last2;
  r->flags = L.flags;
  if (!collisionfree) {
    double Rt[9], Tt[3];
    motion_integrate(g.mo[0], L.toc, Rt, Tt, nullptr);
    for (int i = 0; i < 9; i++) r->R_toc[i] = Rt[i];
    for (int i = 0; i < 3; i++) r->T_toc[i] = Tt[i];
  } else {
    for (int i = 0; i < 9; i++) r->R_toc[i] = 0.0;
    for (int i = 0; i < 3; i++) r->T_toc[i] = 0.0;
  }
  r->num_contact = 0;
  if (!collisionfree && a.with_contacts) {
    const MeshDev M1 = a.meshes[g.mesh[0]];
    const MeshDev M2 = a.meshes[g.mesh[1]];
    double R1[9], T1[3], R2[9], T2[3];
    motion_integrate(g.mo[0], L.toc, R1, T1, nullptr);
    motion_integrate(g.mo[1], L.toc, R2, T2, nullptr);
    L.thr = a.contact_scale * L.D + a.contact_offset;
    L.num_contact = 0;
    start_traversal(g, L, M1, M2, R1, T1, R2, T2);
    g.contact = 1;
    L.state = ST_CONTACT_WALK;
    return;
  }
  L.state = ST_FETCH;
}

// End of a step in the CA loop (C2A.cpp:2029-2123).  Either schedules the
// next step (state ST_STEP) or finishes the query.
template <int K, int CAP>
C2A_HD void after_step(Group<K, CAP>& g, Leader& L, const KArgs& a) {
  if (L.mode == MODE_STEP) {
    // single step: results are the step's distance / mint
    c2a_step_result* r = a.step_out + L.qi;
    r->distance = L.D;
    r->mint = L.M;
    r->abs_err = L.abs_err;
    r->rel_err = L.rel_err;
    for (int i = 0; i < 3; i++) { r->p1[i] = L.p1[i]; r->p2[i] = L.p2[i]; }
    r->bv_tests = L.bv;
    r->tri_tests = L.tri;
    r->last_tri[0] = L.last1;
    r->last_tri[1] = L.last2;
    r->flags = L.flags;
    L.state = ST_FETCH;
    return;
  }
  const double dist = L.D, mint = L.M;
  if (L.numCA == 0) {
    L.numCA = 1;
    L.lastLamda = mint;
    L.nItrs = 0;
  }
  if (L.flags & C2A_FLAG_STACK_OVERFLOW) {
    L.toc = 0;
    finish_query(g, L, a, false);
    return;
  }
  if (dist > a.tol_d) {
    L.nItrs++;
    if (L.nItrs > a.max_iter) {
      L.flags |= C2A_FLAG_ITER_CAP;
    } else if (mint >= 1.0) {
      L.toc = 0;
      finish_query(g, L, a, true);
      return;
    } else {
      const double dlamda = mint;
      if (!(dlamda < a.tol_t)) {
        L.lamda += dlamda;
        if (L.lamda >= 1.0) {
          L.toc = 0;
          finish_query(g, L, a, true);
          return;
        }
        L.lastLamda = L.lamda;
        L.numCA++;
        L.Up = 1.0 - L.lamda;
        L.state = ST_STEP;
        return;
      }
    }
  }
  double toc = L.lastLamda;
  if (toc >= 1 - a.tol_t) toc = 0;
  L.toc = toc;
  finish_query(g, L, a, false);
}

template <int K, int CAP>
C2A_HD bool fetch(Group<K, CAP>& g, Leader& L, const KArgs& a) {
  int64_t qi = next_ticket(a.counter);
  if (qi >= a.nq) {
    L.state = ST_DONE;
    return false;
  }
  L.qi = qi;
  L.flags = 0;
  L.bv = 0;
  L.tri = 0;
  L.num_contact = 0;
  for (int i = 0; i < 3; i++) { L.p1[i] = 0.0; L.p2[i] = 0.0; }
  if (a.mode == MODE_STEP) {
    const StepDev& q = a.step_in[qi];
    g.mo[0] = q.mo[0];
    g.mo[1] = q.mo[1];
    g.mesh[0] = q.mesh[0];
    g.mesh[1] = q.mesh[1];
    L.seed1 = q.seed[0];
    L.seed2 = q.seed[1];
    L.last1 = q.last[0];
    L.last2 = q.last[1];
    L.numCA = q.numCA;
    L.M = q.mint_prev;
    L.Up = q.UpboundTOC;
    L.mode = MODE_STEP;
    const MeshDev M1 = a.meshes[g.mesh[0]];
    const MeshDev M2 = a.meshes[g.mesh[1]];
    start_traversal(g, L, M1, M2, q.R1, q.T1, q.R2, q.T2);
    double p[3], qq[3];
    L.D = tri_distance(g.Rrel, g.Trel, M1.tris + 9 * (int64_t)L.seed1,
                       M2.tris + 9 * (int64_t)L.seed2, p, qq, nullptr);
    if (L.numCA == 0) L.M = 1;
    if (L.M <= 0.005 || L.D <= 0.5 || L.numCA > 5) {
      L.abs_err = 0;
      L.rel_err = 0;
    } else {
      L.abs_err = 1e+30;
      L.rel_err = (L.numCA <= 2) ? 3 : 0.5;
    }
    L.M = 1;
    g.contact = 0;
    L.state = ST_WALK;
    return true;
  }
  if (a.mode == MODE_CONTACT) {
    const ContactDev& c = a.contact_in[qi];
    g.mesh[0] = c.mesh[0];
    g.mesh[1] = c.mesh[1];
    L.mode = MODE_CONTACT;
    L.thr = c.threshold;
    L.num_contact = 0;
    const MeshDev M1 = a.meshes[g.mesh[0]];
    const MeshDev M2 = a.meshes[g.mesh[1]];
    start_traversal(g, L, M1, M2, c.R1, c.T1, c.R2, c.T2);
    g.contact = 1;
    L.state = ST_CONTACT_WALK;
    return true;
  }
  const QueryDev& q = a.q[qi];
  g.mo[0] = q.mo[0];
  g.mo[1] = q.mo[1];
  g.mesh[0] = q.mesh[0];
  g.mesh[1] = q.mesh[1];
  L.seed1 = q.seed[0];
  L.seed2 = q.seed[1];
  L.last1 = q.seed[0];   // model last_tri starts at the seed (demo protocol)
  L.last2 = q.seed[1];
  L.mode = MODE_QUERY;
  L.numCA = 0;
  L.lamda = 0.0;
  L.lastLamda = 0.0;
  L.nItrs = 0;
  L.Up = 1;
  L.M = 1;
  if (q.mo[0].angVel < 1e-8 && q.mo[1].angVel < 1e-8) L.flags |= C2A_FLAG_TRANSLATION_ONLY;
  L.state = ST_STEP;
  return true;
}

// Runs the leader until a round is needed or no work is left.
template <int K, int CAP>
C2A_HD void leader_run(Group<K, CAP>& g, Leader& L, const KArgs& a) {
  g.req = 0;
  for (;;) {
    switch (L.state) {
      case ST_FETCH:
        if (!fetch(g, L, a)) return;
        break;
      case ST_STEP:
        begin_step(g, L, a);
        L.state = ST_WALK;
        break;
      case ST_WALK:
        if (walk(g, L)) return;
        after_step(g, L, a);
        break;
      case ST_CONTACT_WALK:
        if (walk(g, L)) return;
        if (L.mode == MODE_CONTACT) a.counts_out[L.qi] = L.num_contact;
        else a.out[L.qi].num_contact = L.num_contact;
        L.state = ST_FETCH;
        break;
      default:
        return;
    }
  }
}

}  // namespace c2a
