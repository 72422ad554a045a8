// ccd_core.h -- the per-query CA loop and exact-order BVTT walk, shared by the
// gfx950 kernel (ccd_kernel.hip) and the host-side test harness that runs the
// same code with a single lane.
//
// One wavefront per query.  All lanes run the walk of the reference's DFS
// (TOCStepRecurse_Dis, C2A.cpp:1113-1354) in lockstep over a stack in LDS, so
// every prune/descend/leaf decision is made in reference order with uniform
// control flow; stack entries are copied one 8-byte word per lane.  When the
// walk needs pair values it does not have, the lanes compute a speculative
// depth-K subtree below that pair in one round (round_lane): lane j chains the
// child transforms down to heap slot j+1 and evaluates the RSS distance, the
// directional motion bound and, for leaf pairs, the triangle distance.  Pair
// values never depend on the traversal state, so speculation cannot change a
// result; it only lets one round feed up to K levels of the walk.
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

// Makes the LDS writes of all lanes visible to all lanes (one wave per block).
C2A_HD void lane_sync() {
#if defined(__HIP_DEVICE_COMPILE__)
  __syncthreads();
#endif
}

// Node data is read-only global memory; the address-space qualifier lets the
// compiler emit global (not flat) loads.
#if defined(__HIP_DEVICE_COMPILE__)
#define C2A_GLOBAL __attribute__((address_space(1)))
#else
#define C2A_GLOBAL
#endif
typedef const GpuNode C2A_GLOBAL* node_ptr;
typedef const double C2A_GLOBAL* gdouble_ptr;

C2A_HD node_ptr nodes_of(const MeshDev& m) { return (node_ptr)m.nodes; }
C2A_HD gdouble_ptr tris_of(const MeshDev& m) { return (gdouble_ptr)m.tris; }

struct Entry {           // 128 B stack entry: a considered-but-undecided pair
  double R[9], T[3];     // pair transform (model-2 BV in model-1 BV frame)
  double d, mt;          // RSS distance and admissible time of the pair
  int32_t b1, b2;        // node indices
  int32_t sg;            // (generation << 8) | (slot + 1) of cached values, 0 = none
  int32_t flags;         // bit0 leaf1, bit1 leaf2, bit2 root (always visited)
};

struct Slot {            // speculative pair values, heap-indexed below X
  double R[9], T[3];     // same first 14 words as Entry
  double d, mt;
  int32_t b1, b2;        // word 14, same place as in Entry
  int32_t flags;         // bit0 leaf1, bit1 leaf2, bit2 triangle features set
  int32_t valid;
  double dTri, mleaf;    // leaf pairs: triangle distance and leaf admissible time
  double p[3], q[3];     // leaf pairs: closest points in model-1 frame
  int32_t t1, t2;
};

template <int K>
struct Shape {
  static constexpr int NSLOT = (1 << (K + 1)) - 1;
  static constexpr int PER_LANE = (NSLOT - 1 + 63) / 64;
};

template <int K, int CAP>
struct Wave {
  Entry stack[CAP];
  Slot slot[Shape<K>::NSLOT];
  Entry X;                   // root of the pending round
  double Rrel[9], Trel[3];   // model 2 in model 1's frame (leaf tests)
  double r1[9], tt1[3];      // current pose of object 1 (world directions)
  Motion mo[2];
  int32_t mesh[2];
  int32_t contact;           // 1: contact-enumeration traversal (no motion bounds)
  int32_t gen;
};

enum : int { ST_FETCH = 0, ST_STEP = 1, ST_WALK = 2, ST_CONTACT_WALK = 3, ST_DONE = 4 };

// The CA loop and step state of one query; every lane holds the same values.
struct Leader {
  int state;
  int need_round;
  int64_t qi;
  double lamda, lastLamda;
  int nItrs;
  double D, M, abs_err, rel_err, Up;   // res->distance, mint, errors, UpboundTOC
  int numCA;
  int64_t bv, tri;
  int last1, last2, seed1, seed2;
  double p1[3], p2[3];
  int sp;
  double toc;
  int flags;
  int num_contact;
  double thr;
  int mode;
};

C2A_HD int node_leaf(node_ptr n) { return n->first_child < 0; }

// Lane-parallel copy of the 14 value words + node ids of a slot into a stack
// entry; word 15 (sg, flags) is written by one lane.
C2A_HD void slot_to_entry(Entry& dst, const Slot& src, int sg, int flags, int lane, int nl) {
  const uint64_t* s = reinterpret_cast<const uint64_t*>(&src);
  uint64_t* d = reinterpret_cast<uint64_t*>(&dst);
  for (int w = lane; w < 16; w += nl) {
    if (w < 15) d[w] = s[w];
    else { dst.sg = sg; dst.flags = flags; }
  }
}

// Leaf-pair values: C2A.cpp:1145-1166 up to the distance-dependent update.
C2A_HD void leaf_values(Slot& s, const MeshDev& M1, const MeshDev& M2, const double* Rrel,
                        const double* Trel, const double* r1, const double* tt1,
                        const Motion& mo1, const Motion& mo2, int b1, int b2, bool contact) {
  node_ptr A1 = nodes_of(M1) + b1;
  node_ptr A2 = nodes_of(M2) + b2;
  int t1 = -A1->first_child - 1, t2 = -A2->first_child - 1;
  double p[3], q[3], ta[9], tb[9];
  gdouble_ptr g1 = tris_of(M1) + 9 * (int64_t)t1;
  gdouble_ptr g2 = tris_of(M2) + 9 * (int64_t)t2;
  for (int i = 0; i < 9; i++) { ta[i] = g1[i]; tb[i] = g2[i]; }
  int feat[4];
  double dTri = tri_distance(Rrel, Trel, ta, tb, p, q, feat);
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

// Heap slot h of the depth-K subtree below X: children chosen by the descend
// rule (C2A.cpp:1195), transforms (C2A.cpp:1199-1226), values
// (C2A.cpp:1237-1276).
template <int K, int CAP>
C2A_HD void compute_slot(Wave<K, CAP>& g, int h, const MeshDev& M1, const MeshDev& M2,
                         bool contact) {
  node_ptr N1 = nodes_of(M1);
  node_ptr N2 = nodes_of(M2);
  const int lvl = 31 - clz32((unsigned)(h + 1));
  double R[9], T[3];
  for (int i = 0; i < 9; i++) R[i] = g.X.R[i];
  for (int i = 0; i < 3; i++) T[i] = g.X.T[i];
  int b1 = g.X.b1, b2 = g.X.b2;
  Slot& s = g.slot[h];
  for (int lv = 0; lv < lvl; lv++) {
    node_ptr n1 = N1 + b1;
    node_ptr n2 = N2 + b2;
    const int fc1 = n1->first_child, fc2 = n2->first_child;
    const bool l1 = fc1 < 0, l2 = fc2 < 0;
    if (l1 && l2) {
      s.valid = 0;
      return;
    }
    const int bit = ((h + 1) >> (lvl - 1 - lv)) & 1;
    double Rn[9], Tn[3], Tt[3], CR[9], CT[3];
    if (l2 || (!l1 && (n1->size > n2->size))) {
      const int c = fc1 + bit;
      node_ptr cn = N1 + c;
      for (int i = 0; i < 9; i++) CR[i] = cn->R[i];
      for (int i = 0; i < 3; i++) CT[i] = cn->Tr[i];
      mtxm(Rn, CR, R);
      sub3(Tt, T, CT);
      mtxv(Tn, CR, Tt);
      b1 = c;
    } else {
      const int c = fc2 + bit;
      node_ptr cn = N2 + c;
      for (int i = 0; i < 9; i++) CR[i] = cn->R[i];
      for (int i = 0; i < 3; i++) CT[i] = cn->Tr[i];
      mxm(Rn, R, CR);
      mxvpv(Tn, R, CT, T);
      b2 = c;
    }
    for (int i = 0; i < 9; i++) R[i] = Rn[i];
    for (int i = 0; i < 3; i++) T[i] = Tn[i];
  }
  node_ptr A1 = N1 + b1;
  node_ptr A2 = N2 + b2;
  double S[3], la[2] = {A1->l[0], A1->l[1]}, lb[2] = {A2->l[0], A2->l[1]};
  double d = rect_dist(R, T, la, lb, S);
  d -= (A1->r + A2->r);
  d = (d < 0.0) ? 0.0 : d;
  double mt = 0.0;
  if (!contact) {
    if (d != 0.0) {
      double tmp[3], Sw[3], Sn[3], RL[9];
      for (int i = 0; i < 9; i++) RL[i] = A1->Rloc[i];
      mxv(tmp, RL, S);
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

// One lane's share of a round.
template <int K, int CAP>
C2A_HD void round_lane(Wave<K, CAP>& g, int lane, int nl, const KArgs& a) {
  const MeshDev M1 = a.meshes[g.mesh[0]];
  const MeshDev M2 = a.meshes[g.mesh[1]];
  const bool contact = g.contact != 0;
  if ((g.X.flags & 3) == 3) {
    if (lane == 0) {
      Slot& s = g.slot[0];
      s.b1 = g.X.b1;
      s.b2 = g.X.b2;
      s.flags = 3;
      s.valid = 1;
      leaf_values(s, M1, M2, g.Rrel, g.Trel, g.r1, g.tt1, g.mo[0], g.mo[1], g.X.b1, g.X.b2,
                  contact);
    }
    return;
  }
  for (int h = lane + 1; h < Shape<K>::NSLOT; h += nl) compute_slot(g, h, M1, M2, contact);
}

template <int K, int CAP>
C2A_HD void request_round(Wave<K, CAP>& g, Leader& L, int lane, int nl) {
  Entry& e = g.stack[L.sp - 1];
  int gen = g.gen + 1;
  if (gen >= 0x7fffff) {   // generation wrap: forget every cached link
    for (int i = lane; i < L.sp; i += nl) g.stack[i].sg = 0;
    gen = 1;
  }
  const int flags = e.flags;
  const int sg = (gen << 8) | 1;
  lane_sync();
  const uint64_t* s = reinterpret_cast<const uint64_t*>(&e);
  uint64_t* x = reinterpret_cast<uint64_t*>(&g.X);
  for (int w = lane; w < 16; w += nl) x[w] = s[w];
  lane_sync();
  if (lane == 0) {
    e.sg = sg;
    Slot& s0 = g.slot[0];
    s0.b1 = g.X.b1;
    s0.b2 = g.X.b2;
    s0.d = g.X.d;
    s0.mt = g.X.mt;
    s0.flags = flags & 3;
    s0.valid = 1;
    g.gen = gen;
  }
  lane_sync();
  L.need_round = 1;
}

// Root pair of a traversal at absolute poses (C2A.cpp:1793-1806).
template <int K, int CAP>
C2A_HD void start_traversal(Wave<K, CAP>& g, Leader& L, const MeshDev& M1, const MeshDev& M2,
                            const double* R1, const double* T1, const double* R2,
                            const double* T2, int lane) {
  double Tt[3], Rt[9], R[9], T[3], Rrel[9], Trel[3], C1R[9], C2R[9], C1T[3], C2T[3];
  node_ptr r1n = nodes_of(M1);
  node_ptr r2n = nodes_of(M2);
  for (int i = 0; i < 9; i++) { C1R[i] = r1n->R[i]; C2R[i] = r2n->R[i]; }
  for (int i = 0; i < 3; i++) { C1T[i] = r1n->Tr[i]; C2T[i] = r2n->Tr[i]; }
  mtxm(Rrel, R1, R2);
  sub3(Tt, T2, T1);
  mtxv(Trel, R1, Tt);
  mxm(Rt, Rrel, C2R);
  mtxm(R, C1R, Rt);
  mxvpv(Tt, Rrel, C2T, Trel);
  sub3(Tt, Tt, C1T);
  mtxv(T, C1R, Tt);
  lane_sync();
  if (lane == 0) {
    for (int i = 0; i < 9; i++) { g.Rrel[i] = Rrel[i]; g.r1[i] = R1[i]; }
    for (int i = 0; i < 3; i++) { g.Trel[i] = Trel[i]; g.tt1[i] = T1[i]; }
    g.gen = 0;
    Entry& e = g.stack[0];
    for (int i = 0; i < 9; i++) e.R[i] = R[i];
    for (int i = 0; i < 3; i++) e.T[i] = T[i];
    e.d = 0.0;
    e.mt = 0.0;
    e.b1 = 0;
    e.b2 = 0;
    e.sg = 0;
    e.flags = node_leaf(r1n) | (node_leaf(r2n) << 1) | 4;
  }
  lane_sync();
  L.sp = 1;
}

// The reference DFS as a stack machine (C2A.cpp:1113-1354 and, in contact
// mode, 1525-1725).  Uniform across lanes.  Returns true when a round has
// been requested.
template <int K, int CAP>
C2A_HD bool walk(Wave<K, CAP>& g, Leader& L, int lane, int nl) {
  const bool contact = g.contact != 0;
  const int gen = g.gen;
  while (L.sp > 0) {
    const Entry& e = g.stack[L.sp - 1];
    const int flags = e.flags;
    if (!(flags & 4)) {
      const double d = e.d, mt = e.mt;
      bool desc;
      if (contact) desc = d < L.thr;   // abs = rel = 0 (C2A.cpp:1760-1761)
      else desc = mt < L.Up && ((d < (L.D - L.abs_err)) || (d * (1 + L.rel_err) < L.D));
      if (!desc) {
        if (!contact && mt < L.M) L.M = mt;
        L.sp--;
        continue;
      }
    }
    const int sg = e.sg;
    const bool cached = (sg >> 8) == gen && (sg & 0xff) != 0;
    const int s = (sg & 0xff) - 1;
    if ((flags & 3) == 3) {
      if (!cached) {
        request_round(g, L, lane, nl);
        return true;
      }
      const Slot& sl = g.slot[s];
      if (!contact) {
        const double dTri = sl.dTri;
        if (dTri <= L.D) {
          L.D = dTri;
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
    const int lvl = cached ? 31 - clz32((unsigned)(s + 1)) : K;
    if (!cached || lvl >= K) {
      request_round(g, L, lane, nl);
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
    const int sgA = (gen << 8) | (sa + 1), sgC = (gen << 8) | (sc + 1);
    const int fA = A.flags & 3, fC = C.flags & 3;
    // closer child visited first (C2A.cpp:1281): it goes on top
    const bool c_first = C.d < A.d;
    lane_sync();
    if (c_first) {
      slot_to_entry(g.stack[L.sp], A, sgA, fA, lane, nl);
      slot_to_entry(g.stack[L.sp + 1], C, sgC, fC, lane, nl);
    } else {
      slot_to_entry(g.stack[L.sp], C, sgC, fC, lane, nl);
      slot_to_entry(g.stack[L.sp + 1], A, sgA, fA, lane, nl);
    }
    lane_sync();
    L.sp += 2;
  }
  return false;
}

template <int K, int CAP>
C2A_HD void begin_step(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane) {
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
  start_traversal(g, L, M1, M2, R1, T1, R2, T2, lane);
  // initial bound from the caller-seeded pair (C2A.cpp:1812-1858)
  double p[3], q[3], ta[9], tb[9];
  gdouble_ptr g1 = tris_of(M1) + 9 * (int64_t)L.seed1;
  gdouble_ptr g2 = tris_of(M2) + 9 * (int64_t)L.seed2;
  for (int i = 0; i < 9; i++) { ta[i] = g1[i]; tb[i] = g2[i]; }
  double Rrel[9], Trel[3];
  for (int i = 0; i < 9; i++) Rrel[i] = g.Rrel[i];
  for (int i = 0; i < 3; i++) Trel[i] = g.Trel[i];
  L.D = tri_distance(Rrel, Trel, ta, tb, p, q, nullptr);
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
  if (lane == 0) g.contact = 0;
  lane_sync();
}

// Finishes the CA loop (C2A.cpp:2125-2143) and optionally starts the contact
// traversal at the TOC pose (C2A_Solve, C2A.cpp:2411-2435).
template <int K, int CAP>
C2A_HD void finish_query(Wave<K, CAP>& g, Leader& L, const KArgs& a, bool collisionfree,
                         int lane) {
  double Rt[9], Tt[3];
  if (!collisionfree) motion_integrate(g.mo[0], L.toc, Rt, Tt, nullptr);
  if (lane == 0) {
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
    for (int i = 0; i < 9; i++) r->R_toc[i] = collisionfree ? 0.0 : Rt[i];
    for (int i = 0; i < 3; i++) r->T_toc[i] = collisionfree ? 0.0 : Tt[i];
    r->num_contact = 0;
  }
  if (!collisionfree && a.with_contacts) {
    const MeshDev M1 = a.meshes[g.mesh[0]];
    const MeshDev M2 = a.meshes[g.mesh[1]];
    double R1[9], T1[3], R2[9], T2[3];
    motion_integrate(g.mo[0], L.toc, R1, T1, nullptr);
    motion_integrate(g.mo[1], L.toc, R2, T2, nullptr);
    L.thr = a.contact_scale * L.D + a.contact_offset;
    L.num_contact = 0;
    start_traversal(g, L, M1, M2, R1, T1, R2, T2, lane);
    if (lane == 0) g.contact = 1;
    lane_sync();
    L.state = ST_CONTACT_WALK;
    return;
  }
  L.state = ST_FETCH;
}

// End of a step in the CA loop (C2A.cpp:2029-2123): schedules the next step
// (state ST_STEP) or finishes the query.
template <int K, int CAP>
C2A_HD void after_step(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane) {
  if (L.mode == MODE_STEP) {
    if (lane == 0) {
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
    }
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
    finish_query(g, L, a, false, lane);
    return;
  }
  if (dist > a.tol_d) {
    L.nItrs++;
    if (L.nItrs > a.max_iter) {
      L.flags |= C2A_FLAG_ITER_CAP;
    } else if (mint >= 1.0) {
      L.toc = 0;
      finish_query(g, L, a, true, lane);
      return;
    } else {
      const double dlamda = mint;
      if (!(dlamda < a.tol_t)) {
        L.lamda += dlamda;
        if (L.lamda >= 1.0) {
          L.toc = 0;
          finish_query(g, L, a, true, lane);
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
  finish_query(g, L, a, false, lane);
}

// Takes the next query from the work queue (one atomic per wavefront).
C2A_HD int64_t take_ticket(int64_t* counter, int lane) {
#if defined(__HIP_DEVICE_COMPILE__)
  long long t = 0;
  if (lane == 0) t = (long long)atomicAdd((unsigned long long*)counter, 1ull);
  return (int64_t)__shfl(t, 0);
#else
  (void)lane;
  return (*counter)++;
#endif
}

template <int K, int CAP>
C2A_HD bool fetch(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane) {
  int64_t qi = take_ticket(a.counter, lane);
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
  lane_sync();
  if (a.mode == MODE_STEP) {
    const StepDev& q = a.step_in[qi];
    if (lane == 0) {
      g.mo[0] = q.mo[0];
      g.mo[1] = q.mo[1];
      g.mesh[0] = q.mesh[0];
      g.mesh[1] = q.mesh[1];
      g.contact = 0;
    }
    lane_sync();
    L.seed1 = q.seed[0];
    L.seed2 = q.seed[1];
    L.last1 = q.last[0];
    L.last2 = q.last[1];
    L.numCA = q.numCA;
    L.M = q.mint_prev;
    L.Up = q.UpboundTOC;
    L.mode = MODE_STEP;
    const MeshDev M1 = a.meshes[q.mesh[0]];
    const MeshDev M2 = a.meshes[q.mesh[1]];
    double R1[9], T1[3], R2[9], T2[3];
    for (int i = 0; i < 9; i++) { R1[i] = q.R1[i]; R2[i] = q.R2[i]; }
    for (int i = 0; i < 3; i++) { T1[i] = q.T1[i]; T2[i] = q.T2[i]; }
    start_traversal(g, L, M1, M2, R1, T1, R2, T2, lane);
    double p[3], qq[3], ta[9], tb[9], Rrel[9], Trel[3];
    gdouble_ptr g1 = tris_of(M1) + 9 * (int64_t)L.seed1;
    gdouble_ptr g2 = tris_of(M2) + 9 * (int64_t)L.seed2;
    for (int i = 0; i < 9; i++) { ta[i] = g1[i]; tb[i] = g2[i]; Rrel[i] = g.Rrel[i]; }
    for (int i = 0; i < 3; i++) Trel[i] = g.Trel[i];
    L.D = tri_distance(Rrel, Trel, ta, tb, p, qq, nullptr);
    if (L.numCA == 0) L.M = 1;
    if (L.M <= 0.005 || L.D <= 0.5 || L.numCA > 5) {
      L.abs_err = 0;
      L.rel_err = 0;
    } else {
      L.abs_err = 1e+30;
      L.rel_err = (L.numCA <= 2) ? 3 : 0.5;
    }
    L.M = 1;
    L.state = ST_WALK;
    return true;
  }
  if (a.mode == MODE_CONTACT) {
    const ContactDev& c = a.contact_in[qi];
    if (lane == 0) {
      g.mesh[0] = c.mesh[0];
      g.mesh[1] = c.mesh[1];
      g.contact = 1;
    }
    lane_sync();
    L.mode = MODE_CONTACT;
    L.thr = c.threshold;
    L.num_contact = 0;
    const MeshDev M1 = a.meshes[c.mesh[0]];
    const MeshDev M2 = a.meshes[c.mesh[1]];
    double R1[9], T1[3], R2[9], T2[3];
    for (int i = 0; i < 9; i++) { R1[i] = c.R1[i]; R2[i] = c.R2[i]; }
    for (int i = 0; i < 3; i++) { T1[i] = c.T1[i]; T2[i] = c.T2[i]; }
    start_traversal(g, L, M1, M2, R1, T1, R2, T2, lane);
    L.state = ST_CONTACT_WALK;
    return true;
  }
  const QueryDev& q = a.q[qi];
  if (lane == 0) {
    g.mo[0] = q.mo[0];
    g.mo[1] = q.mo[1];
    g.mesh[0] = q.mesh[0];
    g.mesh[1] = q.mesh[1];
  }
  lane_sync();
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

// Advances the query state until a round is needed or no work is left.
template <int K, int CAP>
C2A_HD void leader_run(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane, int nl) {
  L.need_round = 0;
  for (;;) {
    switch (L.state) {
      case ST_FETCH:
        if (!fetch(g, L, a, lane)) return;
        break;
      case ST_STEP:
        begin_step(g, L, a, lane);
        L.state = ST_WALK;
        break;
      case ST_WALK:
        if (walk(g, L, lane, nl)) return;
        after_step(g, L, a, lane);
        break;
      case ST_CONTACT_WALK:
        if (walk(g, L, lane, nl)) return;
        if (lane == 0) {
          if (L.mode == MODE_CONTACT) a.counts_out[L.qi] = L.num_contact;
          else a.out[L.qi].num_contact = L.num_contact;
        }
        L.state = ST_FETCH;
        break;
      default:
        return;
    }
  }
}

}  // namespace c2a
