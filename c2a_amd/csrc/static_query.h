// static_query.h -- static proximity queries, one thread per query: the
// best-first distance search (C2A_Distance with qsize > 2), the distance at
// contact space, the tolerance query and the triangle-contact form of
// C2A_Collide (C2A/src/C2A_PQP.cpp:624-968, :1282-1845).  Shared by the gfx950
// kernel (static_kernel.hip) and the host-side test harness.
//
// These searches carry little state beyond a stack or a bounded priority
// queue, and a batch of them is many short independent traversals, so each
// thread runs one query with its stack / queues in a private slice of a
// global workspace.  The arithmetic is the CCD kernel's (c2a_math.h, FP64,
// same operation order), so the results equal the FP64 oracle bit for bit.
#pragma once

#include <stdint.h>

#include "bvh.h"
#include "c2a_math.h"
#include "ccd_types.h"

#ifndef C2A_CHECK
#define C2A_CHECK(cond) ((void)0)
#endif

namespace c2a {

enum : int32_t {
  SQ_DISTANCE_QUEUE = 0,   // C2A_Distance, qsize > 2 (C2A_PQP.cpp:624-786)
  SQ_DISTANCE_CS = 1,      // C2A_Distance_At_ContactSpace (:1282-1496)
  SQ_TOLERANCE = 2,        // C2A_Tolerance (:1498-1845)
  SQ_COLLIDE = 3,          // C2A_Collide, PQP_CollideResult form (:798-968)
};
// C2A_FLAG_STACK_OVERFLOW: the workspace could not hold the search
constexpr int32_t SQ_FLAG_OVERFLOW = 4;

// One query as the kernel reads it.
struct StaticQueryDev {
  double R1[9], T1[3], R2[9], T2[3];   // absolute poses
  double param;      // SQ_DISTANCE_CS: upper bound; SQ_TOLERANCE: tolerance
  int32_t mesh[2];
  int32_t seed[2];   // SQ_DISTANCE_QUEUE: the models' last_tri (build order)
};

// One query's result (the C-ABI's c2a_static_result, include/c2a_amd.h).
struct StaticOut {
  double distance;
  double p1[3], p2[3];
  int64_t bv_tests, tri_tests;
  int32_t tri[2];    // the pair found / the models' last_tri after the query
  int32_t found;     // tolerance: closer; contact space: found; collide: pairs
  int32_t flags;     // C2A_FLAG_STACK_OVERFLOW: the workspace was too small
};

// A pair under test: transform of model 2's volume in model 1's volume frame
// and, in the queues, its RSS distance.
struct SqEntry {
  double R[9], T[3];
  double d;
  int32_t b1, b2;
};

struct StaticArgs {
  const MeshDev* meshes;
  const StaticQueryDev* q;
  StaticOut* out;
  int32_t* pairs;          // SQ_COLLIDE: nq x pair_cap x 2 build-order triangles
  SqEntry* ws;             // per-thread workspace, ws_entries each
  int64_t* counter;        // work queue ticket
  int64_t nq;
  int64_t ws_entries;
  int32_t kind;
  int32_t qsize;           // SQ_DISTANCE_QUEUE / SQ_TOLERANCE (> 2: best first)
  int32_t flag;            // SQ_COLLIDE: 2 = first contact only
  int32_t pair_cap;
  double rel_err, abs_err;
};

C2A_HD bool sq_leaf(const GpuNode* n) { return n->first_child < 0; }

// Root pair transform and model 2 in model 1's frame (C2A_PQP.cpp:990-1020).
C2A_HD void sq_root(const GpuNode* n1, const GpuNode* n2, const StaticQueryDev& q, double* Rrel,
                    double* Trel, double* R, double* T) {
  double Tt[3], Rt[9], CR[9], CT[3];
  mtxm(Rrel, q.R1, q.R2);
  sub3(Tt, q.T2, q.T1);
  mtxv(Trel, q.R1, Tt);
  for (int i = 0; i < 9; i++) CR[i] = n2->R[i];
  mxm(Rt, Rrel, CR);
  for (int i = 0; i < 9; i++) CR[i] = n1->R[i];
  mtxm(R, CR, Rt);
  for (int i = 0; i < 3; i++) CT[i] = n2->Tr[i];
  mxvpv(Tt, Rrel, CT, Trel);
  for (int i = 0; i < 3; i++) CT[i] = n1->Tr[i];
  sub3(Tt, Tt, CT);
  mtxv(T, CR, Tt);
}

// The two child pairs of (b1, b2) and their transforms: model 1's children
// when model 2's node is a leaf or model 1's is larger, else model 2's
// (C2A_PQP.cpp:527-576).
C2A_HD void sq_expand(const GpuNode* N1, const GpuNode* N2, const SqEntry& p, SqEntry& a,
                      SqEntry& c) {
  const GpuNode* B1 = N1 + p.b1;
  const GpuNode* B2 = N2 + p.b2;
  const bool l1 = sq_leaf(B1), l2 = sq_leaf(B2);
  double CR[9], CT[3], Tt[3];
  if (l2 || (!l1 && (B1->size > B2->size))) {
    SqEntry* out[2] = {&a, &c};
    for (int k = 0; k < 2; k++) {
      const GpuNode* C = N1 + B1->first_child + k;
      for (int i = 0; i < 9; i++) CR[i] = C->R[i];
      for (int i = 0; i < 3; i++) CT[i] = C->Tr[i];
      mtxm(out[k]->R, CR, p.R);
      sub3(Tt, p.T, CT);
      mtxv(out[k]->T, CR, Tt);
      out[k]->b1 = B1->first_child + k;
      out[k]->b2 = p.b2;
    }
  } else {
    SqEntry* out[2] = {&a, &c};
    for (int k = 0; k < 2; k++) {
      const GpuNode* C = N2 + B2->first_child + k;
      for (int i = 0; i < 9; i++) CR[i] = C->R[i];
      for (int i = 0; i < 3; i++) CT[i] = C->Tr[i];
      mxm(out[k]->R, p.R, CR);
      mxvpv(out[k]->T, p.R, CT, p.T);
      out[k]->b1 = p.b1;
      out[k]->b2 = B2->first_child + k;
    }
  }
}

// RSS rectangle distance of a pair (C2ARectDist, C2A_BV.cpp:658-664) and the
// clamped volume distance C2A_BV_Distance (:666-675).
C2A_HD double sq_rect(const GpuNode* N1, const GpuNode* N2, const SqEntry& e) {
  const GpuNode* A = N1 + e.b1;
  const GpuNode* B = N2 + e.b2;
  double S[3], la[2] = {A->l[0], A->l[1]}, lb[2] = {B->l[0], B->l[1]};
  return rect_dist(e.R, e.T, la, lb, S);
}
C2A_HD double sq_bvdist(const GpuNode* N1, const GpuNode* N2, const SqEntry& e) {
  double d = sq_rect(N1, N2, e);
  d -= (N1[e.b1].r + N2[e.b2].r);
  return (d < 0.0) ? 0.0 : d;
}

// Triangle distance of a leaf pair with model 2 mapped into model 1's frame.
C2A_HD double sq_tri(const MeshDev& M1, const MeshDev& M2, const double* Rrel,
                     const double* Trel, int t1, int t2, double* p, double* q) {
  double a[9], b[9];
  for (int i = 0; i < 9; i++) a[i] = M1.tris[9 * (int64_t)t1 + i];
  for (int i = 0; i < 9; i++) b[i] = M2.tris[9 * (int64_t)t2 + i];
  return tri_distance(Rrel, Trel, a, b, p, q, nullptr);
}

C2A_HD void sq_copy(SqEntry& dst, const SqEntry& src) {
  for (int i = 0; i < 9; i++) dst.R[i] = src.R[i];
  for (int i = 0; i < 3; i++) dst.T[i] = src.T[i];
  dst.d = src.d;
  dst.b1 = src.b1;
  dst.b2 = src.b2;
}

// PQP's BVTQ restated (C2A_PQP.cpp:626 and :1610; PQP 1.3 is not vendored):
// a binary min-heap on d, h[1..n] in `heap`.
C2A_HD void heap_add(SqEntry* heap, int& n, int cap, const SqEntry& t) {
  C2A_CHECK(n + 1 <= cap);
  int i = ++n;
  while (i > 1 && heap[i / 2].d > t.d) {
    sq_copy(heap[i], heap[i / 2]);
    i /= 2;
  }
  sq_copy(heap[i], t);
}
C2A_HD void heap_extract_min(SqEntry* heap, int& n, SqEntry& top) {
  sq_copy(top, heap[1]);
  SqEntry last;
  sq_copy(last, heap[n]);
  n--;
  int i = 1;
  for (;;) {
    int c = 2 * i;
    if (c > n) break;
    if (c + 1 <= n && heap[c + 1].d < heap[c].d) c++;
    if (!(heap[c].d < last.d)) break;
    sq_copy(heap[i], heap[c]);
    i = c;
  }
  if (n >= 1) sq_copy(heap[i], last);
}

// Workspace entries one thread needs: a stack of one pending pair per tree
// level for the depth-first searches, one queue of qsize + 1 slots per
// nesting level for the best-first ones (a nested queue's root lies at least
// one level deeper than its parent queue's).
C2A_HD int64_t sq_ws_entries(int kind, int qsize, int depth_sum) {
  const int64_t levels = (int64_t)depth_sum + 2;
  const bool queue = (kind == SQ_DISTANCE_QUEUE) || (kind == SQ_TOLERANCE && qsize > 2);
  return queue ? levels * ((int64_t)qsize + 1) : levels + 1;
}

// Best-first search: C2ADistanceQueueRecurse (C2A_PQP.cpp:624-786) and
// C2AToleranceQueueRecurse (:1606-1759).  A full queue (qsize - 1 tests)
// runs a nested search rooted at the test just taken; the nested queue sits
// after its parent's in the workspace.
C2A_HD void sq_queue_search(const StaticArgs& a, const MeshDev& M1, const MeshDev& M2,
                            const double* Rrel, const double* Trel, const SqEntry& root,
                            SqEntry* ws, StaticOut& o, bool tolerance, double tol) {
  const GpuNode* N1 = M1.nodes;
  const GpuNode* N2 = M2.nodes;
  const int qs = a.qsize;
  const int64_t slot = (int64_t)qs + 1;
  const int max_level = (int)(a.ws_entries / slot) - 1;
  int level = 0;
  int n = 0;   // tests in the current level's queue; a suspended outer
               // queue keeps its count in its slot 0 (the heap uses 1..qsize)
  SqEntry mt;
  sq_copy(mt, root);
  for (;;) {
    C2A_CHECK(level <= max_level);
    SqEntry* heap = ws + (int64_t)level * slot;
    const GpuNode* B1 = N1 + mt.b1;
    const GpuNode* B2 = N2 + mt.b2;
    bool leave = false;
    if (sq_leaf(B1) && sq_leaf(B2)) {
      const int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
      double p[3], q[3];
      o.tri_tests++;
      const double d = sq_tri(M1, M2, Rrel, Trel, t1, t2, p, q);
      if (tolerance) {
        if (d <= tol) {   // :1634-1645: closer than tolerance, done
          o.found = 1;
          o.distance = d;
          for (int i = 0; i < 3; i++) { o.p1[i] = p[i]; o.p2[i] = q[i]; }
          o.tri[0] = t1;
          o.tri[1] = t2;
          return;
        }
      } else if (d < o.distance) {   // :659-668
        o.distance = d;
        for (int i = 0; i < 3; i++) { o.p1[i] = p[i]; o.p2[i] = q[i]; }
        o.tri[0] = t1;
        o.tri[1] = t2;
      }
    } else if (n == qs - 1) {   // :670-676 / :1647-1654: nested search
      if (level >= max_level) {
        o.flags |= SQ_FLAG_OVERFLOW;
        return;
      }
      heap[0].b1 = n;   // suspend this queue's count
      level++;
      n = 0;
      continue;   // the nested search starts with the same test
    } else {   // :677-772 / :1662-1750
      SqEntry t1, t2;
      o.bv_tests += 2;
      sq_expand(N1, N2, mt, t1, t2);
      t1.d = sq_bvdist(N1, N2, t1);
      t2.d = sq_bvdist(N1, N2, t2);
      if (!tolerance || t1.d <= tol) heap_add(heap, n, qs, t1);
      if (!tolerance || t2.d <= tol) heap_add(heap, n, qs, t2);
    }
    // the loop's exit test (:774-785 / :1752-1758), then the suspended
    // outer searches continue at the same test
    for (;;) {
      heap = ws + (int64_t)level * slot;
      if (tolerance) {
        leave = n == 0 || heap[1].d > tol;
        if (!leave) heap_extract_min(heap, n, mt);
      } else {
        leave = n == 0;
        if (!leave) {
          heap_extract_min(heap, n, mt);
          leave = (mt.d + a.abs_err >= o.distance) && ((mt.d * (1 + a.rel_err)) >= o.distance);
        }
      }
      if (!leave) break;
      if (level == 0) return;
      level--;
      n = ws[(int64_t)level * slot].b1;
    }
  }
}

// One query of any kind; writes out[qi] (and the collide pair list).
C2A_HD void static_query(const StaticArgs& a, int64_t qi, SqEntry* ws) {
  const StaticQueryDev& q = a.q[qi];
  const MeshDev& M1 = a.meshes[q.mesh[0]];
  const MeshDev& M2 = a.meshes[q.mesh[1]];
  const GpuNode* N1 = M1.nodes;
  const GpuNode* N2 = M2.nodes;
  StaticOut o;
  o.distance = 0.0;
  for (int i = 0; i < 3; i++) { o.p1[i] = 0.0; o.p2[i] = 0.0; }
  o.bv_tests = 0;
  o.tri_tests = 0;
  o.tri[0] = q.seed[0];
  o.tri[1] = q.seed[1];
  o.found = 0;
  o.flags = 0;
  double Rrel[9], Trel[3];
  SqEntry root;
  sq_root(N1, N2, q, Rrel, Trel, root.R, root.T);
  root.b1 = 0;
  root.b2 = 0;
  root.d = 0.0;
  bool p2_to_frame2 = false;

  if (a.kind == SQ_DISTANCE_QUEUE) {
    // first bound: the models' last_tri pair (C2A_PQP.cpp:998-1005)
    o.distance = sq_tri(M1, M2, Rrel, Trel, q.seed[0], q.seed[1], o.p1, o.p2);
    sq_queue_search(a, M1, M2, Rrel, Trel, root, ws, o, false, 0.0);
    p2_to_frame2 = true;
  } else if (a.kind == SQ_TOLERANCE) {
    const double tol = q.param < 0.0 ? 0.0 : q.param;   // :1786
    o.tri[0] = o.tri[1] = -1;
    // trivial reject on the root pair's distance (:1817-1819)
    if (sq_bvdist(N1, N2, root) <= tol) {
      if (a.qsize > 2) {
        sq_queue_search(a, M1, M2, Rrel, Trel, root, ws, o, true, tol);
      } else {   // C2AToleranceRecurse (:1498-1604)
        int sp = 0;
        const int cap = (int)a.ws_entries;
        SqEntry cur;
        sq_copy(cur, root);
        for (;;) {
          const GpuNode* B1 = N1 + cur.b1;
          const GpuNode* B2 = N2 + cur.b2;
          bool next = true;
          if (sq_leaf(B1) && sq_leaf(B2)) {
            const int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
            double p[3], qq[3];
            o.tri_tests++;
            const double d = sq_tri(M1, M2, Rrel, Trel, t1, t2, p, qq);
            if (d <= tol) {
              o.found = 1;
              o.distance = d;
              for (int i = 0; i < 3; i++) { o.p1[i] = p[i]; o.p2[i] = qq[i]; }
              o.tri[0] = t1;
              o.tri[1] = t2;
              break;
            }
          } else {
            SqEntry e1, e2;
            o.bv_tests += 2;
            sq_expand(N1, N2, cur, e1, e2);
            const double d1 = sq_bvdist(N1, N2, e1), d2 = sq_bvdist(N1, N2, e2);
            // closer pair first; tolerance is fixed, so the second pair's
            // test is decided now (:1590-1603)
            SqEntry& first = (d2 < d1) ? e2 : e1;
            SqEntry& second = (d2 < d1) ? e1 : e2;
            const double df = (d2 < d1) ? d2 : d1, ds = (d2 < d1) ? d1 : d2;
            if (ds <= tol) {
              if (sp >= cap) { o.flags |= SQ_FLAG_OVERFLOW; break; }
              sq_copy(ws[sp++], second);
            }
            if (df <= tol) {
              sq_copy(cur, first);
              next = false;
            }
          }
          if (next) {
            if (sp == 0) break;
            sq_copy(cur, ws[--sp]);
          }
        }
      }
    }
    p2_to_frame2 = o.found != 0;
  } else if (a.kind == SQ_DISTANCE_CS) {
    // C2ADistance_At_ContactSpaceRecurse (:1282-1415): every call returns at
    // once after the first leaf closer than the bound, so until then the
    // bound is the distance and a child's entry test is decided at once
    const double up = q.param;
    o.distance = up;
    int sp = 0;
    const int cap = (int)a.ws_entries;
    SqEntry cur;
    sq_copy(cur, root);
    for (;;) {
      const GpuNode* B1 = N1 + cur.b1;
      const GpuNode* B2 = N2 + cur.b2;
      bool next = true;
      if (sq_leaf(B1) && sq_leaf(B2)) {
        const int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
        double p[3], qq[3];
        o.tri_tests++;
        const double d = sq_tri(M1, M2, Rrel, Trel, t1, t2, p, qq);
        if (d < o.distance) {
          o.distance = d;
          for (int i = 0; i < 3; i++) { o.p1[i] = p[i]; o.p2[i] = qq[i]; }
          o.tri[0] = t1;
          o.tri[1] = t2;
          o.found = 1;
          break;
        }
      } else {
        SqEntry e1, e2;
        o.bv_tests += 2;
        sq_expand(N1, N2, cur, e1, e2);
        const double d1 = sq_bvdist(N1, N2, e1), d2 = sq_bvdist(N1, N2, e2);
        SqEntry& first = (d2 < d1) ? e2 : e1;
        SqEntry& second = (d2 < d1) ? e1 : e2;
        const double df = (d2 < d1) ? d2 : d1, ds = (d2 < d1) ? d1 : d2;
        if (ds < up) {
          if (sp >= cap) { o.flags |= SQ_FLAG_OVERFLOW; break; }
          sq_copy(ws[sp++], second);
        }
        if (df < up) {
          sq_copy(cur, first);
          next = false;
        }
      }
      if (next) {
        if (sp == 0) break;
        sq_copy(cur, ws[--sp]);
      }
    }
    p2_to_frame2 = o.found != 0;
  } else {   // SQ_COLLIDE: CollideRecurse (:798-909)
    int sp = 0;
    const int cap = (int)a.ws_entries;
    int32_t* pl = a.pairs + 2 * (int64_t)a.pair_cap * qi;
    SqEntry cur;
    sq_copy(cur, root);
    for (;;) {
      const GpuNode* B1 = N1 + cur.b1;
      const GpuNode* B2 = N2 + cur.b2;
      bool next = true;
      o.bv_tests++;
      // BV_Overlap for RSS: RectDist <= r1 + r2 (:807)
      if (sq_rect(N1, N2, cur) <= B1->r + B2->r) {
        if (sq_leaf(B1) && sq_leaf(B2)) {
          const int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
          double tb[9], qv[9];
          o.tri_tests++;
          for (int i = 0; i < 9; i++) tb[i] = M2.tris[9 * (int64_t)t2 + i];
          mxvpv(qv + 0, Rrel, tb + 0, Trel);
          mxvpv(qv + 3, Rrel, tb + 3, Trel);
          mxvpv(qv + 6, Rrel, tb + 6, Trel);
          double ta[9];
          for (int i = 0; i < 9; i++) ta[i] = M1.tris[9 * (int64_t)t1 + i];
          if (tri_contact(ta, ta + 3, ta + 6, qv, qv + 3, qv + 6)) {
            if (o.found < a.pair_cap) {
              pl[2 * o.found] = t1;
              pl[2 * o.found + 1] = t2;
            }
            o.found++;
            if (a.flag == 2) break;   // C2A_FIRST_CONTACT (:878, :898)
          }
        } else {
          SqEntry e1, e2;
          sq_expand(N1, N2, cur, e1, e2);
          if (sp >= cap) { o.flags |= SQ_FLAG_OVERFLOW; break; }
          sq_copy(ws[sp++], e2);
          sq_copy(cur, e1);
          next = false;
        }
      }
      if (next) {
        if (sp == 0) break;
        sq_copy(cur, ws[--sp]);
      }
    }
  }
  if (p2_to_frame2) {   // p2 into model 2's frame (C2A_PQP.cpp:1043-1047)
    double u[3];
    sub3(u, o.p2, Trel);
    mtxv(o.p2, Rrel, u);
  }
  a.out[qi] = o;
}

}  // namespace c2a
