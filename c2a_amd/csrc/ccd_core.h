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
//
// Replay: the walk logs every expansion of a CA step.  At the next step the
// lanes first recompute the logged subtree at the new poses, one BVTT level
// at a time (replay), and the walk then takes children from those records;
// only pairs the previous step did not expand need speculative rounds.  The
// records hold exactly the values a round would compute for the same pairs,
// so replay changes the speed of a step and nothing else.
#pragma once

#include <stdint.h>
#include <string.h>

#include "bvh.h"
#include "c2a_math.h"
#include "ccd_types.h"

#ifndef C2A_EMU_COUNT
#define C2A_EMU_COUNT(kind, L) ((void)0)
#endif
// Index checks compiled only into the host-side test harness.
#ifndef C2A_CHECK
#define C2A_CHECK(cond) ((void)0)
#endif

namespace c2a {

#if !defined(__HIP_DEVICE_COMPILE__) && defined(C2A_EMU_SIMT)
// Host-side test harness with 64 lanes (tests/emu/ccd_emu64.cpp): each lane
// runs on its own fiber and the wavefront primitives below are barriers
// that exchange the lanes' values.
namespace simt {
void sync();
uint64_t ballot(bool p);
int bcast(int v, int src);
double bcastd(double v, int src);
int excl_prefix_max(int v);
double mind(double v);
int maxi(int v);
long long lane0(long long v);
}  // namespace simt
#endif

C2A_HD int clz32(unsigned x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __clz((int)x);
#else
  return __builtin_clz(x);
#endif
}

// Makes the LDS writes of all threads of the workgroup visible to all of
// them: drain the LDS counter, then a workgroup barrier.  LDS is coherent
// within the CU, so no cache maintenance is needed, and outstanding global
// stores are not waited for (a full __syncthreads would wait for them).
C2A_HD void lane_sync() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#elif defined(C2A_EMU_SIMT)
  simt::sync();
#endif
}

// Makes global-memory stores of all lanes visible to loads by other lanes of
// the workgroup (the replay passes exchange data through the arena).
C2A_HD void mem_sync() {
#if defined(__HIP_DEVICE_COMPILE__)
  __syncthreads();
#elif defined(C2A_EMU_SIMT)
  simt::sync();
#endif
}

// Fetch-and-add on an LDS counter shared by the lanes of the wavefront.
C2A_HD int lds_add(int32_t* p, int v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(p, v);
#else
  int o = *p;
  *p = o + v;
  return o;
#endif
}

// Values every lane holds identically (read from LDS at one address): moves
// them to scalar registers so the walk branches on scalars.
C2A_HD int32_t uni(int32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_readfirstlane(x);
#else
  return x;
#endif
}
C2A_HD double unid(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  union { double d; int32_t i[2]; } u;
  u.d = x;
  u.i[0] = __builtin_amdgcn_readfirstlane(u.i[0]);
  u.i[1] = __builtin_amdgcn_readfirstlane(u.i[1]);
  return u.d;
#else
  return x;
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

template <class T>
C2A_HD T C2A_GLOBAL* gp(T* p) {
  return (T C2A_GLOBAL*)p;
}

C2A_HD node_ptr nodes_of(const MeshDev& m) { return (node_ptr)m.nodes; }
C2A_HD gdouble_ptr tris_of(const MeshDev& m) { return (gdouble_ptr)m.tris; }

struct Entry {           // 160 B stack entry: a considered-but-undecided pair
  double R[9], T[3];     // pair transform (model-2 BV in model-1 BV frame); for
                         // replayed entries it lives in the arena instead
  double d, mt;          // RSS distance and admissible time of the pair
  int32_t b1, b2;        // node indices (replayed leaf pairs: triangle indices)
  int32_t sg;            // (generation << 8) | (slot + 1) of cached values, 0 = none
  int32_t flags;         // bit0 leaf1, bit1 leaf2, bit2 root (always visited),
                         // bit3 replayed record
  double dTri, mleaf;    // replayed leaf pairs: triangle distance, leaf time
  int32_t ar, cp;        // replayed: record number, expansion index of its children
  int32_t nn, lvl;       // record number in this step's log (-1 = unlogged), depth
};

// Words 12-19 of an Entry: everything the walk decides on, held in registers
// for the pair being processed.
struct Cur {
  double d, mt;
  int32_t b1, b2;
  int32_t sg, flags;
  double dTri, mleaf;
  int32_t ar, cp;
  int32_t nn, lvl;
};

C2A_HD Cur read_cur(const Entry& e) {
  C2A_CHECK(e.flags >= 0);
  Cur c = *reinterpret_cast<const Cur*>(&e.d);
  c.d = unid(c.d);
  c.mt = unid(c.mt);
  c.b1 = uni(c.b1);
  c.b2 = uni(c.b2);
  c.sg = uni(c.sg);
  c.flags = uni(c.flags);
  c.dTri = unid(c.dTri);
  c.mleaf = unid(c.mleaf);
  c.ar = uni(c.ar);
  c.cp = uni(c.cp);
  c.nn = uni(c.nn);
  c.lvl = uni(c.lvl);
  return c;
}

C2A_HD void write_cur(Entry& e, const Cur& c, int lane) {
  if (lane == 0) *reinterpret_cast<Cur*>(&e.d) = c;
}

// A replayed record as the walk holds it: leaf pairs carry their triangle
// indices in b1/b2.
C2A_HD Cur cur_from_hot(const HotRec& h0, int rec, int nn, int lvl) {
  Cur c;
  const int fl = uni(h0.flags);
  const bool leaf = (fl & 3) == 3;
  c.d = unid(h0.d);
  c.mt = unid(h0.mt);
  c.b1 = uni(leaf ? h0.t1 : h0.b1);
  c.b2 = uni(leaf ? h0.t2 : h0.b2);
  c.sg = 0;
  c.flags = (fl & 3) | 8;
  c.dTri = unid(h0.dTri);
  c.mleaf = unid(h0.mleaf);
  c.ar = rec;
  c.cp = uni(h0.cp);
  c.nn = nn;
  c.lvl = lvl;
  return c;
}

constexpr int WIN = 64;  // replayed records staged in LDS for the walk
constexpr int MR_MAX = 16; // misses one round can cover
constexpr int LBUF = 64; // log entries staged in LDS before they go to the arena

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
  MeshDev md[2];             // the two meshes' table entries (set at fetch)
  int32_t contact;           // 1: traversal without motion bounds (contact
                             // enumeration, distance query); 2: a distance
                             // query entering only overlapping RSS pairs
  int32_t gen;
  HotRec win[WIN];           // replayed records [L.wbase, L.wbase + WIN)
  LogEnt lbuf[LBUF];         // log entries [L.nflush, L.nlog) not yet in the arena
  int32_t anc[CAP + 2];      // scan: log index of the last event at each depth
  int32_t lastlane[CAP + 2]; // scan: scratch for the per-depth last event
  int32_t nleafq;            // replay: leaf records queued for the triangle pass
  int32_t next;              // replay: pairs queued for the speculative extension
  int64_t ticket;            // work-queue ticket shared by the wavefronts
  int32_t rarena;            // the pending round writes its subtree to the arena
  int32_t rdepth;            // levels of the slot heap the current round filled
  int32_t rpb;               // ... as pairs [rpb, rpb + 2^K - 1)
  int32_t lh[CAP + 2];       // replay: expansions per level, then fill cursors
  int32_t ls[CAP + 3];       // replay: first sorted expansion of each level
  // Misses ahead in the scan's window share a round: each gets a heap of
  // `mpage` slots (mpage = 0: one heap below X).
  int32_t mpage;             // slots per heap of the current round
  int32_t mrgen;             // generation of the round below the gathered misses
  int32_t nmr;               // misses gathered (then: covered by that round)
  int32_t mrrec[MR_MAX];     // their replayed record numbers, in scan order
  int32_t mrb[MR_MAX][2];    // and node ids
};

// A mesh-table entry held in LDS, made wavefront-uniform so that node and
// triangle addresses use scalar base registers.
template <int K, int CAP>
C2A_HD MeshDev mesh_of(const Wave<K, CAP>& g, int k) {
  MeshDev m = g.md[k];
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t pn = (uint64_t)m.nodes, pt = (uint64_t)m.tris;
  m.nodes = (const GpuNode*)(((uint64_t)(uint32_t)uni((int32_t)(pn >> 32)) << 32) |
                             (uint32_t)uni((int32_t)(uint32_t)pn));
  m.tris = (const double*)(((uint64_t)(uint32_t)uni((int32_t)(pt >> 32)) << 32) |
                           (uint32_t)uni((int32_t)(uint32_t)pt));
  m.ntri = uni(m.ntri);
  m.depth = uni(m.depth);
#endif
  return m;
}

enum : int {
  ST_FETCH = 0, ST_STEP = 1, ST_WALK = 2, ST_CONTACT_WALK = 3, ST_DONE = 4,
  ST_SCAN = 5,   // lane-parallel walk over the replayed subtree
  ST_SUB = 6,    // sequential walk below a pair the scan found unexpanded
  ST_LAYOUT = 7, // replayed records scored: lay them out and scan
  ST_TRANS = 8   // translation-only query (C2A_Solve, both rotations < 1e-8)
};

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
  int nlog, nlog_prev, logbuf;   // expansions logged this step / by the previous step
  int nflush;                    // log entries already copied to the arena
  int pq_pending;                // p1, p2 of the last improving leaf still to compute
  int wbase;                     // first record staged in the LDS window
  int spos, nev, logfull;        // scan position, events, log capacity reached
  int napp;                      // next free record pair of the arena
  int need_score;                // replayed records wait for the scorer
  int kdepth, kmiss, klog;       // this query's speculative depth, depth below
                                 // a miss and replay log capacity
  int kext;                      // levels of speculative extension in the replay
  int kx;                        // ... this step: kext, or 0 after a step with few misses
  int nmiss;                     // misses the scan met in this step
  unsigned char* ab;             // this wavefront's replay arena
};

C2A_HD int node_leaf(node_ptr n) { return n->first_child < 0; }

// Lane-parallel copy of the 14 value words + node ids of a slot into a stack
// entry; the remaining words (sg/flags, replay fields) are set per lane.
C2A_HD void slot_to_entry(Entry& dst, const Slot& src, int sg, int flags, int nn, int lvl,
                          int lane, int nl) {
  const uint64_t* s = reinterpret_cast<const uint64_t*>(&src);
  uint64_t* d = reinterpret_cast<uint64_t*>(&dst);
  for (int w = lane; w < 20; w += nl) {
    if (w < 15) d[w] = s[w];
    else if (w == 15) { dst.sg = sg; dst.flags = flags; }
    else if (w == 16) dst.dTri = 0.0;
    else if (w == 17) dst.mleaf = 0.0;
    else if (w == 18) { dst.ar = -1; dst.cp = -1; }
    else { dst.nn = nn; dst.lvl = lvl; }
  }
}

// Leaf-pair values: C2A.cpp:1145-1166 up to the distance-dependent update.
// Leaf-pair values from the two triangles (9 doubles each, already loaded).
C2A_HD void leaf_values_tris(Slot& s, node_ptr A1, node_ptr A2, const double* ta,
                             const double* tb, int t1, int t2, const double* Rrel,
                             const double* Trel, const double* r1, const double* tt1,
                             const Motion& mo1, const Motion& mo2, bool contact) {
  double p[3], q[3];
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

C2A_HD void leaf_values(Slot& s, const MeshDev& M1, const MeshDev& M2, const double* Rrel,
                        const double* Trel, const double* r1, const double* tt1,
                        const Motion& mo1, const Motion& mo2, int b1, int b2, bool contact) {
  node_ptr A1 = nodes_of(M1) + b1;
  node_ptr A2 = nodes_of(M2) + b2;
  int t1 = -A1->first_child - 1, t2 = -A2->first_child - 1;
  double ta[9], tb[9];
  gdouble_ptr g1 = tris_of(M1) + 9 * (int64_t)t1;
  gdouble_ptr g2 = tris_of(M2) + 9 * (int64_t)t2;
  for (int i = 0; i < 9; i++) { ta[i] = g1[i]; tb[i] = g2[i]; }
  leaf_values_tris(s, A1, A2, ta, tb, t1, t2, Rrel, Trel, r1, tt1, mo1, mo2, contact);
}

// One level of descent (C2A.cpp:1195-1226): replaces (b1, b2) and (R, T) by
// child pair `bit` (0 = the first child node, 1 = its sibling).  Returns
// false, changing nothing, for a leaf pair.
C2A_HD bool descend(node_ptr N1, node_ptr N2, int& b1, int& b2, int bit, double* R, double* T) {
  C2A_CHECK(b1 >= 0 && b2 >= 0);
  node_ptr n1 = N1 + b1;
  node_ptr n2 = N2 + b2;
  const int fc1 = n1->first_child, fc2 = n2->first_child;
  const bool l1 = fc1 < 0, l2 = fc2 < 0;
  if (l1 && l2) return false;
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
  return true;
}

// First-child link and descend-rule size of a node (the fields of its
// second line that a descent reads), carried from the level that fetched the
// node, so that a speculative chain pays one dependent fetch per level: the
// child's transform line and its head are loaded together.
struct NodeHead {
  int32_t fc;
  double size;
};

C2A_HD NodeHead head_of(node_ptr n) {
  NodeHead h;
  h.fc = n->first_child;
  h.size = n->size;
  return h;
}

// descend() with the heads of (b1, b2) given in h1/h2 and updated to the
// heads of the child pair.
C2A_HD bool descend_h(node_ptr N1, node_ptr N2, int& b1, int& b2, NodeHead& h1, NodeHead& h2,
                      int bit, double* R, double* T) {
  C2A_CHECK(b1 >= 0 && b2 >= 0);
  const bool l1 = h1.fc < 0, l2 = h2.fc < 0;
  if (l1 && l2) return false;
  double Rn[9], Tn[3], Tt[3], CR[9], CT[3];
  if (l2 || (!l1 && (h1.size > h2.size))) {
    const int c = h1.fc + bit;
    node_ptr cn = N1 + c;
    for (int i = 0; i < 9; i++) CR[i] = cn->R[i];
    for (int i = 0; i < 3; i++) CT[i] = cn->Tr[i];
    h1 = head_of(cn);
    mtxm(Rn, CR, R);
    sub3(Tt, T, CT);
    mtxv(Tn, CR, Tt);
    b1 = c;
  } else {
    const int c = h2.fc + bit;
    node_ptr cn = N2 + c;
    for (int i = 0; i < 9; i++) CR[i] = cn->R[i];
    for (int i = 0; i < 3; i++) CT[i] = cn->Tr[i];
    h2 = head_of(cn);
    mxm(Rn, R, CR);
    mxvpv(Tn, R, CT, T);
    b2 = c;
  }
  for (int i = 0; i < 9; i++) R[i] = Rn[i];
  for (int i = 0; i < 3; i++) T[i] = Tn[i];
  return true;
}

// RSS distance and admissible time of pair (b1, b2) at transform (R, T)
// (C2A.cpp:1237-1276); returns the leaf flags of the two nodes.
C2A_HD int pair_values(node_ptr N1, node_ptr N2, int b1, int b2, const double* R,
                       const double* T, const double* r1, const Motion& mo1, const Motion& mo2,
                       bool contact, double& d_out, double& mt_out, double* mb_out = nullptr) {
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
      mxv(Sw, r1, tmp);
      neg3(Sn, Sw);
      double mb1 = motion_bound(mo1, A1->angR, Sw);
      double mb2 = motion_bound(mo2, A2->angR, Sn);
      mt = d / (mb1 + mb2);
      if (mt <= 0) mt = 0.0;
      if (mb_out) {
        mb_out[0] = mb1;
        mb_out[1] = mb2;
      }
    }
  }
  d_out = d;
  mt_out = mt;
  return node_leaf(A1) | (node_leaf(A2) << 1);
}

// Scorer features of pair (b1, b2) at transform (R, T), as BF16.  mb: the
// two motion bounds pair_values used (zero when the pair touches).
C2A_HD void pair_feature_row(uint16_t* out, node_ptr N1, node_ptr N2, int b1, int b2,
                             const double* R, const double* T, double d, double mt,
                             const double* mb, const Motion& mo1, const Motion& mo2, double up,
                             int fl) {
  node_ptr A1 = N1 + b1;
  node_ptr A2 = N2 + b2;
  const double l1[2] = {A1->l[0], A1->l[1]}, l2[2] = {A2->l[0], A2->l[1]};
  float f[NF];
  pair_features(f, l1, l2, A1->r, A2->r, R, T, d, mt, mb[0], mb[1], A1->depth, A2->depth,
                A1->angR, A2->angR, mo1, mo2, up, fl & 3, 1.0 / N1->angR);
  for (int k = 0; k < NF; k++) out[k] = f32_to_bf16(f[k]);
}

// Translation-only conservative advancement (C2A_Solve when both rotations
// are below 1e-8): InterpMotion.cpp:289-298 (ConservD, security ratio 0.1),
// 570-655 (CAonRSS), 657-743 (CAonNonAdjacentTriangles).  Only object 1's
// velocity enters; `delta` is C2A_Solve's m_toc_delta (its tolerance_d).
C2A_HD double conserv_d(double d, double delta) { return d - 0.1 * delta; }

// CAonRSS for node pair (A1, A2) at transform (R, T).  Rectangles that
// overlap at the start leave the reference's direction unset; such a pair is
// touching now (time 0, distance 0).
C2A_HD bool ca_on_rss(const Motion& mo1, const double* r1, const double* R, const double* T,
                      node_ptr A1, node_ptr A2, double delta, double& mint, double& distance) {
  double S[3] = {0.0, 0.0, 0.0}, tmp[3], Sw[3], Vel[3], T2[3], RL[9];
  for (int i = 0; i < 9; i++) RL[i] = A1->Rloc[i];
  const double la[2] = {A1->l[0], A1->l[1]}, lb[2] = {A2->l[0], A2->l[1]};
  const double rr = A1->r + A2->r;
  const double angR = A1->angR;
  double d = rect_dist(R, T, la, lb, S);
  d -= rr;
  d = (d < 0.0) ? 0.0 : d;
  mtxv(tmp, r1, mo1.cv);
  mtxv(Vel, RL, tmp);
  if (d == 0.0) {
    mint = 0.0;
    distance = 0.0;
    return true;
  }
  mxv(tmp, RL, S);
  mxv(Sw, r1, tmp);
  double tocf = conserv_d(d, delta) / motion_bound_leaf(mo1, angR, Sw);
  double total = tocf;
  int it = 1;
  while ((d >= delta) && (total <= mint) && (it < 50)) {
    for (int i = 0; i < 3; i++) T2[i] = T[i] + Vel[i] * (-total);
    d = rect_dist(R, T2, la, lb, S);
    d -= rr;
    d = (d < 0.0) ? 0.0 : d;
    if (d == 0) break;
    mxv(tmp, RL, S);
    mxv(Sw, r1, tmp);
    tocf = conserv_d(d, delta) / motion_bound_leaf(mo1, angR, Sw);
    it++;
    if (tocf < delta) break;
    total += tocf;
  }
  if (total < 1.0) {
    mint = total;
    const double dd = (total - delta) * len3(mo1.cv);
    distance = dd < 0 ? 0 : dd;
    return true;
  }
  return false;
}

// CAonNonAdjacentTriangles: triangle A advancing with object 1's velocity
// against triangle B, both in model 1's frame (9 doubles each).
C2A_HD bool ca_on_tris(const Motion& mo1, const double* r1, const double* A, const double* B,
                       double delta, double& mint, double& distance) {
  double Vel[3], p[3], q[3], n[3], At[9];
  mtxv(Vel, r1, mo1.cv);
  double d = tri_dist(p, q, A, B, nullptr);
  sub3(n, q, p);
  normalize3(n);
  double u = dot3(Vel, n);
  if (u <= 0) u = 1e-30;
  if (d == 0) {
    mint = 0.0;
    distance = 0.0;
    return true;
  }
  const double dt = conserv_d(d, delta) / u;
  int it = 1;
  if (dt >= mint) return false;
  double total = dt;
  while ((d > delta) && (total <= mint) && (it < 50)) {
    for (int v = 0; v < 3; v++)
      for (int i = 0; i < 3; i++) At[3 * v + i] = A[3 * v + i] + Vel[i] * total;
    d = tri_dist(p, q, At, B, nullptr);
    if (d == 0.0) break;
    sub3(n, q, p);
    normalize3(n);
    u = dot3(Vel, n);
    if (u <= 0) u = 1e-30;
    const double tofc = conserv_d(d, delta) / u;
    it++;
    if (tofc < delta) break;
    total += tofc;
  }
  if (total <= mint && total >= 0.0) {
    mint = total;
    distance = total * len3(mo1.cv) + delta;
    return true;
  }
  return false;
}

// Heap slot h of the depth-K subtree below X: the pair reached by the path
// bits of h+1, its transform chained from X, and its values.
template <int K, int CAP, class RT>
C2A_HD void compute_slot_below(Wave<K, CAP>& g, const RT* xR, const RT* xT, int b1, int b2, int h,
                               const MeshDev& M1, const MeshDev& M2, bool contact, Slot& s) {
  node_ptr N1 = nodes_of(M1);
  node_ptr N2 = nodes_of(M2);
  const int lvl = 31 - clz32((unsigned)(h + 1));
  double R[9], T[3];
  for (int i = 0; i < 9; i++) R[i] = xR[i];
  for (int i = 0; i < 3; i++) T[i] = xT[i];
  NodeHead h1 = head_of(N1 + b1), h2 = head_of(N2 + b2);
  for (int lv = 0; lv < lvl; lv++) {
    const int bit = ((h + 1) >> (lvl - 1 - lv)) & 1;
    if (!descend_h(N1, N2, b1, b2, h1, h2, bit, R, T)) {
      s.valid = 0;
      return;
    }
  }
  double d, mt;
  C2A_CHECK(b1 >= 0 && b2 >= 0 && b1 < 2 * M1.ntri - 1 && b2 < 2 * M2.ntri - 1);
  // a leaf pair's triangles are fetched before its RSS test, so the fetch
  // overlaps that arithmetic
  const bool leafp = h1.fc < 0 && h2.fc < 0;
  const int t1 = -h1.fc - 1, t2 = -h2.fc - 1;
  double ta[9], tb[9];
  if (leafp) {
    gdouble_ptr g1 = tris_of(M1) + 9 * (int64_t)t1;
    gdouble_ptr g2 = tris_of(M2) + 9 * (int64_t)t2;
    for (int i = 0; i < 9; i++) { ta[i] = g1[i]; tb[i] = g2[i]; }
  }
  const int fl = pair_values(N1, N2, b1, b2, R, T, g.r1, g.mo[0], g.mo[1], contact, d, mt);
  for (int i = 0; i < 9; i++) s.R[i] = R[i];
  for (int i = 0; i < 3; i++) s.T[i] = T[i];
  s.d = d;
  s.mt = mt;
  s.b1 = b1;
  s.b2 = b2;
  s.flags = fl;
  s.valid = 1;
  if (leafp)
    leaf_values_tris(s, N1 + b1, N2 + b2, ta, tb, t1, t2, g.Rrel, g.Trel, g.r1, g.tt1, g.mo[0],
                     g.mo[1], contact);
}

template <int K, int CAP>
C2A_HD void compute_slot_into(Wave<K, CAP>& g, int h, const MeshDev& M1, const MeshDev& M2,
                              bool contact, Slot& s) {
  compute_slot_below(g, g.X.R, g.X.T, g.X.b1, g.X.b2, h, M1, M2, contact, s);
}

template <int K, int CAP>
C2A_HD void compute_slot(Wave<K, CAP>& g, int h, const MeshDev& M1, const MeshDev& M2,
                         bool contact) {
  compute_slot_into(g, h, M1, M2, contact, g.slot[h]);
}

// Heap slot h of the round written as arena record 2 * rpb + h, so that the
// pair stays available for the rest of the step; slots above the last level
// link their children (pair rpb + h).
template <int K, int CAP>
C2A_HD void compute_slot_arena(Wave<K, CAP>& g, int h, const MeshDev& M1, const MeshDev& M2,
                               const ArenaView& av) {
  Slot s;
  compute_slot_into(g, h, M1, M2, false, s);
  if (!s.valid) return;   // below a leaf pair: never referenced
  const int rpb = g.rpb;
  const int rec = 2 * rpb + h;
  C2A_CHECK(rec > 0);
  const int lvl = 31 - clz32((unsigned)(h + 1));
  double C2A_GLOBAL* xf = gp(av.xf) + 12 * (int64_t)rec;
  for (int i = 0; i < 9; i++) xf[i] = s.R[i];
  for (int i = 0; i < 3; i++) xf[9 + i] = s.T[i];
  HotRec hr;
  const bool leaf = (s.flags & 3) == 3;
  hr.d = s.d;
  hr.mt = s.mt;
  hr.dTri = leaf ? s.dTri : 0.0;
  hr.mleaf = leaf ? s.mleaf : 0.0;
  hr.b1 = s.b1;
  hr.b2 = s.b2;
  hr.flags = s.flags & 3;
  hr.cp = (!leaf && lvl < K) ? rpb + h : -1;
  hr.t1 = leaf ? s.t1 : -1;
  hr.t2 = leaf ? s.t2 : -1;
  hr.score = 0.0f;
  hr.pad = 0;
  gp(av.hot)[rec] = hr;
}

// One lane's share of a round.
template <int K, int CAP>
C2A_HD void round_lane(Wave<K, CAP>& g, const Leader& L, int lane, int nl, const KArgs& a) {
  const MeshDev M1 = mesh_of(g, 0);
  const MeshDev M2 = mesh_of(g, 1);
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
  if (g.rarena) {
    const ArenaView av = arena_view(L.ab, a.logcap);
    for (int h = lane + 1; h < Shape<K>::NSLOT; h += nl) compute_slot_arena(g, h, M1, M2, av);
    return;
  }
  const int mp = g.mpage;
  if (mp) {
    // heap r (slots [r * mp, r * mp + mp - 1)) below gathered miss r, whose
    // transform is its replayed record's
    const ArenaView av = arena_view(L.ab, a.logcap);
    const int nr = g.nmr;
    for (int h = lane; h < nr * mp; h += nl) {
      const int r = h / mp, l = h - r * mp;
      if (l == 0 || l == mp - 1) continue;   // the miss itself; padding
      const double C2A_GLOBAL* xf = gp(av.xf) + 12 * (int64_t)g.mrrec[r];
      compute_slot_below(g, xf, xf + 9, g.mrb[r][0], g.mrb[r][1], l, M1, M2, contact, g.slot[h]);
    }
    return;
  }
  const int nslot = (2 << g.rdepth) - 1;   // levels 0..rdepth of the heap
  for (int h = lane + 1; h < nslot; h += nl) compute_slot(g, h, M1, M2, contact);
}

template <int K, int CAP>
C2A_HD void request_round(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane, int nl) {
  Entry& e = g.stack[L.sp - 1];
  int gen = g.gen + 1;
  if (gen >= 0x7fffff) {   // generation wrap: forget every cached link
    for (int i = lane; i < L.sp; i += nl) g.stack[i].sg = 0;
    gen = 1;
    if (lane == 0) g.mrgen = -1;
  }
  const int flags = e.flags;
  const int ar = e.ar;
  const int sg = (gen << 8) | 1;
  lane_sync();
  const uint64_t* s = reinterpret_cast<const uint64_t*>(&e);
  uint64_t* x = reinterpret_cast<uint64_t*>(&g.X);
  if (flags & 8) {
    C2A_CHECK(ar >= 0 && ar < arena_records(a.logcap));
    // replayed pair: the transform is in the arena
    const uint64_t C2A_GLOBAL* xf = gp(reinterpret_cast<const uint64_t*>(
        arena_view(L.ab, a.logcap).xf + 12 * (int64_t)ar));
    for (int w = lane; w < 16; w += nl) x[w] = w < 12 ? xf[w] : s[w];
  } else {
    for (int w = lane; w < 16; w += nl) x[w] = s[w];
  }
  lane_sync();
  // internal pairs of a CA step go to the arena while it has room
  constexpr int NPAIR = (1 << K) - 1;
  const bool to_arena = a.arena_rounds && !a.scorer.on && a.logcap > 0 && !g.contact &&
                        (flags & 3) != 3 && L.napp + NPAIR <= 2 * a.logcap;
  // a miss the scan found with more misses after it in its window: one
  // shallow heap below each (the same slots a round below each would fill)
  const int mpage = 2 << L.kmiss;
  const bool multi = !to_arena && (flags & 8) && (flags & 3) != 3 && L.kmiss > 0 &&
                     L.kmiss < L.kdepth && 2 * mpage <= Shape<K>::NSLOT + 1 && g.nmr > 1 &&
                     ar == g.mrrec[0];
  lane_sync();
  if (lane == 0) {
    if (multi) {
      g.rarena = 0;
      e.sg = sg;               // heap 0, slot 0: the miss being walked
      g.gen = gen;
      g.mrgen = gen;
      g.mpage = mpage;
      g.rdepth = L.kmiss;
      const int cap = (Shape<K>::NSLOT + 1) / mpage;
      if (g.nmr > cap) g.nmr = cap;
    } else if (to_arena) {
      g.rarena = 1;
      g.rpb = L.napp;
      e.cp = L.napp;
      e.flags = flags | 8;
    } else {
      g.rarena = 0;
      g.mpage = 0;
      e.sg = sg;
      Slot& s0 = g.slot[0];
      s0.b1 = g.X.b1;
      s0.b2 = g.X.b2;
      s0.d = g.X.d;
      s0.mt = g.X.mt;
      s0.flags = flags & 3;
      s0.valid = 1;
      g.gen = gen;
      // a replayed pair that must now be expanded (a miss) usually needs
      // only a level or two: a shallow first round has a shorter chain of
      // dependent node loads
      g.rdepth = ((flags & 8) && L.kmiss > 0 && L.kmiss < L.kdepth) ? L.kmiss : L.kdepth;
    }
  }
  if (to_arena) {
    L.napp += NPAIR;
    L.wbase = -(1 << 30);   // the window may hold stale copies of these records
  }
  lane_sync();
  L.need_round = 1;
}

// Root pair of a traversal at absolute poses (C2A.cpp:1793-1806).  `nn` is
// the root's record number in this step's log (0), or -1 when the traversal
// is not logged.
template <int K, int CAP>
C2A_HD void start_traversal(Wave<K, CAP>& g, Leader& L, const MeshDev& M1, const MeshDev& M2,
                            const double* R1, const double* T1, const double* R2,
                            const double* T2, int nn, int lane) {
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
    g.mpage = 0;
    g.mrgen = -1;
    g.nmr = 0;
    Entry& e = g.stack[0];
    for (int i = 0; i < 9; i++) e.R[i] = R[i];
    for (int i = 0; i < 3; i++) e.T[i] = T[i];
    e.d = 0.0;
    e.mt = 0.0;
    e.b1 = 0;
    e.b2 = 0;
    e.sg = 0;
    e.flags = node_leaf(r1n) | (node_leaf(r2n) << 1) | 4;
    e.dTri = 0.0;
    e.mleaf = 0.0;
    e.ar = -1;
    e.cp = -1;
    e.nn = nn;
    e.lvl = 0;
  }
  lane_sync();
  L.sp = 1;
}

// RSS distance of the two root volumes at absolute poses: start_traversal's
// root transform and pair_values' distance, evaluated by the host when it
// sets up a C2A_Collide query (its root overlap test, C2A_PQP.cpp:1068).
inline double root_rss_distance(const GpuNode* r1n, const GpuNode* r2n, const double* R1,
                                const double* T1, const double* R2, const double* T2) {
  double Tt[3], Rt[9], R[9], T[3], Rrel[9], Trel[3], C1R[9], C2R[9], C1T[3], C2T[3];
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
  Motion none;
  memset(&none, 0, sizeof(none));
  double d, mt;
  (void)pair_values((node_ptr)r1n, (node_ptr)r2n, 0, 0, R, T, R1, none, none, true, d, mt);
  return d;
}

// Wavefront primitives of the scan (the host harness runs one lane).
C2A_HD uint64_t wave_ballot(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint64_t)__ballot(p);
#elif defined(C2A_EMU_SIMT)
  return simt::ballot(p);
#else
  return p ? 1u : 0u;
#endif
}
C2A_HD int lanes_below(uint64_t m, int lane) {
#if defined(__HIP_DEVICE_COMPILE__)
  (void)lane;
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
#else
  return __builtin_popcountll(m & ((1ull << lane) - 1));
#endif
}
C2A_HD int lane_bcast(int v, int src) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __shfl(v, src);
#elif defined(C2A_EMU_SIMT)
  return simt::bcast(v, src);
#else
  (void)src;
  return v;
#endif
}
C2A_HD double lane_bcastd(double v, int src) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __shfl(v, src);
#elif defined(C2A_EMU_SIMT)
  return simt::bcastd(v, src);
#else
  (void)src;
  return v;
#endif
}
// max over lanes below this one (0 for lane 0)
C2A_HD int excl_prefix_max(int v, int lane) {
#if defined(__HIP_DEVICE_COMPILE__)
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, (unsigned)o);
    if (lane >= o) v = v > u ? v : u;
  }
  const int x = __shfl_up(v, 1u);
  return lane == 0 ? 0 : x;
#elif defined(C2A_EMU_SIMT)
  (void)lane;
  return simt::excl_prefix_max(v);
#else
  (void)v;
  (void)lane;
  return 0;
#endif
}
C2A_HD double wave_mind(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  for (int o = 32; o >= 1; o >>= 1) {
    const double u = __shfl_xor(v, o);
    v = u < v ? u : v;
  }
#elif defined(C2A_EMU_SIMT)
  v = simt::mind(v);
#endif
  return v;
}
C2A_HD int wave_maxi(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
  for (int o = 32; o >= 1; o >>= 1) {
    const int u = __shfl_xor(v, o);
    v = u > v ? u : v;
  }
#elif defined(C2A_EMU_SIMT)
  v = simt::maxi(v);
#endif
  return v;
}
C2A_HD float unif(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
#else
  return x;
#endif
}
C2A_HD unsigned long long global_add_u64(unsigned long long* p, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(p, v);
#else
  const unsigned long long o = *p;
  *p = o + v;
  return o;
#endif
}
C2A_HD long long lane0_bcast64(long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
  const unsigned long long u = (unsigned long long)v;
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
#elif defined(C2A_EMU_SIMT)
  return simt::lane0(v);
#else
  return v;
#endif
}
C2A_HD int lds_max(int32_t* p, int v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicMax(p, v);
#else
  int o = *p;
  if (v > o) *p = v;
  return o;
#endif
}

// Training samples of one replayed step: every record below the root whose
// replayed subtree holds a leaf pair, selected by hashing its key, goes out
// as its features and log(closest replayed triangle distance + 1e-4).
// Wavefront-aggregated: one atomic per 64 records.
template <class LeaderT>
C2A_HD void write_samples(const LeaderT& L, const KArgs& a, const ArenaView& av, int nrec,
                          double inv_s, int lane, int nl) {
  const double C2A_GLOBAL* ml = gp(av.minleaf);
  const uint16_t C2A_GLOBAL* fe = gp(av.feat);
  for (int base = 1; base < nrec; base += nl) {
    const int r = base + lane;
    const int64_t key = ((int64_t)L.qi << 34) | ((int64_t)(L.numCA & 0xff) << 26) | (r & 0x3ffffff);
    const bool take = r < nrec && ml[r] < INFINITY && (mix64((uint64_t)key) >> 32) < a.sample_thr;
    const uint64_t m = wave_ballot(take);
    if (!m) continue;
    long long start = 0;
    if ((lane & 63) == 0)   // one claim per wavefront
      start = (long long)global_add_u64(a.sample_count, (unsigned long long)__builtin_popcountll(m));
    start = lane0_bcast64(start);
    const long long slot = start + lanes_below(m, lane);
    if (take && slot < a.sample_cap) {
      float C2A_GLOBAL* o = gp(a.samples) + slot * (NF + 1);
      for (int k = 0; k < NF; k++) o[k] = bf16_to_f32(fe[(int64_t)r * NF + k]);
      o[NF] = slog(ml[r] * inv_s);
      gp(a.sample_keys)[slot] = key;
    }
  }
}

// One child record of expansion e (side 0: first child node, 1: its
// sibling) at this step's pose: transform from the parent record, pair
// values, features, and a queue entry for the triangle pass if it is a leaf
// pair.  `fresh`: the record lies beyond the previous step's records (its
// child link starts empty).
template <int K, int CAP>
C2A_HD void replay_record(Wave<K, CAP>& g, const Leader& L, const KArgs& a, const ArenaView& av,
                          node_ptr N1, node_ptr N2, const LogEnt le, int e, int side, bool fresh) {
  double C2A_GLOBAL* xf = gp(av.xf);
  int32_t C2A_GLOBAL* cpa = gp(av.cpa);
  HotRec C2A_GLOBAL* hot = gp(av.hot);
  int32_t C2A_GLOBAL* leafq = gp(av.pos);
  const int rec = 2 * e + 1 + side;
  double R[9], T[3];
  const double C2A_GLOBAL* pxf = xf + 12 * (int64_t)le.rec;
  for (int i = 0; i < 9; i++) R[i] = pxf[i];
  for (int i = 0; i < 3; i++) T[i] = pxf[9 + i];
  int b1 = le.b1, b2 = le.b2;
  NodeHead h1 = head_of(N1 + b1), h2 = head_of(N2 + b2);
  descend_h(N1, N2, b1, b2, h1, h2, side, R, T);   // an expanded pair is never a leaf pair
  double d, mt, mb[2] = {0.0, 0.0};
  const int fl = pair_values(N1, N2, b1, b2, R, T, g.r1, g.mo[0], g.mo[1], false, d, mt, mb);
  if (a.feats) {
    uint16_t f[NF];
    pair_feature_row(f, N1, N2, b1, b2, R, T, d, mt, mb, g.mo[0], g.mo[1], L.Up, fl);
    uint16_t C2A_GLOBAL* of = gp(av.feat) + (int64_t)rec * NF;
    for (int k = 0; k < NF; k++) of[k] = f[k];
  }
  double C2A_GLOBAL* oxf = xf + 12 * (int64_t)rec;
  for (int i = 0; i < 9; i++) oxf[i] = R[i];
  for (int i = 0; i < 3; i++) oxf[9 + i] = T[i];
  HotRec h;
  h.d = d;
  h.mt = mt;
  h.b1 = b1;
  h.b2 = b2;
  h.flags = fl;
  if (fresh) cpa[rec] = -1;
  h.cp = fresh ? -1 : cpa[rec];
  h.dTri = 0.0;
  h.mleaf = 0.0;
  h.t1 = -1;
  h.t2 = -1;
  h.score = 0.0f;
  h.pad = 0;
  hot[rec] = h;
  if (fl == 3) leafq[lds_add(&g.nleafq, 1)] = rec;
}

// Recomputes, at this step's poses, every pair the previous step's walk
// computed as a child of an expansion (C2A.cpp:1195-1276 per pair), level by
// level so that each record's parent transform is ready before it is needed.
// Lanes take one record each.
template <int K, int CAP>
C2A_HD void replay(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane, int nl) {
  const ArenaView av = arena_view(L.ab, a.logcap, a.feats != 0);
  const LogEnt C2A_GLOBAL* lg = gp(L.logbuf ? av.log0 : av.log1);
  double C2A_GLOBAL* xf = gp(av.xf);
  int32_t C2A_GLOBAL* cpa = gp(av.cpa);
  int32_t C2A_GLOBAL* order = gp(av.order);
  HotRec C2A_GLOBAL* hot = gp(av.hot);
  const int n = L.nlog_prev;
  const int nrec = 1 + 2 * n;
  const MeshDev M1 = mesh_of(g, 0);
  const MeshDev M2 = mesh_of(g, 1);
  node_ptr N1 = nodes_of(M1);
  node_ptr N2 = nodes_of(M2);
  mem_sync();   // the previous walk's log stores
  // level histogram, children links
  int32_t C2A_GLOBAL* leafq = gp(av.pos);   // free until the layout pass
  for (int i = lane; i < CAP + 2; i += nl) g.lh[i] = 0;
  if (lane == 0) g.nleafq = 0;
  for (int i = lane; i < nrec; i += nl) cpa[i] = -1;
  mem_sync();
  for (int e = lane; e < n; e += nl) lds_add(&g.lh[lg[e].level], 1);
  for (int e = lane; e < n; e += nl) cpa[lg[e].rec] = e;
  lane_sync();
  if (lane == 0) {
    int acc = 0, top = 0;
    for (int v = 0; v < CAP + 2; v++) {
      g.ls[v] = acc;
      const int c = g.lh[v];
      if (c) top = v + 1;
      g.lh[v] = acc;
      acc += c;
    }
    g.ls[CAP + 2] = top;   // number of levels in use
    // root record: the transform start_traversal placed in the root entry
    for (int i = 0; i < 9; i++) xf[i] = g.stack[0].R[i];
    for (int i = 0; i < 3; i++) xf[9 + i] = g.stack[0].T[i];
  }
  lane_sync();
  for (int e = lane; e < n; e += nl) {
    const int pos = lds_add(&g.lh[lg[e].level], 1);
    order[pos] = e;
  }
  mem_sync();
  const int top = g.ls[CAP + 2];
  for (int v = 0; v < top; v++) {
    const int s0 = g.ls[v];
    const int cnt = (v + 1 < CAP + 2 ? g.ls[v + 1] : n) - s0;
    for (int j = lane; j < 2 * cnt; j += nl) {
      const int e = order[s0 + (j >> 1)];
      replay_record(g, L, a, av, N1, N2, lg[e], e, j & 1, false);
    }
    mem_sync();
  }
  // Speculative extension: a replayed pair the previous step did not expand
  // but whose values now pass the descend test at this step's initial bound
  // may be expanded by this step (the bound only tightens, so pairs that fail
  // it now will be pruned).  Their children -- and, level by level, the
  // children of those that pass too -- are computed here in parallel as
  // pseudo-expansions appended to the log, so that the scan finds them
  // replayed instead of walking below them one round at a time.  Values never
  // depend on the bound, so this changes no result.
  int ntot = n;
  if (L.kx > 0) {
    LogEnt C2A_GLOBAL* lgw = const_cast<LogEnt C2A_GLOBAL*>(lg);
    int lo = 1, hi = nrec;
    for (int pass = 0; pass < L.kx; pass++) {
      if (lane == 0) g.next = 0;
      lane_sync();
      const int room = L.klog - ntot;
      for (int r = lo + lane; r < hi; r += nl) {
        const HotRec h = hot[r];
        if ((h.flags & 3) == 3 || cpa[r] >= 0) continue;
        const bool desc =
            h.mt < L.Up && ((h.d < (L.D - L.abs_err)) || (h.d * (1 + L.rel_err) < L.D));
        if (!desc) continue;
        const int k = lds_add(&g.next, 1);
        if (k >= room) continue;   // no room: the scan walks below it
        const int e = ntot + k;
        LogEnt v;
        v.rec = r;
        v.b1 = h.b1;
        v.b2 = h.b2;
        v.level = lgw[(r - 1) >> 1].level + 1;   // depth of record r
        lgw[e] = v;
        cpa[r] = e;
        hot[r].cp = e;
      }
      mem_sync();
      const int cnt = uni(g.next) < room ? uni(g.next) : room;
      if (cnt <= 0) break;
      for (int j = lane; j < 2 * cnt; j += nl) {
        const int e = ntot + (j >> 1);
        replay_record(g, L, a, av, N1, N2, lgw[e], e, j & 1, true);
      }
      mem_sync();
      lo = 2 * ntot + 1;
      hi = 2 * (ntot + cnt) + 1;
      ntot += cnt;
    }
    if (ntot > n) {
      // level order over the logged and the pseudo-expansions for the layout
      for (int i = lane; i < CAP + 2; i += nl) g.lh[i] = 0;
      lane_sync();
      for (int e = lane; e < ntot; e += nl) lds_add(&g.lh[lgw[e].level], 1);
      lane_sync();
      if (lane == 0) {
        int acc = 0, tp = 0;
        for (int v = 0; v < CAP + 2; v++) {
          g.ls[v] = acc;
          const int c = g.lh[v];
          if (c) tp = v + 1;
          g.lh[v] = acc;
          acc += c;
        }
        g.ls[CAP + 2] = tp;
      }
      lane_sync();
      for (int e = lane; e < ntot; e += nl) order[lds_add(&g.lh[lgw[e].level], 1)] = e;
      mem_sync();
      L.nlog_prev = ntot;   // the layout and the scan cover the extension too
      L.napp = ntot;        // arena rounds append after the extension
    }
  }
  // triangle distances of the leaf pairs, in dense chunks (C2A.cpp:1141-1166)
  const int nleaf = g.nleafq;
  for (int k = lane; k < nleaf; k += nl) {
    const int rec = leafq[k];
    HotRec C2A_GLOBAL* hr = hot + rec;
    Slot tmp;
    tmp.flags = 3;
    leaf_values(tmp, M1, M2, g.Rrel, g.Trel, g.r1, g.tt1, g.mo[0], g.mo[1], hr->b1, hr->b2, false);
    hr->dTri = tmp.dTri;
    hr->mleaf = tmp.mleaf;
    hr->t1 = tmp.t1;
    hr->t2 = tmp.t2;
  }
  mem_sync();
}

// Scan layout of the replayed records: replayed subtree sizes bottom-up,
// then depth-first positions top-down in this step's visiting order (closer
// child first, C2A.cpp:1281, or the lower score under the learned order),
// then the records scattered into that order.  Training samples of the
// replayed pairs are written on the way up.
template <int K, int CAP>
C2A_HD void replay_layout(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane, int nl) {
  const ArenaView av = arena_view(L.ab, a.logcap, a.feats != 0);
  const LogEnt C2A_GLOBAL* lg = gp(L.logbuf ? av.log0 : av.log1);
  int32_t C2A_GLOBAL* order = gp(av.order);
  HotRec C2A_GLOBAL* hot = gp(av.hot);
  const int n = L.nlog_prev;
  const int nrec = 1 + 2 * n;
  const int top = g.ls[CAP + 2];
  const bool learned = a.scorer.on != 0;
  const bool sampling = a.sample_thr != 0;
  int32_t C2A_GLOBAL* sz = gp(av.sz);
  int32_t C2A_GLOBAL* pos = gp(av.pos);
  Event C2A_GLOBAL* ev = gp(av.ev);
  double C2A_GLOBAL* ml = gp(av.minleaf);
  for (int i = lane; i < nrec; i += nl) {
    sz[i] = 1;
    if (sampling) {
      const HotRec h = hot[i];
      ml[i] = (i > 0 && (h.flags & 3) == 3) ? h.dTri : INFINITY;
    }
  }
  mem_sync();
  for (int v = top - 1; v >= 0; v--) {
    const int s0 = g.ls[v];
    const int cnt = (v + 1 < CAP + 2 ? g.ls[v + 1] : n) - s0;
    for (int j = lane; j < cnt; j += nl) {
      const int e = order[s0 + j];
      sz[lg[e].rec] = 1 + sz[2 * e + 1] + sz[2 * e + 2];
      if (sampling) {
        const double u = ml[2 * e + 1], w = ml[2 * e + 2];
        ml[lg[e].rec] = u < w ? u : w;
      }
    }
    mem_sync();
  }
  if (sampling) write_samples(L, a, av, nrec, 1.0 / nodes_of(mesh_of(g, 0))->angR, lane, nl);
  if (lane == 0) pos[0] = -1;
  mem_sync();
  for (int v = 0; v < top; v++) {
    const int s0 = g.ls[v];
    const int cnt = (v + 1 < CAP + 2 ? g.ls[v + 1] : n) - s0;
    for (int j = lane; j < cnt; j += nl) {
      const int e = order[s0 + j];
      const int p0 = pos[lg[e].rec] + 1;
      const int rA = 2 * e + 1, rC = 2 * e + 2;
      const bool c_first = learned ? hot[rC].score < hot[rA].score : hot[rC].d < hot[rA].d;
      const int rf = c_first ? rC : rA, rs = c_first ? rA : rC;
      pos[rf] = p0;
      pos[rs] = p0 + sz[rf];
    }
    mem_sync();
  }
  for (int r = lane + 1; r < nrec; r += nl) {
    const LogEnt le = lg[(r - 1) >> 1];
    const HotRec h = hot[r];
    const bool leaf = (h.flags & 3) == 3;
    Event E;
    E.d = h.d;
    E.mt = h.mt;
    E.dTri = h.dTri;
    E.mleaf = h.mleaf;
    E.b1 = leaf ? h.t1 : h.b1;
    E.b2 = leaf ? h.t2 : h.b2;
    E.rec = r;
    E.size = sz[r];
    E.ppos = pos[le.rec];
    E.info = (h.flags & 3) | (h.cp >= 0 ? 4 : 0) | (((r - 1) & 1) << 3) | ((le.level + 1) << 8);
    ev[pos[r]] = E;
  }
  mem_sync();
  L.wbase = -(1 << 30);
  lane_sync();
}

// Starts the scan of a replayed step: the root pair is always expanded
// (C2A.cpp:1920 -> 1279) and logged first.
template <int K, int CAP>
C2A_HD void start_scan(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane, int nl) {
  const ArenaView av = arena_view(L.ab, a.logcap);
  if (lane == 0) {
    LogEnt v;
    v.rec = 0;
    v.b1 = 0;
    v.b2 = 0;
    v.level = 0;
    gp(L.logbuf ? av.log1 : av.log0)[0] = v;
  }
  for (int i = lane; i < CAP + 2; i += nl) {
    g.anc[i] = i == 0 ? 0 : -1;
    g.lastlane[i] = -1;
  }
  lane_sync();
  L.nlog = 1;
  L.nflush = 1;
  L.logfull = 0;
  L.bv += 2;
  L.spos = 0;
  L.nev = 2 * L.nlog_prev;
}

// The reference DFS over the replayed subtree, 64 decisions at a time.  The
// events are the replayed pairs in this step's depth-first order; each lane
// decides one with the current bound (C2A.cpp:1284-1349).  The running
// distance only changes at an improving leaf, so every decision before the
// first improving leaf is final; a pair the previous step did not expand is
// handed to the sequential walk.  A pruned pair the previous step expanded
// skips its replayed subtree.  Returns 1 when such a pair waits on the
// stack, 0 when the step is complete.
template <int K, int CAP>
C2A_HD int scan(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane, int nl) {
  const ArenaView av = arena_view(L.ab, a.logcap);
  const Event C2A_GLOBAL* ev = gp(av.ev);
  LogEnt C2A_GLOBAL* lw = gp(L.logbuf ? av.log1 : av.log0);
  const double inf = 1.0 / 0.0;
  // every wavefront of the workgroup decides the same 64 events (its lane j
  // takes event t0 + j) and so holds the same state
  const int cw = nl < 64 ? nl : 64;
  const int j = lane & 63;
  while (L.spos < L.nev) {
    const int t0 = L.spos;
    const int n = (L.nev - t0) < cw ? (L.nev - t0) : cw;
    const bool valid = j < n;
    Event E;
    if (valid) {
      E = ev[t0 + j];
    } else {
      E.d = E.mt = E.dTri = E.mleaf = 0.0;
      E.b1 = E.b2 = E.rec = E.ppos = 0;
      E.size = 1;
      E.info = 3;
    }
    const int lev = E.info >> 8;
    C2A_CHECK(!valid || (lev >= 1 && lev < CAP + 2 && E.size >= 1));
    const bool internal = (E.info & 3) != 3;
    const bool hasch = (E.info & 4) != 0;
    const bool desc =
        valid && E.mt < L.Up && ((E.d < (L.D - L.abs_err)) || (E.d * (1 + L.rel_err) < L.D));
    const bool skip = valid && internal && !desc && hasch;
    const int covered = excl_prefix_max(skip ? j + E.size : 0, j);
    const bool active = valid && j >= covered;
    const bool miss = active && internal && desc && !hasch;
    const bool tested = active && !internal && desc;
    const bool improving = tested && E.dTri <= L.D;
    const uint64_t stopm = wave_ballot(miss || improving);
    const int stop = stopm ? (int)__builtin_ctzll(stopm) : n;
    const bool in = active && j < stop;
    const bool expand = in && internal && desc;
    // pruned decisions lower mint (C2A.cpp:1284-1349, `else if`); in strict
    // mode so do the tested leaves before the stop (none of them improves)
    const double mc = wave_mind((in && !desc)                ? E.mt
                                : (a.strict && in && tested) ? leaf_time(E.dTri, E.mleaf, 1)
                                                             : inf);
    if (mc < L.M) L.M = mc;
    const uint64_t em = wave_ballot(expand);
    const int nexp = (int)__builtin_popcountll(em);
    L.bv += 2 * nexp;
    L.tri += (int)__builtin_popcountll(wave_ballot(in && tested));
    // log the expansions in depth-first order (next step's replay)
    if (!L.logfull && L.nlog + nexp > L.klog) L.logfull = 1;
    const int en = (expand && !L.logfull) ? L.nlog + lanes_below(em, j) : -1;
    const int pl = E.ppos - t0;
    const int pen_in = lane_bcast(en, pl >= 0 ? pl : 0);
    const int pen = E.ppos < 0 ? 0 : (pl >= 0 ? pen_in : g.anc[lev - 1 < 0 ? 0 : lev - 1]);
    const int nn = (L.logfull || pen < 0) ? -1 : 2 * pen + 1 + ((E.rec - 1) & 1);
    if (expand && en >= 0 && lane < 64) {
      LogEnt v;
      v.rec = nn;
      v.b1 = E.b1;
      v.b2 = E.b2;
      v.level = lev;
      lw[en] = v;
    }
    if (!L.logfull) L.nlog += nexp;
    L.nflush = L.nlog;
    // the last decided event at each depth is the parent of later events
    lane_sync();
    if (in) lds_max(&g.lastlane[lev], j);
    lane_sync();
    const bool last = in && g.lastlane[lev] == j;
    lane_sync();
    if (last) {
      g.anc[lev] = en;
      g.lastlane[lev] = -1;
    }
    lane_sync();
    if (stop < n) {
      const double sd = lane_bcastd(E.d, stop), smt = lane_bcastd(E.mt, stop);
      const double sdt = lane_bcastd(E.dTri, stop), sml = lane_bcastd(E.mleaf, stop);
      const int sb1 = lane_bcast(E.b1, stop), sb2 = lane_bcast(E.b2, stop);
      const int sinfo = lane_bcast(E.info, stop), srec = lane_bcast(E.rec, stop);
      const int snn = lane_bcast(nn, stop);
      L.spos = t0 + stop + 1;
      if ((sinfo & 3) == 3) {
        // improving leaf (C2A.cpp:1141-1183)
        L.D = sdt;
        L.pq_pending = 1;
        const double ml = leaf_time(sdt, sml, a.strict);
        if (ml <= L.M) L.M = ml;
        L.last1 = sb1;
        L.last2 = sb2;
        L.tri++;
        if (lane == 0) g.anc[sinfo >> 8] = -1;
        lane_sync();
        continue;
      }
      // a pair the previous step did not expand: sequential walk below it.
      L.nmiss++;
      // If the last round was the one below an earlier miss of this window
      // and covered this one too, its heap is still in the slots; otherwise
      // this miss and the misses after it in the window are gathered for
      // the next round.
      int csg = 0;
      const int mgen = g.mrgen, nmr = g.nmr, mpg = g.mpage;
      if (mgen == g.gen)
        for (int r = 1; r < nmr; r++)
          if (g.mrrec[r] == srec) csg = (mgen << 8) | (r * mpg + 1);
      lane_sync();
      if (csg == 0) {
        const bool gather = miss && j >= stop;
        const uint64_t mm = wave_ballot(gather);
        const int r = lanes_below(mm, j);
        if (gather && r < MR_MAX && lane < 64) {
          g.mrrec[r] = E.rec;
          g.mrb[r][0] = E.b1;
          g.mrb[r][1] = E.b2;
        }
        if (lane == 0) {
          const int cnt = (int)__builtin_popcountll(mm);
          g.nmr = cnt < MR_MAX ? cnt : MR_MAX;
          g.mrgen = -1;
        }
      }
      Cur c;
      c.d = sd;
      c.mt = smt;
      c.b1 = sb1;
      c.b2 = sb2;
      c.sg = csg;
      c.flags = (sinfo & 3) | 8;
      c.dTri = sdt;
      c.mleaf = sml;
      c.ar = srec;
      c.cp = -1;
      c.nn = snn;
      c.lvl = sinfo >> 8;
      if (lane == 0) g.anc[sinfo >> 8] = -1;
      write_cur(g.stack[0], c, lane);
      lane_sync();
      L.sp = 1;
      return 1;
    }
    const int reach = wave_maxi((in && skip) ? j + E.size : 0);
    L.spos = t0 + (reach > n ? reach : n);
  }
  return 0;
}

// Stages replayed records [rec, rec + WIN) in LDS.
template <int K, int CAP>
C2A_HD void load_window(Wave<K, CAP>& g, Leader& L, const ArenaView& av, int rec, int nrec_max,
                        int lane, int nl) {
  const HotRec C2A_GLOBAL* hot = gp(av.hot);
  C2A_CHECK(rec >= 0 && rec + 1 < nrec_max);
  lane_sync();
  for (int i = lane; i < WIN; i += nl)
    if (rec + i < nrec_max) g.win[i] = hot[rec + i];
  L.wbase = rec;
  lane_sync();
}

// Copies the staged log entries to the arena (lane-parallel).  Keeping
// global stores out of the per-expansion path matters: a store's completion
// is waited for wherever the walk next waits on vector memory.
template <int K, int CAP>
C2A_HD void flush_log(Wave<K, CAP>& g, Leader& L, const ArenaView& av, int lane, int nl) {
  const int n = L.nlog - L.nflush;
  if (n <= 0) return;
  LogEnt C2A_GLOBAL* lw = gp(L.logbuf ? av.log1 : av.log0) + L.nflush;
  lane_sync();
  for (int i = lane; i < n; i += nl) lw[i] = g.lbuf[i];
  L.nflush = L.nlog;
  lane_sync();
}

// One contact record (C2A.cpp:1561-1631) of triangle pair (t1, t2) at the
// contact traversal's pose: the features and closest points of the same
// triangle distance the leaf test made.  Feature ids are the vertex slots
// 0..2 of the triangle; the host maps them to the caller's vertex ids.
template <int K, int CAP>
C2A_HD void record_contact(const Wave<K, CAP>& g, c2a_contact* out, int t1, int t2) {
  const MeshDev M1 = mesh_of(g, 0);
  const MeshDev M2 = mesh_of(g, 1);
  double ta[9], tb[9], Rrel[9], Trel[3], p[3], q[3], u[3], pb[3];
  gdouble_ptr g1 = tris_of(M1) + 9 * (int64_t)t1;
  gdouble_ptr g2 = tris_of(M2) + 9 * (int64_t)t2;
  for (int i = 0; i < 9; i++) { ta[i] = g1[i]; tb[i] = g2[i]; Rrel[i] = g.Rrel[i]; }
  for (int i = 0; i < 3; i++) Trel[i] = g.Trel[i];
  int feat[4];
  const double d = tri_distance(Rrel, Trel, ta, tb, p, q, feat);
  sub3(u, q, Trel);
  mtxv(pb, Rrel, u);
  c2a_contact c;
  c.tri[0] = t1;
  c.tri[1] = t2;
  c.type[0] = feat[0] + 1;
  c.type[1] = feat[2] + 1;
  for (int k = 0; k < 3; k++) { c.fid[0][k] = -1; c.fid[1][k] = -1; }
  c.fid[0][0] = feat[1];
  c.fid[1][0] = feat[3];
  for (int i = 0; i < 3; i++) { c.pa[i] = p[i]; c.pb[i] = pb[i]; }
  c.distance = d;
  *gp(out) = c;
}

// Puts the pair being processed back on the stack (a round will be rooted
// at it).  where: 0 = its full entry is still at stack[sp] (it was just
// popped), 1 = it is slot `csl` of the current round, 2 = a replayed record.
template <int K, int CAP>
C2A_HD void push_back(Wave<K, CAP>& g, Leader& L, const Cur& c, int where, int csl, int lane,
                      int nl) {
  if (where == 1) {
    lane_sync();
    slot_to_entry(g.stack[L.sp], g.slot[csl], c.sg, c.flags, c.nn, c.lvl, lane, nl);
  } else if (where == 2) {
    lane_sync();
    write_cur(g.stack[L.sp], c, lane);
  }
  lane_sync();
  L.sp++;
}

// The reference DFS (C2A.cpp:1113-1354 and, in contact mode, 1525-1725) as
// a stack machine.  The pair being processed is held in registers; an
// expansion pushes the child visited second and continues with the first,
// so the stack holds exactly the reference's pending siblings.  Uniform
// across lanes.  Returns true when a round has been requested.
template <int K, int CAP>
C2A_HD bool walk(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane, int nl) {
  const int cmode = g.contact;
  const bool contact = cmode != 0;
  const bool learned = a.scorer.on && !contact;
  const int gen = uni(g.gen);
  const int mp = uni(g.mpage);
  const ArenaView av = arena_view(L.ab, a.logcap);
  const int nrec_max = arena_records(a.logcap);
  Cur c;
  int where = 0, csl = 0;
  bool have = false;
  for (;;) {
    if (!have) {
      if (L.sp == 0) return false;
      L.sp--;
      c = read_cur(g.stack[L.sp]);
      where = 0;
    }
    have = false;
    const int flags = c.flags;
    if (!(flags & 4)) {
      bool desc;
      if (contact)
        desc = (L.mode == MODE_DISTANCE)   // C2A_PQP.cpp:594-617
                   ? ((c.d < (L.D - L.abs_err)) || (c.d * (1 + L.rel_err) < L.D)) &&
                         // C2A_Collide's overlap test (C2A_PQP.cpp:1068-1070):
                         // RectDist <= r1 + r2 exactly when the clamped RSS
                         // distance is 0
                         (cmode != 2 || c.d == 0.0)
                   : c.d < L.thr;   // abs = rel = 0 (C2A.cpp:1760-1761)
      else desc = c.mt < L.Up && ((c.d < (L.D - L.abs_err)) || (c.d * (1 + L.rel_err) < L.D));
      if (!desc) {
        if (!contact && c.mt < L.M) L.M = c.mt;
        continue;
      }
    }
    const int sg = c.sg;
    const bool cached = (sg >> 8) == gen && (sg & 0xff) != 0;
    const int s = (sg & 0xff) - 1;
    if ((flags & 3) == 3) {
      if (flags & 8) {
        // replayed leaf pair (C2A.cpp:1141-1183); p1/p2 are recomputed at the
        // end of the step if this stays the last improving leaf
        const bool imp = c.dTri <= L.D;
        const double ml = leaf_time(c.dTri, c.mleaf, a.strict);
        if ((imp || a.strict) && ml <= L.M) L.M = ml;
        if (imp) {
          L.D = c.dTri;
          L.pq_pending = 1;
          L.last1 = c.b1;
          L.last2 = c.b2;
        }
        L.tri++;
        continue;
      }
      if (!cached) {
        C2A_EMU_COUNT(0, L);
        push_back(g, L, c, where, csl, lane, nl);
        request_round(g, L, a, lane, nl);
        return true;
      }
      const Slot& sl = g.slot[s];
      if (!contact) {
        const double dTri = unid(sl.dTri);
        const bool imp = dTri <= L.D;
        const double ml = leaf_time(dTri, unid(sl.mleaf), a.strict);
        if ((imp || a.strict) && ml <= L.M) L.M = ml;
        if (imp) {
          L.D = dTri;
          L.pq_pending = 0;
          for (int i = 0; i < 3; i++) { L.p1[i] = unid(sl.p[i]); L.p2[i] = unid(sl.q[i]); }
          L.last1 = uni(sl.t1);
          L.last2 = uni(sl.t2);
        }
        L.tri++;
      } else if (L.mode == MODE_DISTANCE) {   // C2A_PQP.cpp:493-520 (strict <)
        const double dTri = unid(sl.dTri);
        if (dTri < L.D) {
          L.D = dTri;
          for (int i = 0; i < 3; i++) { L.p1[i] = unid(sl.p[i]); L.p2[i] = unid(sl.q[i]); }
          L.last1 = uni(sl.t1);
          L.last2 = uni(sl.t2);
        }
        L.tri++;
      } else {
        if ((uni(sl.flags) & 4) && unid(sl.dTri) <= L.thr) {
          if (a.contacts && L.num_contact < a.contact_cap && lane == 0)
            record_contact(g, a.contacts + L.qi * (int64_t)a.contact_cap + L.num_contact,
                           uni(sl.t1), uni(sl.t2));
          L.num_contact++;
        }
      }
      continue;
    }
    const bool from_arena = (flags & 8) && c.cp >= 0;
    const int loc = mp ? (s & (mp - 1)) : s;   // index within its heap
    const int lvl = cached ? 31 - clz32((unsigned)(loc + 1)) : K;
    if (!from_arena && (!cached || lvl >= uni(g.rdepth))) {
      C2A_EMU_COUNT((flags & 8) ? 1 : (cached ? 2 : 3), L);
      push_back(g, L, c, where, csl, lane, nl);
      request_round(g, L, a, lane, nl);
      return true;
    }
    if (L.sp + 2 > CAP) {
      L.flags |= C2A_FLAG_STACK_OVERFLOW;
      L.sp = 0;
      return false;
    }
    if (!contact || L.mode == MODE_DISTANCE) L.bv += 2;
    // log the expansion for the next step's replay
    int nnA = -1;
    if (c.nn >= 0 && L.nlog < L.klog) {
      const int en = L.nlog++;
      if (lane == 0) {
        LogEnt v;
        v.rec = c.nn;
        v.b1 = c.b1;
        v.b2 = c.b2;
        v.level = c.lvl;
        g.lbuf[en - L.nflush] = v;
      }
      nnA = 2 * en + 1;
      if (L.nlog - L.nflush == LBUF) flush_log(g, L, av, lane, nl);
    }
    const int nnC = nnA < 0 ? -1 : nnA + 1;
    const int clvl = c.lvl + 1;
    if (from_arena) {
      const int ra = 2 * c.cp + 1;
      if (ra < L.wbase || ra + 1 >= L.wbase + WIN) load_window(g, L, av, ra, nrec_max, lane, nl);
      C2A_CHECK(ra - L.wbase >= 0 && ra + 1 - L.wbase < WIN && L.sp < CAP);
      const Cur cA = cur_from_hot(g.win[ra - L.wbase], ra, nnA, clvl);
      const Cur cC = cur_from_hot(g.win[ra + 1 - L.wbase], ra + 1, nnC, clvl);
      // closer child visited first (C2A.cpp:1281), or the lower score under
      // the learned order (replayed pairs are scored); the other one waits
      const bool c_first = learned ? unif(g.win[ra + 1 - L.wbase].score) <
                                         unif(g.win[ra - L.wbase].score)
                                   : cC.d < cA.d;
      lane_sync();
      write_cur(g.stack[L.sp], c_first ? cA : cC, lane);
      lane_sync();
      L.sp++;
      c = c_first ? cC : cA;
      where = 2;
      have = true;
      continue;
    }
    const int sa = s + loc + 1, sc = sa + 1;   // children in the same heap
    // pairs first reached in this step keep the closer-child-first order:
    // scoring a round's slots costs more than the better order saves
    const double dA = unid(g.slot[sa].d), dC = unid(g.slot[sc].d);
    const bool c_first = dC < dA;
    const int sf = c_first ? sc : sa, ss = c_first ? sa : sc;
    const int nf = c_first ? nnC : nnA, ns = c_first ? nnA : nnC;
    const Slot& F = g.slot[sf];
    lane_sync();
    slot_to_entry(g.stack[L.sp], g.slot[ss], (gen << 8) | (ss + 1), uni(g.slot[ss].flags) & 3,
                  ns, clvl, lane, nl);
    lane_sync();
    L.sp++;
    c.d = c_first ? dC : dA;
    c.mt = unid(F.mt);
    c.b1 = uni(F.b1);
    c.b2 = uni(F.b2);
    c.sg = (gen << 8) | (sf + 1);
    c.flags = uni(F.flags) & 3;
    c.dTri = 0.0;
    c.mleaf = 0.0;
    c.ar = -1;
    c.cp = -1;
    c.nn = nf;
    c.lvl = clvl;
    where = 1;
    csl = sf;
    have = true;
  }
}

// After a step's walk: closest points of a replayed last improving leaf
// (the same triangle-distance call a round would have made), and the log
// hand-over to the next step.
template <int K, int CAP>
C2A_HD void end_walk(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane, int nl) {
  if (L.pq_pending) {
    const MeshDev M1 = mesh_of(g, 0);
    const MeshDev M2 = mesh_of(g, 1);
    double ta[9], tb[9], Rrel[9], Trel[3];
    gdouble_ptr g1 = tris_of(M1) + 9 * (int64_t)L.last1;
    gdouble_ptr g2 = tris_of(M2) + 9 * (int64_t)L.last2;
    for (int i = 0; i < 9; i++) { ta[i] = g1[i]; tb[i] = g2[i]; Rrel[i] = g.Rrel[i]; }
    for (int i = 0; i < 3; i++) Trel[i] = g.Trel[i];
    (void)tri_distance(Rrel, Trel, ta, tb, L.p1, L.p2, nullptr);
    L.pq_pending = 0;
  }
  if (a.logcap > 0) {
    flush_log(g, L, arena_view(L.ab, a.logcap), lane, nl);
    L.nlog_prev = L.nlog;
    L.logbuf ^= 1;
  }
}

template <int K, int CAP>
C2A_HD void begin_step(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane) {
  const MeshDev M1 = mesh_of(g, 0);
  const MeshDev M2 = mesh_of(g, 1);
  if (L.mode != MODE_DISTANCE) {   // a distance query's traversal starts in fetch
    double R1[9], T1[3], R2[9], T2[3];
    if (L.numCA == 0) {
      for (int i = 0; i < 9; i++) { R1[i] = g.mo[0].R0[i]; R2[i] = g.mo[1].R0[i]; }
      for (int i = 0; i < 3; i++) { T1[i] = g.mo[0].T0[i]; T2[i] = g.mo[1].T0[i]; }
    } else {
      motion_integrate(g.mo[0], L.lamda, R1, T1, nullptr);
      motion_integrate(g.mo[1], L.lamda, R2, T2, nullptr);
    }
    start_traversal(g, L, M1, M2, R1, T1, R2, T2, a.logcap > 0 ? 0 : -1, lane);
  }
  L.nlog = 0;
  L.nflush = 0;
  L.pq_pending = 0;
  L.napp = a.logcap > 0 ? L.nlog_prev : 0;   // round records follow the replayed ones
  // The replay's extension removes misses at the price of records the step
  // may not expand.  Adaptive (as the library runs it): it stays off for a
  // step after one whose misses were fewer than one in 64 logged expansions
  // (a step without a scan, the first one, counts as many).  Values never
  // depend on it.
  L.kx = (!a.ext_adaptive || L.nmiss < 0 || 64 * L.nmiss >= L.nlog_prev) ? L.kext : 0;
  L.nmiss = (a.logcap > 0 && L.nlog_prev > 0) ? 0 : -1;
  L.wbase = -(1 << 30);
  // initial bound from the caller-seeded pair (C2A.cpp:1812-1858)
  double p[3], q[3], ta[9], tb[9];
  gdouble_ptr g1 = tris_of(M1) + 9 * (int64_t)L.seed1;
  gdouble_ptr g2 = tris_of(M2) + 9 * (int64_t)L.seed2;
  for (int i = 0; i < 9; i++) { ta[i] = g1[i]; tb[i] = g2[i]; }
  double Rrel[9], Trel[3];
  for (int i = 0; i < 9; i++) Rrel[i] = g.Rrel[i];
  for (int i = 0; i < 3; i++) Trel[i] = g.Trel[i];
  L.D = tri_distance(Rrel, Trel, ta, tb, p, q, nullptr);
  if (L.mode == MODE_DISTANCE) {   // C2A_Distance's first bound (C2A_PQP.cpp:998-1005)
    for (int i = 0; i < 3; i++) { L.p1[i] = p[i]; L.p2[i] = q[i]; }
    L.state = ST_CONTACT_WALK;
    return;
  }
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
    for (int i = 0; i < 9; i++) r->R[i] = g.Rrel[i];
    for (int i = 0; i < 3; i++) r->T[i] = g.Trel[i];
    r->abs_err = L.abs_err;
    r->rel_err = L.rel_err;
    r->UpboundTOC = L.Up;
  }
  if (!collisionfree && a.with_contacts) {
    const MeshDev M1 = mesh_of(g, 0);
    const MeshDev M2 = mesh_of(g, 1);
    double R1[9], T1[3], R2[9], T2[3];
    motion_integrate(g.mo[0], L.toc, R1, T1, nullptr);
    motion_integrate(g.mo[1], L.toc, R2, T2, nullptr);
    L.thr = a.contact_scale * L.D + a.contact_offset;
    L.num_contact = 0;
    start_traversal(g, L, M1, M2, R1, T1, R2, T2, -1, lane);
    if (lane == 0) g.contact = 1;
    lane_sync();
    L.state = ST_CONTACT_WALK;
    return;
  }
  L.state = ST_FETCH;
}

// Saves the CA-loop state of a query about to start its next step and hands
// the query to the tail pass (the next step starts there without a replay:
// pair values do not depend on it).
C2A_HD void hand_over(Leader& L, const KArgs& a, int lane) {
  if (lane == 0) {
    TailRec& t = a.tail_rec[L.qi];
    t.lamda = L.lamda;
    t.lastLamda = L.lastLamda;
    t.Up = L.Up;
    t.M = L.M;
    for (int i = 0; i < 3; i++) { t.p1[i] = L.p1[i]; t.p2[i] = L.p2[i]; }
    t.bv = L.bv;
    t.tri = L.tri;
    t.nItrs = L.nItrs;
    t.numCA = L.numCA;
    t.flags = L.flags;
    t.last1 = L.last1;
    t.last2 = L.last2;
#if defined(__HIP_DEVICE_COMPILE__)
    a.tail_list[atomicAdd(a.tail_count, 1)] = (int32_t)L.qi;
#else
    a.tail_list[(*a.tail_count)++] = (int32_t)L.qi;
#endif
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
      for (int i = 0; i < 9; i++) r->R[i] = g.Rrel[i];
      for (int i = 0; i < 3; i++) r->T[i] = g.Trel[i];
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
        if (a.tail_budget > 0 && !a.tail_pass && L.bv >= a.tail_budget) hand_over(L, a, lane);
        return;
      }
    }
  }
  double toc = L.lastLamda;
  if (toc >= 1 - a.tol_t) toc = 0;
  L.toc = toc;
  finish_query(g, L, a, false, lane);
}

// Takes the next query from the work queue (one atomic per workgroup).
template <int K, int CAP>
C2A_HD int64_t take_ticket(Wave<K, CAP>& g, int64_t* counter, int lane) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (lane == 0) g.ticket = (int64_t)atomicAdd((unsigned long long*)counter, 1ull);
#else
  if (lane == 0) g.ticket = (*counter)++;
#endif
  lane_sync();
  const int64_t t = g.ticket;
  lane_sync();
  return t;
}

template <int K, int CAP>
C2A_HD bool fetch(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane) {
  int64_t qi = take_ticket(g, a.counter, lane);
  if (a.tail_pass) {   // the first launch's hand-overs
    if (qi >= *a.tail_count) {
      L.state = ST_DONE;
      return false;
    }
    qi = a.tail_list[qi];
  } else if (qi >= a.nq) {
    L.state = ST_DONE;
    return false;
  }
  L.qi = qi;
  L.kdepth = K;
  L.kmiss = a.miss_depth;
  L.klog = a.logcap;
  L.kext = a.ext_levels;
  L.nmiss = -1;
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
      g.md[0] = a.meshes[q.mesh[0]];
      g.md[1] = a.meshes[q.mesh[1]];
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
    start_traversal(g, L, M1, M2, R1, T1, R2, T2, -1, lane);
    L.nlog = 0;
    L.nflush = 0;
    L.pq_pending = 0;
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
  if (a.mode == MODE_CONTACT || a.mode == MODE_DISTANCE) {
    // contact enumeration, or C2A_Distance with qsize <= 2 (C2A_PQP.cpp:970-1057):
    // both walk pair distances without motion bounds
    const ContactDev& c = a.contact_in[qi];
    if (lane == 0) {
      g.mesh[0] = c.mesh[0];
      g.mesh[1] = c.mesh[1];
      g.md[0] = a.meshes[c.mesh[0]];
      g.md[1] = a.meshes[c.mesh[1]];
      g.contact = (a.mode == MODE_DISTANCE && c.overlap_only) ? 2 : 1;
    }
    lane_sync();
    L.mode = a.mode;
    L.thr = c.threshold;
    L.num_contact = 0;
    L.rel_err = c.rel_err;
    L.abs_err = c.abs_err;
    L.seed1 = c.seed[0];
    L.seed2 = c.seed[1];
    L.last1 = c.seed[0];
    L.last2 = c.seed[1];
    const MeshDev M1 = a.meshes[c.mesh[0]];
    const MeshDev M2 = a.meshes[c.mesh[1]];
    double R1[9], T1[3], R2[9], T2[3];
    for (int i = 0; i < 9; i++) { R1[i] = c.R1[i]; R2[i] = c.R2[i]; }
    for (int i = 0; i < 3; i++) { T1[i] = c.T1[i]; T2[i] = c.T2[i]; }
    start_traversal(g, L, M1, M2, R1, T1, R2, T2, -1, lane);
    // C2A_Collide tests the root pair too (C2A_PQP.cpp:1068); the host made
    // that test when it set up the query
    if (c.overlap_only == 2) L.sp = 0;
    // a distance query takes its first bound in begin_step
    L.state = a.mode == MODE_DISTANCE ? ST_STEP : ST_CONTACT_WALK;
    return true;
  }
  int s1, s2;
  if (a.qmot) {
    // device-resident records: mesh ids and seeds are checked here, since
    // the host never sees them; a bad query is flagged and skipped
    const int m1 = a.qpairs[2 * qi], m2 = a.qpairs[2 * qi + 1];
    bool bad = m1 < 0 || m1 >= a.nmesh || m2 < 0 || m2 >= a.nmesh;
    s1 = a.qseeds ? a.qseeds[2 * qi] : 0;
    s2 = a.qseeds ? a.qseeds[2 * qi + 1] : 0;
    s1 = s1 < 0 ? 0 : s1;
    s2 = s2 < 0 ? 0 : s2;
    if (!bad) bad = s1 >= a.meshes[m1].ntri || s2 >= a.meshes[m2].ntri;
    if (bad) {
      if (lane == 0) {
        c2a_result* r = a.out + qi;
        memset(r, 0, sizeof(*r));
        r->flags = C2A_FLAG_BAD_INPUT;
        r->last_tri[0] = r->last_tri[1] = -1;
      }
      L.state = ST_FETCH;
      return true;
    }
    if (lane == 0) {
      for (int k = 0; k < 2; k++) {
        const double* src = a.qmot + (2 * qi + k) * 24;
        double* dst = reinterpret_cast<double*>(&g.mo[k]);
        for (int w = 0; w < 24; w++) dst[w] = src[w];
        Motion& mo = g.mo[k];
        if (a.strict) mxv(mo.bax, mo.R0, mo.axis);
        else for (int i = 0; i < 3; i++) mo.bax[i] = mo.axis[i];
      }
      g.mesh[0] = m1;
      g.mesh[1] = m2;
      g.md[0] = a.meshes[m1];
      g.md[1] = a.meshes[m2];
    }
  } else {
    const QueryDev& q = a.q[qi];
    if (lane == 0) {
      g.mo[0] = q.mo[0];
      g.mo[1] = q.mo[1];
      g.mesh[0] = q.mesh[0];
      g.mesh[1] = q.mesh[1];
      g.md[0] = a.meshes[q.mesh[0]];
      g.md[1] = a.meshes[q.mesh[1]];
    }
    s1 = q.seed[0];
    s2 = q.seed[1];
  }
  lane_sync();
  L.seed1 = s1;
  L.seed2 = s2;
  L.last1 = s1;   // model last_tri starts at the seed (demo protocol)
  L.last2 = s2;
  if (a.ntune > 0) {   // the size bucket of the larger mesh
    const int n1 = uni(g.md[0].ntri), n2 = uni(g.md[1].ntri);
    const int big = n1 > n2 ? n1 : n2;
    int b = 0;
    while (b + 1 < a.ntune && big > a.tune_ntri[b]) b++;
    L.kdepth = a.tune_depth[b] < K ? a.tune_depth[b] : K;
    L.kmiss = a.tune_miss[b];
    L.klog = a.tune_log[b] < a.logcap ? a.tune_log[b] : a.logcap;
    L.kext = a.tune_ext[b];
  }
  L.mode = MODE_QUERY;
  L.nlog_prev = 0;
  L.logbuf = 0;
  L.numCA = 0;
  L.lamda = 0.0;
  L.lastLamda = 0.0;
  L.nItrs = 0;
  L.Up = 1;
  L.M = 1;
  if (unid(g.mo[0].angVel) < 1e-8 && unid(g.mo[1].angVel) < 1e-8)
    L.flags |= C2A_FLAG_TRANSLATION_ONLY;
  L.state = (a.translation && (L.flags & C2A_FLAG_TRANSLATION_ONLY)) ? ST_TRANS : ST_STEP;
  if (a.tail_pass) {   // continue where the first launch stopped
    const TailRec& t = a.tail_rec[qi];
    L.lamda = t.lamda;
    L.lastLamda = t.lastLamda;
    L.Up = t.Up;
    L.M = t.M;
    for (int i = 0; i < 3; i++) { L.p1[i] = t.p1[i]; L.p2[i] = t.p2[i]; }
    L.bv = t.bv;
    L.tri = t.tri;
    L.nItrs = t.nItrs;
    L.numCA = t.numCA;
    L.flags = t.flags;
    L.last1 = t.last1;
    L.last2 = t.last2;
    L.state = ST_STEP;
  }
  return true;
}

// Triangle t of a mesh (9 doubles), and the same in model 1's frame.
C2A_HD void load_tri(double* out, const MeshDev& M, int t) {
  gdouble_ptr g = tris_of(M) + 9 * (int64_t)t;
  for (int i = 0; i < 9; i++) out[i] = g[i];
}

// The translation-only query (C2A.cpp:1818-1852 seed bound, 1362-1521
// traversal, 1900-1917 final pose, 2032-2049 toc = mint): one traversal from
// the start poses, depth first with the reference's order and tests, over
// the LDS stack.  Every lane runs the same sequential code.  Pair values
// depend on the current mint (CAonRSS stops at it), so they are computed when
// their parent is expanded, as in the reference.  The reference leaves
// res->p1/p2 uninitialised on this branch; here they are the closest points
// of the final triangle distance.  The models' last_tri are not updated.
template <int K, int CAP>
C2A_HD void translation_query(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane) {
  const MeshDev M1 = mesh_of(g, 0);
  const MeshDev M2 = mesh_of(g, 1);
  node_ptr N1 = nodes_of(M1);
  node_ptr N2 = nodes_of(M2);
  const double delta = a.tol_d;
  double R1[9], T1[3], R2[9], T2[3];
  for (int i = 0; i < 9; i++) { R1[i] = g.mo[0].R0[i]; R2[i] = g.mo[1].R0[i]; }
  for (int i = 0; i < 3; i++) { T1[i] = g.mo[0].T0[i]; T2[i] = g.mo[1].T0[i]; }
  start_traversal(g, L, M1, M2, R1, T1, R2, T2, -1, lane);
  double Rrel[9], Trel[3], r1[9];
  for (int i = 0; i < 9; i++) { Rrel[i] = g.Rrel[i]; r1[i] = R1[i]; }
  for (int i = 0; i < 3; i++) Trel[i] = g.Trel[i];
  int la = L.seed1, lb = L.seed2;
  {
    double A[9], tb[9], B[9];
    load_tri(A, M1, la);
    load_tri(tb, M2, lb);
    for (int v = 0; v < 3; v++) mxvpv(B + 3 * v, Rrel, tb + 3 * v, Trel);
    double mint = 1.0, dTri = 0.0;
    if (ca_on_tris(g.mo[0], r1, A, B, delta, mint, dTri)) {
      L.M = mint;
      L.D = dTri;
    } else {
      L.M = 1.0;
      L.D = 1e+30;
    }
  }
  L.abs_err = 0;
  L.rel_err = 0;
  while (L.sp > 0) {
    L.sp--;
    const Entry& e = g.stack[L.sp];
    double R[9], T[3];
    for (int i = 0; i < 9; i++) R[i] = unid(e.R[i]);
    for (int i = 0; i < 3; i++) T[i] = unid(e.T[i]);
    const int b1 = uni(e.b1), b2 = uni(e.b2), flags = uni(e.flags);
    const double dv = unid(e.d), mv = unid(e.mt);
    lane_sync();
    if (!(flags & 4)) {   // the parent's deferred test (C2A.cpp:1477-1513)
      if (!(mv < L.M && ((dv < (L.D - L.abs_err)) || (dv * (1 + L.rel_err) < L.D)))) continue;
    }
    node_ptr B1 = N1 + b1;
    node_ptr B2 = N2 + b2;
    const bool l1 = node_leaf(B1), l2 = node_leaf(B2);
    if (l1 && l2) {   // C2A.cpp:1395-1425
      const int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
      double A[9], tb[9], B[9];
      load_tri(A, M1, t1);
      load_tri(tb, M2, t2);
      for (int v = 0; v < 3; v++) mxvpv(B + 3 * v, Rrel, tb + 3 * v, Trel);
      if (ca_on_tris(g.mo[0], r1, A, B, delta, L.M, L.D)) {
        la = t1;
        lb = t2;
      }
      L.tri++;
      continue;
    }
    if (L.sp + 2 > CAP) {
      L.flags |= C2A_FLAG_STACK_OVERFLOW;
      break;
    }
    // children (C2A.cpp:1435-1466) and their local advancement (1470-1474)
    double RA[9], TA[3], RC[9], TC[3];
    for (int i = 0; i < 9; i++) { RA[i] = R[i]; RC[i] = R[i]; }
    for (int i = 0; i < 3; i++) { TA[i] = T[i]; TC[i] = T[i]; }
    int a1 = b1, a2 = b2, c1 = b1, c2 = b2;
    descend(N1, N2, a1, a2, 0, RA, TA);
    descend(N1, N2, c1, c2, 1, RC, TC);
    double dA = 1e+30, dC = 1e+30, mA = L.M, mC = L.M;
    ca_on_rss(g.mo[0], r1, RA, TA, N1 + a1, N2 + a2, delta, mA, dA);
    ca_on_rss(g.mo[0], r1, RC, TC, N1 + c1, N2 + c2, delta, mC, dC);
    L.bv += 2;
    // closer child first (C2A.cpp:1478); both tests wait for their turn
    const bool c_first = dC < dA;
    if (lane == 0) {
      for (int k = 0; k < 2; k++) {
        const bool second = k == 0;   // pushed first, popped last
        const bool isC = second ? !c_first : c_first;
        Entry& o = g.stack[L.sp + k];
        const double* RR = isC ? RC : RA;
        const double* TT = isC ? TC : TA;
        for (int i = 0; i < 9; i++) o.R[i] = RR[i];
        for (int i = 0; i < 3; i++) o.T[i] = TT[i];
        o.d = isC ? dC : dA;
        o.mt = isC ? mC : mA;
        o.b1 = isC ? c1 : a1;
        o.b2 = isC ? c2 : a2;
        o.flags = 0;
        o.sg = 0;
      }
    }
    lane_sync();
    L.sp += 2;
  }
  L.sp = 0;
  // distance at the advanced pose (C2A.cpp:1906-1916)
  motion_integrate(g.mo[0], L.M, R1, T1, nullptr);
  motion_integrate(g.mo[1], L.M, R2, T2, nullptr);
  double Tt[3];
  mtxm(Rrel, R1, R2);
  sub3(Tt, T2, T1);
  mtxv(Trel, R1, Tt);
  {
    double ta[9], tb[9];
    load_tri(ta, M1, la);
    load_tri(tb, M2, lb);
    L.D = tri_distance(Rrel, Trel, ta, tb, L.p1, L.p2, nullptr);
  }
  lane_sync();
  if (lane == 0) {   // res->R, res->T at the advanced pose (C2A.cpp:1902-1904)
    for (int i = 0; i < 9; i++) g.Rrel[i] = Rrel[i];
    for (int i = 0; i < 3; i++) g.Trel[i] = Trel[i];
  }
  lane_sync();
  // this branch leaves the models' last_tri alone and updates the result's
  // last_triA/B to the last improving pair (C2A.cpp:1413-1414): the result
  // reports that pair
  L.last1 = la;
  L.last2 = lb;
  L.toc = L.M;
  L.numCA = 0;
  finish_query(g, L, a, L.toc >= 1.0, lane);
}

// Advances the query state until a round is needed or no work is left.
template <int K, int CAP>
C2A_HD void leader_run(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane, int nl) {
  L.need_round = 0;
  L.need_score = 0;
  for (;;) {
    switch (L.state) {
      case ST_FETCH:
        if (!fetch(g, L, a, lane)) return;
        break;
      case ST_STEP:
        begin_step(g, L, a, lane);
        if (L.state == ST_CONTACT_WALK) break;   // distance query
        if (a.logcap > 0 && L.nlog_prev > 0) {
          replay(g, L, a, lane, nl);
          L.state = ST_LAYOUT;
          if (a.scorer.on) {   // the caller scores the records, then comes back
            L.need_score = 1;
            return;
          }
        } else {
          L.state = ST_WALK;
        }
        break;
      case ST_TRANS:
        translation_query(g, L, a, lane);
        break;
      case ST_LAYOUT:
        replay_layout(g, L, a, lane, nl);
        start_scan(g, L, a, lane, nl);
        L.state = ST_SCAN;
        break;
      case ST_SCAN:
        if (scan(g, L, a, lane, nl)) {
          L.state = ST_SUB;
          break;
        }
        end_walk(g, L, a, lane, nl);
        after_step(g, L, a, lane);
        break;
      case ST_SUB:
        if (walk(g, L, a, lane, nl)) return;
        flush_log(g, L, arena_view(L.ab, a.logcap), lane, nl);
        if (L.flags & C2A_FLAG_STACK_OVERFLOW) {
          end_walk(g, L, a, lane, nl);
          after_step(g, L, a, lane);
        } else {
          L.state = ST_SCAN;
        }
        break;
      case ST_WALK:
        if (walk(g, L, a, lane, nl)) return;
        end_walk(g, L, a, lane, nl);
        after_step(g, L, a, lane);
        break;
      case ST_CONTACT_WALK:
        if (walk(g, L, a, lane, nl)) return;
        if (lane == 0) {
          if (L.mode == MODE_DISTANCE) {
            // p2 back into model 2's frame (C2A_PQP.cpp:1043-1047)
            c2a_distance_result* r = a.dist_out + L.qi;
            double u[3], Rr[9], Tr[3], q2[3];
            for (int i = 0; i < 9; i++) Rr[i] = g.Rrel[i];
            for (int i = 0; i < 3; i++) Tr[i] = g.Trel[i];
            sub3(u, L.p2, Tr);
            mtxv(q2, Rr, u);
            r->distance = L.D;
            for (int i = 0; i < 3; i++) { r->p1[i] = L.p1[i]; r->p2[i] = q2[i]; }
            r->bv_tests = L.bv;
            r->tri_tests = L.tri;
            r->last_tri[0] = L.last1;
            r->last_tri[1] = L.last2;
            r->flags = L.flags;
            r->pad = 0;
          } else if (L.mode == MODE_CONTACT)
            a.counts_out[L.qi] = L.num_contact;
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
