// ccd_types.h -- device-side data of the CCD kernel (plain structs shared by
// the kernel, the host launcher and the host-side test harness).
#pragma once

#include <stdint.h>

#include "../../include/c2a_amd.h"
#include "bvh.h"
#include "c2a_math.h"
#include "scorer.h"

namespace c2a {


enum : int { MODE_QUERY = 0, MODE_STEP = 1, MODE_CONTACT = 2, MODE_DISTANCE = 3 };

constexpr int C2A_TUNE_MAX = 8;   // size buckets of the autotuned knobs

struct MeshDev {
  const GpuNode* nodes;
  const double* tris;      // build order, 9 doubles each
  int32_t ntri;
  int32_t depth;
};

// One query as the kernel reads it: both motions precomputed on the host.
struct QueryDev {
  Motion mo[2];
  int32_t mesh[2];
  int32_t seed[2];
};

// The CA-loop state of a query between two steps (C2A.cpp:2075-2123): what
// the tail pass needs to continue it.  Everything else of the next step is
// recomputed from the query and these values.
struct TailRec {
  double lamda, lastLamda, Up, M;   // M: the last step's mint (error schedule)
  double p1[3], p2[3];              // closest points of the last improving leaf
  int64_t bv, tri;
  int32_t nItrs, numCA, flags, last1, last2, pad[3];
};

// One explicit CA step (backs C2A_TimeOfContactStep).
struct StepDev {
  Motion mo[2];
  double R1[9], T1[3], R2[9], T2[3];
  double mint_prev, UpboundTOC;
  int32_t mesh[2];
  int32_t seed[2];
  int32_t last[2];
  int32_t numCA;
  int32_t pad;
};

// One contact enumeration at explicit poses (backs C2A_QueryContact), or
// one static distance query (backs C2A_Distance: seed pair, error bounds).
struct ContactDev {
  double R1[9], T1[3], R2[9], T2[3];
  double threshold;
  int32_t mesh[2];
  int32_t seed[2];
  int32_t overlap_only;   // C2A_Collide: enter only pairs whose RSS overlap;
                          // 2 = the root volumes are already disjoint
  int32_t pad;
  double rel_err, abs_err;
};

// Replay of the previous CA step.  Consecutive steps of one query visit
// nearly the same pairs, so the walk logs every expansion and the next step
// recomputes the logged subtree in parallel, level by level, before it walks.
// Record numbering of a step: the root pair is record 0 and the two children
// of logged expansion e are records 2e+1 (first child node) and 2e+2 (its
// sibling).
struct LogEnt {          // one expansion of a walk
  int32_t rec;           // record number of the expanded pair
  int32_t b1, b2;        // its node ids
  int32_t level;         // its BVTT depth (root 0)
};

struct HotRec {          // what the walk reads of a replayed record (64 B)
  double d, mt;          // RSS distance and admissible time
  double dTri, mleaf;    // leaf pairs: triangle distance and leaf admissible time
  int32_t b1, b2;        // node ids
  int32_t flags;         // bit0 leaf node of model 1, bit1 leaf node of model 2
  int32_t cp;            // expansion whose children are this pair's children, -1 none
  int32_t t1, t2;        // leaf pairs: triangle indices
  float score;           // learned order: the scorer's output for the pair
  int32_t pad;
};

// One replayed pair as the lane-parallel scan reads it: the records of the
// previous step's subtree laid out in this step's depth-first order (64 B).
struct Event {
  double d, mt;          // RSS distance and admissible time
  double dTri, mleaf;    // leaf pairs: triangle distance and leaf admissible time
  int32_t b1, b2;        // node ids (leaf pairs: triangle indices)
  int32_t rec;           // record number
  int32_t size;          // events in its replayed subtree, itself included
  int32_t ppos;          // scan position of its parent, -1 for the root
  int32_t info;          // bit0/1 leaf nodes, bit2 has replayed children,
                         // bit3 second child node, bits 8+ depth
};

// Per-wavefront replay storage, carved from one device allocation.
struct ArenaView {
  LogEnt* log0;          // written by the walk / read by the next replay, alternating
  LogEnt* log1;
  double* xf;            // nrec x 12: pair transform R[9], T[3]
  HotRec* hot;           // nrec
  int32_t* cpa;          // nrec: expansion index of each record's children
  int32_t* order;        // logcap: expansions sorted by level
  int32_t* sz;           // nrec: replayed subtree size of each record
  int32_t* pos;          // nrec: scan position of each record
  Event* ev;             // nrec - 1: the scan order
  // learned ordering / training samples (only when the arena has room for them)
  uint16_t* feat;        // nrec x NF pair features (BF16)
  double* minleaf;       // nrec closest replayed triangle distance in the subtree
};

// Records: the replayed ones [0, 1 + 2 * logcap), then room for as many
// again computed by speculative rounds during the step.
C2A_HD int arena_records(int logcap) { return 1 + 4 * logcap; }

C2A_HD size_t arena_bytes(int logcap, bool feats = false) {
  const size_t nrec = 1 + 2 * (size_t)logcap;
  const size_t nall = (size_t)arena_records(logcap);
  size_t b = 2 * (size_t)logcap * sizeof(LogEnt) + nall * 12 * sizeof(double) +
             nall * sizeof(HotRec) + nrec * sizeof(int32_t) + (size_t)logcap * sizeof(int32_t) +
             2 * nrec * sizeof(int32_t) + nrec * sizeof(Event);
  // features: 64 rows of slack so that the scorer may read whole 64-row tiles
  if (feats) b += 16 + nrec * sizeof(double) + (nrec + 64) * NF * sizeof(uint16_t);
  return (b + 16 + 255) & ~(size_t)255;
}

C2A_HD ArenaView arena_view(unsigned char* base, int logcap, bool feats = false) {
  const size_t nrec = 1 + 2 * (size_t)logcap;
  const size_t nall = (size_t)arena_records(logcap);
  ArenaView v;
  unsigned char* p = base;
  v.log0 = (LogEnt*)p;
  p += (size_t)logcap * sizeof(LogEnt);
  v.log1 = (LogEnt*)p;
  p += (size_t)logcap * sizeof(LogEnt);
  v.xf = (double*)p;
  p += nall * 12 * sizeof(double);
  v.hot = (HotRec*)p;
  p += nall * sizeof(HotRec);
  v.cpa = (int32_t*)p;
  p += nrec * sizeof(int32_t);
  v.order = (int32_t*)p;
  p += (size_t)logcap * sizeof(int32_t);
  v.sz = (int32_t*)p;
  p += nrec * sizeof(int32_t);
  v.pos = (int32_t*)p;
  p += nrec * sizeof(int32_t);
  p = (unsigned char*)(((uintptr_t)p + 15) & ~(uintptr_t)15);
  v.ev = (Event*)p;
  p += nrec * sizeof(Event);
  v.feat = nullptr;
  v.minleaf = nullptr;
  if (feats) {
    p = (unsigned char*)(((uintptr_t)p + 15) & ~(uintptr_t)15);
    v.minleaf = (double*)p;
    p += nrec * sizeof(double);
    v.feat = (uint16_t*)p;
  }
  return v;
}

struct KArgs {
  const MeshDev* meshes;
  const QueryDev* q;
  // queries as device-resident motion records instead of q (the
  // allocation-free entry point): nq x 2 x 24 doubles laid out as the first
  // 24 words of Motion, mesh pairs and seed pairs (nq x 2; seeds may be
  // null), validated against nmesh and the meshes' triangle counts
  const double* qmot;
  const int32_t* qpairs;
  const int32_t* qseeds;
  int32_t nmesh;
  int32_t pad0;
  const StepDev* step_in;
  c2a_result* out;
  c2a_step_result* step_out;
  const ContactDev* contact_in;
  int32_t* counts_out;
  c2a_contact* contacts;   // contact records, contact_cap per query (or null)
  int32_t contact_cap;
  int32_t pad1;
  c2a_distance_result* dist_out;
  int64_t* counter;        // work-queue ticket, zeroed before each launch
  int64_t nq;
  double tol_d, tol_t;
  double contact_scale, contact_offset;
  int32_t max_iter;
  int32_t with_contacts;
  int32_t mode;
  int32_t logcap;          // expansions logged per step per wavefront, 0 = no replay
  int32_t arena_rounds;    // speculative rounds write their subtree to the arena
  unsigned char* arena;    // one arena_bytes(logcap, feats) slice per workgroup
  int32_t feats;           // the arena holds features (learned order or samples)
  int32_t miss_depth;      // depth of the first round below a miss (0 = K)
  int32_t ext_levels;      // replay: levels of speculative extension below
                           // pairs that newly pass the descend test
  int32_t ext_adaptive;    // 1 (the library): a step after one with few misses
                           // skips it; 0 (test harness): every step
  int32_t translation;     // C2A_Solve's translation-only branch for queries
                           // whose two rotations are below 1e-8
  int32_t strict;          // strict (conservative) mode: every tested leaf
                           // pair bounds mint (motions carry the world axis)
  // Per-size-bucket traversal knobs (autotuned): a CA query takes the first
  // bucket whose triangle bound covers its larger mesh and uses that
  // bucket's speculative depth (<= K), depth below a miss, replay log
  // (<= logcap) and replay extension levels.  ntune == 0: every query uses
  // K, miss_depth, logcap and ext_levels.
  int32_t ntune;
  int32_t tune_ntri[C2A_TUNE_MAX];
  int32_t tune_depth[C2A_TUNE_MAX];
  int32_t tune_miss[C2A_TUNE_MAX];
  int32_t tune_log[C2A_TUNE_MAX];
  int32_t tune_ext[C2A_TUNE_MAX];
  ScorerDev scorer;        // learned child order when scorer.on
  // training samples of replayed pairs: a pair is logged when the top 32
  // bits of mix64(key) are below sample_thr (2^32 = every pair)
  uint64_t sample_thr;
  float* samples;          // sample_cap x (NF + 1): features, then the label
  int64_t* sample_keys;    // sample_cap: query << 34 | CA step << 26 | record
  int64_t sample_cap;
  unsigned long long* sample_count;   // rows claimed (may exceed sample_cap)
  // Tail pass: the first launch hands a query whose BV tests have reached
  // tail_budget at the end of a CA step over to a second launch, which
  // continues it from its saved CA-loop state with several wavefronts per
  // query.  The batch time is its slowest query's; the second launch gives
  // the few long queries the lanes the others no longer use.
  int64_t tail_budget;     // 0: one launch
  int32_t tail_pass;       // 1 in the second launch: queries come from tail_list
  int32_t pad2;
  TailRec* tail_rec;       // per query
  int32_t* tail_list;      // queries handed over, in hand-over order
  int32_t* tail_count;     // their number
};

// splitmix64 finaliser: the sample selection hash.
C2A_HD uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace c2a
