// ccd_emu.cpp -- TEST HARNESS: runs the GPU kernel's per-group logic
// (c2a_amd/csrc/ccd_core.h) on the CPU, one group and one lane at a time, so
// the walk / speculation logic can be checked against the oracle without a
// GPU.  The product library never links this file.
#include <string.h>

#include <algorithm>
#include <memory>
#include <vector>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
static int64_t g_kind[2][4];
#define C2A_CHECK(cond)                                                      \
  do {                                                                       \
    if (!(cond)) {                                                           \
      fprintf(stderr, "ccd_emu: check failed at %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      abort();                                                               \
    }                                                                        \
  } while (0)   // rounds by reason, first step / later steps
#ifdef C2A_EMU_SIMT
// 64 lanes on fibers (simt.cpp); rounds are counted once, by lane 0
namespace c2a { namespace simt { int lane_id(); int launch(void (*)(int, void*), void*); } }
#define C2A_EMU_COUNT(kind, L) \
  (c2a::simt::lane_id() == 0 ? (void)g_kind[(L).numCA > 1 ? 1 : 0][kind]++ : (void)0)
#else
#define C2A_EMU_COUNT(kind, L) (g_kind[(L).numCA > 1 ? 1 : 0][kind]++)
#endif
#include "../../c2a_amd/csrc/bvh.h"
#include "../../c2a_amd/csrc/ccd_core.h"

using namespace c2a;

static int64_t g_rounds = 0;   // speculative rounds of the last batch
extern "C" int64_t emu_rounds() { return g_rounds; }
static int emu_arena_rounds = 1;   // the CPU harness exercises the arena-round path
extern "C" void emu_set_arena_rounds(int on) { emu_arena_rounds = on; }
// Tail pass: queries reaching this many BV tests at the end of a CA step are
// handed over and continued by a second run (0 = one run).
static int64_t emu_tail_budget = 0;
static int64_t emu_tail_handed = 0;
extern "C" void emu_set_tail_budget(int64_t b) { emu_tail_budget = b; }
// The extension switched per step by the previous step's misses (the
// library's mode for the autotuned levels).
static int emu_ext_adaptive = 0;
extern "C" void emu_set_ext_adaptive(int on) { emu_ext_adaptive = on; }
extern "C" int64_t emu_tail_count() { return emu_tail_handed; }
extern "C" int64_t emu_round_kind(int later, int kind) { return g_kind[later][kind]; }

// Learned order on the CPU: a host copy of a scorer (scalar forward pass, so
// scores may differ from the matrix cores' in the last bits).
static std::vector<uint16_t> g_w1t, g_w2t;
static std::vector<float> g_b1, g_b2, g_w3;
static ScorerDev g_scorer;
extern "C" void emu_set_scorer(const float* params) {
  memset(&g_scorer, 0, sizeof(g_scorer));
  if (!params) return;
  g_w1t.assign(NF * NH, 0);
  g_w2t.assign(NH * NH, 0);
  for (int k = 0; k < NF; k++)
    for (int j = 0; j < NH; j++) g_w1t[j * NF + k] = f32_to_bf16(params[P_W1 + k * NH + j]);
  for (int k = 0; k < NH; k++)
    for (int j = 0; j < NH; j++) g_w2t[j * NH + k] = f32_to_bf16(params[P_W2 + k * NH + j]);
  g_b1.assign(params + P_B1, params + P_B1 + NH);
  g_b2.assign(params + P_B2, params + P_B2 + NH);
  g_w3.assign(params + P_W3, params + P_W3 + NH);
  g_scorer.w1t = g_w1t.data();
  g_scorer.w2t = g_w2t.data();
  g_scorer.b1 = g_b1.data();
  g_scorer.b2 = g_b2.data();
  g_scorer.w3 = g_w3.data();
  g_scorer.b3 = params[P_B3];
  g_scorer.on = 1;
}

// Training-sample logging on the CPU.
static double g_rate = 0.0;
static std::vector<float> g_samples;
static std::vector<int64_t> g_keys;
static unsigned long long g_scount = 0;
extern "C" void emu_set_sampling(double rate, int64_t cap) {
  g_rate = rate;
  g_samples.assign((size_t)(cap > 0 ? cap : 0) * (NF + 1), 0.0f);
  g_keys.assign((size_t)(cap > 0 ? cap : 0), 0);
}
extern "C" int64_t emu_samples(float* rows, int64_t* keys, int64_t cap) {
  const int64_t n = std::min<int64_t>((int64_t)g_scount, (int64_t)g_keys.size());
  for (int64_t i = 0; i < n && i < cap; i++) {
    memcpy(rows + i * (NF + 1), &g_samples[(size_t)i * (NF + 1)], (NF + 1) * sizeof(float));
    keys[i] = g_keys[(size_t)i];
  }
  return (int64_t)g_scount;
}

#ifdef C2A_EMU_SIMT
// One wavefront of 64 lanes running the kernel's loop (ccd_kernel.hip), the
// learned order's scores computed per lane by the scalar forward pass.
template <int K, int CAP>
struct SimtCtx {
  const KArgs* a;
  Wave<K, CAP>* g;
  unsigned char* ab;
};

template <int K, int CAP>
static void simt_body(int lane, void* p) {
  SimtCtx<K, CAP>& c = *(SimtCtx<K, CAP>*)p;
  const KArgs& a = *c.a;
  Wave<K, CAP>& g = *c.g;
  Leader L;
  memset(&L, 0, sizeof(L));
  L.state = ST_FETCH;
  L.ab = c.ab;
  for (;;) {
    leader_run(g, L, a, lane, 64);
    if (L.state == ST_DONE) break;
    if (L.need_score) {
      const ArenaView av = arena_view(L.ab, a.logcap, true);
      for (int r = 1 + lane; r < 1 + 2 * L.nlog_prev; r += 64)
        av.hot[r].score = score_dev_scalar(a.scorer, av.feat + (int64_t)r * NF);
      mem_sync();
      continue;
    }
    if (lane == 0) g_rounds++;
    round_lane(g, L, lane, 64, a);
    if (g.rarena) mem_sync();
    else lane_sync();
  }
}

template <int K, int CAP>
static int run(const KArgs& a) {
  std::vector<unsigned char> arena(a.logcap > 0 ? arena_bytes(a.logcap, a.feats != 0) : 1);
  std::unique_ptr<Wave<K, CAP>> gp(new Wave<K, CAP>());
  memset(gp.get(), 0, sizeof(Wave<K, CAP>));
  g_rounds = 0;
  memset(g_kind, 0, sizeof(g_kind));
  g_scount = 0;
  SimtCtx<K, CAP> ctx{&a, gp.get(), a.logcap > 0 ? arena.data() : nullptr};
  return simt::launch(simt_body<K, CAP>, &ctx) ? 10 : 0;
}
#else
template <int K, int CAP>
static int run(const KArgs& a) {
  std::vector<unsigned char> arena(a.logcap > 0 ? arena_bytes(a.logcap, a.feats != 0) : 1);
  using W = Wave<K, CAP>;
  std::unique_ptr<W> gp(new W());
  W& g = *gp;
  memset(&g, 0, sizeof(W));
  Leader L;
  memset(&L, 0, sizeof(L));
  L.state = ST_FETCH;
  L.ab = a.logcap > 0 ? arena.data() : nullptr;
  g_rounds = 0;
  memset(g_kind, 0, sizeof(g_kind));
  g_scount = 0;
  for (;;) {
    leader_run(g, L, a, 0, 1);
    if (L.state == ST_DONE) break;
    if (L.need_score) {
      const ArenaView av = arena_view(L.ab, a.logcap, true);
      for (int r = 1; r < 1 + 2 * L.nlog_prev; r++)
        av.hot[r].score = score_dev_scalar(a.scorer, av.feat + (int64_t)r * NF);
      continue;
    }
    g_rounds++;
    round_lane(g, L, 0, 1, a);
  }
  return 0;
}
#endif

extern "C" int emu_ccd_batch(const double* tris1, int64_t n1, const double* tris2, int64_t n2,
                             const c2a_query* q, int64_t nq, const c2a_params* p, c2a_result* out,
                             int K) {
  Bvh b1, b2;
  if (build_bvh(tris1, nullptr, nullptr, n1, &b1) || build_bvh(tris2, nullptr, nullptr, n2, &b2))
    return 1;
  MeshDev tab[2];
  tab[0].nodes = b1.nodes.data(); tab[0].tris = b1.tris.data();
  tab[0].ntri = (int32_t)n1; tab[0].depth = b1.max_depth;
  tab[1].nodes = b2.nodes.data(); tab[1].tris = b2.tris.data();
  tab[1].ntri = (int32_t)n2; tab[1].depth = b2.max_depth;
  std::vector<QueryDev> hq((size_t)nq);
  for (int64_t i = 0; i < nq; i++) {
    for (int k = 0; k < 2; k++) {
      hq[i].mesh[k] = q[i].mesh[k];
      hq[i].seed[k] = q[i].seed_tri[k] < 0 ? 0 : q[i].seed_tri[k];
      double R0[9], T0[3], R1[9], T1[3];
      memcpy(R0, q[i].pose[2 * k], 72); memcpy(T0, q[i].pose[2 * k] + 9, 24);
      memcpy(R1, q[i].pose[2 * k + 1], 72); memcpy(T1, q[i].pose[2 * k + 1] + 9, 24);
      motion_init(&hq[i].mo[k], R0, T0, R1, T1, p->strict != 0);
    }
  }
  int64_t counter = 0;
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.meshes = tab;
  a.q = hq.data();
  a.out = out;
  a.counter = &counter;
  a.nq = nq;
  a.tol_d = p->tol_d;
  a.tol_t = p->tol_t;
  a.contact_scale = p->contact_scale;
  a.contact_offset = p->contact_offset;
  a.max_iter = p->max_iter;
  a.with_contacts = p->with_contacts;
  a.mode = MODE_QUERY;
  a.logcap = p->replay_log;
  if (a.logcap < 0) {   // automatic capacity, as the library picks it
    a.logcap = 8192;
    while (a.logcap < std::max(n1, n2) / 8 && a.logcap < (1 << 20)) a.logcap <<= 1;
  }
  a.arena_rounds = emu_arena_rounds;
  a.miss_depth = p->miss_depth;
  a.ext_levels = p->ext_levels < 0 ? 0 : p->ext_levels;
  a.ext_adaptive = emu_ext_adaptive;
  a.strict = p->strict != 0;
  a.translation = p->translation && !a.strict;
  a.scorer = g_scorer;
  if (!g_keys.empty() && a.logcap > 0) {
    a.sample_thr = (uint64_t)((g_rate < 0 ? 0 : g_rate > 1 ? 1 : g_rate) * 4294967296.0);
    a.samples = g_samples.data();
    a.sample_keys = g_keys.data();
    a.sample_cap = (int64_t)g_keys.size();
    a.sample_count = &g_scount;
  }
  a.feats = (a.scorer.on || a.sample_cap > 0) ? 1 : 0;
  // stack capacity as the library picks it (the deepest one at depth 5)
  const int need = b1.max_depth + b2.max_depth + 2;
  const bool small = need <= 48;
  if (need > 96) K = 0;
  auto run_shape = [&](const KArgs& ka) {
    if (K == 0) return need <= 192 ? run<5, 192>(ka) : 2;
    switch (K) {
      case 1: return small ? run<1, 48>(ka) : run<1, 96>(ka);
      case 2: return small ? run<2, 48>(ka) : run<2, 96>(ka);
      case 3: return small ? run<3, 48>(ka) : run<3, 96>(ka);
      case 4: return small ? run<4, 48>(ka) : run<4, 96>(ka);
      case 5: return small ? run<5, 48>(ka) : run<5, 96>(ka);
      default: return small ? run<6, 48>(ka) : run<6, 96>(ka);
    }
  };
  emu_tail_handed = 0;
  if (emu_tail_budget <= 0) return run_shape(a);
  // two runs, as the library's two launches
  std::vector<TailRec> rec((size_t)nq);
  std::vector<int32_t> list((size_t)nq + 1);
  int32_t count = 0;
  a.tail_budget = emu_tail_budget;
  a.tail_rec = rec.data();
  a.tail_list = list.data();
  a.tail_count = &count;
  int rc = run_shape(a);
  if (rc) return rc;
  emu_tail_handed = count;
  counter = 0;
  a.tail_pass = 1;
  return run_shape(a);
}

// Static distance queries (MODE_DISTANCE) between two meshes, one lane.
extern "C" int emu_distance(const double* tris1, int64_t n1, const double* tris2, int64_t n2,
                            const double* poses, const int32_t* seeds, double rel_err,
                            double abs_err, int64_t nq, c2a_distance_result* out,
                            int overlap_only) {
  Bvh b1, b2;
  if (build_bvh(tris1, nullptr, nullptr, n1, &b1) || build_bvh(tris2, nullptr, nullptr, n2, &b2))
    return 1;
  MeshDev tab[2];
  tab[0].nodes = b1.nodes.data(); tab[0].tris = b1.tris.data();
  tab[0].ntri = (int32_t)n1; tab[0].depth = b1.max_depth;
  tab[1].nodes = b2.nodes.data(); tab[1].tris = b2.tris.data();
  tab[1].ntri = (int32_t)n2; tab[1].depth = b2.max_depth;
  std::vector<ContactDev> hc((size_t)nq);
  for (int64_t i = 0; i < nq; i++) {
    ContactDev& c = hc[(size_t)i];
    memset(&c, 0, sizeof(c));
    c.mesh[0] = 0;
    c.mesh[1] = 1;
    c.seed[0] = seeds ? seeds[2 * i] : 0;
    c.seed[1] = seeds ? seeds[2 * i + 1] : 0;
    memcpy(c.R1, poses + 24 * i, 72); memcpy(c.T1, poses + 24 * i + 9, 24);
    memcpy(c.R2, poses + 24 * i + 12, 72); memcpy(c.T2, poses + 24 * i + 21, 24);
    c.rel_err = rel_err;
    c.abs_err = abs_err;
    c.overlap_only = overlap_only;
    if (overlap_only &&
        root_rss_distance(b1.nodes.data(), b2.nodes.data(), c.R1, c.T1, c.R2, c.T2) != 0.0)
      c.overlap_only = 2;
  }
  int64_t counter = 0;
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.meshes = tab;
  a.contact_in = hc.data();
  a.dist_out = out;
  a.counter = &counter;
  a.nq = nq;
  a.mode = MODE_DISTANCE;
  const bool small = b1.max_depth + b2.max_depth + 2 <= 48;
  return small ? run<5, 48>(a) : run<5, 96>(a);
}

// Static queries of static_query.h (one thread's code, run query by query)
// between two meshes; pairs_out as c2a_collide_pairs.
#include "../../c2a_amd/csrc/static_query.h"
extern "C" int emu_static(const double* tris1, int64_t n1, const double* tris2, int64_t n2,
                          int kind, const double* poses, const int32_t* seeds,
                          const double* params, double rel_err, double abs_err, int qsize,
                          int flag, int64_t nq, int pair_cap, int32_t* pairs_out,
                          c2a::StaticOut* out) {
  Bvh b1, b2;
  if (build_bvh(tris1, nullptr, nullptr, n1, &b1) || build_bvh(tris2, nullptr, nullptr, n2, &b2))
    return 1;
  MeshDev tab[2];
  tab[0].nodes = b1.nodes.data(); tab[0].tris = b1.tris.data();
  tab[0].ntri = (int32_t)n1; tab[0].depth = b1.max_depth;
  tab[1].nodes = b2.nodes.data(); tab[1].tris = b2.tris.data();
  tab[1].ntri = (int32_t)n2; tab[1].depth = b2.max_depth;
  std::vector<StaticQueryDev> hq((size_t)nq);
  for (int64_t i = 0; i < nq; i++) {
    StaticQueryDev& q = hq[(size_t)i];
    memset(&q, 0, sizeof(q));
    q.mesh[0] = 0;
    q.mesh[1] = 1;
    q.seed[0] = seeds ? seeds[2 * i] : 0;
    q.seed[1] = seeds ? seeds[2 * i + 1] : 0;
    memcpy(q.R1, poses + 24 * i, 72); memcpy(q.T1, poses + 24 * i + 9, 24);
    memcpy(q.R2, poses + 24 * i + 12, 72); memcpy(q.T2, poses + 24 * i + 21, 24);
    q.param = params ? params[i] : 0.0;
  }
  const int64_t per = sq_ws_entries(kind, qsize, b1.max_depth + b2.max_depth);
  std::vector<SqEntry> ws((size_t)per);
  StaticArgs a;
  memset(&a, 0, sizeof(a));
  a.meshes = tab;
  a.q = hq.data();
  a.out = out;
  a.pairs = pairs_out;
  a.ws = ws.data();
  a.nq = nq;
  a.ws_entries = per;
  a.kind = kind;
  a.qsize = qsize;
  a.flag = flag;
  a.pair_cap = pair_cap;
  a.rel_err = rel_err;
  a.abs_err = abs_err;
  for (int64_t i = 0; i < nq; i++) static_query(a, i, ws.data());
  return 0;
}
