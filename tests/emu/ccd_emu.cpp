// ccd_emu.cpp -- TEST HARNESS: runs the GPU kernel's per-group logic
// (c2a_amd/csrc/ccd_core.h) on the CPU, one group and one lane at a time, so
// the walk / speculation logic can be checked against the oracle without a
// GPU.  The product library never links this file.
#include <string.h>

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
#define C2A_EMU_COUNT(kind, L) (g_kind[(L).numCA > 1 ? 1 : 0][kind]++)
#include "../../c2a_amd/csrc/bvh.h"
#include "../../c2a_amd/csrc/ccd_core.h"

using namespace c2a;

static int64_t g_rounds = 0;   // speculative rounds of the last batch
extern "C" int64_t emu_rounds() { return g_rounds; }
static int emu_arena_rounds = 1;   // the CPU harness exercises the arena-round path
extern "C" void emu_set_arena_rounds(int on) { emu_arena_rounds = on; }
extern "C" int64_t emu_round_kind(int later, int kind) { return g_kind[later][kind]; }

template <int K, int CAP>
static int run(const KArgs& a) {
  std::vector<unsigned char> arena(a.logcap > 0 ? arena_bytes(a.logcap) : 1);
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
  for (;;) {
    leader_run(g, L, a, 0, 1);
    if (L.state == ST_DONE) break;
    g_rounds++;
    round_lane(g, L, 0, 1, a);
  }
  return 0;
}

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
      motion_init(&hq[i].mo[k], R0, T0, R1, T1);
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
  a.arena_rounds = emu_arena_rounds;
  // stack capacity as the library picks it
  const bool small = b1.max_depth + b2.max_depth + 2 <= 48;
  switch (K) {
    case 1: return small ? run<1, 48>(a) : run<1, 96>(a);
    case 2: return small ? run<2, 48>(a) : run<2, 96>(a);
    case 3: return small ? run<3, 48>(a) : run<3, 96>(a);
    case 4: return small ? run<4, 48>(a) : run<4, 96>(a);
    case 5: return small ? run<5, 48>(a) : run<5, 96>(a);
    default: return small ? run<6, 48>(a) : run<6, 96>(a);
  }
}
