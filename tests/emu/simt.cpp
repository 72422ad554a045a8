// simt.cpp -- TEST HARNESS: 64 lanes of one wavefront on the CPU.
//
// Each lane runs the kernel's per-lane code on its own fiber (ucontext).
// The wavefront primitives of ccd_core.h (lane_sync, mem_sync, ballot,
// shuffles, cross-lane reductions) are barriers: a lane that reaches one
// parks; when every live lane has parked, the barrier's results are
// computed from all lanes' inputs and the lanes run on, in lane order, to
// their next barrier.  Between barriers the lanes therefore run one after
// another, not in lockstep, so code that passes data between lanes through
// memory without a barrier reads stale values here; lanes that reach
// different primitives, or finish while others wait, are reported.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ucontext.h>

namespace c2a {
namespace simt {

namespace {

constexpr int NL = 64;
constexpr size_t kStack = 8u << 20;
enum Kind : int { NONE = 0, SYNC, BALLOT, BCAST, BCASTD, PMAX, MIND, MAXI, LANE0, DONE };
const char* const kName[] = {"none", "sync", "ballot", "bcast", "bcastd",
                             "excl_prefix_max", "mind", "maxi", "lane0", "done"};

struct Lane {
  ucontext_t ctx;
  char* stack = nullptr;
  int kind = NONE;
  int64_t iv = 0, oi = 0;
  double dv = 0.0, od = 0.0;
  int src = 0;
};

Lane lanes[NL];
ucontext_t sched;
int cur = -1;
void (*g_body)(int, void*) = nullptr;
void* g_arg = nullptr;
int64_t g_barriers = 0;

void arrive(int k) {
  Lane& l = lanes[cur];
  l.kind = k;
  swapcontext(&l.ctx, &sched);
}

void trampoline(int lane) {
  g_body(lane, g_arg);
  lanes[lane].kind = DONE;
}

void resolve(int k) {
  g_barriers++;
  switch (k) {
    case BALLOT: {
      uint64_t m = 0;
      for (int i = 0; i < NL; i++)
        if (lanes[i].iv) m |= 1ull << i;
      for (int i = 0; i < NL; i++) lanes[i].oi = (int64_t)m;
      break;
    }
    case BCAST:
      for (int i = 0; i < NL; i++) lanes[i].oi = lanes[lanes[i].src & (NL - 1)].iv;
      break;
    case BCASTD:
      for (int i = 0; i < NL; i++) lanes[i].od = lanes[lanes[i].src & (NL - 1)].dv;
      break;
    case PMAX: {
      int64_t run = 0;
      for (int i = 0; i < NL; i++) {
        const int64_t v = lanes[i].iv;
        lanes[i].oi = i == 0 ? 0 : run;
        run = (i == 0 || v > run) ? v : run;
      }
      break;
    }
    case MIND: {
      double m = lanes[0].dv;
      for (int i = 1; i < NL; i++) m = lanes[i].dv < m ? lanes[i].dv : m;
      for (int i = 0; i < NL; i++) lanes[i].od = m;
      break;
    }
    case MAXI: {
      int64_t m = lanes[0].iv;
      for (int i = 1; i < NL; i++) m = lanes[i].iv > m ? lanes[i].iv : m;
      for (int i = 0; i < NL; i++) lanes[i].oi = m;
      break;
    }
    case LANE0:
      for (int i = 0; i < NL; i++) lanes[i].oi = lanes[0].iv;
      break;
    default:
      break;
  }
  for (int i = 0; i < NL; i++) lanes[i].kind = NONE;
}

}  // namespace

int lane_id() { return cur; }
int64_t barriers() { return g_barriers; }

void sync() { arrive(SYNC); }
uint64_t ballot(bool p) {
  lanes[cur].iv = p ? 1 : 0;
  arrive(BALLOT);
  return (uint64_t)lanes[cur].oi;
}
int bcast(int v, int src) {
  lanes[cur].iv = v;
  lanes[cur].src = src;
  arrive(BCAST);
  return (int)lanes[cur].oi;
}
double bcastd(double v, int src) {
  lanes[cur].dv = v;
  lanes[cur].src = src;
  arrive(BCASTD);
  return lanes[cur].od;
}
int excl_prefix_max(int v) {
  lanes[cur].iv = v;
  arrive(PMAX);
  return (int)lanes[cur].oi;
}
double mind(double v) {
  lanes[cur].dv = v;
  arrive(MIND);
  return lanes[cur].od;
}
int maxi(int v) {
  lanes[cur].iv = v;
  arrive(MAXI);
  return (int)lanes[cur].oi;
}
long long lane0(long long v) {
  lanes[cur].iv = v;
  arrive(LANE0);
  return (long long)lanes[cur].oi;
}

// Runs body(lane, arg) on 64 lanes; 0 when every lane finished together,
// else the lanes diverged (the message names the primitives).
int launch(void (*body)(int, void*), void* arg) {
  g_body = body;
  g_arg = arg;
  g_barriers = 0;
  for (int i = 0; i < NL; i++) {
    Lane& l = lanes[i];
    if (!l.stack) l.stack = (char*)malloc(kStack);
    if (!l.stack) return 3;
    getcontext(&l.ctx);
    l.ctx.uc_stack.ss_sp = l.stack;
    l.ctx.uc_stack.ss_size = kStack;
    l.ctx.uc_link = &sched;
    makecontext(&l.ctx, (void (*)())trampoline, 1, i);
    l.kind = NONE;
  }
  for (;;) {
    for (int i = 0; i < NL; i++) {
      if (lanes[i].kind != NONE) continue;
      cur = i;
      swapcontext(&sched, &lanes[i].ctx);
    }
    cur = -1;
    int ndone = 0, k = -1;
    for (int i = 0; i < NL; i++) {
      if (lanes[i].kind == DONE) {
        ndone++;
        continue;
      }
      if (k < 0) k = lanes[i].kind;
      else if (lanes[i].kind != k) {
        fprintf(stderr, "simt: lane %d at %s while lane 0.. wait at %s (barrier %lld)\n", i,
                kName[lanes[i].kind], kName[k], (long long)g_barriers);
        return 1;
      }
    }
    if (ndone == NL) return 0;
    if (ndone > 0) {
      fprintf(stderr, "simt: %d lanes finished while %d wait at %s (barrier %lld)\n", ndone,
              NL - ndone, kName[k], (long long)g_barriers);
      return 2;
    }
    resolve(k);
  }
}

}  // namespace simt
}  // namespace c2a
