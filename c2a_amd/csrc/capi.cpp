// capi.cpp -- the extern "C" boundary (include/c2a_amd.h).
//
// Host responsibilities: build the RSS hierarchy (bvh.cpp), precompute each
// query's two linear motions (InterpMotion.cpp:148-270; acos runs here),
// upload, pick the kernel shape, launch on the caller's stream and copy
// results back.  No exception crosses this file's entry points.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/c2a_amd.h"
#include "bvh.h"
#include "c2a_math.h"
#include "ccd_core.h"
#include "ccd_kernel.h"
#include "scorer.h"
#include "static_kernel.h"
#include "bvh_gpu.h"
#include "tuning_table.h"

namespace c2a {
ScorerDev scorer_view(const c2a_scorer* s);   // scorer.hip
}

using namespace c2a;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return C2A_ERR_HIP;
}

#define HIP_TRY(expr)                                  \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_fail(_e, #expr);  \
  } while (0)

int check_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) return fail(C2A_ERR_NO_DEVICE, "no HIP device available");
  return C2A_OK;
}

// A device allocation owned by the library.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    if (n <= bytes && p) return C2A_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, n ? n : 1);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    bytes = n;
    return C2A_OK;
  }
};

void unpack_pose(const double* p12, double* R, double* T) {
  for (int i = 0; i < 9; i++) R[i] = p12[i];
  for (int i = 0; i < 3; i++) T[i] = p12[9 + i];
}

}  // namespace

namespace c2a {
// Status with the thread's c2a_last_error text (used by scorer.hip too).
int set_error(int code, const char* msg) { return fail(code, msg); }
}  // namespace c2a

struct c2a_mesh {
  Bvh host;
  DevBuf nodes, tris;
  int64_t ntri = 0;
  uint64_t src_digest = 0;   // the triangles, ids and vertex indices it was made from
};

struct c2a_batch {
  std::vector<MeshDev> mtab;
  DevBuf d_mesh, d_q, d_out, d_counter, d_arena, d_samples, d_keys, d_scount;
  DevBuf d_tail, d_tlist, d_tctr;         // tail pass: saved states, hand-overs, counters
  int64_t nq = 0;
  KArgs args{};
  KArgs args2{};                         // the tail pass's arguments (tail = true)
  bool tail = false;
  int blocks2 = 0;
  int K = 5, CAP = 48, NW = 4, blocks = 1;
  int logcap = 0;
  const c2a_scorer* scorer = nullptr;   // learned order
  int64_t sample_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  hipGraphExec_t graph = nullptr;        // captured launch sequence (c2a_batch_run_graph)
  KArgs tune{};                          // the autotuned bucket knobs (ntune, tune_*)
  ~c2a_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

namespace {

// Uploads a mesh's hierarchy and triangles the first time a launch needs it.
int ensure_uploaded(c2a_mesh* m) {
  if (m->nodes.p) return C2A_OK;
  size_t nb = m->host.nodes.size() * sizeof(GpuNode), tb = m->host.tris.size() * sizeof(double);
  int rc;
  if ((rc = m->nodes.alloc(nb)) || (rc = m->tris.alloc(tb))) return rc;
  HIP_TRY(hipMemcpy(m->nodes.p, m->host.nodes.data(), nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(m->tris.p, m->host.tris.data(), tb, hipMemcpyHostToDevice));
  return C2A_OK;
}

int make_mesh_table(c2a_mesh* const* meshes, int32_t nmesh, std::vector<MeshDev>& tab) {
  if (!meshes || nmesh <= 0) return fail(C2A_ERR_BAD_ARG, "empty mesh table");
  tab.resize((size_t)nmesh);
  for (int i = 0; i < nmesh; i++) {
    if (!meshes[i]) return fail(C2A_ERR_BAD_ARG, "null mesh in table");
    int rc = ensure_uploaded(meshes[i]);
    if (rc) return rc;
    tab[(size_t)i].nodes = (const GpuNode*)meshes[i]->nodes.p;
    tab[(size_t)i].tris = (const double*)meshes[i]->tris.p;
    tab[(size_t)i].ntri = (int32_t)meshes[i]->ntri;
    tab[(size_t)i].depth = meshes[i]->host.max_depth;
  }
  return C2A_OK;
}

// Traversal stack capacity: the DFS holds at most one pending pair per BVTT
// level plus the current one, and the BVTT is at most depth1+depth2 deep.
int pick_cap(const std::vector<MeshDev>& tab, const int32_t* pairs, int64_t n, int stride) {
  int need = 0;
  for (int64_t i = 0; i < n; i++) {
    int a = pairs[i * stride], b = pairs[i * stride + 1];
    need = std::max(need, tab[(size_t)a].depth + tab[(size_t)b].depth + 2);
  }
  if (need <= 48) return 48;
  if (need <= 96) return 96;
  if (need <= 192) return 192;
  return -1;
}

int pick_shape(const c2a_params* p, int& K, int& NW) {
  K = (p && p->spec_depth > 0) ? p->spec_depth : 5;
  NW = (p && p->waves_per_query > 0) ? p->waves_per_query : 1;
  const bool learned = p && p->order == C2A_ORDER_LEARNED;
  if (p && p->order != C2A_ORDER_PLAIN && !learned)
    return fail(C2A_ERR_BAD_ARG, "order must be C2A_ORDER_PLAIN or C2A_ORDER_LEARNED");
  if (learned && !p->scorer) return fail(C2A_ERR_BAD_ARG, "learned order needs a scorer");
  if (K < 1 || K > 6) return fail(C2A_ERR_BAD_ARG, "spec_depth must be 1..6");
  if (NW != 1 || !kernel_shape_ok(K, 48, NW, false))
    return fail(C2A_ERR_BAD_ARG, "waves_per_query must be 0 or 1");
  if (learned && !kernel_shape_ok(K, 48, NW, true))
    return fail(C2A_ERR_BAD_ARG, "learned order runs at spec_depth 0 or 5");
  return C2A_OK;
}

// Replay log capacity from the largest mesh: a CA step whose expansions
// overflow the log walks the rest sequentially at the next step.  The
// heaviest steps measured expand about ntri / 20 pairs (11.5k at 262,144
// triangles), so the capacity is the next power of two of ntri / 8, at
// least 8192 (arena ≈ 1 KB per expansion per resident workgroup).
int auto_replay_log(const std::vector<MeshDev>& tab) {
  int64_t big = 0;
  for (const MeshDev& m : tab) big = std::max<int64_t>(big, m.ntri);
  int cap = 8192;
  while (cap < big / 8 && cap < (1 << 20)) cap <<= 1;
  return cap;
}

// Resident workgroups: the kernel is persistent, so launching more than fit
// at once only adds workgroups that find the queue empty.
int grid_blocks(int K, int CAP, int NW, int64_t nq, bool learned = false) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  int cus = 256;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  int per_cu = kernel_blocks_per_cu(K, CAP, NW, learned);
  if (per_cu < 1) per_cu = 1;
  const int64_t cap = (int64_t)cus * per_cu;
  return (int)std::max<int64_t>(1, std::min<int64_t>(nq, cap));
}

int prepare_queries(const std::vector<MeshDev>& tab, const c2a_query* q, int64_t nq,
                    bool strict, std::vector<QueryDev>& out, std::vector<int32_t>& pairs) {
  out.resize((size_t)nq);
  pairs.resize((size_t)nq * 2);
  for (int64_t i = 0; i < nq; i++) {
    const c2a_query& x = q[i];
    QueryDev& d = out[(size_t)i];
    for (int k = 0; k < 2; k++) {
      int m = x.mesh[k];
      if (m < 0 || m >= (int)tab.size()) return fail(C2A_ERR_BAD_ARG, "query mesh index out of range");
      int s = x.seed_tri[k] < 0 ? 0 : x.seed_tri[k];
      if (s >= tab[(size_t)m].ntri) return fail(C2A_ERR_BAD_ARG, "seed triangle out of range");
      d.mesh[k] = m;
      d.seed[k] = s;
      pairs[(size_t)i * 2 + k] = m;
      double R0[9], T0[3], R1[9], T1[3];
      unpack_pose(x.pose[2 * k], R0, T0);
      unpack_pose(x.pose[2 * k + 1], R1, T1);
      motion_init(&d.mo[k], R0, T0, R1, T1, strict);
    }
  }
  return C2A_OK;
}

// Kernel arguments that come from the query parameters.
void params_to_args(KArgs& a, const c2a_params* p) {
  a.tol_d = p->tol_d;
  a.tol_t = p->tol_t;
  a.contact_scale = p->contact_scale;
  a.contact_offset = p->contact_offset;
  a.max_iter = p->max_iter;
  a.with_contacts = p->with_contacts;
  a.mode = MODE_QUERY;
  a.miss_depth = p->miss_depth < 0 ? 0 : p->miss_depth;
  a.ext_levels = p->ext_levels < 0 ? 0 : (p->ext_levels > 8 ? 8 : p->ext_levels);
  a.ext_adaptive = 1;   // switched per step by the previous step's misses
  a.strict = p->strict != 0 ? 1 : 0;
  a.translation = (p->translation != 0 && !a.strict) ? 1 : 0;
}

// The autotuned per-size-bucket knobs (tuning_table.h, generated from
// c2a_amd/tuning.json by the autotuner): the template depth is the deepest
// among the buckets the batch's queries use (5 for the learned order, its one
// instantiated depth), and the replay arena is sized for their largest log
// (an explicit replay_log >= 0 or ext_levels >= 0 overrides every bucket's).
// The extension levels are tuned for the plain order only: under the learned
// order, and when the batch logs training samples (which are drawn from the
// replayed records), they stay at 0 unless given.
void apply_tuning(c2a_batch* b, const std::vector<int32_t>& pairs, int replay_log,
                  int ext_levels) {
  if (kTuneRows <= 0) return;
  const int rows = std::min(kTuneRows, C2A_TUNE_MAX);
  KArgs& t = b->tune;
  t.ntune = rows;
  for (int r = 0; r < rows; r++) {
    t.tune_ntri[r] = kTune[r][0];
    t.tune_depth[r] = std::min(6, std::max(1, kTune[r][1]));
    t.tune_miss[r] = std::max(0, kTune[r][2]);
    t.tune_log[r] = replay_log >= 0 ? replay_log : kTune[r][3];
    t.tune_ext[r] = ext_levels >= 0 ? std::min(8, ext_levels)
                                    : ((b->scorer || b->sample_cap > 0)
                                           ? 0
                                           : std::min(8, std::max(0, kTune[r][4])));
  }
  // the kernel's depth and arena are sized for the buckets the batch uses
  int kmax = 1, logmax = 0;
  for (size_t i = 0; i + 1 < pairs.size(); i += 2) {
    const int big = std::max(b->mtab[(size_t)pairs[i]].ntri, b->mtab[(size_t)pairs[i + 1]].ntri);
    int r = 0;
    while (r + 1 < rows && big > t.tune_ntri[r]) r++;
    kmax = std::max(kmax, t.tune_depth[r]);
    logmax = std::max(logmax, t.tune_log[r]);
  }
  b->K = b->scorer ? 5 : kmax;
  if (replay_log < 0) b->logcap = logmax;
}

int fill_args(c2a_batch* b, const c2a_params* p) {
  c2a_params def;
  c2a_params_default(&def);
  if (!p) p = &def;
  KArgs& a = b->args;
  memset(&a, 0, sizeof(a));
  a.meshes = (const MeshDev*)b->d_mesh.p;
  a.q = (const QueryDev*)b->d_q.p;
  a.out = (c2a_result*)b->d_out.p;
  a.counter = (int64_t*)b->d_counter.p;
  a.nq = b->nq;
  params_to_args(a, p);
  a.ntune = b->tune.ntune;
  for (int r = 0; r < C2A_TUNE_MAX; r++) {
    a.tune_ntri[r] = b->tune.tune_ntri[r];
    a.tune_depth[r] = b->tune.tune_depth[r];
    a.tune_miss[r] = b->tune.tune_miss[r];
    a.tune_log[r] = b->tune.tune_log[r];
    a.tune_ext[r] = b->tune.tune_ext[r];
  }
  a.logcap = b->logcap;
  a.arena = (unsigned char*)b->d_arena.p;
  a.feats = (b->scorer || b->sample_cap > 0) ? 1 : 0;
  if (b->sample_cap > 0) {
    const double r = p->sample_rate < 0 ? 0.0 : (p->sample_rate > 1 ? 1.0 : p->sample_rate);
    a.sample_thr = (uint64_t)(r * 4294967296.0);
    a.samples = (float*)b->d_samples.p;
    a.sample_keys = (int64_t*)b->d_keys.p;
    a.sample_cap = b->sample_cap;
    a.sample_count = (unsigned long long*)b->d_scount.p;
  }
  return C2A_OK;
}

// The tail pass's kernel shape.
constexpr int kTailK = 5;
constexpr int kTailWaves = 4;

__global__ void zero_counter(int64_t* c) { *c = 0; }
__global__ void zero_tail(int64_t* c, int32_t* n) {
  *c = 0;
  *n = 0;
}

int run_kernel(const KArgs& a, int K, int CAP, int NW, int blocks, hipStream_t s) {
  // counter cleared by a kernel (graph-capture safe, no memset node)
  hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, s, a.counter);
  HIP_TRY(hipGetLastError());
  hipError_t e = launch_ccd(a, K, CAP, NW, blocks, s);
  if (e != hipSuccess) return hip_fail(e, "launch_ccd");
  return C2A_OK;
}

// A batch's launches: the CCD kernel, then (tail pass) the second kernel
// over the queries the first one handed over, four wavefronts per query.
int run_batch(const c2a_batch* b, hipStream_t s) {
  if (b->tail) {
    hipLaunchKernelGGL(zero_tail, dim3(1), dim3(1), 0, s, b->args2.counter, b->args.tail_count);
    HIP_TRY(hipGetLastError());
  }
  int rc = run_kernel(b->args, b->K, b->CAP, b->NW, b->blocks, s);
  if (rc || !b->tail) return rc;
  const hipError_t e = launch_ccd(b->args2, kTailK, b->CAP, kTailWaves, b->blocks2, s);
  if (e != hipSuccess) return hip_fail(e, "launch_ccd (tail pass)");
  return C2A_OK;
}

// Sets up the tail pass when the batch can take it: plain order, no sample
// logging, a stack the four-wavefront shapes cover.
int setup_tail(c2a_batch* b, const c2a_params* p) {
  const int64_t budget = p->tail_budget;
  if (budget <= 0 || b->nq < 2 || b->scorer || b->sample_cap > 0 ||
      !kernel_shape_ok(kTailK, b->CAP, kTailWaves, false))
    return C2A_OK;
  int rc;
  if ((rc = b->d_tail.alloc((size_t)b->nq * sizeof(TailRec))) ||
      (rc = b->d_tlist.alloc((size_t)b->nq * sizeof(int32_t))) ||
      (rc = b->d_tctr.alloc(16)))
    return rc;
  KArgs& a = b->args;
  a.tail_budget = budget;
  a.tail_rec = (TailRec*)b->d_tail.p;
  a.tail_list = (int32_t*)b->d_tlist.p;
  a.tail_count = (int32_t*)((char*)b->d_tctr.p + 8);
  b->args2 = a;
  b->args2.tail_pass = 1;
  b->args2.counter = (int64_t*)b->d_tctr.p;
  // one replay arena per workgroup: no more workgroups than the first launch
  b->blocks2 = std::min(grid_blocks(kTailK, b->CAP, kTailWaves, b->nq), b->blocks);
  b->tail = true;
  return C2A_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free batch (c2a_ccd_batch_noalloc).  Its device memory beyond
// the caller's inputs and outputs -- the work-queue counter, the mesh table
// and the replay arenas -- is this file-scope array, one copy per device,
// held until the process ends.  Host code never passes the array's name to
// the device: each device's address comes from hipGetSymbolAddress, called
// once per device with that device current.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchBytes = (size_t)C2A_NOALLOC_SCRATCH_MIB << 20;
constexpr size_t kScratchTable = 256;    // mesh table offset (the counter is at 0)
constexpr size_t kScratchArena = 1024;   // first replay arena
constexpr int kMaxDevices = 64;
static_assert(kScratchTable + C2A_NOALLOC_MAX_MESHES * sizeof(MeshDev) <= kScratchArena,
              "mesh table overlaps the arenas");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct NoallocTable {
  MeshDev m[C2A_NOALLOC_MAX_MESHES];
};

// The mesh table travels by value as a kernel argument (no host buffer has
// to outlive the call) and is written into the scratch by this kernel.
__global__ void write_mesh_table(MeshDev* dst, NoallocTable t, int n) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = t.m[i];
}

std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxDevices];   // each device's scratch address

// The current device's scratch address, looked up on the first call there.
int device_scratch(unsigned char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(C2A_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = (unsigned char*)p;
  }
  *out = g_scratch_dev[dev];
  return C2A_OK;
}

}  // namespace

extern "C" {

int c2a_ccd_batch_noalloc(c2a_mesh* const* meshes, int32_t nmesh, const double* motions,
                          const int32_t* mesh_pairs, const int32_t* seed_tris, int64_t nq,
                          const c2a_params* params, c2a_result* out, void* stream) {
  if (nq < 0) return fail(C2A_ERR_BAD_ARG, "nq must be >= 0");
  if (nq == 0) return C2A_OK;
  if (!motions || !mesh_pairs || !out) return fail(C2A_ERR_BAD_ARG, "null device array");
  if (!meshes || nmesh <= 0) return fail(C2A_ERR_BAD_ARG, "empty mesh table");
  if (nmesh > C2A_NOALLOC_MAX_MESHES)
    return fail(C2A_ERR_BAD_ARG, "the allocation-free batch takes at most " +
                                     std::to_string(C2A_NOALLOC_MAX_MESHES) + " meshes");
  c2a_params def;
  c2a_params_default(&def);
  const c2a_params* p = params ? params : &def;
  if (p->order != C2A_ORDER_PLAIN)
    return fail(C2A_ERR_BAD_ARG, "the allocation-free batch runs the plain order only");
  if (p->sample_cap != 0)
    return fail(C2A_ERR_BAD_ARG, "the allocation-free batch logs no training samples");
  if (p->strict != 0 && p->strict != 1) return fail(C2A_ERR_BAD_ARG, "strict must be 0 or 1");
  int K = 5, NW = 1, rc;
  if ((rc = pick_shape(p, K, NW))) return rc;
  if ((rc = check_device())) return rc;
  NoallocTable tab;
  memset(&tab, 0, sizeof(tab));
  int maxd = 0;
  for (int i = 0; i < nmesh; i++) {
    const c2a_mesh* m = meshes[i];
    if (!m) return fail(C2A_ERR_BAD_ARG, "null mesh in table");
    if (!m->nodes.p || !m->tris.p)
      return fail(C2A_ERR_BAD_ARG, "mesh " + std::to_string(i) +
                                       " is not on the device yet: run it once through "
                                       "c2a_ccd_batch or c2a_batch_create first");
    tab.m[i].nodes = (const GpuNode*)m->nodes.p;
    tab.m[i].tris = (const double*)m->tris.p;
    tab.m[i].ntri = (int32_t)m->ntri;
    tab.m[i].depth = m->host.max_depth;
    maxd = std::max(maxd, m->host.max_depth);
  }
  // any two meshes of the table may meet: stack capacity from the deepest
  const int need = 2 * maxd + 2;
  const int CAP = need <= 48 ? 48 : (need <= 96 ? 96 : (need <= 192 ? 192 : -1));
  if (CAP < 0) return fail(C2A_ERR_UNSUPPORTED, "hierarchies too deep for the traversal stack");
  if (CAP == 192) K = 5;   // the deepest stack's one depth (no result depends on it)
  const int logcap = p->replay_log == -1 ? 1024 : p->replay_log;
  if (logcap < 0) return fail(C2A_ERR_BAD_ARG, "replay_log must be -1 (1024) or >= 0");
  const size_t room = kScratchBytes - kScratchArena;
  const size_t slice = logcap > 0 ? arena_bytes(logcap) : 0;
  if (slice > room)
    return fail(C2A_ERR_BAD_ARG, "replay_log " + std::to_string(logcap) + " needs a " +
                                     std::to_string(slice) + "-byte arena; the scratch holds " +
                                     std::to_string(room) + " bytes");
  unsigned char* sp = nullptr;
  if ((rc = device_scratch(&sp))) return rc;
  int blocks = grid_blocks(K, CAP, NW, nq);
  if (slice) blocks = (int)std::min<size_t>((size_t)blocks, room / slice);
  KArgs a;
  memset(&a, 0, sizeof(a));
  params_to_args(a, p);
  a.meshes = (const MeshDev*)(sp + kScratchTable);
  a.qmot = motions;
  a.qpairs = mesh_pairs;
  a.qseeds = seed_tris;
  a.nmesh = nmesh;
  a.out = out;
  a.counter = (int64_t*)sp;
  a.nq = nq;
  a.logcap = logcap;
  a.arena = logcap > 0 ? sp + kScratchArena : nullptr;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(write_mesh_table, dim3(1), dim3(C2A_NOALLOC_MAX_MESHES), 0, s,
                     (MeshDev*)(sp + kScratchTable), tab, nmesh);
  HIP_TRY(hipGetLastError());
  return run_kernel(a, K, CAP, NW, blocks, s);
}

void c2a_params_default(c2a_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->tol_d = 1e-4;
  p->tol_t = 1e-4;
  p->max_iter = 150;
  p->with_contacts = 1;
  p->contact_scale = 2.0;
  p->contact_offset = 0.001;
  p->spec_depth = 0;
  p->waves_per_query = 0;
  p->replay_log = -1;
  p->order = C2A_ORDER_PLAIN;
  p->scorer = nullptr;
  p->sample_rate = 0.0;
  p->sample_cap = 0;
  p->miss_depth = 2;
  p->ext_levels = -1;
  p->translation = 1;
  p->strict = 0;
  p->tail_waves = 0;
  p->tail_budget = 0;
}

namespace {

// 64-bit digest of a byte range, eight bytes at a time (splitmix64 of the
// running state xor each word; the tail is zero-padded).
uint64_t digest_bytes(uint64_t h, const void* data, size_t n) {
  const unsigned char* p = (const unsigned char*)data;
  auto mix = [](uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
  };
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    memcpy(&w, p + i, 8);
    h = mix(h ^ w);
  }
  if (i < n) {
    uint64_t w = 0;
    memcpy(&w, p + i, n - i);
    h = mix(h ^ w);
  }
  return mix(h ^ (uint64_t)n);
}

uint64_t source_digest(const double* xyz9, const int32_t* ids, const int32_t* vidx3, int64_t n) {
  uint64_t h = digest_bytes(0x43324142u, xyz9, (size_t)n * 9 * sizeof(double));
  h = ids ? digest_bytes(h, ids, (size_t)n * 4) : digest_bytes(h, "noids", 5);
  h = vidx3 ? digest_bytes(h, vidx3, (size_t)n * 12) : digest_bytes(h, "novidx", 6);
  return h;
}

uint64_t payload_digest(const Bvh& b) {
  uint64_t h = digest_bytes(0x50415931u, b.nodes.data(), b.nodes.size() * sizeof(GpuNode));
  h = digest_bytes(h, b.tris.data(), b.tris.size() * sizeof(double));
  h = digest_bytes(h, b.tri_id.data(), b.tri_id.size() * sizeof(int32_t));
  h = digest_bytes(h, b.tri_vidx.data(), b.tri_vidx.size() * sizeof(int32_t));
  return digest_bytes(h, b.src_index.data(), b.src_index.size() * sizeof(int32_t));
}

const char kCacheMagic[8] = {'C', '2', 'A', 'B', 'V', 'H', 0, 1};
constexpr uint32_t kCacheVersion = 2;   // 2: the source index of each triangle
struct CacheHeader {
  char magic[8];
  uint32_t version, node_bytes;
  int64_t ntri;
  int32_t max_depth, pad;
  uint64_t src_digest, payload_digest;
};
static_assert(sizeof(CacheHeader) == 48, "cache header layout");

int check_mesh_input(const double* tri_xyz9, int64_t ntri) {
  if (ntri <= 0) return fail(C2A_ERR_EMPTY_MODEL, "model has no triangles");
  if (!tri_xyz9) return fail(C2A_ERR_BAD_ARG, "null triangle array");
  if (ntri > (1 << 28)) return fail(C2A_ERR_BAD_ARG, "too many triangles");
  return C2A_OK;
}

}  // namespace

int c2a_mesh_create(const double* tri_xyz9, const int32_t* tri_ids, const int32_t* vert_idx3,
                    int64_t ntri, c2a_mesh** out) {
  if (!out) return fail(C2A_ERR_BAD_ARG, "null output handle");
  *out = nullptr;
  const int rc = check_mesh_input(tri_xyz9, ntri);
  if (rc) return rc;
  c2a_mesh* m = new (std::nothrow) c2a_mesh();
  if (!m) return fail(C2A_ERR_OOM, "host allocation failed");
  try {
    if (build_bvh(tri_xyz9, tri_ids, vert_idx3, ntri, &m->host) != 0) {
      delete m;
      return fail(C2A_ERR_EMPTY_MODEL, "model has no triangles");
    }
  } catch (const std::bad_alloc&) {
    delete m;
    return fail(C2A_ERR_OOM, "host allocation failed during the build");
  }
  m->ntri = ntri;
  m->src_digest = source_digest(tri_xyz9, tri_ids, vert_idx3, ntri);
  *out = m;
  return C2A_OK;
}

int c2a_mesh_create_device(const double* tri_xyz9, const int32_t* tri_ids,
                           const int32_t* vert_idx3, int64_t ntri, void* stream,
                           c2a_mesh** out) {
  if (!out) return fail(C2A_ERR_BAD_ARG, "null output handle");
  *out = nullptr;
  int rc = check_mesh_input(tri_xyz9, ntri);
  if (rc) return rc;
  if ((rc = check_device())) return rc;
  c2a_mesh* m = new (std::nothrow) c2a_mesh();
  if (!m) return fail(C2A_ERR_OOM, "host allocation failed");
  void *nodes = nullptr, *tris = nullptr;
  hipError_t e;
  try {
    e = build_bvh_device(tri_xyz9, tri_ids, vert_idx3, ntri, &m->host, &nodes, &tris,
                         (hipStream_t)stream);
  } catch (const std::bad_alloc&) {
    delete m;
    return fail(C2A_ERR_OOM, "host allocation failed during the build");
  }
  if (e != hipSuccess) {
    delete m;
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? fail(C2A_ERR_OOM, "device memory exhausted by the build")
                                    : hip_fail(e, "device hierarchy build");
  }
  m->nodes.p = nodes;
  m->nodes.bytes = m->host.nodes.size() * sizeof(GpuNode);
  m->tris.p = tris;
  m->tris.bytes = m->host.tris.size() * sizeof(double);
  m->ntri = ntri;
  m->src_digest = source_digest(tri_xyz9, tri_ids, vert_idx3, ntri);
  *out = m;
  return C2A_OK;
}

int c2a_mesh_save(const c2a_mesh* m, const char* path) {
  if (!m || !path) return fail(C2A_ERR_BAD_ARG, "null argument");
  CacheHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, kCacheMagic, 8);
  h.version = kCacheVersion;
  h.node_bytes = (uint32_t)sizeof(GpuNode);
  h.ntri = m->ntri;
  h.max_depth = m->host.max_depth;
  h.src_digest = m->src_digest;
  h.payload_digest = payload_digest(m->host);
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
    if (!f) return fail(C2A_ERR_BAD_ARG, "cannot write " + tmp);
    f.write((const char*)&h, sizeof(h));
    f.write((const char*)m->host.nodes.data(),
            (std::streamsize)(m->host.nodes.size() * sizeof(GpuNode)));
    f.write((const char*)m->host.tris.data(),
            (std::streamsize)(m->host.tris.size() * sizeof(double)));
    f.write((const char*)m->host.tri_id.data(), (std::streamsize)(m->host.tri_id.size() * 4));
    f.write((const char*)m->host.tri_vidx.data(), (std::streamsize)(m->host.tri_vidx.size() * 4));
    f.write((const char*)m->host.src_index.data(), (std::streamsize)(m->host.src_index.size() * 4));
    if (!f) return fail(C2A_ERR_BAD_ARG, "write failed: " + tmp);
  }
  // written under a temporary name, then moved: readers never see half a file
  if (rename(tmp.c_str(), path) != 0)
    return fail(C2A_ERR_BAD_ARG, std::string("cannot rename to ") + path);
  return C2A_OK;
}

int c2a_mesh_load(const char* path, const double* tri_xyz9, const int32_t* tri_ids,
                  const int32_t* vert_idx3, int64_t ntri, c2a_mesh** out) {
  if (!out || !path) return fail(C2A_ERR_BAD_ARG, "null argument");
  *out = nullptr;
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(C2A_ERR_CACHE, std::string("cannot open ") + path);
  CacheHeader h;
  if (!f.read((char*)&h, sizeof(h)) || memcmp(h.magic, kCacheMagic, 8) != 0 || h.version != kCacheVersion ||
      h.node_bytes != sizeof(GpuNode) || h.ntri <= 0 || h.ntri > (1 << 28))
    return fail(C2A_ERR_CACHE, std::string("not a hierarchy file: ") + path);
  if (tri_xyz9 &&
      (ntri != h.ntri || source_digest(tri_xyz9, tri_ids, vert_idx3, ntri) != h.src_digest))
    return fail(C2A_ERR_CACHE, std::string("hierarchy file built from other triangles: ") + path);
  c2a_mesh* m = new (std::nothrow) c2a_mesh();
  if (!m) return fail(C2A_ERR_OOM, "host allocation failed");
  const int64_t n = h.ntri;
  try {
    m->host.nodes.resize((size_t)(2 * n - 1));
    m->host.tris.resize((size_t)n * 9);
    m->host.tri_id.resize((size_t)n);
    m->host.tri_vidx.resize((size_t)n * 3);
    m->host.src_index.resize((size_t)n);
  } catch (const std::bad_alloc&) {
    delete m;
    return fail(C2A_ERR_OOM, "host allocation failed");
  }
  f.read((char*)m->host.nodes.data(), (std::streamsize)(m->host.nodes.size() * sizeof(GpuNode)));
  f.read((char*)m->host.tris.data(), (std::streamsize)(m->host.tris.size() * sizeof(double)));
  f.read((char*)m->host.tri_id.data(), (std::streamsize)(m->host.tri_id.size() * 4));
  f.read((char*)m->host.tri_vidx.data(), (std::streamsize)(m->host.tri_vidx.size() * 4));
  f.read((char*)m->host.src_index.data(), (std::streamsize)(m->host.src_index.size() * 4));
  bool ok = (bool)f;
  char extra;
  if (ok && f.read(&extra, 1)) ok = false;   // trailing bytes: not this format
  if (!ok || payload_digest(m->host) != h.payload_digest) {
    delete m;
    return fail(C2A_ERR_CACHE, std::string("damaged hierarchy file: ") + path);
  }
  m->host.max_depth = h.max_depth;
  m->ntri = n;
  m->src_digest = h.src_digest;
  *out = m;
  return C2A_OK;
}

int c2a_mesh_refit(c2a_mesh* m, const double* tri_xyz9, int32_t on_device, void* stream) {
  if (!m || !tri_xyz9) return fail(C2A_ERR_BAD_ARG, "null argument");
  if ((int64_t)m->host.src_index.size() != m->ntri)
    return fail(C2A_ERR_BAD_ARG, "mesh has no source order");
  int rc;
  if (on_device) {
    if ((rc = check_device()) || (rc = ensure_uploaded(m))) return rc;
    hipError_t e;
    try {
      e = refit_bvh_device(tri_xyz9, &m->host, m->nodes.p, m->tris.p, (hipStream_t)stream);
    } catch (const std::bad_alloc&) {
      return fail(C2A_ERR_OOM, "host allocation failed during the refit");
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return e == hipErrorOutOfMemory ? fail(C2A_ERR_OOM, "device memory exhausted by the refit")
                                      : hip_fail(e, "device refit");
    }
  } else {
    if (refit_bvh(&m->host, tri_xyz9) != 0) return fail(C2A_ERR_BAD_ARG, "refit failed");
    if (m->nodes.p) {   // keep the device copy current
      HIP_TRY(hipMemcpy(m->nodes.p, m->host.nodes.data(), m->host.nodes.size() * sizeof(GpuNode),
                        hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(m->tris.p, m->host.tris.data(), m->host.tris.size() * sizeof(double),
                        hipMemcpyHostToDevice));
    }
  }
  // a refitted hierarchy is not the build of these triangles: its cache
  // file matches no build input
  m->src_digest = digest_bytes(m->src_digest ^ 0x5245464954ull, tri_xyz9,
                               (size_t)m->ntri * 9 * sizeof(double));
  return C2A_OK;
}

int c2a_mesh_destroy(c2a_mesh* m) {
  delete m;
  return C2A_OK;
}

int c2a_mesh_info(const c2a_mesh* m, int64_t* ntri, int64_t* nbv, int32_t* max_depth) {
  if (!m) return fail(C2A_ERR_BAD_ARG, "null mesh");
  if (ntri) *ntri = m->ntri;
  if (nbv) *nbv = (int64_t)m->host.nodes.size();
  if (max_depth) *max_depth = m->host.max_depth;
  return C2A_OK;
}

int c2a_mesh_tri_ids(const c2a_mesh* m, int32_t* ids_out) {
  if (!m || !ids_out) return fail(C2A_ERR_BAD_ARG, "null argument");
  memcpy(ids_out, m->host.tri_id.data(), m->host.tri_id.size() * sizeof(int32_t));
  return C2A_OK;
}

int c2a_mesh_export(const c2a_mesh* m, double* nodes32, double* tris9) {
  if (!m) return fail(C2A_ERR_BAD_ARG, "null mesh");
  if (nodes32) {
    for (size_t i = 0; i < m->host.nodes.size(); i++) {
      const GpuNode& n = m->host.nodes[i];
      double* o = nodes32 + 32 * i;
      for (int k = 0; k < 9; k++) o[k] = n.R[k];
      for (int k = 0; k < 3; k++) o[9 + k] = n.Tr[k];
      o[12] = n.l[0];
      o[13] = n.l[1];
      o[14] = n.r;
      o[15] = n.angR;
      for (int k = 0; k < 9; k++) o[16 + k] = n.Rloc[k];
      o[25] = n.size;
      o[26] = (double)n.first_child;
      o[27] = (double)n.depth;
      o[28] = o[29] = o[30] = o[31] = 0.0;
    }
  }
  if (tris9) memcpy(tris9, m->host.tris.data(), m->host.tris.size() * sizeof(double));
  return C2A_OK;
}

int c2a_batch_create(c2a_mesh* const* meshes, int32_t nmesh, const c2a_query* q, int64_t nq,
                     const c2a_params* params, c2a_batch** out) {
  if (!out) return fail(C2A_ERR_BAD_ARG, "null output handle");
  *out = nullptr;
  if (nq < 0 || (nq > 0 && !q)) return fail(C2A_ERR_BAD_ARG, "bad query array");
  int rc = check_device();
  if (rc) return rc;
  c2a_batch* b = new (std::nothrow) c2a_batch();
  if (!b) return fail(C2A_ERR_OOM, "host allocation failed");
  std::vector<QueryDev> hq;
  std::vector<int32_t> pairs;
  if (params && params->strict != 0 && params->strict != 1) {
    delete b;
    return fail(C2A_ERR_BAD_ARG, "strict must be 0 or 1");
  }
  if (params && params->tail_waves != 0 && params->tail_waves != kTailWaves) {
    delete b;
    return fail(C2A_ERR_BAD_ARG, "tail_waves must be 0 or 4");
  }
  try {
    if ((rc = make_mesh_table(meshes, nmesh, b->mtab)) ||
        (rc = prepare_queries(b->mtab, q, nq, params && params->strict, hq, pairs))) {
      delete b;
      return rc;
    }
  } catch (const std::bad_alloc&) {
    delete b;
    return fail(C2A_ERR_OOM, "host allocation failed");
  }
  b->nq = nq;
  if ((rc = pick_shape(params, b->K, b->NW))) {
    delete b;
    return rc;
  }
  b->CAP = nq ? pick_cap(b->mtab, pairs.data(), nq, 2) : 48;
  if (b->CAP < 0) {
    delete b;
    return fail(C2A_ERR_UNSUPPORTED, "hierarchies too deep for the traversal stack");
  }
  {
    c2a_params def;
    c2a_params_default(&def);
    const c2a_params* pp = params ? params : &def;
    if (pp->replay_log < -1 || pp->replay_log > (1 << 24)) {
      delete b;
      return fail(C2A_ERR_BAD_ARG, "replay_log must be -1 (auto) or 0..2^24");
    }
    if (pp->sample_cap < 0 || (pp->sample_cap > 0 && pp->replay_log == 0)) {
      delete b;
      return fail(C2A_ERR_BAD_ARG, "sample_cap must be >= 0 and needs replay_log != 0");
    }
    b->logcap = pp->replay_log >= 0 ? pp->replay_log : auto_replay_log(b->mtab);
    b->scorer = pp->order == C2A_ORDER_LEARNED ? pp->scorer : nullptr;
    b->sample_cap = pp->sample_cap;
    // spec_depth 0: the autotuned knobs of each query's size bucket
    if (pp->spec_depth == 0) apply_tuning(b, pairs, pp->replay_log, pp->ext_levels);
  }
  // the deepest stack is instantiated at one depth; the depth changes no result
  if (b->CAP == 192) b->K = 5;
  b->blocks = grid_blocks(b->K, b->CAP, b->NW, nq, b->scorer != nullptr);
  const bool feats = b->scorer || b->sample_cap > 0;
  if (b->logcap > 0 && nq > 0) {
    // one replay arena per resident workgroup: fewer workgroups when the
    // device has no room for all of them
    const size_t slice = arena_bytes(b->logcap, feats);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0) {
      const size_t fit = (size_t)(0.8 * (double)free_b) / slice;
      if (fit < (size_t)b->blocks) b->blocks = (int)std::max<size_t>(1, fit);
    }
    while ((rc = b->d_arena.alloc((size_t)b->blocks * slice)) && b->blocks > 1) {
      (void)hipGetLastError();
      b->blocks /= 2;
    }
    if (rc) {
      delete b;
      return rc;
    }
  }
  if ((b->sample_cap > 0 &&
       ((rc = b->d_samples.alloc((size_t)b->sample_cap * (NF + 1) * sizeof(float))) ||
        (rc = b->d_keys.alloc((size_t)b->sample_cap * sizeof(int64_t))) ||
        (rc = b->d_scount.alloc(sizeof(int64_t))))) ||
      (rc = b->d_mesh.alloc(b->mtab.size() * sizeof(MeshDev))) ||
      (rc = b->d_q.alloc((size_t)std::max<int64_t>(nq, 1) * sizeof(QueryDev))) ||
      (rc = b->d_out.alloc((size_t)std::max<int64_t>(nq, 1) * sizeof(c2a_result))) ||
      (rc = b->d_counter.alloc(sizeof(int64_t)))) {
    delete b;
    return rc;
  }
  hipError_t e = hipMemcpy(b->d_mesh.p, b->mtab.data(), b->mtab.size() * sizeof(MeshDev),
                           hipMemcpyHostToDevice);
  if (e == hipSuccess && nq)
    e = hipMemcpy(b->d_q.p, hq.data(), (size_t)nq * sizeof(QueryDev), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e != hipSuccess) {
    delete b;
    return hip_fail(e, "batch upload");
  }
  fill_args(b, params);
  {
    c2a_params def;
    c2a_params_default(&def);
    if ((rc = setup_tail(b, params ? params : &def))) {
      delete b;
      return rc;
    }
  }
  *out = b;
  return C2A_OK;
}

int c2a_batch_run(c2a_batch* b, void* stream) {
  if (!b) return fail(C2A_ERR_BAD_ARG, "null batch");
  if (b->nq == 0) return C2A_OK;
  hipStream_t s = (hipStream_t)stream;
  if (b->scorer) b->args.scorer = scorer_view(b->scorer);   // current parameters
  if (b->sample_cap > 0) {
    hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, s, (int64_t*)b->d_scount.p);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(b->ev0, s));
  int rc = run_batch(b, s);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(b->ev1, s));
  b->timed = true;
  return C2A_OK;
}

int c2a_batch_run_graph(c2a_batch* b, void* stream) {
  if (!b) return fail(C2A_ERR_BAD_ARG, "null batch");
  if (b->nq == 0) return C2A_OK;
  if (b->scorer || b->sample_cap > 0)
    return fail(C2A_ERR_BAD_ARG,
                "graph replay takes batches without a scorer or sample logging");
  hipStream_t s = (hipStream_t)stream;
  if (!b->graph) {
    // capture the launch sequence (counter reset by a kernel, then the CCD
    // kernel: no memset node) on a private stream, once per batch
    hipStream_t cs = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = C2A_OK;
    if (e == hipSuccess) {
      rc = run_batch(b, cs);
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(cs, &g);
      if (!rc && e2 != hipSuccess) rc = hip_fail(e2, "hipStreamEndCapture");
      if (!rc) {
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
          b->graph = nullptr;
          rc = hip_fail(e, "hipGraphInstantiate");
        }
      }
      if (g) (void)hipGraphDestroy(g);
    } else {
      rc = hip_fail(e, "hipStreamBeginCapture");
    }
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(b->ev0, s));
  HIP_TRY(hipGraphLaunch(b->graph, s));
  HIP_TRY(hipEventRecord(b->ev1, s));
  b->timed = true;
  return C2A_OK;
}

int c2a_batch_tail_count(c2a_batch* b, int64_t* n) {
  if (!b || !n) return fail(C2A_ERR_BAD_ARG, "null argument");
  *n = 0;
  if (!b->tail) return C2A_OK;
  int32_t v = 0;
  HIP_TRY(hipMemcpy(&v, b->args.tail_count, sizeof(v), hipMemcpyDeviceToHost));
  *n = v;
  return C2A_OK;
}

int c2a_batch_fetch(c2a_batch* b, c2a_result* out, void* stream) {
  if (!b || (!out && b->nq)) return fail(C2A_ERR_BAD_ARG, "null argument");
  if (b->nq == 0) return C2A_OK;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpyAsync(out, b->d_out.p, (size_t)b->nq * sizeof(c2a_result),
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return C2A_OK;
}

int c2a_batch_last_ms(c2a_batch* b, float* ms) {
  if (!b || !ms) return fail(C2A_ERR_BAD_ARG, "null argument");
  if (!b->timed) return fail(C2A_ERR_BAD_ARG, "batch has not run");
  HIP_TRY(hipEventSynchronize(b->ev1));
  HIP_TRY(hipEventElapsedTime(ms, b->ev0, b->ev1));
  return C2A_OK;
}

int c2a_batch_samples(c2a_batch* b, float* feats, float* labels, int64_t* keys, int64_t cap,
                      int64_t* n_out, int64_t* n_written) {
  if (!b || cap < 0 || (cap > 0 && (!feats || !labels || !keys)))
    return fail(C2A_ERR_BAD_ARG, "bad sample buffers");
  if (b->sample_cap <= 0) return fail(C2A_ERR_BAD_ARG, "the batch logs no samples");
  unsigned long long cnt = 0;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(&cnt, b->d_scount.p, sizeof(cnt), hipMemcpyDeviceToHost));
  const int64_t have = std::min<int64_t>((int64_t)cnt, b->sample_cap);
  std::vector<float> rows;
  std::vector<int64_t> k;
  try {
    rows.resize((size_t)have * (NF + 1));
    k.resize((size_t)have);
  } catch (const std::bad_alloc&) {
    return fail(C2A_ERR_OOM, "host allocation failed");
  }
  if (have > 0) {
    HIP_TRY(hipMemcpy(rows.data(), b->d_samples.p, rows.size() * sizeof(float),
                      hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(k.data(), b->d_keys.p, k.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  }
  std::vector<int64_t> idx((size_t)have);
  for (int64_t i = 0; i < have; i++) idx[(size_t)i] = i;
  std::sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return k[(size_t)x] < k[(size_t)y]; });
  const int64_t w = std::min(have, cap);
  for (int64_t i = 0; i < w; i++) {
    const int64_t j = idx[(size_t)i];
    memcpy(feats + i * NF, &rows[(size_t)j * (NF + 1)], NF * sizeof(float));
    labels[i] = rows[(size_t)j * (NF + 1) + NF];
    keys[i] = k[(size_t)j];
  }
  if (n_out) *n_out = (int64_t)cnt;
  if (n_written) *n_written = w;
  return C2A_OK;
}

int c2a_batch_destroy(c2a_batch* b) {
  delete b;
  return C2A_OK;
}

int c2a_ccd_batch(c2a_mesh* const* meshes, int32_t nmesh, const c2a_query* q, int64_t nq,
                  const c2a_params* params, c2a_result* out, void* stream) {
  if (nq > 0 && !out) return fail(C2A_ERR_BAD_ARG, "null result array");
  c2a_batch* b = nullptr;
  int rc = c2a_batch_create(meshes, nmesh, q, nq, params, &b);
  if (rc) return rc;
  rc = c2a_batch_run(b, stream);
  if (!rc) rc = c2a_batch_fetch(b, out, stream);
  c2a_batch_destroy(b);
  return rc;
}

int c2a_ccd_step(c2a_mesh* const* meshes, int32_t nmesh, const c2a_step_query* q, int64_t nq,
                 c2a_step_result* out, void* stream) {
  if (nq < 0 || (nq > 0 && (!q || !out))) return fail(C2A_ERR_BAD_ARG, "bad arrays");
  if (nq == 0) return C2A_OK;
  int rc = check_device();
  if (rc) return rc;
  std::vector<MeshDev> tab;
  if ((rc = make_mesh_table(meshes, nmesh, tab))) return rc;
  std::vector<StepDev> hs((size_t)nq);
  std::vector<int32_t> pairs((size_t)nq * 2);
  for (int64_t i = 0; i < nq; i++) {
    const c2a_step_query& x = q[i];
    StepDev& d = hs[(size_t)i];
    memset(&d, 0, sizeof(d));
    for (int k = 0; k < 2; k++) {
      int m = x.mesh[k];
      if (m < 0 || m >= nmesh) return fail(C2A_ERR_BAD_ARG, "mesh index out of range");
      int s = x.seed_tri[k] < 0 ? 0 : x.seed_tri[k];
      if (s >= tab[(size_t)m].ntri) return fail(C2A_ERR_BAD_ARG, "seed triangle out of range");
      d.mesh[k] = m;
      d.seed[k] = s;
      const int l = x.last_tri[k];
      if (l >= tab[(size_t)m].ntri) return fail(C2A_ERR_BAD_ARG, "last triangle out of range");
      d.last[k] = l < 0 ? s : l;
      pairs[(size_t)i * 2 + k] = m;
      double R0[9], T0[3], R1[9], T1[3];
      unpack_pose(x.pose_s[k], R0, T0);
      unpack_pose(x.pose_e[k], R1, T1);
      motion_init(&d.mo[k], R0, T0, R1, T1);
    }
    memcpy(d.R1, x.R1, sizeof(d.R1));
    memcpy(d.T1, x.T1, sizeof(d.T1));
    memcpy(d.R2, x.R2, sizeof(d.R2));
    memcpy(d.T2, x.T2, sizeof(d.T2));
    d.numCA = x.numCA;
    d.mint_prev = x.mint_prev;
    d.UpboundTOC = x.UpboundTOC;
  }
  int CAP = pick_cap(tab, pairs.data(), nq, 2);
  if (CAP < 0) return fail(C2A_ERR_UNSUPPORTED, "hierarchies too deep for the traversal stack");
  const int K = 5;
  DevBuf dm, din, dout, dc;
  if ((rc = dm.alloc(tab.size() * sizeof(MeshDev))) || (rc = din.alloc((size_t)nq * sizeof(StepDev))) ||
      (rc = dout.alloc((size_t)nq * sizeof(c2a_step_result))) || (rc = dc.alloc(sizeof(int64_t))))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpy(dm.p, tab.data(), tab.size() * sizeof(MeshDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(din.p, hs.data(), (size_t)nq * sizeof(StepDev), hipMemcpyHostToDevice));
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.meshes = (const MeshDev*)dm.p;
  a.step_in = (const StepDev*)din.p;
  a.step_out = (c2a_step_result*)dout.p;
  a.counter = (int64_t*)dc.p;
  a.nq = nq;
  a.mode = MODE_STEP;
  a.max_iter = 150;
  if ((rc = run_kernel(a, K, CAP, 1, grid_blocks(K, CAP, 1, nq), s))) return rc;
  HIP_TRY(hipMemcpyAsync(out, dout.p, (size_t)nq * sizeof(c2a_step_result), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return C2A_OK;
}

static int contact_queries(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                           const int32_t* mesh_pairs, const double* thresholds, int64_t nq,
                           int32_t cap, c2a_contact* out, int32_t* counts_out, void* stream) {
  if (nq < 0 || (nq > 0 && (!poses || !mesh_pairs || !thresholds || !counts_out)))
    return fail(C2A_ERR_BAD_ARG, "bad arrays");
  if (cap < 0 || (cap > 0 && !out)) return fail(C2A_ERR_BAD_ARG, "bad contact buffer");
  if (nq == 0) return C2A_OK;
  int rc = check_device();
  if (rc) return rc;
  std::vector<MeshDev> tab;
  if ((rc = make_mesh_table(meshes, nmesh, tab))) return rc;
  std::vector<ContactDev> hc((size_t)nq);
  for (int64_t i = 0; i < nq; i++) {
    ContactDev& c = hc[(size_t)i];
    memset(&c, 0, sizeof(c));
    for (int k = 0; k < 2; k++) {
      int m = mesh_pairs[2 * i + k];
      if (m < 0 || m >= nmesh) return fail(C2A_ERR_BAD_ARG, "mesh index out of range");
      c.mesh[k] = m;
    }
    unpack_pose(poses + 24 * i, c.R1, c.T1);
    unpack_pose(poses + 24 * i + 12, c.R2, c.T2);
    c.threshold = thresholds[i];
  }
  int CAP = pick_cap(tab, mesh_pairs, nq, 2);
  if (CAP < 0) return fail(C2A_ERR_UNSUPPORTED, "hierarchies too deep for the traversal stack");
  const int K = 5;
  DevBuf dm, din, dout, dc, drec;
  if ((rc = dm.alloc(tab.size() * sizeof(MeshDev))) ||
      (rc = din.alloc((size_t)nq * sizeof(ContactDev))) ||
      (rc = dout.alloc((size_t)nq * sizeof(int32_t))) || (rc = dc.alloc(sizeof(int64_t))) ||
      (cap > 0 && (rc = drec.alloc((size_t)nq * cap * sizeof(c2a_contact)))))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpy(dm.p, tab.data(), tab.size() * sizeof(MeshDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(din.p, hc.data(), (size_t)nq * sizeof(ContactDev), hipMemcpyHostToDevice));
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.meshes = (const MeshDev*)dm.p;
  a.contact_in = (const ContactDev*)din.p;
  a.counts_out = (int32_t*)dout.p;
  a.contacts = cap > 0 ? (c2a_contact*)drec.p : nullptr;
  a.contact_cap = cap;
  a.counter = (int64_t*)dc.p;
  a.nq = nq;
  a.mode = MODE_CONTACT;
  if ((rc = run_kernel(a, K, CAP, 1, grid_blocks(K, CAP, 1, nq), s))) return rc;
  HIP_TRY(hipMemcpyAsync(counts_out, dout.p, (size_t)nq * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (cap > 0)
    HIP_TRY(hipMemcpyAsync(out, drec.p, (size_t)nq * cap * sizeof(c2a_contact),
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  // feature slots -> the caller's vertex ids (AddTri i1..i3, build order),
  // as C2A.cpp:1565-1604 reads them from the triangles' index triples
  for (int64_t i = 0; cap > 0 && i < nq; i++) {
    const int n = std::min<int>(counts_out[i], cap);
    for (int k = 0; k < n; k++) {
      c2a_contact& c = out[i * (int64_t)cap + k];
      for (int side = 0; side < 2; side++) {
        const Bvh& h = meshes[mesh_pairs[2 * i + side]]->host;
        const int32_t* v = &h.tri_vidx[3 * (size_t)c.tri[side]];
        const int f = c.fid[side][0];
        for (int j = 0; j < 3; j++) c.fid[side][j] = -1;
        if (c.type[side] == 1) c.fid[side][0] = v[f];
        else if (c.type[side] == 2) {
          c.fid[side][0] = v[f];
          c.fid[side][1] = v[(f + 1) % 3];
        } else if (c.type[side] == 3) {
          for (int j = 0; j < 3; j++) c.fid[side][j] = v[j];
        }
      }
    }
  }
  return C2A_OK;
}

int c2a_contact_count(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                      const int32_t* mesh_pairs, const double* thresholds, int64_t nq,
                      int32_t* counts_out, void* stream) {
  return contact_queries(meshes, nmesh, poses, mesh_pairs, thresholds, nq, 0, nullptr,
                         counts_out, stream);
}

int c2a_contact_list(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                     const int32_t* mesh_pairs, const double* thresholds, int64_t nq,
                     int32_t cap, c2a_contact* out, int32_t* counts_out, void* stream) {
  return contact_queries(meshes, nmesh, poses, mesh_pairs, thresholds, nq, cap, out, counts_out,
                         stream);
}

static int distance_queries(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                            const int32_t* mesh_pairs, const int32_t* seed_tris, double rel_err,
                            double abs_err, int64_t nq, c2a_distance_result* out, void* stream,
                            int overlap_only) {
  if (nq < 0 || (nq > 0 && (!poses || !mesh_pairs || !out)))
    return fail(C2A_ERR_BAD_ARG, "bad arrays");
  if (!(rel_err >= 0) || !(abs_err >= 0))
    return fail(C2A_ERR_BAD_ARG, "error bounds must be >= 0");
  if (nq == 0) return C2A_OK;
  int rc = check_device();
  if (rc) return rc;
  std::vector<MeshDev> tab;
  if ((rc = make_mesh_table(meshes, nmesh, tab))) return rc;
  std::vector<ContactDev> hc((size_t)nq);
  for (int64_t i = 0; i < nq; i++) {
    ContactDev& c = hc[(size_t)i];
    memset(&c, 0, sizeof(c));
    for (int k = 0; k < 2; k++) {
      int m = mesh_pairs[2 * i + k];
      if (m < 0 || m >= nmesh) return fail(C2A_ERR_BAD_ARG, "mesh index out of range");
      c.mesh[k] = m;
      int t = seed_tris ? seed_tris[2 * i + k] : 0;
      if (t < 0) t = 0;
      if (t >= tab[(size_t)m].ntri) return fail(C2A_ERR_BAD_ARG, "seed triangle out of range");
      c.seed[k] = t;
    }
    unpack_pose(poses + 24 * i, c.R1, c.T1);
    unpack_pose(poses + 24 * i + 12, c.R2, c.T2);
    c.rel_err = rel_err;
    c.abs_err = abs_err;
    c.overlap_only = overlap_only;
    // C2A_Collide's root overlap test (C2A_PQP.cpp:1068), once per query
    if (overlap_only &&
        root_rss_distance(meshes[c.mesh[0]]->host.nodes.data(),
                          meshes[c.mesh[1]]->host.nodes.data(), c.R1, c.T1, c.R2,
                          c.T2) != 0.0)
      c.overlap_only = 2;
  }
  int CAP = pick_cap(tab, mesh_pairs, nq, 2);
  if (CAP < 0) return fail(C2A_ERR_UNSUPPORTED, "hierarchies too deep for the traversal stack");
  const int K = 5;
  DevBuf dm, din, dout, dc;
  if ((rc = dm.alloc(tab.size() * sizeof(MeshDev))) ||
      (rc = din.alloc((size_t)nq * sizeof(ContactDev))) ||
      (rc = dout.alloc((size_t)nq * sizeof(c2a_distance_result))) ||
      (rc = dc.alloc(sizeof(int64_t))))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpy(dm.p, tab.data(), tab.size() * sizeof(MeshDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(din.p, hc.data(), (size_t)nq * sizeof(ContactDev), hipMemcpyHostToDevice));
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.meshes = (const MeshDev*)dm.p;
  a.contact_in = (const ContactDev*)din.p;
  a.dist_out = (c2a_distance_result*)dout.p;
  a.counter = (int64_t*)dc.p;
  a.nq = nq;
  a.mode = MODE_DISTANCE;
  if ((rc = run_kernel(a, K, CAP, 1, grid_blocks(K, CAP, 1, nq), s))) return rc;
  HIP_TRY(hipMemcpyAsync(out, dout.p, (size_t)nq * sizeof(c2a_distance_result),
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return C2A_OK;
}

int c2a_distance(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                 const int32_t* mesh_pairs, const int32_t* seed_tris, double rel_err,
                 double abs_err, int64_t nq, c2a_distance_result* out, void* stream) {
  return distance_queries(meshes, nmesh, poses, mesh_pairs, seed_tris, rel_err, abs_err, nq, out,
                          stream, 0);
}

int c2a_collide_distance(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                         const int32_t* mesh_pairs, const int32_t* seed_tris, double rel_err,
                         double abs_err, int64_t nq, c2a_distance_result* out, void* stream) {
  return distance_queries(meshes, nmesh, poses, mesh_pairs, seed_tris, rel_err, abs_err, nq, out,
                          stream, 1);
}

namespace {

static_assert(sizeof(StaticOut) == sizeof(c2a_static_result), "static result layout");

// The static queries of static_query.h: one thread per query over a
// workspace of per-thread stacks / queues sized for the deepest pair of
// hierarchies, with fewer threads when that would not fit in a quarter of
// the free device memory.
int static_queries(c2a_mesh* const* meshes, int32_t nmesh, int32_t kind, const double* poses,
                   const int32_t* mesh_pairs, const int32_t* seed_tris, const double* params,
                   double rel_err, double abs_err, int32_t qsize, int32_t flag, int64_t nq,
                   int32_t pair_cap, int32_t* pairs_out, void* out, void* stream) {
  if (nq < 0 || (nq > 0 && (!poses || !mesh_pairs || !out)))
    return fail(C2A_ERR_BAD_ARG, "bad arrays");
  if (nq > 0 && (kind == SQ_DISTANCE_CS || kind == SQ_TOLERANCE) && !params)
    return fail(C2A_ERR_BAD_ARG, "null per-query bounds");
  if (kind == SQ_COLLIDE && (pair_cap < 0 || (pair_cap > 0 && nq > 0 && !pairs_out)))
    return fail(C2A_ERR_BAD_ARG, "bad pair list");
  if (kind == SQ_COLLIDE && flag != 1 && flag != 2)
    return fail(C2A_ERR_BAD_ARG, "flag must be 1 (all contacts) or 2 (first contact)");
  if (qsize > (1 << 20)) return fail(C2A_ERR_BAD_ARG, "qsize above 2^20");
  if (kind == SQ_DISTANCE_QUEUE && !(rel_err >= 0 && abs_err >= 0))
    return fail(C2A_ERR_BAD_ARG, "error bounds must be >= 0");
  if (nq == 0) return C2A_OK;
  int rc = check_device();
  if (rc) return rc;
  std::vector<MeshDev> tab;
  if ((rc = make_mesh_table(meshes, nmesh, tab))) return rc;
  std::vector<StaticQueryDev> hq((size_t)nq);
  int depth_sum = 0;
  for (int64_t i = 0; i < nq; i++) {
    StaticQueryDev& q = hq[(size_t)i];
    memset(&q, 0, sizeof(q));
    for (int k = 0; k < 2; k++) {
      const int m = mesh_pairs[2 * i + k];
      if (m < 0 || m >= nmesh) return fail(C2A_ERR_BAD_ARG, "mesh index out of range");
      q.mesh[k] = m;
      int t = seed_tris ? seed_tris[2 * i + k] : 0;
      if (t < 0) t = 0;
      if (t >= tab[(size_t)m].ntri) return fail(C2A_ERR_BAD_ARG, "seed triangle out of range");
      q.seed[k] = t;
    }
    depth_sum = std::max(depth_sum, tab[(size_t)q.mesh[0]].depth + tab[(size_t)q.mesh[1]].depth);
    unpack_pose(poses + 24 * i, q.R1, q.T1);
    unpack_pose(poses + 24 * i + 12, q.R2, q.T2);
    q.param = params ? params[i] : 0.0;
  }
  const int64_t per_thread = sq_ws_entries(kind, qsize, depth_sum);
  const size_t thread_bytes = (size_t)per_thread * sizeof(SqEntry);
  size_t free_b = 0, total_b = 0;
  size_t budget = (size_t)1 << 30;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0) budget = free_b / 4;
  int64_t blocks = (nq + kStaticBlock - 1) / kStaticBlock;
  blocks = std::min<int64_t>(blocks, 4096);
  while (blocks > 1 && (size_t)blocks * kStaticBlock * thread_bytes > budget) blocks /= 2;
  if ((size_t)blocks * kStaticBlock * thread_bytes > budget)
    return fail(C2A_ERR_OOM, "search workspace for one workgroup exceeds the device budget");
  const size_t out_bytes = kind == SQ_DISTANCE_QUEUE ? sizeof(StaticOut) : sizeof(c2a_static_result);
  DevBuf dm, dq, dout, dc, dws, dp;
  if ((rc = dm.alloc(tab.size() * sizeof(MeshDev))) ||
      (rc = dq.alloc((size_t)nq * sizeof(StaticQueryDev))) ||
      (rc = dout.alloc((size_t)nq * out_bytes)) || (rc = dc.alloc(sizeof(int64_t))) ||
      (rc = dws.alloc((size_t)blocks * kStaticBlock * thread_bytes)))
    return rc;
  const size_t pair_bytes = kind == SQ_COLLIDE ? (size_t)nq * pair_cap * 2 * sizeof(int32_t) : 0;
  if (pair_bytes && (rc = dp.alloc(pair_bytes))) return rc;
  hipStream_t s = (hipStream_t)stream;
  HIP_TRY(hipMemcpy(dm.p, tab.data(), tab.size() * sizeof(MeshDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dq.p, hq.data(), (size_t)nq * sizeof(StaticQueryDev), hipMemcpyHostToDevice));
  StaticArgs a;
  memset(&a, 0, sizeof(a));
  a.meshes = (const MeshDev*)dm.p;
  a.q = (const StaticQueryDev*)dq.p;
  a.out = (StaticOut*)dout.p;
  a.pairs = pair_bytes ? (int32_t*)dp.p : nullptr;
  a.ws = (SqEntry*)dws.p;
  a.counter = (int64_t*)dc.p;
  a.nq = nq;
  a.ws_entries = per_thread;
  a.kind = kind;
  a.qsize = qsize;
  a.flag = flag;
  a.pair_cap = kind == SQ_COLLIDE ? pair_cap : 0;
  a.rel_err = rel_err;
  a.abs_err = abs_err;
  HIP_TRY(launch_static(a, (int)blocks, s));
  if (kind == SQ_DISTANCE_QUEUE) {
    // the distance-query record: StaticOut's tri is the models' last_tri
    std::vector<StaticOut> h((size_t)nq);
    HIP_TRY(hipMemcpyAsync(h.data(), dout.p, (size_t)nq * sizeof(StaticOut),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    c2a_distance_result* r = (c2a_distance_result*)out;
    for (int64_t i = 0; i < nq; i++) {
      const StaticOut& o = h[(size_t)i];
      memset(&r[i], 0, sizeof(r[i]));
      r[i].distance = o.distance;
      for (int k = 0; k < 3; k++) {
        r[i].p1[k] = o.p1[k];
        r[i].p2[k] = o.p2[k];
      }
      r[i].bv_tests = o.bv_tests;
      r[i].tri_tests = o.tri_tests;
      r[i].last_tri[0] = o.tri[0];
      r[i].last_tri[1] = o.tri[1];
      r[i].flags = o.flags;
    }
    return C2A_OK;
  }
  HIP_TRY(hipMemcpyAsync(out, dout.p, (size_t)nq * sizeof(c2a_static_result),
                         hipMemcpyDeviceToHost, s));
  if (pair_bytes)
    HIP_TRY(hipMemcpyAsync(pairs_out, dp.p, pair_bytes, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return C2A_OK;
}

}  // namespace

int c2a_distance_queue(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                       const int32_t* mesh_pairs, const int32_t* seed_tris, double rel_err,
                       double abs_err, int32_t qsize, int64_t nq, c2a_distance_result* out,
                       void* stream) {
  if (qsize <= 2)
    return c2a_distance(meshes, nmesh, poses, mesh_pairs, seed_tris, rel_err, abs_err, nq, out,
                        stream);
  return static_queries(meshes, nmesh, SQ_DISTANCE_QUEUE, poses, mesh_pairs, seed_tris, nullptr,
                        rel_err, abs_err, qsize, 0, nq, 0, nullptr, out, stream);
}

int c2a_distance_contact_space(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                               const int32_t* mesh_pairs, const int32_t* seed_tris,
                               const double* upper_bounds, int64_t nq, c2a_static_result* out,
                               void* stream) {
  return static_queries(meshes, nmesh, SQ_DISTANCE_CS, poses, mesh_pairs, seed_tris,
                        upper_bounds, 0.0, 0.0, 2, 0, nq, 0, nullptr, out, stream);
}

int c2a_tolerance(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                  const int32_t* mesh_pairs, const double* tolerances, int32_t qsize,
                  int64_t nq, c2a_static_result* out, void* stream) {
  return static_queries(meshes, nmesh, SQ_TOLERANCE, poses, mesh_pairs, nullptr, tolerances, 0.0,
                        0.0, qsize, 0, nq, 0, nullptr, out, stream);
}

int c2a_collide_pairs(c2a_mesh* const* meshes, int32_t nmesh, const double* poses,
                      const int32_t* mesh_pairs, int32_t flag, int64_t nq, int32_t pair_cap,
                      int32_t* pairs_out, c2a_static_result* out, void* stream) {
  return static_queries(meshes, nmesh, SQ_COLLIDE, poses, mesh_pairs, nullptr, nullptr, 0.0, 0.0,
                        2, flag, nq, pair_cap, pairs_out, out, stream);
}

int c2a_motion_prepare(const double* pose_start, const double* pose_end, double* motion24) {
  if (!pose_start || !pose_end || !motion24) return fail(C2A_ERR_BAD_ARG, "null argument");
  double R0[9], T0[3], R1[9], T1[3];
  unpack_pose(pose_start, R0, T0);
  unpack_pose(pose_end, R1, T1);
  Motion m;
  motion_init(&m, R0, T0, R1, T1);
  memcpy(motion24, &m, sizeof(double) * 24);
  return C2A_OK;
}

int c2a_motion_pose(const double* motion24, double t, double* pose12, double* quat4) {
  if (!motion24 || !pose12) return fail(C2A_ERR_BAD_ARG, "null argument");
  Motion m;
  memset(&m, 0, sizeof(m));
  memcpy(&m, motion24, sizeof(double) * 24);
  motion_integrate(m, t, pose12, pose12 + 9, quat4);
  return C2A_OK;
}

const char* c2a_status_string(int status) {
  switch (status) {
    case C2A_OK: return "ok";
    case C2A_ERR_BAD_ARG: return "bad argument";
    case C2A_ERR_OOM: return "out of memory";
    case C2A_ERR_HIP: return "HIP runtime error";
    case C2A_ERR_NO_DEVICE: return "no GPU device";
    case C2A_ERR_UNSUPPORTED: return "unsupported input";
    case C2A_ERR_EMPTY_MODEL: return "empty model";
    case C2A_ERR_CACHE: return "hierarchy file not usable";
    default: return "unknown status";
  }
}

const char* c2a_last_error(void) { return g_err.c_str(); }

}  // extern "C"
