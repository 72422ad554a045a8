// ccd_kernel.hip -- batched conservative-advancement CCD on gfx950.
// Compiled once per kernel shape (ccd_shapes.h); the dispatcher is
// ccd_dispatch.cpp.
//
// One workgroup of NW wavefronts per query.  The kernel is persistent:
// workgroups take queries from a device work queue until it drains.  Each
// loop iteration advances the query's walk (ccd_core.h; every wavefront
// makes the same decisions) until it needs pair values, then every thread
// computes its slots of the speculative subtree; the replay of a CA step
// uses all threads of the workgroup.  All arithmetic is FP64 in the reference's operation
// order, so results match the host oracle bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "c2a_math.h"
#include "bvh.h"
#include "ccd_core.h"
#include "ccd_shapes.h"
#include "mlp_dev.h"

namespace c2a {

// Learned order: scores of the replayed records (features in the arena,
// 64 records per matrix-core pass); the layout then orders siblings by them.
template <int K, int CAP>
__device__ void score_replay(Wave<K, CAP>& g, Leader& L, const KArgs& a, int lane) {
  const ArenaView av = arena_view(L.ab, a.logcap, true);
  const int nrec = 1 + 2 * L.nlog_prev;
  for (int base = 1; base < nrec; base += 64) {
    // rows up to base + 63 stay inside the feature array's 64-row slack
    const float y = score64(a.scorer, av.feat + (int64_t)base * NF, lane);
    if (base + lane < nrec) gp(av.hot)[base + lane].score = y;
  }
  mem_sync();
}

template <int K, int CAP, int NW, bool LRN>
__global__ void __launch_bounds__(64 * NW) ccd_kernel(KArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Wave<K, CAP>& g = *reinterpret_cast<Wave<K, CAP>*>(smem_raw);
  const int lane = threadIdx.x;
  Leader L;
  L.state = ST_FETCH;
  L.sp = 0;
  L.ab = a.logcap > 0 ? a.arena + (size_t)blockIdx.x * arena_bytes(a.logcap, a.feats != 0)
                      : nullptr;
  for (;;) {
    leader_run(g, L, a, lane, 64 * NW);
    if (L.state == ST_DONE) break;
    if constexpr (LRN) {
      if (L.need_score) {
        score_replay(g, L, a, lane);
        continue;
      }
    }
    round_lane(g, L, lane, 64 * NW, a);
    if (g.rarena) mem_sync();   // the walk reads the new records through the LDS window
    else lane_sync();
  }
}

// ---------------------------------------------------------------------------
// host-side launch of this unit's shape
// ---------------------------------------------------------------------------
template <int K, int CAP, int NW, bool LRN>
hipError_t launch_shape(const KArgs& a, int blocks, hipStream_t s) {
  const size_t lds = sizeof(Wave<K, CAP>);
  auto fn = ccd_kernel<K, CAP, NW, LRN>;
  hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(64 * NW), lds, s, a);
  return hipGetLastError();
}

template <int K, int CAP, int NW, bool LRN>
int occupancy_shape() {
  const size_t lds = sizeof(Wave<K, CAP>);
  auto fn = ccd_kernel<K, CAP, NW, LRN>;
  if (hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
      hipSuccess)
    return 0;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64 * NW, lds) != hipSuccess) return 0;
  return n;
}

#if !defined(C2A_SK) || !defined(C2A_SCAP) || !defined(C2A_SNW) || !defined(C2A_SLRN)
#error "ccd_kernel.hip is compiled once per shape: define C2A_SK, C2A_SCAP, C2A_SNW, C2A_SLRN"
#endif
template hipError_t launch_shape<C2A_SK, C2A_SCAP, C2A_SNW, (C2A_SLRN != 0)>(const KArgs&, int,
                                                                            hipStream_t);
template int occupancy_shape<C2A_SK, C2A_SCAP, C2A_SNW, (C2A_SLRN != 0)>();

}  // namespace c2a
