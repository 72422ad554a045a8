// ccd_kernel.hip -- batched conservative-advancement CCD on gfx950.
//
// One wavefront (one 64-lane workgroup) per query.  The kernel is persistent:
// wavefronts take queries from a device work queue until it drains.  Each
// loop iteration advances the query's walk in lockstep (ccd_core.h) until it
// needs pair values, then every lane computes its slots of the speculative
// subtree.  All arithmetic is FP64 in the reference's operation
// order, so results match the host oracle bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "c2a_math.h"
#include "bvh.h"
#include "ccd_kernel.h"
#include "ccd_core.h"

namespace c2a {

template <int K, int CAP>
__global__ void __launch_bounds__(64) ccd_kernel(KArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Wave<K, CAP>& g = *reinterpret_cast<Wave<K, CAP>*>(smem_raw);
  const int lane = threadIdx.x;
  Leader L;
  L.state = ST_FETCH;
  L.sp = 0;
  L.ab = a.logcap > 0 ? a.arena + (size_t)blockIdx.x * arena_bytes(a.logcap) : nullptr;
  for (;;) {
    leader_run(g, L, a, lane, 64);
    if (L.state == ST_DONE) break;
    round_lane(g, lane, 64, a);
    lane_sync();
  }
}

// ---------------------------------------------------------------------------
// host-side launch
// ---------------------------------------------------------------------------
template <int K, int CAP>
static hipError_t launch_one(const KArgs& a, int blocks, hipStream_t s) {
  const size_t lds = sizeof(Wave<K, CAP>);
  auto fn = ccd_kernel<K, CAP>;
  hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(64), lds, s, a);
  return hipGetLastError();
}

size_t kernel_lds_bytes(int K, int CAP) {
#define C2A_LDS_CASE(k, cap) \
  if (K == k && CAP == cap) return sizeof(Wave<k, cap>);
  C2A_LDS_CASE(1, 48) C2A_LDS_CASE(2, 48) C2A_LDS_CASE(3, 48) C2A_LDS_CASE(4, 48)
  C2A_LDS_CASE(5, 48) C2A_LDS_CASE(6, 48)
  C2A_LDS_CASE(1, 96) C2A_LDS_CASE(2, 96) C2A_LDS_CASE(3, 96) C2A_LDS_CASE(4, 96)
  C2A_LDS_CASE(5, 96) C2A_LDS_CASE(6, 96)
#undef C2A_LDS_CASE
  return 0;
}

hipError_t launch_ccd(const KArgs& a, int K, int CAP, int blocks, hipStream_t s) {
#define C2A_LAUNCH_CASE(k, cap) \
  if (K == k && CAP == cap) return launch_one<k, cap>(a, blocks, s);
  C2A_LAUNCH_CASE(1, 48) C2A_LAUNCH_CASE(2, 48) C2A_LAUNCH_CASE(3, 48) C2A_LAUNCH_CASE(4, 48)
  C2A_LAUNCH_CASE(5, 48) C2A_LAUNCH_CASE(6, 48)
  C2A_LAUNCH_CASE(1, 96) C2A_LAUNCH_CASE(2, 96) C2A_LAUNCH_CASE(3, 96) C2A_LAUNCH_CASE(4, 96)
  C2A_LAUNCH_CASE(5, 96) C2A_LAUNCH_CASE(6, 96)
#undef C2A_LAUNCH_CASE
  return hipErrorInvalidValue;
}

}  // namespace c2a
