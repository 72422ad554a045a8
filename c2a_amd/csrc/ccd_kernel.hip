// ccd_kernel.hip -- batched conservative-advancement CCD on gfx950.
//
// A wavefront holds NG lane groups of G lanes, one query per group.  The
// kernel is persistent: groups take queries from a device work queue until it
// drains.  Each loop iteration has two phases separated by barriers: the
// group leaders advance their walks (ccd_core.h) until each needs a round,
// then every lane of a requesting group computes its slot of that group's
// speculative subtree.  All arithmetic is FP64 in the reference's operation
// order, so results match the host oracle bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "c2a_math.h"
#include "bvh.h"
#include "ccd_kernel.h"
#include "ccd_core.h"

namespace c2a {

template <int K, int NG, int CAP>
__global__ void __launch_bounds__(64) ccd_kernel(KArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using Grp = Group<K, CAP>;
  constexpr int G = Shape<K>::G;
  Grp* groups = reinterpret_cast<Grp*>(smem_raw);
  const int lane = threadIdx.x;
  const int gi = lane / G;
  const int gl = lane % G;
  const bool member = gi < NG;
  const bool leader = member && gl == 0;
  Leader L;
  L.state = ST_FETCH;
  L.sp = 0;
  if (leader) groups[gi].req = 0;
  __syncthreads();
  for (;;) {
    if (leader) leader_run(groups[gi], L, a);
    const int alive = __syncthreads_or(leader && L.state != ST_DONE);
    if (!alive) break;
    if (member && groups[gi].req) round_lane(groups[gi], gl, a);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host-side launch
// ---------------------------------------------------------------------------
template <int K, int NG, int CAP>
static hipError_t launch_one(const KArgs& a, int blocks, hipStream_t s) {
  const size_t lds = sizeof(Group<K, CAP>) * NG;
  auto fn = ccd_kernel<K, NG, CAP>;
  hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(64), lds, s, a);
  return hipGetLastError();
}

size_t kernel_lds_bytes(int K, int NG, int CAP) {
#define C2A_LDS_CASE(k, cap) \
  if (K == k && CAP == cap) return sizeof(Group<k, cap>) * (size_t)NG;
  C2A_LDS_CASE(1, 48) C2A_LDS_CASE(2, 48) C2A_LDS_CASE(3, 48) C2A_LDS_CASE(4, 48) C2A_LDS_CASE(5, 48)
  C2A_LDS_CASE(1, 96) C2A_LDS_CASE(2, 96) C2A_LDS_CASE(3, 96) C2A_LDS_CASE(4, 96) C2A_LDS_CASE(5, 96)
#undef C2A_LDS_CASE
  return 0;
}

int kernel_groups(int K) {
  switch (K) {
    case 1: return 64 / Shape<1>::G;
    case 2: return 64 / Shape<2>::G;
    case 3: return 64 / Shape<3>::G;
    case 4: return 64 / Shape<4>::G;
    default: return 64 / Shape<5>::G;
  }
}

hipError_t launch_ccd(const KArgs& a, int K, int NG, int CAP, int blocks, hipStream_t s) {
#define C2A_LAUNCH_CASE(k, ng, cap) \
  if (K == k && NG == ng && CAP == cap) return launch_one<k, ng, cap>(a, blocks, s);
  C2A_LAUNCH_CASE(1, 8, 48) C2A_LAUNCH_CASE(2, 8, 48) C2A_LAUNCH_CASE(3, 4, 48)
  C2A_LAUNCH_CASE(4, 2, 48) C2A_LAUNCH_CASE(5, 1, 48)
  C2A_LAUNCH_CASE(1, 8, 96) C2A_LAUNCH_CASE(2, 8, 96) C2A_LAUNCH_CASE(3, 4, 96)
  C2A_LAUNCH_CASE(4, 2, 96) C2A_LAUNCH_CASE(5, 1, 96)
#undef C2A_LAUNCH_CASE
  return hipErrorInvalidValue;
}

}  // namespace c2a
