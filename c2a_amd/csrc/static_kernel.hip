// static_kernel.hip -- the static proximity queries of static_query.h on
// gfx950: one thread per query, threads take queries from a device work
// queue until it drains, each with its own slice of the workspace.
#include <hip/hip_runtime.h>

#include "static_query.h"
#include "static_kernel.h"

namespace c2a {

__global__ void static_zero_counter(int64_t* c) { *c = 0; }

__global__ void __launch_bounds__(kStaticBlock) static_kernel(StaticArgs a) {
  const int64_t tid = (int64_t)blockIdx.x * kStaticBlock + threadIdx.x;
  SqEntry* ws = a.ws + tid * a.ws_entries;
  for (;;) {
    const int64_t qi =
        (int64_t)atomicAdd(reinterpret_cast<unsigned long long*>(a.counter), 1ull);
    if (qi >= a.nq) break;
    static_query(a, qi, ws);
  }
}

hipError_t launch_static(const StaticArgs& a, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(static_zero_counter, dim3(1), dim3(1), 0, s, a.counter);
  hipLaunchKernelGGL(static_kernel, dim3(blocks), dim3(kStaticBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace c2a
