// ccd_dispatch.cpp -- selects the compiled CCD kernel shape at run time
// (the shapes are listed in ccd_shapes.h, each built in its own unit).
#include <hip/hip_runtime.h>

#include "ccd_core.h"
#include "ccd_kernel.h"
#include "ccd_shapes.h"

namespace c2a {

bool kernel_shape_ok(int K, int CAP, int NW, bool LRN) {
#define C2A_OK_CASE(k, cap, nw, lrn) \
  if (K == k && CAP == cap && NW == nw && LRN == (lrn != 0)) return true;
  C2A_SHAPES(C2A_OK_CASE)
#undef C2A_OK_CASE
  return false;
}

size_t kernel_lds_bytes(int K, int CAP) {
#define C2A_LDS_CASE(k, cap, nw, lrn) \
  if (K == k && CAP == cap) return sizeof(Wave<k, cap>);
  C2A_SHAPES(C2A_LDS_CASE)
#undef C2A_LDS_CASE
  return 0;
}

int kernel_blocks_per_cu(int K, int CAP, int NW, bool LRN) {
#define C2A_OCC_CASE(k, cap, nw, lrn) \
  if (K == k && CAP == cap && NW == nw && LRN == (lrn != 0)) \
    return occupancy_shape<k, cap, nw, (lrn != 0)>();
  C2A_SHAPES(C2A_OCC_CASE)
#undef C2A_OCC_CASE
  return 0;
}

hipError_t launch_ccd(const KArgs& a, int K, int CAP, int NW, int blocks, hipStream_t s) {
  const bool LRN = a.scorer.on != 0;
#define C2A_LAUNCH_CASE(k, cap, nw, lrn)                       \
  if (K == k && CAP == cap && NW == nw && LRN == (lrn != 0)) \
    return launch_shape<k, cap, nw, (lrn != 0)>(a, blocks, s);
  C2A_SHAPES(C2A_LAUNCH_CASE)
#undef C2A_LAUNCH_CASE
  return hipErrorInvalidValue;
}

}  // namespace c2a
