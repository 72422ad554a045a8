// ccd_shapes.h -- the instantiated shapes of the CCD kernel.
//
// Each shape is compiled in its own translation unit (ccd_kernel.hip built
// once per shape with C2A_SK / C2A_SCAP / C2A_SNW / C2A_SLRN set), so the
// shapes compile in parallel and the compiler sees one kernel per unit.
// The dispatcher (ccd_dispatch.cpp) selects among them at run time.
#pragma once

#include <hip/hip_runtime.h>

#include "ccd_types.h"

// X(K, CAP, NW, LRN): K = speculative subtree depth, CAP = DFS stack entries,
// NW = wavefronts per query, LRN = learned child order.  CAP 192 (hierarchy
// pairs up to 190 levels deep, 52 KB of LDS) is instantiated at K = 5 only;
// NW = 4 (the tail pass) at K = 5, plain order, CAP 48 and 96.
#define C2A_SHAPES(X)                                                                  \
  X(1, 48, 1, 0) X(2, 48, 1, 0) X(3, 48, 1, 0) X(4, 48, 1, 0) X(5, 48, 1, 0)            \
  X(6, 48, 1, 0) X(1, 96, 1, 0) X(2, 96, 1, 0) X(3, 96, 1, 0) X(4, 96, 1, 0)            \
  X(5, 96, 1, 0) X(6, 96, 1, 0) X(5, 48, 1, 1) X(5, 96, 1, 1) X(5, 192, 1, 0)          \
  X(5, 192, 1, 1) X(5, 48, 4, 0) X(5, 96, 4, 0)

namespace c2a {

// Defined (and explicitly instantiated for one shape) by each compilation
// of ccd_kernel.hip.
template <int K, int CAP, int NW, bool LRN>
hipError_t launch_shape(const KArgs& a, int blocks, hipStream_t s);
template <int K, int CAP, int NW, bool LRN>
int occupancy_shape();

}  // namespace c2a
