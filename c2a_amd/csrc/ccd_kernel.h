// ccd_kernel.h -- launcher of the gfx950 CCD kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "ccd_types.h"

namespace c2a {

// K = speculative subtree depth (1..6), CAP = DFS stack capacity in entries
// (48 or 96), NW = wavefronts per query (1), LRN = learned child order
// (instantiated for K = 5).  launch_ccd takes LRN from a.scorer.on.
bool kernel_shape_ok(int K, int CAP, int NW, bool LRN);
hipError_t launch_ccd(const KArgs& a, int K, int CAP, int NW, int blocks, hipStream_t s);
size_t kernel_lds_bytes(int K, int CAP);
int kernel_blocks_per_cu(int K, int CAP, int NW, bool LRN);

}  // namespace c2a
