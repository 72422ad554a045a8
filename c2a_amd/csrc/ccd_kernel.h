// ccd_kernel.h -- launcher of the gfx950 CCD kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "ccd_types.h"

namespace c2a {

// K = speculative subtree depth (1..6), CAP = DFS stack capacity in entries,
// NW = wavefronts per query (1, 2 or 4; 2 and 4 for K = 3..5 only).
bool kernel_shape_ok(int K, int CAP, int NW);
hipError_t launch_ccd(const KArgs& a, int K, int CAP, int NW, int blocks, hipStream_t s);
size_t kernel_lds_bytes(int K, int CAP);
int kernel_blocks_per_cu(int K, int CAP, int NW);

}  // namespace c2a
