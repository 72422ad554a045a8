// ccd_kernel.h -- launcher of the gfx950 CCD kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "ccd_types.h"

namespace c2a {

// K = speculative subtree depth, NG = query groups per wavefront,
// CAP = DFS stack capacity (entries) per group.
hipError_t launch_ccd(const KArgs& a, int K, int NG, int CAP, int blocks, hipStream_t s);
size_t kernel_lds_bytes(int K, int NG, int CAP);
int kernel_groups(int K);

}  // namespace c2a
