// ccd_kernel.h -- launcher of the gfx950 CCD kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "ccd_types.h"

namespace c2a {

// K = speculative subtree depth (1..6), CAP = DFS stack capacity in entries.
hipError_t launch_ccd(const KArgs& a, int K, int CAP, int blocks, hipStream_t s);
size_t kernel_lds_bytes(int K, int CAP);

}  // namespace c2a
