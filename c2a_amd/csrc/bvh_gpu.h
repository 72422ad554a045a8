// bvh_gpu.h -- the RSS hierarchy built on the device (bvh_gpu.hip), equal
// bit for bit to build_bvh's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bvh.h"

namespace c2a {

// Builds on the device and copies the hierarchy back into *out.  The device
// copies of the nodes (2n - 1 GpuNode) and of the reordered triangles (n x 9
// doubles) are handed to the caller in *nodes_dev / *tris_dev (hipFree).
hipError_t build_bvh_device(const double* xyz9, const int32_t* ids, const int32_t* vidx3,
                            int64_t n, Bvh* out, void** nodes_dev, void** tris_dev,
                            hipStream_t s);

// Refits the hierarchy *b (host copy) and its device copies to the
// triangles xyz9 (host, the caller's order of the build): refit_bvh's
// result, computed on the device, copied back into *b.
hipError_t refit_bvh_device(const double* xyz9, Bvh* b, void* nodes_dev, void* tris_dev,
                            hipStream_t s);

}  // namespace c2a
