// static_kernel.h -- launcher of the static-query kernel (static_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "static_query.h"

namespace c2a {

// Threads per workgroup; the workspace holds blocks * kStaticBlock slices.
constexpr int kStaticBlock = 256;

hipError_t launch_static(const StaticArgs& a, int blocks, hipStream_t s);

}  // namespace c2a
