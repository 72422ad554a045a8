// bvh.h -- host-side RSS bounding-volume hierarchy, built PQP-style
// (C2A_Model::EndModel, C2A_PQP.cpp:331-417 -> C2A_BuildModel,
// C2A_Build.cpp:393-574) and laid out for the GPU.
#pragma once

#include <stdint.h>

#include <vector>

namespace c2a {

// One node = two 128-byte lines.  Line 0 carries what the child transform and
// the rectangle distance read (R, Tr, l, r) plus the angular radius used by
// the motion bound; line 1 carries the model-frame orientation (motion-bound
// direction), the descend-rule size and the child link.
struct alignas(128) GpuNode {
  double R[9];        // parent-relative orientation
  double Tr[3];       // parent-relative rectangle corner
  double l[2];        // rectangle side lengths
  double r;           // sphere-sweep radius
  double angR;        // max |vertex| from the model origin
  double Rloc[9];     // model-frame orientation
  double size;        // sqrt(l0^2 + l1^2) + 2r (descend rule)
  int32_t first_child;  // >= 0: children at first_child, first_child+1; < 0: leaf, tri = -fc-1
  int32_t depth;
  int32_t pad[10];
};
static_assert(sizeof(GpuNode) == 256, "GpuNode must be 256 bytes");

struct Bvh {
  std::vector<GpuNode> nodes;     // 2n-1 nodes, root at 0
  std::vector<double> tris;       // n * 9, reordered by the build
  std::vector<int32_t> tri_id;    // caller id of each reordered triangle
  std::vector<int32_t> tri_vidx;  // n * 3 vertex indices (reordered)
  std::vector<int32_t> src_index; // caller's index of each reordered triangle
  int max_depth = 0;
};

// Builds the hierarchy.  Returns 0, or a negative code for an empty model.
int build_bvh(const double* xyz9, const int32_t* ids, const int32_t* vidx3, int64_t n, Bvh* out);

// Refit: the tree, the triangle order and the node numbers stay; every
// node's volume is fitted again, with the build's fit, to the triangles
// xyz9 (n x 9, the caller's order of the build).  Each node is fitted over
// its triangles in their final order.  Returns 0, or -1 without a
// src_index.
int refit_bvh(Bvh* b, const double* xyz9);

// First triangle position and triangle count of every node's subtree.
void node_ranges(const std::vector<GpuNode>& nodes, std::vector<int32_t>& first,
                 std::vector<int32_t>& num, std::vector<int32_t>& depth);

// Jacobi eigen-decomposition used by the builder (exported for tests).
void jacobi_eigen3(const double* A, double* V, double* d);

}  // namespace c2a
