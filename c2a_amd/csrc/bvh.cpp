// bvh.cpp -- PQP-style RSS hierarchy builder (host).
//
// Same arithmetic, in the same order, as the reference build path
// (C2A_Build.cpp:253-574, C2A_BV.cpp:71-111 and 354-635), so the node values
// are reproducible bit for bit.  The recursion is replaced by an explicit
// work list that visits nodes in the same order (first child's subtree
// before the second's), which fixes the node numbering.
#include "bvh.h"

#include <math.h>
#include <string.h>

#include <algorithm>

#include "bvh_fit.h"
#include "c2a_math.h"

namespace c2a {

namespace {

struct Tri {
  double p[9];
  int32_t id;
  int32_t vidx[3];
  int32_t src;   // caller's index
};

void make_parent_relative_nodes(std::vector<GpuNode>& nodes);

// vertex covariance (C2A_Build.cpp:279-335); acc[0..2] is the vertex sum
void covariance(double* M, double* acc, const Tri* t, int n) {
  for (int k = 0; k < 9; k++) acc[k] = 0.0;
  for (int i = 0; i < n; i++) cov_add(acc, t[i].p);
  cov_finish(acc, n, M);
}

// RSS fit of n triangles in orientation O (C2A_BV.cpp:354-635) and the
// angular radius about the model origin (C2A_BV.cpp:71-111).
void fit_rss(GpuNode& nd, const double* O, const Tri* t, int n, std::vector<double>& P) {
  double angR = 0.0;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      const double d = vertex_radius(t[i].p + 3 * k);
      if (d > angR) angR = d;
    }
  nd.angR = angR;
  for (int i = 0; i < 9; i++) nd.R[i] = O[i];
  const int np = 3 * n;
  P.resize((size_t)np * 3);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) mtxv(&P[(size_t)(3 * i + k) * 3], nd.R, t[i].p + 3 * k);
  auto X = [&](int i) { return P[(size_t)i * 3]; };
  auto Y = [&](int i) { return P[(size_t)i * 3 + 1]; };
  auto Z = [&](int i) { return P[(size_t)i * 3 + 2]; };

  double minz = Z(0), maxz = Z(0);
  for (int i = 1; i < np; i++) {
    if (Z(i) < minz) minz = Z(i);
    else if (Z(i) > maxz) maxz = Z(i);
  }
  const double r = 0.5 * (maxz - minz);
  const double radsqr = r * r;
  const double cz = 0.5 * (maxz + minz);

  // extent along one in-plane axis, grown so every point's sphere fits
  auto extent = [&](auto coord, double& lo, double& hi) {
    int mn = 0, mx = 0;
    for (int i = 1; i < np; i++) {
      if (coord(i) < coord(mn)) mn = i;
      else if (coord(i) > coord(mx)) mx = i;
    }
    lo = extent_lo(coord(mn), Z(mn), cz, radsqr);
    hi = extent_hi(coord(mx), Z(mx), cz, radsqr);
    for (int i = 0; i < np; i++)
      if (coord(i) < lo) {
        const double x = extent_lo(coord(i), Z(i), cz, radsqr);
        if (x < lo) lo = x;
      }
    for (int i = 0; i < np; i++)
      if (coord(i) > hi) {
        const double x = extent_hi(coord(i), Z(i), cz, radsqr);
        if (x > hi) hi = x;
      }
  };
  double minx, maxx, miny, maxy;
  extent(X, minx, maxx);
  extent(Y, miny, maxy);

  // corners: grow along the diagonal where a corner cylinder misses a point
  for (int i = 0; i < np; i++) corner_grow(X(i), Y(i), Z(i), cz, radsqr, minx, maxx, miny, maxy);
  fit_finish(nd.R, minx, maxx, miny, maxy, cz, r, nd.Tr, nd.l, &nd.r);
}

struct Builder {
  std::vector<Tri> tris;
  std::vector<GpuNode> nodes;
  std::vector<double> scratch;
  int num_nodes = 1;
  int max_depth = 0;

  // one node: orientation from the covariance eigenvectors (major, middle,
  // cross product), RSS fit, then the mean split along the major axis
  // (C2A_Build.cpp:393-488); returns the sizes of the two halves
  void fit_and_split(int bn, int first, int num, int depth, int& nfirst) {
    GpuNode& b = nodes[bn];
    double C[9], acc[9], E[9], s[3], R[9];
    covariance(C, acc, &tris[first], num);
    jacobi3(C, E, s);
    orient_from_eigen(E, s, R);
    fit_rss(b, R, &tris[first], num, scratch);
    for (int i = 0; i < 9; i++) b.Rloc[i] = b.R[i];
    b.depth = depth;
    if (depth > max_depth) max_depth = depth;
    nfirst = 0;
    if (num == 1) {
      b.first_child = -(first + 1);
      return;
    }
    b.first_child = num_nodes;
    num_nodes += 2;
    const double axis[3] = {R[0], R[3], R[6]};
    const double coord = split_plane(acc, num, axis);
    // split_tris (C2A_Build.cpp:344-385)
    int c1 = 0;
    for (int i = 0; i < num; i++) {
      if (split_coord(tris[first + i].p, axis) <= coord) {
        std::swap(tris[first + i], tris[first + c1]);
        c1++;
      }
    }
    if (c1 == 0 || c1 == num) c1 = num / 2;
    nfirst = c1;
  }

  void build(int n) {
    struct Job { int bn, first, num, depth; };
    std::vector<Job> stack;
    stack.push_back({0, 0, n, 0});
    while (!stack.empty()) {
      Job j = stack.back();
      stack.pop_back();
      int nfirst;
      fit_and_split(j.bn, j.first, j.num, j.depth, nfirst);
      if (j.num > 1) {
        int fc = nodes[j.bn].first_child;
        // second half pushed first so the first half's subtree is built first
        stack.push_back({fc + 1, j.first + nfirst, j.num - nfirst, j.depth + 1});
        stack.push_back({fc, j.first, nfirst, j.depth + 1});
      }
    }
  }

  void make_parent_relative() { make_parent_relative_nodes(nodes); }
};

// make_parent_relative (C2A_Build.cpp:493-544): post-order, each node made
// relative to its parent's model-frame R/Tr.
void make_parent_relative_nodes(std::vector<GpuNode>& nodes) {
  {
    struct Item { int bn; bool expanded; double pR[9], pT[3]; };
    std::vector<Item> st;
    Item root{0, false, {1, 0, 0, 0, 1, 0, 0, 0, 1}, {0, 0, 0}};
    st.push_back(root);
    while (!st.empty()) {
      Item& it = st.back();
      GpuNode& b = nodes[it.bn];
      if (!it.expanded && b.first_child >= 0) {
        it.expanded = true;
        Item c2{b.first_child + 1, false, {}, {}}, c1{b.first_child, false, {}, {}};
        for (int i = 0; i < 9; i++) c1.pR[i] = c2.pR[i] = b.R[i];
        for (int i = 0; i < 3; i++) c1.pT[i] = c2.pT[i] = b.Tr[i];
        st.push_back(c2);
        st.push_back(c1);
        continue;
      }
      parent_relative(it.pR, it.pT, b.R, b.Tr);
      st.pop_back();
    }
  }
}

}  // namespace

void jacobi_eigen3(const double* A, double* V, double* d) { jacobi3(A, V, d); }

int build_bvh(const double* xyz9, const int32_t* ids, const int32_t* vidx3, int64_t n, Bvh* out) {
  if (n <= 0) return -1;
  Builder B;
  B.tris.resize((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    Tri& t = B.tris[(size_t)i];
    for (int k = 0; k < 9; k++) t.p[k] = xyz9[9 * i + k];
    t.id = ids ? ids[i] : (int32_t)i;
    t.src = (int32_t)i;
    for (int k = 0; k < 3; k++) t.vidx[k] = vidx3 ? vidx3[3 * i + k] : (int32_t)(3 * i + k);
  }
  B.nodes.assign((size_t)(2 * n - 1), GpuNode{});
  for (auto& nd : B.nodes) memset(&nd, 0, sizeof(nd));
  B.build((int)n);
  B.make_parent_relative();
  for (auto& nd : B.nodes) nd.size = node_size(nd.l, nd.r);
  out->nodes = std::move(B.nodes);
  out->tris.resize((size_t)n * 9);
  out->tri_id.resize((size_t)n);
  out->tri_vidx.resize((size_t)n * 3);
  out->src_index.resize((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    const Tri& t = B.tris[(size_t)i];
    for (int k = 0; k < 9; k++) out->tris[(size_t)i * 9 + k] = t.p[k];
    out->tri_id[(size_t)i] = t.id;
    for (int k = 0; k < 3; k++) out->tri_vidx[(size_t)i * 3 + k] = t.vidx[k];
    out->src_index[(size_t)i] = t.src;
  }
  out->max_depth = B.max_depth;
  return 0;
}

void node_ranges(const std::vector<GpuNode>& nodes, std::vector<int32_t>& first,
                 std::vector<int32_t>& num, std::vector<int32_t>& depth) {
  const size_t nn = nodes.size();
  first.assign(nn, 0);
  num.assign(nn, 0);
  depth.assign(nn, 0);
  if (nn == 0) return;
  // pre-order for the depths, then children after parents reversed for the
  // ranges (a child's number is always above its parent's)
  for (size_t v = 0; v < nn; v++) {
    const int fc = nodes[v].first_child;
    if (fc >= 0) depth[(size_t)fc] = depth[(size_t)fc + 1] = depth[v] + 1;
  }
  for (size_t v = nn; v-- > 0;) {
    const int fc = nodes[v].first_child;
    if (fc < 0) {
      first[v] = -fc - 1;
      num[v] = 1;
    } else {
      first[v] = first[(size_t)fc];
      num[v] = num[(size_t)fc] + num[(size_t)fc + 1];
    }
  }
}

int refit_bvh(Bvh* b, const double* xyz9) {
  const size_t n = b->tri_id.size();
  if (n == 0 || b->src_index.size() != n) return -1;
  std::vector<Tri> tris(n);
  for (size_t i = 0; i < n; i++) {
    const int64_t s = b->src_index[i];
    for (int k = 0; k < 9; k++) tris[i].p[k] = xyz9[9 * s + k];
  }
  std::vector<int32_t> first, num, depth;
  node_ranges(b->nodes, first, num, depth);
  std::vector<double> scratch;
  for (size_t v = 0; v < b->nodes.size(); v++) {
    GpuNode& nd = b->nodes[v];
    double C[9], acc[9], E[9], sv[3], R[9];
    covariance(C, acc, &tris[(size_t)first[v]], num[v]);
    jacobi3(C, E, sv);
    orient_from_eigen(E, sv, R);
    fit_rss(nd, R, &tris[(size_t)first[v]], num[v], scratch);
    for (int k = 0; k < 9; k++) nd.Rloc[k] = nd.R[k];
  }
  make_parent_relative_nodes(b->nodes);
  for (auto& nd : b->nodes) nd.size = node_size(nd.l, nd.r);
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 9; k++) b->tris[i * 9 + k] = tris[i].p[k];
  return 0;
}

}  // namespace c2a
