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

#include "c2a_math.h"

namespace c2a {

namespace {

struct Tri {
  double p[9];
  int32_t id;
  int32_t vidx[3];
};

// Cyclic Jacobi sweeps with the Numerical Recipes threshold schedule
// (0.2*S/n^2 in sweeps 1-3, negligible off-diagonals zeroed after sweep 4,
// at most 50 sweeps); the eigen-solver PQP's Meigen implements
// (called at C2A_Build.cpp:416).
void jacobi(const double* A, double* V, double* d) {
  double a[3][3], v[3][3], b[3], z[3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      a[i][j] = A[3 * i + j];
      v[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int p = 0; p < 3; p++) {
    b[p] = d[p] = a[p][p];
    z[p] = 0.0;
  }
  auto rot = [](double m[3][3], int i, int j, int k, int l, double s, double tau) {
    double g = m[i][j], h = m[k][l];
    m[i][j] = g - s * (h + g * tau);
    m[k][l] = h + s * (g - h * tau);
  };
  for (int sweep = 1; sweep <= 50; sweep++) {
    double sm = 0.0;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) sm += fabs(a[p][q]);
    if (sm == 0.0) break;
    double tresh = (sweep < 4) ? 0.2 * sm / (3 * 3) : 0.0;
    for (int p = 0; p < 2; p++) {
      for (int q = p + 1; q < 3; q++) {
        double g = 100.0 * fabs(a[p][q]);
        if (sweep > 4 && fabs(d[p]) + g == fabs(d[p]) && fabs(d[q]) + g == fabs(d[q])) {
          a[p][q] = 0.0;
        } else if (fabs(a[p][q]) > tresh) {
          double h = d[q] - d[p], t;
          if (fabs(h) + g == fabs(h)) {
            t = (a[p][q]) / h;
          } else {
            double theta = 0.5 * h / (a[p][q]);
            t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          double c = 1.0 / sqrt(1 + t * t);
          double s = t * c;
          double tau = s / (1.0 + c);
          h = t * a[p][q];
          z[p] -= h;
          z[q] += h;
          d[p] -= h;
          d[q] += h;
          a[p][q] = 0.0;
          for (int j = 0; j < p; j++) rot(a, j, p, j, q, s, tau);
          for (int j = p + 1; j < q; j++) rot(a, p, j, j, q, s, tau);
          for (int j = q + 1; j < 3; j++) rot(a, p, j, q, j, s, tau);
          for (int j = 0; j < 3; j++) rot(v, j, p, j, q, s, tau);
        }
      }
    }
    for (int p = 0; p < 3; p++) {
      b[p] += z[p];
      d[p] = b[p];
      z[p] = 0.0;
    }
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) V[3 * i + j] = v[i][j];
}

// vertex covariance (C2A_Build.cpp:279-335)
void covariance(double* M, const Tri* t, int n) {
  double S1[3] = {0, 0, 0};
  double s00 = 0, s11 = 0, s22 = 0, s01 = 0, s02 = 0, s12 = 0;
  for (int i = 0; i < n; i++) {
    const double *p1 = t[i].p, *p2 = t[i].p + 3, *p3 = t[i].p + 6;
    S1[0] += p1[0] + p2[0] + p3[0];
    S1[1] += p1[1] + p2[1] + p3[1];
    S1[2] += p1[2] + p2[2] + p3[2];
    s00 += (p1[0] * p1[0] + p2[0] * p2[0] + p3[0] * p3[0]);
    s11 += (p1[1] * p1[1] + p2[1] * p2[1] + p3[1] * p3[1]);
    s22 += (p1[2] * p1[2] + p2[2] * p2[2] + p3[2] * p3[2]);
    s01 += (p1[0] * p1[1] + p2[0] * p2[1] + p3[0] * p3[1]);
    s02 += (p1[0] * p1[2] + p2[0] * p2[2] + p3[0] * p3[2]);
    s12 += (p1[1] * p1[2] + p2[1] * p2[2] + p3[1] * p3[2]);
  }
  double cnt = (double)(3 * n);
  M[0] = s00 - S1[0] * S1[0] / cnt;
  M[4] = s11 - S1[1] * S1[1] / cnt;
  M[8] = s22 - S1[2] * S1[2] / cnt;
  M[1] = s01 - S1[0] * S1[1] / cnt;
  M[5] = s12 - S1[1] * S1[2] / cnt;
  M[2] = s02 - S1[0] * S1[2] / cnt;
  M[3] = M[1];
  M[6] = M[2];
  M[7] = M[5];
}

inline double max2(double a, double b) { return a > b ? a : b; }

// RSS fit of n triangles in orientation O (C2A_BV.cpp:354-635) and the
// angular radius about the model origin (C2A_BV.cpp:71-111).
void fit_rss(GpuNode& nd, const double* O, const Tri* t, int n, std::vector<double>& P) {
  double angR = 0.0;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      double pr[3] = {t[i].p[3 * k] - 0.0, t[i].p[3 * k + 1] - 0.0, t[i].p[3 * k + 2] - 0.0};
      double d = len3(pr);
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
    double dz = Z(mn) - cz;
    lo = coord(mn) + sqrt(max2(radsqr - dz * dz, 0));
    dz = Z(mx) - cz;
    hi = coord(mx) - sqrt(max2(radsqr - dz * dz, 0));
    for (int i = 0; i < np; i++)
      if (coord(i) < lo) {
        dz = Z(i) - cz;
        double x = coord(i) + sqrt(max2(radsqr - dz * dz, 0));
        if (x < lo) lo = x;
      }
    for (int i = 0; i < np; i++)
      if (coord(i) > hi) {
        dz = Z(i) - cz;
        double x = coord(i) - sqrt(max2(radsqr - dz * dz, 0));
        if (x > hi) hi = x;
      }
  };
  double minx, maxx, miny, maxy;
  extent(X, minx, maxx);
  extent(Y, miny, maxy);

  // corners: grow along the diagonal where a corner cylinder misses a point
  const double a = sqrt(0.5);
  for (int i = 0; i < np; i++) {
    double dx, dy, u, tt;
    const double px = X(i), py = Y(i), pz = Z(i);
    if (px > maxx) {
      if (py > maxy) {
        dx = px - maxx; dy = py - maxy;
        u = dx * a + dy * a;
        tt = (a * u - dx) * (a * u - dx) + (a * u - dy) * (a * u - dy) + (cz - pz) * (cz - pz);
        u = u - sqrt(max2(radsqr - tt, 0));
        if (u > 0) { maxx += u * a; maxy += u * a; }
      } else if (py < miny) {
        dx = px - maxx; dy = py - miny;
        u = dx * a - dy * a;
        tt = (a * u - dx) * (a * u - dx) + (-a * u - dy) * (-a * u - dy) + (cz - pz) * (cz - pz);
        u = u - sqrt(max2(radsqr - tt, 0));
        if (u > 0) { maxx += u * a; miny -= u * a; }
      }
    } else if (px < minx) {
      if (py > maxy) {
        dx = px - minx; dy = py - maxy;
        u = dy * a - dx * a;
        tt = (-a * u - dx) * (-a * u - dx) + (a * u - dy) * (a * u - dy) + (cz - pz) * (cz - pz);
        u = u - sqrt(max2(radsqr - tt, 0));
        if (u > 0) { minx -= u * a; maxy += u * a; }
      } else if (py < miny) {
        dx = px - minx; dy = py - miny;
        u = -dx * a - dy * a;
        tt = (-a * u - dx) * (-a * u - dx) + (-a * u - dy) * (-a * u - dy) + (cz - pz) * (cz - pz);
        u = u - sqrt(max2(radsqr - tt, 0));
        if (u > 0) { minx -= u * a; miny -= u * a; }
      }
    }
  }
  const double c0[3] = {minx, miny, cz};
  mxv(nd.Tr, nd.R, c0);
  nd.l[0] = maxx - minx;
  if (nd.l[0] < 0) nd.l[0] = 0;
  nd.l[1] = maxy - miny;
  if (nd.l[1] < 0) nd.l[1] = 0;
  nd.r = r;
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
    double C[9], E[9], s[3], R[9];
    covariance(C, &tris[first], num);
    jacobi(C, E, s);
    int mn, mid, mx;
    if (s[0] > s[1]) { mx = 0; mn = 1; }
    else { mn = 0; mx = 1; }
    if (s[2] < s[mn]) { mid = mn; mn = 2; }
    else if (s[2] > s[mx]) { mid = mx; mx = 2; }
    else { mid = 2; }
    for (int i = 0; i < 3; i++) {
      R[3 * i + 0] = E[3 * i + mx];
      R[3 * i + 1] = E[3 * i + mid];
    }
    R[2] = E[3 + mx] * E[6 + mid] - E[3 + mid] * E[6 + mx];
    R[5] = E[mid] * E[6 + mx] - E[mx] * E[6 + mid];
    R[8] = E[mx] * E[3 + mid] - E[mid] * E[3 + mx];
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
    double mean[3] = {0, 0, 0};
    for (int i = 0; i < num; i++) {
      const double* p = tris[first + i].p;
      mean[0] += p[0] + p[3] + p[6];
      mean[1] += p[1] + p[4] + p[7];
      mean[2] += p[2] + p[5] + p[8];
    }
    const double cnt = (double)(3 * num);
    mean[0] /= cnt; mean[1] /= cnt; mean[2] /= cnt;
    const double coord = dot3(axis, mean);
    // split_tris (C2A_Build.cpp:344-385)
    int c1 = 0;
    for (int i = 0; i < num; i++) {
      Tri& ti = tris[first + i];
      double p[3];
      add3(p, ti.p, ti.p + 3);
      add3(p, p, ti.p + 6);
      double x = dot3(p, axis);
      x /= 3.0;
      if (x <= coord) {
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

  // make_parent_relative (C2A_Build.cpp:493-544): post-order, each node made
  // relative to its parent's model-frame R/Tr.
  void make_parent_relative() {
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
      double Rpc[9], Tpc[3];
      mtxm(Rpc, it.pR, b.R);
      for (int i = 0; i < 9; i++) b.R[i] = Rpc[i];
      sub3(Tpc, b.Tr, it.pT);
      mtxv(b.Tr, it.pR, Tpc);
      st.pop_back();
    }
  }
};

}  // namespace

void jacobi_eigen3(const double* A, double* V, double* d) { jacobi(A, V, d); }

int build_bvh(const double* xyz9, const int32_t* ids, const int32_t* vidx3, int64_t n, Bvh* out) {
  if (n <= 0) return -1;
  Builder B;
  B.tris.resize((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    Tri& t = B.tris[(size_t)i];
    for (int k = 0; k < 9; k++) t.p[k] = xyz9[9 * i + k];
    t.id = ids ? ids[i] : (int32_t)i;
    for (int k = 0; k < 3; k++) t.vidx[k] = vidx3 ? vidx3[3 * i + k] : (int32_t)(3 * i + k);
  }
  B.nodes.assign((size_t)(2 * n - 1), GpuNode{});
  for (auto& nd : B.nodes) memset(&nd, 0, sizeof(nd));
  B.build((int)n);
  B.make_parent_relative();
  for (auto& nd : B.nodes) nd.size = sqrt(nd.l[0] * nd.l[0] + nd.l[1] * nd.l[1]) + 2 * nd.r;
  out->nodes = std::move(B.nodes);
  out->tris.resize((size_t)n * 9);
  out->tri_id.resize((size_t)n);
  out->tri_vidx.resize((size_t)n * 3);
  for (int64_t i = 0; i < n; i++) {
    const Tri& t = B.tris[(size_t)i];
    for (int k = 0; k < 9; k++) out->tris[(size_t)i * 9 + k] = t.p[k];
    out->tri_id[(size_t)i] = t.id;
    for (int k = 0; k < 3; k++) out->tri_vidx[(size_t)i * 3 + k] = t.vidx[k];
  }
  out->max_depth = B.max_depth;
  return 0;
}

}  // namespace c2a
