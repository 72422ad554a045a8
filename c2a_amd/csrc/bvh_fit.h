// bvh_fit.h -- the per-node arithmetic of the RSS hierarchy build, shared by
// the host builder (bvh.cpp) and the device builder (bvh_gpu.hip) so both
// evaluate every expression in the same order (C2A_Build.cpp:279-488,
// C2A_BV.cpp:71-111 and 354-635).  The sums over a node's triangles are
// sequential on both sides; minima and maxima may be reduced in any order,
// since they are exact.
#pragma once

#include <math.h>
#include <stdint.h>

#include "c2a_math.h"

namespace c2a {

// Cyclic Jacobi sweeps with the Numerical Recipes threshold schedule
// (0.2*S/n^2 in sweeps 1-3, negligible off-diagonals zeroed after sweep 4,
// at most 50 sweeps); the eigen-solver PQP's Meigen implements
// (called at C2A_Build.cpp:416).
C2A_HD void jacobi_rot(double m[3][3], int i, int j, int k, int l, double s, double tau) {
  const double g = m[i][j], h = m[k][l];
  m[i][j] = g - s * (h + g * tau);
  m[k][l] = h + s * (g - h * tau);
}

C2A_HD void jacobi3(const double* A, double* V, double* d) {
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
  for (int sweep = 1; sweep <= 50; sweep++) {
    double sm = 0.0;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) sm += fabs(a[p][q]);
    if (sm == 0.0) break;
    const double tresh = (sweep < 4) ? 0.2 * sm / (3 * 3) : 0.0;
    for (int p = 0; p < 2; p++) {
      for (int q = p + 1; q < 3; q++) {
        const double g = 100.0 * fabs(a[p][q]);
        if (sweep > 4 && fabs(d[p]) + g == fabs(d[p]) && fabs(d[q]) + g == fabs(d[q])) {
          a[p][q] = 0.0;
        } else if (fabs(a[p][q]) > tresh) {
          double h = d[q] - d[p], t;
          if (fabs(h) + g == fabs(h)) {
            t = (a[p][q]) / h;
          } else {
            const double theta = 0.5 * h / (a[p][q]);
            t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          const double c = 1.0 / sqrt(1 + t * t);
          const double s = t * c;
          const double tau = s / (1.0 + c);
          h = t * a[p][q];
          z[p] -= h;
          z[q] += h;
          d[p] -= h;
          d[q] += h;
          a[p][q] = 0.0;
          for (int j = 0; j < p; j++) jacobi_rot(a, j, p, j, q, s, tau);
          for (int j = p + 1; j < q; j++) jacobi_rot(a, p, j, j, q, s, tau);
          for (int j = q + 1; j < 3; j++) jacobi_rot(a, p, j, q, j, s, tau);
          for (int j = 0; j < 3; j++) jacobi_rot(v, j, p, j, q, s, tau);
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

// Vertex covariance accumulation of one triangle (C2A_Build.cpp:279-335):
// the nine terms t = S1[3], then s00, s11, s22, s01, s02, s12.  S1 is also the split's
// vertex sum (C2A_Build.cpp:461-472 sums the same terms in the same order).
C2A_HD void cov_terms(const double* p, double* t) {
  const double *p1 = p, *p2 = p + 3, *p3 = p + 6;
  t[0] = p1[0] + p2[0] + p3[0];
  t[1] = p1[1] + p2[1] + p3[1];
  t[2] = p1[2] + p2[2] + p3[2];
  t[3] = (p1[0] * p1[0] + p2[0] * p2[0] + p3[0] * p3[0]);
  t[4] = (p1[1] * p1[1] + p2[1] * p2[1] + p3[1] * p3[1]);
  t[5] = (p1[2] * p1[2] + p2[2] * p2[2] + p3[2] * p3[2]);
  t[6] = (p1[0] * p1[1] + p2[0] * p2[1] + p3[0] * p3[1]);
  t[7] = (p1[0] * p1[2] + p2[0] * p2[2] + p3[0] * p3[2]);
  t[8] = (p1[1] * p1[2] + p2[1] * p2[2] + p3[1] * p3[2]);
}

// acc[k] += t[k]: the reference adds each right-hand side to its sum, so a
// sum's terms can be computed anywhere as long as they are added in order.
C2A_HD void cov_add(double* acc, const double* p) {
  double t[9];
  cov_terms(p, t);
  for (int k = 0; k < 9; k++) acc[k] += t[k];
}

C2A_HD void cov_finish(const double* acc, int n, double* M) {
  const double cnt = (double)(3 * n);
  M[0] = acc[3] - acc[0] * acc[0] / cnt;
  M[4] = acc[4] - acc[1] * acc[1] / cnt;
  M[8] = acc[5] - acc[2] * acc[2] / cnt;
  M[1] = acc[6] - acc[0] * acc[1] / cnt;
  M[5] = acc[8] - acc[1] * acc[2] / cnt;
  M[2] = acc[7] - acc[0] * acc[2] / cnt;
  M[3] = M[1];
  M[6] = M[2];
  M[7] = M[5];
}

// Orientation from the eigenvectors: major, middle, then their cross
// product (C2A_Build.cpp:416-436).
C2A_HD void orient_from_eigen(const double* E, const double* s, double* R) {
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
}

C2A_HD double fit_max2(double a, double b) { return a > b ? a : b; }

// |vertex| about the model origin (C2A_BV.cpp:71-111)
C2A_HD double vertex_radius(const double* v) {
  const double pr[3] = {v[0] - 0.0, v[1] - 0.0, v[2] - 0.0};
  return len3(pr);
}

// In-plane extent candidates of point (c, z) about the slab centre cz
// (C2A_BV.cpp:420-520): the low side moves up by the sphere's half chord,
// the high side down.
C2A_HD double extent_lo(double c, double z, double cz, double radsqr) {
  const double dz = z - cz;
  return c + sqrt(fit_max2(radsqr - dz * dz, 0));
}
C2A_HD double extent_hi(double c, double z, double cz, double radsqr) {
  const double dz = z - cz;
  return c - sqrt(fit_max2(radsqr - dz * dz, 0));
}

// Corner growth for one projected point (C2A_BV.cpp:522-606): where a corner
// cylinder misses the point, grow the rectangle along the diagonal.
C2A_HD void corner_grow(double px, double py, double pz, double cz, double radsqr, double& minx,
                        double& maxx, double& miny, double& maxy) {
  const double a = sqrt(0.5);
  double dx, dy, u, tt;
  if (px > maxx) {
    if (py > maxy) {
      dx = px - maxx; dy = py - maxy;
      u = dx * a + dy * a;
      tt = (a * u - dx) * (a * u - dx) + (a * u - dy) * (a * u - dy) + (cz - pz) * (cz - pz);
      u = u - sqrt(fit_max2(radsqr - tt, 0));
      if (u > 0) { maxx += u * a; maxy += u * a; }
    } else if (py < miny) {
      dx = px - maxx; dy = py - miny;
      u = dx * a - dy * a;
      tt = (a * u - dx) * (a * u - dx) + (-a * u - dy) * (-a * u - dy) + (cz - pz) * (cz - pz);
      u = u - sqrt(fit_max2(radsqr - tt, 0));
      if (u > 0) { maxx += u * a; miny -= u * a; }
    }
  } else if (px < minx) {
    if (py > maxy) {
      dx = px - minx; dy = py - maxy;
      u = dy * a - dx * a;
      tt = (-a * u - dx) * (-a * u - dx) + (a * u - dy) * (a * u - dy) + (cz - pz) * (cz - pz);
      u = u - sqrt(fit_max2(radsqr - tt, 0));
      if (u > 0) { minx -= u * a; maxy += u * a; }
    } else if (py < miny) {
      dx = px - minx; dy = py - miny;
      u = -dx * a - dy * a;
      tt = (-a * u - dx) * (-a * u - dx) + (-a * u - dy) * (-a * u - dy) + (cz - pz) * (cz - pz);
      u = u - sqrt(fit_max2(radsqr - tt, 0));
      if (u > 0) { minx -= u * a; miny -= u * a; }
    }
  }
}

// A point can ever grow a corner only if it lies outside the rectangle in
// both in-plane directions at the start of the corner pass (the rectangle
// only grows, so the tests only get harder).
C2A_HD bool corner_candidate(double px, double py, double minx, double maxx, double miny,
                             double maxy) {
  return (px > maxx || px < minx) && (py > maxy || py < miny);
}

// Rectangle corner, sides and radius of the fit (C2A_BV.cpp:608-635).
C2A_HD void fit_finish(const double* R, double minx, double maxx, double miny, double maxy,
                       double cz, double r, double* Tr, double* l, double* rad) {
  const double c0[3] = {minx, miny, cz};
  mxv(Tr, R, c0);
  l[0] = maxx - minx;
  if (l[0] < 0) l[0] = 0;
  l[1] = maxy - miny;
  if (l[1] < 0) l[1] = 0;
  *rad = r;
}

// Split coordinate of a triangle along the axis (split_tris,
// C2A_Build.cpp:344-385): the centroid's projection.
C2A_HD double split_coord(const double* p, const double* axis) {
  double c[3];
  add3(c, p, p + 3);
  add3(c, c, p + 6);
  double x = dot3(c, axis);
  x /= 3.0;
  return x;
}

// The split plane: the vertex mean projected on the major axis
// (C2A_Build.cpp:461-472), from the covariance pass's vertex sum.
C2A_HD double split_plane(const double* S1, int n, const double* axis) {
  const double cnt = (double)(3 * n);
  const double mean[3] = {S1[0] / cnt, S1[1] / cnt, S1[2] / cnt};
  return dot3(axis, mean);
}

// make_parent_relative (C2A_Build.cpp:493-544) for one node: R and Tr from
// the model frame into the parent's (model-frame) R and Tr.
C2A_HD void parent_relative(const double* pR, const double* pT, double* R, double* Tr) {
  double Rpc[9], Tpc[3];
  mtxm(Rpc, pR, R);
  for (int i = 0; i < 9; i++) R[i] = Rpc[i];
  sub3(Tpc, Tr, pT);
  mtxv(Tr, pR, Tpc);
}

// The descend rule's node size (PQP BV::GetSize for RSS).
C2A_HD double node_size(const double* l, double r) { return sqrt(l[0] * l[0] + l[1] * l[1]) + 2 * r; }

// Index of a node's first child when nodes are numbered in the order of a
// depth-first build that takes the first half first (C2A_Build.cpp:455-457):
// the internal nodes before it in that order are its `depth` ancestors and
// the internal nodes of the first-half subtrees left of its path, whose
// triangles are the `first` before its range, in `rights` subtrees.
C2A_HD int32_t dfs_first_child(int depth, int64_t first, int rights) {
  return (int32_t)(1 + 2 * ((int64_t)depth + first - rights));
}

}  // namespace c2a
