// host_check.cpp -- TEST PROGRAM: exercises the parts of the source-compatible
// interface that need no GPU: C2A_Model building (triangle order and ids after
// EndModel, last_tri) and CInterpMotion_Linear integration.
//
// input : int32 n, n*9 doubles (one model), int32 nm, nm*24 doubles (start and
//         end pose of each motion), int32 nt, nt doubles (times)
// output: int32 n, n*9 doubles (triangles in hierarchy order), n int32 ids,
//         int32 last_tri index, then per motion: axis (3), angle (1) and per
//         time the pose (12) and (w, x, y, z, t) from integrate (7)
#include <math.h>
#include <stdio.h>

#include <vector>

#include "c2a/C2A.h"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int n = 0;
  if (fread(&n, 4, 1, f) != 1) return 3;
  std::vector<double> t((size_t)n * 9);
  if (fread(t.data(), 8, t.size(), f) != t.size()) return 3;
  int nm = 0;
  if (fread(&nm, 4, 1, f) != 1) return 3;
  std::vector<double> mp((size_t)nm * 24);
  if (fread(mp.data(), 8, mp.size(), f) != mp.size()) return 3;
  int nt = 0;
  if (fread(&nt, 4, 1, f) != 1) return 3;
  std::vector<double> ts((size_t)nt);
  if (fread(ts.data(), 8, ts.size(), f) != ts.size()) return 3;
  fclose(f);
  {
    // LinearMath's operators (LinearMath.h:453-493 and their definitions)
    auto near = [](double a, double b) { return fabs(a - b) <= 1e-12; };
    Coord3D a(1, 2, 3), b(-2, 0.5, 4);
    if (a * b != 1 * -2 + 2 * 0.5 + 3 * 4.0) return 7;
    const Coord3D c = a % b;
    if (!near(c * a, 0) || !near(c * b, 0) || !((a + b) - b == a) || !(-a == Coord3D(-1, -2, -3)))
      return 7;
    Matrix3x3 M;
    M.Set_Value(0.3, -0.2, 0.7);   // Rz (Ry Rx)
    const Matrix3x3 I1 = M % M, I2 = M / M, I3 = M.Inverse() * M;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        const double e = i == j ? 1.0 : 0.0;
        if (!near(I1[i][j], e) || !near(I2[i][j], e) || !near(I3[i][j], e)) return 7;
      }
    const Coord3D v = M * (M % a);
    if (!near(v[0], 1) || !near(v[1], 2) || !near(v[2], 3)) return 7;
    Matrix3x3 Mt = M;
    Mt.Transpose();
    if (Mt[0][1] != M[1][0] || Mt[2][0] != M[0][2]) return 7;
    Quaternion q(Coord3D(1, 1, 0), 0.8);   // 0.8 rad about (1, 1, 0)
    Matrix3x3 Mq;
    Mq.Set_Value(q);
    Quaternion q2 = Mq.Quaternion_();
    const double sgn = q2.W() * q.W() < 0 ? -1.0 : 1.0;
    for (int i = 0; i < 4; i++)
      if (!near(sgn * q2[i], q[i])) return 7;
    Quaternion qc = q;
    qc.Invert();
    const Quaternion e = q % qc;   // q q* = |q|^2
    if (!near(e.W(), q * q) || !near(e.X(), 0) || !near(e.Y(), 0) || !near(e.Z(), 0)) return 7;
    Transform X(M, Coord3D(0.5, -1, 2)), Xi;
    Xi.Invert(X);
    const Coord3D p = Xi * (X * a), r = X & a;
    if (!near(p[0], 1) || !near(p[1], 2) || !near(p[2], 3)) return 7;
    const Coord3D r2 = M * a;
    if (!(r == r2)) return 7;
    Coord3D w = a;
    w ^= X;
    if (!(w == X * a)) return 7;
  }
  {
    // ContactF as C2A.h:116-153 / ListTri.h:43-61 declare it
    int fa[3] = {1, 2, 3}, fb[3] = {4, -1, -1};
    PQP_REAL pa[3] = {0.5, 0, 0}, pb[3] = {0, 0.25, 0};
    ContactF c1(3, 1, fa, fb, 10, 20, pa, pb, 0.75), c2(2, 2, fb, fa, 11, 21, pb, pa, 0.5);
    if (!(c2 < c1) || !(c1 > c2) || c1 < c2) return 6;
    if (c1.FeatureID_A[2] != 3 || c1.TriangleID_B != 20 || c1.P_B[1] != 0.25) return 6;
    ContactFList l;
    l.push_front(c1);
    l.push_front(c2);
    l.sort();
    if (l.front().Distance != 0.5) return 6;
  }
  C2A_Model m;
  if (m.BeginModel() != PQP_OK) return 4;
  for (int i = 0; i < n; i++) m.AddTri(&t[9 * i], &t[9 * i + 3], &t[9 * i + 6], 1000 + i);
  if (m.EndModel() != PQP_OK) return 4;
  if (m.EndModel() != PQP_ERR_BUILD_OUT_OF_SEQUENCE) return 5;
  FILE* o = fopen(argv[2], "wb");
  fwrite(&m.num_tris, 4, 1, o);
  for (int i = 0; i < m.num_tris; i++) {
    fwrite(m.tris[i].p1, 8, 3, o);
    fwrite(m.tris[i].p2, 8, 3, o);
    fwrite(m.tris[i].p3, 8, 3, o);
  }
  for (int i = 0; i < m.num_tris; i++) fwrite(&m.tris[i].id, 4, 1, o);
  int lt = m.tri_index(m.last_tri);
  fwrite(&lt, 4, 1, o);
  for (int k = 0; k < nm; k++) {
    const double* a = &mp[(size_t)k * 24];
    PQP_REAL R0[3][3], T0[3], R1[3][3], T1[3];
    for (int i = 0; i < 3; i++) {
      T0[i] = a[9 + i];
      T1[i] = a[21 + i];
      for (int j = 0; j < 3; j++) {
        R0[i][j] = a[3 * i + j];
        R1[i][j] = a[12 + 3 * i + j];
      }
    }
    CInterpMotion_Linear mo(R0, T0, R1, T1);
    double ax[4] = {mo.m_axis.X(), mo.m_axis.Y(), mo.m_axis.Z(), mo.m_angVel};
    fwrite(ax, 8, 4, o);
    for (int s = 0; s < nt; s++) {
      PQP_REAL R[3][3], T[3], q[7];
      mo.integrate(ts[s], R, T);
      mo.integrate(ts[s], q);
      fwrite(R, 8, 9, o);
      fwrite(T, 8, 3, o);
      fwrite(q, 8, 7, o);
    }
  }
  fclose(o);
  return 0;
}
