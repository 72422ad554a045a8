// c2a_math.h -- FP64 geometry kernels shared by the HIP device code and the
// host-side builder / API layer.
//
// Every routine evaluates its expressions in a fixed order and the whole
// library is compiled with -ffp-contract=off, so host and device produce the
// same bits.  Reference citations name the EwhaGlab/C2A file:line whose
// semantics each routine implements (PQP routines are cited at their C2A
// call sites; PQP itself is not vendored).
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define C2A_HD __host__ __device__ inline
#else
#define C2A_HD inline
#endif

namespace c2a {

struct V3 {
  double x[3];
};

C2A_HD double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
C2A_HD void sub3(double* r, const double* a, const double* b) {
  r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
C2A_HD void add3(double* r, const double* a, const double* b) {
  r[0] = a[0] + b[0]; r[1] = a[1] + b[1]; r[2] = a[2] + b[2];
}
C2A_HD void cross3(double* r, const double* a, const double* b) {
  double r0 = a[1] * b[2] - a[2] * b[1];
  double r1 = a[2] * b[0] - a[0] * b[2];
  double r2 = a[0] * b[1] - a[1] * b[0];
  r[0] = r0; r[1] = r1; r[2] = r2;
}
C2A_HD void neg3(double* r, const double* a) { r[0] = a[0] * -1; r[1] = a[1] * -1; r[2] = a[2] * -1; }
// r = a + b*s
C2A_HD void axpy3(double* r, const double* a, const double* b, double s) {
  r[0] = a[0] + b[0] * s; r[1] = a[1] + b[1] * s; r[2] = a[2] + b[2] * s;
}
C2A_HD double len3(const double* a) { return sqrt(dot3(a, a)); }
C2A_HD double dist2_3(const double* a, const double* b) {
  double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  return d0 * d0 + d1 * d1 + d2 * d2;
}
// multiply by the reciprocal length; no zero guard (a zero vector gives NaN)
C2A_HD void normalize3(double* v) {
  double d = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] *= d; v[1] *= d; v[2] *= d;
}
// Matrices are row-major double[9].
C2A_HD void mxv(double* r, const double* M, const double* v) {
  double t0 = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
  double t1 = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
  double t2 = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
  r[0] = t0; r[1] = t1; r[2] = t2;
}
C2A_HD void mtxv(double* r, const double* M, const double* v) {
  double t0 = M[0] * v[0] + M[3] * v[1] + M[6] * v[2];
  double t1 = M[1] * v[0] + M[4] * v[1] + M[7] * v[2];
  double t2 = M[2] * v[0] + M[5] * v[1] + M[8] * v[2];
  r[0] = t0; r[1] = t1; r[2] = t2;
}
C2A_HD void mxvpv(double* r, const double* M, const double* v, const double* w) {
  double t0 = M[0] * v[0] + M[1] * v[1] + M[2] * v[2] + w[0];
  double t1 = M[3] * v[0] + M[4] * v[1] + M[5] * v[2] + w[1];
  double t2 = M[6] * v[0] + M[7] * v[1] + M[8] * v[2] + w[2];
  r[0] = t0; r[1] = t1; r[2] = t2;
}
// R = A B
C2A_HD void mxm(double* R, const double* A, const double* B) {
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[3 * i + j] = A[3 * i + 0] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
  for (int i = 0; i < 9; i++) R[i] = t[i];
}
// R = A^T B
C2A_HD void mtxm(double* R, const double* A, const double* B) {
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
  for (int i = 0; i < 9; i++) R[i] = t[i];
}

// ---------------------------------------------------------------------------
// sin/cos of the interpolation angle.  Cody-Waite reduction by pi/2 (three
// 33-bit parts, exact products for |k| < 2^20) and Taylor polynomials on
// |r| <= pi/4 (sin through r^17, cos through r^18; truncation < 1e-19).
// Identical arithmetic on host and device, so poses integrated on the GPU
// match host-side checks bit for bit.
// ---------------------------------------------------------------------------
C2A_HD void sincos_ca(double x, double* s, double* c) {
  double k = rint(x * 0.6366197723675814);
  double r = x - k * 1.5707963267341256;
  r = r - k * 6.077100506303966e-11;
  r = r - k * 2.0222662487959506e-21;
  double r2 = r * r;
  double ps = 2.8114572543455207632e-15;
  ps = ps * r2 + -7.6471637318198164759e-13;
  ps = ps * r2 + 1.6059043836821614599e-10;
  ps = ps * r2 + -2.5052108385441718775e-08;
  ps = ps * r2 + 2.7557319223985890653e-06;
  ps = ps * r2 + -1.9841269841269841270e-04;
  ps = ps * r2 + 8.3333333333333333333e-03;
  ps = ps * r2 + -1.6666666666666666667e-01;
  double sv = r + (r * r2) * ps;
  double pc = 1.5619206968586225796e-16;
  pc = pc * r2 + -4.7794773323873852974e-14;
  pc = pc * r2 + 1.1470745597729724714e-11;
  pc = pc * r2 + -2.0876756987868098979e-09;
  pc = pc * r2 + 2.7557319223985890653e-07;
  pc = pc * r2 + -2.4801587301587301587e-05;
  pc = pc * r2 + 1.3888888888888888889e-03;
  pc = pc * r2 + -4.1666666666666666667e-02;
  double cv = (1.0 - 0.5 * r2) - (r2 * r2) * pc;
  long q = ((long)k) & 3;
  if (q == 0) { *s = sv; *c = cv; }
  else if (q == 1) { *s = cv; *c = -sv; }
  else if (q == 2) { *s = -sv; *c = -cv; }
  else { *s = -cv; *c = sv; }
}

// ---------------------------------------------------------------------------
// Rotation conventions (LinearMath.h): quaternions (x, y, z, w).
// ---------------------------------------------------------------------------
// Matrix3x3::Quaternion_ (LinearMath.h:759-793)
C2A_HD void quat_from_matrix(const double* v, double* q) {
  double trace = v[0] + v[4] + v[8];
  if (trace > 0.0) {
    double s = sqrt(trace + 1.0);
    q[3] = s * 0.5;
    s = 0.5 / s;
    q[0] = (v[7] - v[5]) * s;
    q[1] = (v[2] - v[6]) * s;
    q[2] = (v[3] - v[1]) * s;
  } else {
    int i = v[0] < v[4] ? (v[4] < v[8] ? 2 : 1) : (v[0] < v[8] ? 2 : 0);
    int j = (i + 1) % 3, k = (i + 2) % 3;
    double s = sqrt(v[i * 3 + i] - v[j * 3 + j] - v[k * 3 + k] + 1.0);
    q[i] = s * 0.5;
    s = 0.5 / s;
    q[3] = (v[k * 3 + j] - v[j * 3 + k]) * s;
    q[j] = (v[j * 3 + i] + v[i * 3 + j]) * s;
    q[k] = (v[k * 3 + i] + v[i * 3 + k]) * s;
  }
}
// Matrix3x3::Set_Value(const Quaternion&) (LinearMath.h:809-831)
C2A_HD void matrix_from_quat(const double* q, double* R) {
  double d = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  double s = 2.0 / d;
  double xs = q[0] * s, ys = q[1] * s, zs = q[2] * s;
  double wx = q[3] * xs, wy = q[3] * ys, wz = q[3] * zs;
  double xx = q[0] * xs, xy = q[0] * ys, xz = q[0] * zs;
  double yy = q[1] * ys, yz = q[1] * zs, zz = q[2] * zs;
  R[0] = 1.0 - (yy + zz); R[1] = xy - wz; R[2] = xz + wy;
  R[3] = xy + wz; R[4] = 1.0 - (xx + zz); R[5] = yz - wx;
  R[6] = xz - wy; R[7] = yz + wx; R[8] = 1.0 - (xx + yy);
}
// Hamilton product (LinearMath.h:1018-1025)
C2A_HD void quat_mul(const double* a, const double* b, double* r) {
  double x = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
  double y = a[3] * b[1] + a[1] * b[3] + a[2] * b[0] - a[0] * b[2];
  double z = a[3] * b[2] + a[2] * b[3] + a[0] * b[1] - a[1] * b[0];
  double w = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
  r[0] = x; r[1] = y; r[2] = z; r[3] = w;
}

// Linear motion of one object over t in [0,1] (CInterpMotion_Linear).
struct Motion {
  double R0[9], T0[3];
  double q0[4];
  double cv[3];
  double axis[3];
  double angVel;
  double pad;
};

// CInterpMotion_Linear ctor + velocity() + LinearAngularVelocity
// (InterpMotion.cpp:148-168, 486-491, 228-270).  Host only (acos).
inline void motion_init(Motion* m, const double* R0, const double* T0, const double* R1,
                        const double* T1) {
  for (int i = 0; i < 9; i++) m->R0[i] = R0[i];
  for (int i = 0; i < 3; i++) m->T0[i] = T0[i];
  sub3(m->cv, T1, T0);
  double qs[4], qt[4], qc[4], qr[4];
  quat_from_matrix(R0, qs);
  quat_from_matrix(R1, qt);
  for (int i = 0; i < 4; i++) m->q0[i] = qs[i];
  qc[0] = -qs[0]; qc[1] = -qs[1]; qc[2] = -qs[2]; qc[3] = qs[3];
  quat_mul(qc, qt, qr);
  double s = 1 < qr[3] ? 1 : qr[3];
  double sign = s < 0 ? -1 : 1;
  double a = (fabs(s - 1) <= 1e-40 || fabs(s + 1) <= 1e-40)
                 ? (2 * sign)
                 : (sign * acos(2 * s * s - 1) / sqrt(1 - s * s));
  double tg[3] = {a * qr[0], a * qr[1], a * qr[2]};
  m->angVel = len3(tg);
  double len = tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2];
  if (len < (double)0.00000001f) {
    m->axis[0] = 1.0; m->axis[1] = 0.0; m->axis[2] = 0.0;
  } else {
    double inv = 1.0 / sqrt(len);
    m->axis[0] = tg[0] * inv; m->axis[1] = tg[1] * inv; m->axis[2] = tg[2] * inv;
  }
  m->pad = 0.0;
}

// CInterpMotion_Linear::integrate (InterpMotion.cpp:516-568): pose at t
// (t clamped to 1); q (optional) receives the quaternion (x,y,z,w).
C2A_HD void motion_integrate(const Motion& m, double t, double* R, double* T, double* q_out) {
  double dt = t;
  if (dt > 1) dt = 1;
  T[0] = m.T0[0] + dt * m.cv[0];
  T[1] = m.T0[1] + dt * m.cv[1];
  T[2] = m.T0[2] + dt * m.cv[2];
  double ang = 0.5 * m.angVel * dt;
  double sa, ca;
  sincos_ca(ang, &sa, &ca);
  double drt[4] = {sa * m.axis[0], sa * m.axis[1], sa * m.axis[2], ca};
  double q[4];
  quat_mul(m.q0, drt, q);
  matrix_from_quat(q, R);
  if (q_out) { q_out[0] = q[0]; q_out[1] = q[1]; q_out[2] = q[2]; q_out[3] = q[3]; }
}

// computeTOC_MotionBound (InterpMotion.cpp:831-881); N is normalised in place.
C2A_HD double motion_bound(const Motion& m, double angR, double* N) {
  double cr[3];
  normalize3(N);
  cross3(cr, m.axis, N);
  double w = angR * len3(cr) * m.angVel;
  double v = dot3(m.cv, N);
  if (v < 0) v = 0;
  double path = v + w;
  if (path <= 0) path = 1e-30;
  return path;
}

// computeTOC(d, r1, S) leaf bound (InterpMotion.cpp:746-828); S normalised in place.
C2A_HD double motion_bound_leaf(const Motion& m, double angR, double* S) {
  normalize3(S);
  double w, v;
  if (m.angVel == 0) {
    w = 0;
    v = dot3(m.cv, S);
    if (v < 0) v = 0;
  } else {
    double cwc[3] = {m.axis[0] * m.angVel, m.axis[1] * m.angVel, m.axis[2] * m.angVel};
    double cr[3];
    cross3(cr, cwc, S);
    w = angR * len3(cr);
    v = dot3(m.cv, S);
    if (v < 0) v = 0;
  }
  double path = w + v;
  if (path == 0) path = 1e-30;
  return path;
}

// ---------------------------------------------------------------------------
// RSS rectangle distance (C2A_RectDist.h:44-934).  A = [0,a0]x[0,a1] in its
// own frame, B = Tab + Rab*[0,b0]x[0,b1].  Returns the distance and S = Q-P
// (A's frame); S is not written when the rectangles overlap.
// ---------------------------------------------------------------------------
C2A_HD void clampd(double& v, double lo, double hi) {
  if (v < lo) v = lo;
  else if (v > hi) v = hi;
}
// Lumelsky segment parameters (C2A_RectDist.h:81-111)
C2A_HD void seg_params(double& t, double& u, double a, double b, double AB, double AT, double BT) {
  double denom = 1 - AB * AB;
  if (denom == 0) t = 0;
  else {
    t = (AT - BT * AB) / denom;
    clampd(t, 0, a);
  }
  u = t * AB - BT;
  if (u < 0) {
    u = 0;
    t = AT;
    clampd(t, 0, a);
  } else if (u > b) {
    u = b;
    t = u * AB + AT;
    clampd(t, 0, a);
  }
}
// InsideVoronoi (C2A_RectDist.h:123-154)
C2A_HD bool voronoi(double a, double b, double AnB, double AnT, double AB, double AT, double BT) {
  if (fabs(AnB) < 1e-7) return false;
  double u = -AnT / AnB;
  clampd(u, 0, b);
  double t = u * AB + AT;
  clampd(t, 0, a);
  double v = t * AB - BT;
  if (AnB > 0) {
    if (v > (u + 1e-7)) return true;
  } else {
    if (v < (u - 1e-7)) return true;
  }
  return false;
}
// edge-pair case epilogue: P on A's edge, Q = Tab + col_f*lf + col_v*u
C2A_HD double rect_edge_tail(double P0, double P1, const double* Rab, const double* Tab, int cf,
                             double lf, bool has_f, int cv, double u, double* S) {
  double Q[3];
  for (int i = 0; i < 3; i++) {
    if (has_f) Q[i] = Tab[i] + Rab[3 * i + cf] * lf + Rab[3 * i + cv] * u;
    else Q[i] = Tab[i] + Rab[3 * i + cv] * u;
  }
  S[0] = Q[0] - P0; S[1] = Q[1] - P1; S[2] = Q[2] - 0;
  return sqrt(dot3(S, S));
}

C2A_HD double rect_dist(const double* Rab, const double* Tab, const double* a, const double* b,
                        double* S) {
  const double AB00 = Rab[0], AB01 = Rab[1], AB10 = Rab[3], AB11 = Rab[4];
  const double aAB00 = a[0] * AB00, aAB01 = a[0] * AB01;
  const double aAB10 = a[1] * AB10, aAB11 = a[1] * AB11;
  const double bAB00 = b[0] * AB00, bAB10 = b[0] * AB10;
  const double bAB01 = b[1] * AB01, bAB11 = b[1] * AB11;
  double Tba[3];
  mtxv(Tba, Rab, Tab);
  double t, u;
  const double ALLx = -Tba[0];
  const double ALUx = ALLx + aAB10;
  const double AULx = ALLx + aAB00;
  const double AUUx = ALUx + aAB00;
  double LA1lx, LA1ux, UA1lx, UA1ux;
  if (ALLx < ALUx) { LA1lx = ALLx; LA1ux = ALUx; UA1lx = AULx; UA1ux = AUUx; }
  else { LA1lx = ALUx; LA1ux = ALLx; UA1lx = AUUx; UA1ux = AULx; }
  const double BLLx = Tab[0];
  const double BLUx = BLLx + bAB01;
  const double BULx = BLLx + bAB00;
  const double BUUx = BLUx + bAB00;
  double LB1lx, LB1ux, UB1lx, UB1ux;
  if (BLLx < BLUx) { LB1lx = BLLx; LB1ux = BLUx; UB1lx = BULx; UB1ux = BUUx; }
  else { LB1lx = BLUx; LB1ux = BLLx; UB1lx = BUUx; UB1ux = BULx; }

  if ((UA1ux > b[0]) && (UB1ux > a[0]) &&
      ((UA1lx > b[0]) || voronoi(b[1], a[1], AB10, aAB00 - b[0] - Tba[0], AB11, aAB01 - Tba[1],
                                 -Tab[1] - bAB10)) &&
      ((UB1lx > a[0]) || voronoi(a[1], b[1], AB01, Tab[0] + bAB00 - a[0], AB11, Tab[1] + bAB10,
                                 Tba[1] - aAB01))) {
    seg_params(t, u, a[1], b[1], AB11, Tab[1] + bAB10, Tba[1] - aAB01);
    return rect_edge_tail(a[0], t, Rab, Tab, 0, b[0], true, 1, u, S);
  }
  if ((UA1lx < 0) && (LB1ux > a[0]) &&
      ((UA1ux < 0) || voronoi(b[1], a[1], -AB10, Tba[0] - aAB00, AB11, aAB01 - Tba[1], -Tab[1])) &&
      ((LB1lx > a[0]) || voronoi(a[1], b[1], AB01, Tab[0] - a[0], AB11, Tab[1], Tba[1] - aAB01))) {
    seg_params(t, u, a[1], b[1], AB11, Tab[1], Tba[1] - aAB01);
    return rect_edge_tail(a[0], t, Rab, Tab, 0, 0.0, false, 1, u, S);
  }
  if ((LA1ux > b[0]) && (UB1lx < 0) &&
      ((LA1lx > b[0]) || voronoi(b[1], a[1], AB10, -Tba[0] - b[0], AB11, -Tba[1], -Tab[1] - bAB10)) &&
      ((UB1ux < 0) || voronoi(a[1], b[1], -AB01, -Tab[0] - bAB00, AB11, Tab[1] + bAB10, Tba[1]))) {
    seg_params(t, u, a[1], b[1], AB11, Tab[1] + bAB10, Tba[1]);
    return rect_edge_tail(0, t, Rab, Tab, 0, b[0], true, 1, u, S);
  }
  if ((LA1lx < 0) && (LB1lx < 0) &&
      ((LA1ux < 0) || voronoi(b[1], a[1], -AB10, Tba[0], AB11, -Tba[1], -Tab[1])) &&
      ((LB1ux < 0) || voronoi(a[1], b[1], -AB01, -Tab[0], AB11, Tab[1], Tba[1]))) {
    seg_params(t, u, a[1], b[1], AB11, Tab[1], Tba[1]);
    return rect_edge_tail(0, t, Rab, Tab, 0, 0.0, false, 1, u, S);
  }

  const double ALLy = -Tba[1];
  const double ALUy = ALLy + aAB11;
  const double AULy = ALLy + aAB01;
  const double AUUy = ALUy + aAB01;
  double LA1ly, LA1uy, UA1ly, UA1uy, LB0lx, LB0ux, UB0lx, UB0ux;
  if (ALLy < ALUy) { LA1ly = ALLy; LA1uy = ALUy; UA1ly = AULy; UA1uy = AUUy; }
  else { LA1ly = ALUy; LA1uy = ALLy; UA1ly = AUUy; UA1uy = AULy; }
  if (BLLx < BULx) { LB0lx = BLLx; LB0ux = BULx; UB0lx = BLUx; UB0ux = BUUx; }
  else { LB0lx = BULx; LB0ux = BLLx; UB0lx = BUUx; UB0ux = BLUx; }

  if ((UA1uy > b[1]) && (UB0ux > a[0]) &&
      ((UA1ly > b[1]) || voronoi(b[0], a[1], AB11, aAB01 - Tba[1] - b[1], AB10, aAB00 - Tba[0],
                                 -Tab[1] - bAB11)) &&
      ((UB0lx > a[0]) || voronoi(a[1], b[0], AB00, Tab[0] - a[0] + bAB01, AB10, Tab[1] + bAB11,
                                 Tba[0] - aAB00))) {
    seg_params(t, u, a[1], b[0], AB10, Tab[1] + bAB11, Tba[0] - aAB00);
    return rect_edge_tail(a[0], t, Rab, Tab, 1, b[1], true, 0, u, S);
  }
  if ((UA1ly < 0) && (LB0ux > a[0]) &&
      ((UA1uy < 0) || voronoi(b[0], a[1], -AB11, Tba[1] - aAB01, AB10, aAB00 - Tba[0], -Tab[1])) &&
      ((LB0lx > a[0]) || voronoi(a[1], b[0], AB00, Tab[0] - a[0], AB10, Tab[1], Tba[0] - aAB00))) {
    seg_params(t, u, a[1], b[0], AB10, Tab[1], Tba[0] - aAB00);
    return rect_edge_tail(a[0], t, Rab, Tab, 1, 0.0, false, 0, u, S);
  }
  if ((LA1uy > b[1]) && (UB0lx < 0) &&
      ((LA1ly > b[1]) || voronoi(b[0], a[1], AB11, -Tba[1] - b[1], AB10, -Tba[0], -Tab[1] - bAB11)) &&
      ((UB0ux < 0) || voronoi(a[1], b[0], -AB00, -Tab[0] - bAB01, AB10, Tab[1] + bAB11, Tba[0]))) {
    seg_params(t, u, a[1], b[0], AB10, Tab[1] + bAB11, Tba[0]);
    return rect_edge_tail(0, t, Rab, Tab, 1, b[1], true, 0, u, S);
  }
  if ((LA1ly < 0) && (LB0lx < 0) &&
      ((LA1uy < 0) || voronoi(b[0], a[1], -AB11, Tba[1], AB10, -Tba[0], -Tab[1])) &&
      ((LB0ux < 0) || voronoi(a[1], b[0], -AB00, -Tab[0], AB10, Tab[1], Tba[0]))) {
    seg_params(t, u, a[1], b[0], AB10, Tab[1], Tba[0]);
    return rect_edge_tail(0, t, Rab, Tab, 1, 0.0, false, 0, u, S);
  }

  const double BLLy = Tab[1];
  const double BLUy = BLLy + bAB11;
  const double BULy = BLLy + bAB10;
  const double BUUy = BLUy + bAB10;
  double LA0lx, LA0ux, UA0lx, UA0ux, LB1ly, LB1uy, UB1ly, UB1uy;
  if (ALLx < AULx) { LA0lx = ALLx; LA0ux = AULx; UA0lx = ALUx; UA0ux = AUUx; }
  else { LA0lx = AULx; LA0ux = ALLx; UA0lx = AUUx; UA0ux = ALUx; }
  if (BLLy < BLUy) { LB1ly = BLLy; LB1uy = BLUy; UB1ly = BULy; UB1uy = BUUy; }
  else { LB1ly = BLUy; LB1uy = BLLy; UB1ly = BUUy; UB1uy = BULy; }

  if ((UA0ux > b[0]) && (UB1uy > a[1]) &&
      ((UA0lx > b[0]) || voronoi(b[1], a[0], AB00, aAB10 - Tba[0] - b[0], AB01, aAB11 - Tba[1],
                                 -Tab[0] - bAB00)) &&
      ((UB1ly > a[1]) || voronoi(a[0], b[1], AB11, Tab[1] - a[1] + bAB10, AB01, Tab[0] + bAB00,
                                 Tba[1] - aAB11))) {
    seg_params(t, u, a[0], b[1], AB01, Tab[0] + bAB00, Tba[1] - aAB11);
    return rect_edge_tail(t, a[1], Rab, Tab, 0, b[0], true, 1, u, S);
  }
  if ((UA0lx < 0) && (LB1uy > a[1]) &&
      ((UA0ux < 0) || voronoi(b[1], a[0], -AB00, Tba[0] - aAB10, AB01, aAB11 - Tba[1], -Tab[0])) &&
      ((LB1ly > a[1]) || voronoi(a[0], b[1], AB11, Tab[1] - a[1], AB01, Tab[0], Tba[1] - aAB11))) {
    seg_params(t, u, a[0], b[1], AB01, Tab[0], Tba[1] - aAB11);
    return rect_edge_tail(t, a[1], Rab, Tab, 0, 0.0, false, 1, u, S);
  }
  if ((LA0ux > b[0]) && (UB1ly < 0) &&
      ((LA0lx > b[0]) || voronoi(b[1], a[0], AB00, -b[0] - Tba[0], AB01, -Tba[1], -bAB00 - Tab[0])) &&
      ((UB1uy < 0) || voronoi(a[0], b[1], -AB11, -Tab[1] - bAB10, AB01, Tab[0] + bAB00, Tba[1]))) {
    seg_params(t, u, a[0], b[1], AB01, Tab[0] + bAB00, Tba[1]);
    return rect_edge_tail(t, 0, Rab, Tab, 0, b[0], true, 1, u, S);
  }
  if ((LA0lx < 0) && (LB1ly < 0) &&
      ((LA0ux < 0) || voronoi(b[1], a[0], -AB00, Tba[0], AB01, -Tba[1], -Tab[0])) &&
      ((LB1uy < 0) || voronoi(a[0], b[1], -AB11, -Tab[1], AB01, Tab[0], Tba[1]))) {
    seg_params(t, u, a[0], b[1], AB01, Tab[0], Tba[1]);
    return rect_edge_tail(t, 0, Rab, Tab, 0, 0.0, false, 1, u, S);
  }

  double LA0ly, LA0uy, UA0ly, UA0uy, LB0ly, LB0uy, UB0ly, UB0uy;
  if (ALLy < AULy) { LA0ly = ALLy; LA0uy = AULy; UA0ly = ALUy; UA0uy = AUUy; }
  else { LA0ly = AULy; LA0uy = ALLy; UA0ly = AUUy; UA0uy = ALUy; }
  if (BLLy < BULy) { LB0ly = BLLy; LB0uy = BULy; UB0ly = BLUy; UB0uy = BUUy; }
  else { LB0ly = BULy; LB0uy = BLLy; UB0ly = BUUy; UB0uy = BLUy; }

  if ((UA0uy > b[1]) && (UB0uy > a[1]) &&
      ((UA0ly > b[1]) || voronoi(b[0], a[0], AB01, aAB11 - Tba[1] - b[1], AB00, aAB10 - Tba[0],
                                 -Tab[0] - bAB01)) &&
      ((UB0ly > a[1]) || voronoi(a[0], b[0], AB10, Tab[1] - a[1] + bAB11, AB00, Tab[0] + bAB01,
                                 Tba[0] - aAB10))) {
    seg_params(t, u, a[0], b[0], AB00, Tab[0] + bAB01, Tba[0] - aAB10);
    return rect_edge_tail(t, a[1], Rab, Tab, 1, b[1], true, 0, u, S);
  }
  if ((UA0ly < 0) && (LB0uy > a[1]) &&
      ((UA0uy < 0) || voronoi(b[0], a[0], -AB01, Tba[1] - aAB11, AB00, aAB10 - Tba[0], -Tab[0])) &&
      ((LB0ly > a[1]) || voronoi(a[0], b[0], AB10, Tab[1] - a[1], AB00, Tab[0], Tba[0] - aAB10))) {
    seg_params(t, u, a[0], b[0], AB00, Tab[0], Tba[0] - aAB10);
    return rect_edge_tail(t, a[1], Rab, Tab, 1, 0.0, false, 0, u, S);
  }
  if ((LA0uy > b[1]) && (UB0ly < 0) &&
      ((LA0ly > b[1]) || voronoi(b[0], a[0], AB01, -Tba[1] - b[1], AB00, -Tba[0], -Tab[0] - bAB01)) &&
      ((UB0uy < 0) || voronoi(a[0], b[0], -AB10, -Tab[1] - bAB11, AB00, Tab[0] + bAB01, Tba[0]))) {
    seg_params(t, u, a[0], b[0], AB00, Tab[0] + bAB01, Tba[0]);
    return rect_edge_tail(t, 0, Rab, Tab, 1, b[1], true, 0, u, S);
  }
  if ((LA0ly < 0) && (LB0ly < 0) &&
      ((LA0uy < 0) || voronoi(b[0], a[0], -AB01, Tba[1], AB00, -Tba[0], -Tab[0])) &&
      ((LB0uy < 0) || voronoi(a[0], b[0], -AB10, -Tab[1], AB00, Tab[0], Tba[0]))) {
    seg_params(t, u, a[0], b[0], AB00, Tab[0], Tba[0]);
    return rect_edge_tail(t, 0, Rab, Tab, 1, 0.0, false, 0, u, S);
  }

  // face-normal separation (C2A_RectDist.h:855-933)
  double sep1, sep2;
  if (Tab[2] > 0.0) {
    sep1 = Tab[2];
    if (Rab[6] < 0.0) sep1 += b[0] * Rab[6];
    if (Rab[7] < 0.0) sep1 += b[1] * Rab[7];
  } else {
    sep1 = -Tab[2];
    if (Rab[6] > 0.0) sep1 -= b[0] * Rab[6];
    if (Rab[7] > 0.0) sep1 -= b[1] * Rab[7];
  }
  if (Tba[2] < 0) {
    sep2 = -Tba[2];
    if (Rab[2] < 0.0) sep2 += a[0] * Rab[2];
    if (Rab[5] < 0.0) sep2 += a[1] * Rab[5];
  } else {
    sep2 = Tba[2];
    if (Rab[2] > 0.0) sep2 -= a[0] * Rab[2];
    if (Rab[5] > 0.0) sep2 -= a[1] * Rab[5];
  }
  if (sep1 >= sep2 && sep1 >= 0) {
    S[0] = 0.0 - 0.0;
    S[1] = 0.0 - 0.0;
    S[2] = ((Tab[2] > 0.0) ? sep1 : -sep1) - 0.0;
  }
  if (sep2 >= sep1 && sep2 >= 0) {
    double P[3];
    if (Tba[2] < 0) {
      P[0] = Rab[2] * sep2 + Tab[0];
      P[1] = Rab[5] * sep2 + Tab[1];
      P[2] = Rab[8] * sep2 + Tab[2];
    } else {
      P[0] = -Rab[2] * sep2 + Tab[0];
      P[1] = -Rab[5] * sep2 + Tab[1];
      P[2] = -Rab[8] * sep2 + Tab[2];
    }
    S[0] = Tab[0] - P[0]; S[1] = Tab[1] - P[1]; S[2] = Tab[2] - P[2];
  }
  double sep = (sep1 > sep2 ? sep1 : sep2);
  return (sep > 0 ? sep : 0);
}

// ---------------------------------------------------------------------------
// Triangle-triangle distance (C2A.cpp:59-405 restating PQP TriDist).
// Feature codes: type 0 vertex, 1 edge, 2 face; -1 unset.
// ---------------------------------------------------------------------------
C2A_HD void seg_points(double* VEC, double* X, double* Y, const double* P, const double* A,
                       const double* Q, const double* B) {
  double T[3], TMP[3];
  sub3(T, Q, P);
  double AA = dot3(A, A), BB = dot3(B, B), AB = dot3(A, B);
  double AT = dot3(A, T), BT = dot3(B, T);
  double denom = AA * BB - AB * AB;
  double t = (AT * BB - BT * AB) / denom;
  if ((t < 0) || isnan(t)) t = 0;
  else if (t > 1) t = 1;
  double u = (t * AB - BT) / BB;
  if ((u <= 0) || isnan(u)) {
    Y[0] = Q[0]; Y[1] = Q[1]; Y[2] = Q[2];
    t = AT / AA;
    if ((t <= 0) || isnan(t)) {
      X[0] = P[0]; X[1] = P[1]; X[2] = P[2];
      sub3(VEC, Q, P);
    } else if (t >= 1) {
      add3(X, P, A);
      sub3(VEC, Q, X);
    } else {
      axpy3(X, P, A, t);
      cross3(TMP, T, A);
      cross3(VEC, A, TMP);
    }
  } else if (u >= 1) {
    add3(Y, Q, B);
    t = (AB + AT) / AA;
    if ((t <= 0) || isnan(t)) {
      X[0] = P[0]; X[1] = P[1]; X[2] = P[2];
      sub3(VEC, Y, P);
    } else if (t >= 1) {
      add3(X, P, A);
      sub3(VEC, Y, X);
    } else {
      axpy3(X, P, A, t);
      sub3(T, Y, P);
      cross3(TMP, T, A);
      cross3(VEC, A, TMP);
    }
  } else {
    axpy3(Y, Q, B, u);
    if ((t <= 0) || isnan(t)) {
      X[0] = P[0]; X[1] = P[1]; X[2] = P[2];
      cross3(TMP, T, B);
      cross3(VEC, B, TMP);
    } else if (t >= 1) {
      add3(X, P, A);
      sub3(T, Q, X);
      cross3(TMP, T, B);
      cross3(VEC, B, TMP);
    } else {
      axpy3(X, P, A, t);
      cross3(VEC, A, B);
      if (dot3(VEC, T) < 0) { VEC[0] = VEC[0] * -1; VEC[1] = VEC[1] * -1; VEC[2] = VEC[2] * -1; }
    }
  }
}

// vertex of triangle `tt` closest to the face of `ts` (or -1)
C2A_HD int tri_vertex_face(const double* ts, const double* sv, const double* tt, double* Nrm,
                           double& Nl, double* proj, int& disjoint) {
  double V[3], Z[3];
  cross3(Nrm, sv + 0, sv + 3);
  Nl = dot3(Nrm, Nrm);
  if (!(Nl > 1e-15)) return -1;
  sub3(V, ts, tt + 0); proj[0] = dot3(V, Nrm);
  sub3(V, ts, tt + 3); proj[1] = dot3(V, Nrm);
  sub3(V, ts, tt + 6); proj[2] = dot3(V, Nrm);
  int point = -1;
  if ((proj[0] > 0) && (proj[1] > 0) && (proj[2] > 0)) {
    point = (proj[0] < proj[1]) ? 0 : 1;
    if (proj[2] < proj[point]) point = 2;
  } else if ((proj[0] < 0) && (proj[1] < 0) && (proj[2] < 0)) {
    point = (proj[0] > proj[1]) ? 0 : 1;
    if (proj[2] > proj[point]) point = 2;
  }
  if (point < 0) return -1;
  disjoint = 1;
  for (int e = 0; e < 3; e++) {
    sub3(V, tt + 3 * point, ts + 3 * e);
    cross3(Z, Nrm, sv + 3 * e);
    if (!(dot3(V, Z) > 0)) return -1;
  }
  return point;
}

// S, T: 3 vertices each (9 doubles).  P, Q: closest points.
C2A_HD double tri_dist(double* P, double* Q, const double* S, const double* T, int* feat) {
  double Sv[9], Tv[9], VEC[3], V[3], Z[3], minP[3], minQ[3];
  int ta = -1, fa = -1, tb = -1, fb = -1;
  sub3(Sv + 0, S + 3, S + 0); sub3(Sv + 3, S + 6, S + 3); sub3(Sv + 6, S + 0, S + 6);
  sub3(Tv + 0, T + 3, T + 0); sub3(Tv + 3, T + 6, T + 3); sub3(Tv + 6, T + 0, T + 6);
  int shown_disjoint = 0;
  double mindd = dist2_3(S, T) + 1;
  double result = -1;
  bool done = false;
  for (int i = 0; i < 3 && !done; i++) {
    for (int j = 0; j < 3; j++) {
      seg_points(VEC, P, Q, S + 3 * i, Sv + 3 * i, T + 3 * j, Tv + 3 * j);
      sub3(V, Q, P);
      double dd = dot3(V, V);
      if (dd <= mindd) {
        minP[0] = P[0]; minP[1] = P[1]; minP[2] = P[2];
        minQ[0] = Q[0]; minQ[1] = Q[1]; minQ[2] = Q[2];
        mindd = dd;
        ta = 1; fa = i; tb = 1; fb = j;
        sub3(Z, S + 3 * ((i + 2) % 3), P);
        double a = dot3(Z, VEC);
        sub3(Z, T + 3 * ((j + 2) % 3), Q);
        double b = dot3(Z, VEC);
        if ((a <= 0) && (b >= 0)) { result = sqrt(dd); done = true; break; }
        double p = dot3(V, VEC);
        if (a < 0) a = 0;
        if (b > 0) b = 0;
        if ((p - a + b) > 0) shown_disjoint = 1;
      }
    }
  }
  if (!done) {
    double Nrm[3], Nl, proj[3];
    int pt = tri_vertex_face(S, Sv, T, Nrm, Nl, proj, shown_disjoint);
    if (pt >= 0) {
      ta = 2; fa = -1; tb = 0; fb = pt;
      axpy3(P, T + 3 * pt, Nrm, proj[pt] / Nl);
      Q[0] = T[3 * pt]; Q[1] = T[3 * pt + 1]; Q[2] = T[3 * pt + 2];
      result = sqrt(dist2_3(P, Q));
    } else {
      pt = tri_vertex_face(T, Tv, S, Nrm, Nl, proj, shown_disjoint);
      if (pt >= 0) {
        ta = 0; fa = pt; tb = 2; fb = -1;
        P[0] = S[3 * pt]; P[1] = S[3 * pt + 1]; P[2] = S[3 * pt + 2];
        axpy3(Q, S + 3 * pt, Nrm, proj[pt] / Nl);
        result = sqrt(dist2_3(P, Q));
      } else if (shown_disjoint) {
        P[0] = minP[0]; P[1] = minP[1]; P[2] = minP[2];
        Q[0] = minQ[0]; Q[1] = minQ[1]; Q[2] = minQ[2];
        result = sqrt(mindd);
      } else {
        result = 0;
      }
    }
  }
  if (feat) { feat[0] = ta; feat[1] = fa; feat[2] = tb; feat[3] = fb; }
  return result;
}

// C2A_TriDistance: triangle 2 mapped into model 1's frame by (R, T)
C2A_HD double tri_distance(const double* R, const double* T, const double* t1, const double* t2,
                           double* p, double* q, int* feat) {
  double b[9];
  mxvpv(b + 0, R, t2 + 0, T);
  mxvpv(b + 3, R, t2 + 3, T);
  mxvpv(b + 6, R, t2 + 6, T);
  return tri_dist(p, q, t1, b, feat);
}

}  // namespace c2a
