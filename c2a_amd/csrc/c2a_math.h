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
// Forced inlining: the CCD kernel is one persistent loop, and an outlined
// call in it costs a scratch-resident frame and register save/restore.
#define C2A_HD __host__ __device__ inline __attribute__((always_inline))
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
  // The axis the motion bounds cross with world-frame directions: `axis`
  // itself (the reference, InterpMotion.cpp:838-845, 784-791) or, in strict
  // mode, the world-frame axis R0 * axis (q(t) = q0 * q_rel(t) turns about
  // the body-frame `axis`, i.e. about R0 * axis in world coordinates).
  double bax[3];
};

// CInterpMotion_Linear ctor + velocity() + LinearAngularVelocity
// (InterpMotion.cpp:148-168, 486-491, 228-270).  Host only (acos).
inline void motion_init(Motion* m, const double* R0, const double* T0, const double* R1,
                        const double* T1, bool strict = false) {
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
  if (strict) mxv(m->bax, m->R0, m->axis);
  else for (int i = 0; i < 3; i++) m->bax[i] = m->axis[i];
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

// Admissible time a tested leaf pair contributes to mint.  Reference: only
// an improving leaf (dTri <= distance, C2A.cpp:1150) contributes, and a
// touching one (dTri == 0: zero-length normalise, NaN) does not.  Strict:
// every tested leaf contributes, a touching one with 0.
C2A_HD double leaf_time(double dTri, double mleaf, int strict) {
  return (strict && dTri == 0.0) ? 0.0 : mleaf;
}

// computeTOC_MotionBound (InterpMotion.cpp:831-881); N is normalised in place.
C2A_HD double motion_bound(const Motion& m, double angR, double* N) {
  double cr[3];
  normalize3(N);
  cross3(cr, m.bax, N);
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
    double cwc[3] = {m.bax[0] * m.angVel, m.bax[1] * m.angVel, m.bax[2] * m.angVel};
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

// True when any active lane of the wavefront holds p (host: the lane itself).
C2A_HD bool wave_any(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __ballot(p) != 0;
#else
  return p;
#endif
}

// Ordered (lo, hi) of two values the way the reference builds its edge
// extents: hi/lo of (x, y) plus the matching pair (z, w).
C2A_HD void order2(double x, double y, double z, double w, double& lo, double& hi, double& zlo,
                   double& zhi) {
  if (x < y) { lo = x; hi = y; zlo = z; zhi = w; }
  else { lo = y; hi = x; zlo = w; zhi = z; }
}

// The 16 edge-pair cases are tested in the reference's order; the first one
// whose extent pre-test and two Voronoi tests hold gives the closest points.
// The loop over cases is uniform across the wavefront (each case's
// parameters are formed by a wave-uniform switch) and stops once every lane
// has its case, so lanes working on different pairs do not serialize the
// sixteen case bodies.  Each case's expressions are the reference's.
C2A_HD double rect_dist(const double* Rab, const double* Tab, const double* a, const double* b,
                        double* S) {
  const double AB00 = Rab[0], AB01 = Rab[1], AB10 = Rab[3], AB11 = Rab[4];
  const double aAB00 = a[0] * AB00, aAB01 = a[0] * AB01;
  const double aAB10 = a[1] * AB10, aAB11 = a[1] * AB11;
  const double bAB00 = b[0] * AB00, bAB10 = b[0] * AB10;
  const double bAB01 = b[1] * AB01, bAB11 = b[1] * AB11;
  double Tba[3];
  mtxv(Tba, Rab, Tab);
  // corner coordinates of A in B's frame and of B in A's frame
  const double ALLx = -Tba[0];
  const double ALUx = ALLx + aAB10;
  const double AULx = ALLx + aAB00;
  const double AUUx = ALUx + aAB00;
  const double BLLx = Tab[0];
  const double BLUx = BLLx + bAB01;
  const double BULx = BLLx + bAB00;
  const double BUUx = BLUx + bAB00;
  const double ALLy = -Tba[1];
  const double ALUy = ALLy + aAB11;
  const double AULy = ALLy + aAB01;
  const double AUUy = ALUy + aAB01;
  const double BLLy = Tab[1];
  const double BLUy = BLLy + bAB11;
  const double BULy = BLLy + bAB10;
  const double BUUy = BLUy + bAB10;

  int win = -1;
  double g_a = 0, g_b = 0, g_AB = 0, g_AT = 0, g_BT = 0;   // segment parameters
  double g_p = 0;                                          // fixed coordinate of P
  int g_tx = 0;                                            // 1: P = (t, p), 0: P = (p, t)
  int g_cf = 0, g_cv = 0;
  double g_lf = 0;
  bool g_hf = false;
  for (int c = 0; c < 16; c++) {
    if (!wave_any(win < 0)) break;
    double lo1, hi1, lo2, hi2, ulo1, uhi1, ulo2, uhi2;
    bool pre = false, s1 = false, s2 = false;
    double v1[7], v2[7], sg[5];
    double pfix = 0;
    int tx = 0, cf = 0, cv = 0;
    double lf = 0;
    bool hf = false;
    if (c < 4) {
      order2(ALLx, ALUx, AULx, AUUx, lo1, hi1, ulo1, uhi1);   // LA1 x-extent, UA1 x-extent
      order2(BLLx, BLUx, BULx, BUUx, lo2, hi2, ulo2, uhi2);   // LB1, UB1
    } else if (c < 8) {
      order2(ALLy, ALUy, AULy, AUUy, lo1, hi1, ulo1, uhi1);   // LA1 y, UA1 y
      order2(BLLx, BULx, BLUx, BUUx, lo2, hi2, ulo2, uhi2);   // LB0 x, UB0 x
    } else if (c < 12) {
      order2(ALLx, AULx, ALUx, AUUx, lo1, hi1, ulo1, uhi1);   // LA0 x, UA0 x
      order2(BLLy, BLUy, BULy, BUUy, lo2, hi2, ulo2, uhi2);   // LB1 y, UB1 y
    } else {
      order2(ALLy, AULy, ALUy, AUUy, lo1, hi1, ulo1, uhi1);   // LA0 y, UA0 y
      order2(BLLy, BULy, BLUy, BUUy, lo2, hi2, ulo2, uhi2);   // LB0 y, UB0 y
    }
#define C2A_V(o, A0, A1, A2, A3, A4, A5, A6) \
  o[0] = (A0); o[1] = (A1); o[2] = (A2); o[3] = (A3); o[4] = (A4); o[5] = (A5); o[6] = (A6)
#define C2A_G(A0, A1, A2, A3, A4) sg[0] = (A0); sg[1] = (A1); sg[2] = (A2); sg[3] = (A3); sg[4] = (A4)
    switch (c) {
      case 0:   // A x=a0 edge, B b0 edge (both along the 1-axes)
        pre = (uhi1 > b[0]) && (uhi2 > a[0]); s1 = (ulo1 > b[0]); s2 = (ulo2 > a[0]);
        C2A_V(v1, b[1], a[1], AB10, aAB00 - b[0] - Tba[0], AB11, aAB01 - Tba[1], -Tab[1] - bAB10);
        C2A_V(v2, a[1], b[1], AB01, Tab[0] + bAB00 - a[0], AB11, Tab[1] + bAB10, Tba[1] - aAB01);
        C2A_G(a[1], b[1], AB11, Tab[1] + bAB10, Tba[1] - aAB01);
        pfix = a[0]; tx = 0; cf = 0; lf = b[0]; hf = true; cv = 1;
        break;
      case 1:
        pre = (ulo1 < 0) && (hi2 > a[0]); s1 = (uhi1 < 0); s2 = (lo2 > a[0]);
        C2A_V(v1, b[1], a[1], -AB10, Tba[0] - aAB00, AB11, aAB01 - Tba[1], -Tab[1]);
        C2A_V(v2, a[1], b[1], AB01, Tab[0] - a[0], AB11, Tab[1], Tba[1] - aAB01);
        C2A_G(a[1], b[1], AB11, Tab[1], Tba[1] - aAB01);
        pfix = a[0]; tx = 0; cv = 1;
        break;
      case 2:
        pre = (hi1 > b[0]) && (ulo2 < 0); s1 = (lo1 > b[0]); s2 = (uhi2 < 0);
        C2A_V(v1, b[1], a[1], AB10, -Tba[0] - b[0], AB11, -Tba[1], -Tab[1] - bAB10);
        C2A_V(v2, a[1], b[1], -AB01, -Tab[0] - bAB00, AB11, Tab[1] + bAB10, Tba[1]);
        C2A_G(a[1], b[1], AB11, Tab[1] + bAB10, Tba[1]);
        pfix = 0; tx = 0; cf = 0; lf = b[0]; hf = true; cv = 1;
        break;
      case 3:
        pre = (lo1 < 0) && (lo2 < 0); s1 = (hi1 < 0); s2 = (hi2 < 0);
        C2A_V(v1, b[1], a[1], -AB10, Tba[0], AB11, -Tba[1], -Tab[1]);
        C2A_V(v2, a[1], b[1], -AB01, -Tab[0], AB11, Tab[1], Tba[1]);
        C2A_G(a[1], b[1], AB11, Tab[1], Tba[1]);
        pfix = 0; tx = 0; cv = 1;
        break;
      case 4:   // A 1-axis edges against B 0-axis edges
        pre = (uhi1 > b[1]) && (uhi2 > a[0]); s1 = (ulo1 > b[1]); s2 = (ulo2 > a[0]);
        C2A_V(v1, b[0], a[1], AB11, aAB01 - Tba[1] - b[1], AB10, aAB00 - Tba[0], -Tab[1] - bAB11);
        C2A_V(v2, a[1], b[0], AB00, Tab[0] - a[0] + bAB01, AB10, Tab[1] + bAB11, Tba[0] - aAB00);
        C2A_G(a[1], b[0], AB10, Tab[1] + bAB11, Tba[0] - aAB00);
        pfix = a[0]; tx = 0; cf = 1; lf = b[1]; hf = true; cv = 0;
        break;
      case 5:
        pre = (ulo1 < 0) && (hi2 > a[0]); s1 = (uhi1 < 0); s2 = (lo2 > a[0]);
        C2A_V(v1, b[0], a[1], -AB11, Tba[1] - aAB01, AB10, aAB00 - Tba[0], -Tab[1]);
        C2A_V(v2, a[1], b[0], AB00, Tab[0] - a[0], AB10, Tab[1], Tba[0] - aAB00);
        C2A_G(a[1], b[0], AB10, Tab[1], Tba[0] - aAB00);
        pfix = a[0]; tx = 0; cv = 0;
        break;
      case 6:
        pre = (hi1 > b[1]) && (ulo2 < 0); s1 = (lo1 > b[1]); s2 = (uhi2 < 0);
        C2A_V(v1, b[0], a[1], AB11, -Tba[1] - b[1], AB10, -Tba[0], -Tab[1] - bAB11);
        C2A_V(v2, a[1], b[0], -AB00, -Tab[0] - bAB01, AB10, Tab[1] + bAB11, Tba[0]);
        C2A_G(a[1], b[0], AB10, Tab[1] + bAB11, Tba[0]);
        pfix = 0; tx = 0; cf = 1; lf = b[1]; hf = true; cv = 0;
        break;
      case 7:
        pre = (lo1 < 0) && (lo2 < 0); s1 = (hi1 < 0); s2 = (hi2 < 0);
        C2A_V(v1, b[0], a[1], -AB11, Tba[1], AB10, -Tba[0], -Tab[1]);
        C2A_V(v2, a[1], b[0], -AB00, -Tab[0], AB10, Tab[1], Tba[0]);
        C2A_G(a[1], b[0], AB10, Tab[1], Tba[0]);
        pfix = 0; tx = 0; cv = 0;
        break;
      case 8:   // A 0-axis edges against B 1-axis edges
        pre = (uhi1 > b[0]) && (uhi2 > a[1]); s1 = (ulo1 > b[0]); s2 = (ulo2 > a[1]);
        C2A_V(v1, b[1], a[0], AB00, aAB10 - Tba[0] - b[0], AB01, aAB11 - Tba[1], -Tab[0] - bAB00);
        C2A_V(v2, a[0], b[1], AB11, Tab[1] - a[1] + bAB10, AB01, Tab[0] + bAB00, Tba[1] - aAB11);
        C2A_G(a[0], b[1], AB01, Tab[0] + bAB00, Tba[1] - aAB11);
        pfix = a[1]; tx = 1; cf = 0; lf = b[0]; hf = true; cv = 1;
        break;
      case 9:
        pre = (ulo1 < 0) && (hi2 > a[1]); s1 = (uhi1 < 0); s2 = (lo2 > a[1]);
        C2A_V(v1, b[1], a[0], -AB00, Tba[0] - aAB10, AB01, aAB11 - Tba[1], -Tab[0]);
        C2A_V(v2, a[0], b[1], AB11, Tab[1] - a[1], AB01, Tab[0], Tba[1] - aAB11);
        C2A_G(a[0], b[1], AB01, Tab[0], Tba[1] - aAB11);
        pfix = a[1]; tx = 1; cv = 1;
        break;
      case 10:
        pre = (hi1 > b[0]) && (ulo2 < 0); s1 = (lo1 > b[0]); s2 = (uhi2 < 0);
        C2A_V(v1, b[1], a[0], AB00, -b[0] - Tba[0], AB01, -Tba[1], -bAB00 - Tab[0]);
        C2A_V(v2, a[0], b[1], -AB11, -Tab[1] - bAB10, AB01, Tab[0] + bAB00, Tba[1]);
        C2A_G(a[0], b[1], AB01, Tab[0] + bAB00, Tba[1]);
        pfix = 0; tx = 1; cf = 0; lf = b[0]; hf = true; cv = 1;
        break;
      case 11:
        pre = (lo1 < 0) && (lo2 < 0); s1 = (hi1 < 0); s2 = (hi2 < 0);
        C2A_V(v1, b[1], a[0], -AB00, Tba[0], AB01, -Tba[1], -Tab[0]);
        C2A_V(v2, a[0], b[1], -AB11, -Tab[1], AB01, Tab[0], Tba[1]);
        C2A_G(a[0], b[1], AB01, Tab[0], Tba[1]);
        pfix = 0; tx = 1; cv = 1;
        break;
      case 12:  // A 0-axis edges against B 0-axis edges
        pre = (uhi1 > b[1]) && (uhi2 > a[1]); s1 = (ulo1 > b[1]); s2 = (ulo2 > a[1]);
        C2A_V(v1, b[0], a[0], AB01, aAB11 - Tba[1] - b[1], AB00, aAB10 - Tba[0], -Tab[0] - bAB01);
        C2A_V(v2, a[0], b[0], AB10, Tab[1] - a[1] + bAB11, AB00, Tab[0] + bAB01, Tba[0] - aAB10);
        C2A_G(a[0], b[0], AB00, Tab[0] + bAB01, Tba[0] - aAB10);
        pfix = a[1]; tx = 1; cf = 1; lf = b[1]; hf = true; cv = 0;
        break;
      case 13:
        pre = (ulo1 < 0) && (hi2 > a[1]); s1 = (uhi1 < 0); s2 = (lo2 > a[1]);
        C2A_V(v1, b[0], a[0], -AB01, Tba[1] - aAB11, AB00, aAB10 - Tba[0], -Tab[0]);
        C2A_V(v2, a[0], b[0], AB10, Tab[1] - a[1], AB00, Tab[0], Tba[0] - aAB10);
        C2A_G(a[0], b[0], AB00, Tab[0], Tba[0] - aAB10);
        pfix = a[1]; tx = 1; cv = 0;
        break;
      case 14:
        pre = (hi1 > b[1]) && (ulo2 < 0); s1 = (lo1 > b[1]); s2 = (uhi2 < 0);
        C2A_V(v1, b[0], a[0], AB01, -Tba[1] - b[1], AB00, -Tba[0], -Tab[0] - bAB01);
        C2A_V(v2, a[0], b[0], -AB10, -Tab[1] - bAB11, AB00, Tab[0] + bAB01, Tba[0]);
        C2A_G(a[0], b[0], AB00, Tab[0] + bAB01, Tba[0]);
        pfix = 0; tx = 1; cf = 1; lf = b[1]; hf = true; cv = 0;
        break;
      default:
        pre = (lo1 < 0) && (lo2 < 0); s1 = (hi1 < 0); s2 = (hi2 < 0);
        C2A_V(v1, b[0], a[0], -AB01, Tba[1], AB00, -Tba[0], -Tab[0]);
        C2A_V(v2, a[0], b[0], -AB10, -Tab[1], AB00, Tab[0], Tba[0]);
        C2A_G(a[0], b[0], AB00, Tab[0], Tba[0]);
        pfix = 0; tx = 1; cv = 0;
        break;
    }
#undef C2A_V
#undef C2A_G
    if (win < 0 && pre) {
      bool ok = s1 || voronoi(v1[0], v1[1], v1[2], v1[3], v1[4], v1[5], v1[6]);
      ok = ok && (s2 || voronoi(v2[0], v2[1], v2[2], v2[3], v2[4], v2[5], v2[6]));
      if (ok) {
        win = c;
        g_a = sg[0]; g_b = sg[1]; g_AB = sg[2]; g_AT = sg[3]; g_BT = sg[4];
        g_p = pfix; g_tx = tx; g_cf = cf; g_cv = cv; g_lf = lf; g_hf = hf;
      }
    }
  }
  if (win >= 0) {
    double t, u;
    seg_params(t, u, g_a, g_b, g_AB, g_AT, g_BT);
    return g_tx ? rect_edge_tail(t, g_p, Rab, Tab, g_cf, g_lf, g_hf, g_cv, u, S)
                : rect_edge_tail(g_p, t, Rab, Tab, g_cf, g_lf, g_hf, g_cv, u, S);
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

// PQP's TriContact (declared at C2A_PQP.cpp:441, called by CollideRecurse
// :831; PQP itself is not vendored): the separating-axis test over the two
// face normals, the nine edge-edge cross products and the six in-plane edge
// normals, vertices taken relative to P1.  1 when the projections overlap
// on every axis (touching counts as contact).
C2A_HD bool project6(const double* ax, const double* p1, const double* p2, const double* p3,
                     const double* q1, const double* q2, const double* q3) {
  const double P1 = dot3(ax, p1), P2 = dot3(ax, p2), P3 = dot3(ax, p3);
  const double Q1 = dot3(ax, q1), Q2 = dot3(ax, q2), Q3 = dot3(ax, q3);
  double mx1 = P1 > P2 ? P1 : P2, mn1 = P1 < P2 ? P1 : P2;
  mx1 = mx1 > P3 ? mx1 : P3;
  mn1 = mn1 < P3 ? mn1 : P3;
  double mx2 = Q1 > Q2 ? Q1 : Q2, mn2 = Q1 < Q2 ? Q1 : Q2;
  mx2 = mx2 > Q3 ? mx2 : Q3;
  mn2 = mn2 < Q3 ? mn2 : Q3;
  return !(mn1 > mx2) && !(mn2 > mx1);
}

C2A_HD bool tri_contact(const double* P1, const double* P2, const double* P3, const double* Q1,
                        const double* Q2, const double* Q3) {
  const double p1[3] = {0.0, 0.0, 0.0};
  double p2[3], p3[3], q1[3], q2[3], q3[3];
  double e[3][3], f[3][3], n1[3], m1[3], g[3][3], h[3][3], ef[3];
  sub3(p2, P2, P1);
  sub3(p3, P3, P1);
  sub3(q1, Q1, P1);
  sub3(q2, Q2, P1);
  sub3(q3, Q3, P1);
  sub3(e[0], p2, p1);
  sub3(e[1], p3, p2);
  sub3(e[2], p1, p3);
  sub3(f[0], q2, q1);
  sub3(f[1], q3, q2);
  sub3(f[2], q1, q3);
  cross3(n1, e[0], e[1]);
  cross3(m1, f[0], f[1]);
  for (int i = 0; i < 3; i++) {
    cross3(g[i], e[i], n1);
    cross3(h[i], f[i], m1);
  }
  if (!project6(n1, p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(m1, p1, p2, p3, q1, q2, q3)) return false;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cross3(ef, e[i], f[j]);
      if (!project6(ef, p1, p2, p3, q1, q2, q3)) return false;
    }
  for (int i = 0; i < 3; i++) {
    if (!project6(g[i], p1, p2, p3, q1, q2, q3)) return false;
    if (!project6(h[i], p1, p2, p3, q1, q2, q3)) return false;
  }
  return true;
}

}  // namespace c2a
