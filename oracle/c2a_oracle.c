/*
 * c2a_oracle.c -- FP64 CPU oracle (test infrastructure only; see header).
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off).  Contraction is off so
 * every a*b+c rounds twice, exactly as the GPU path (also built with
 * -ffp-contract=off) evaluates the same expressions.
 */
#include "c2a_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* PQP MatVec operations (PQP is not vendored; these restate its published
 * MatVec.h semantics as used at the C2A call sites, e.g. C2A.cpp:1204-1226,
 * C2A_Build.cpp:426-537).  Sums are evaluated left to right.               */
/* ------------------------------------------------------------------------ */
typedef double m3[3][3];

static inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void sub3(double r[3], const double a[3], const double b[3]) {
  r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
static inline void add3(double r[3], const double a[3], const double b[3]) {
  r[0] = a[0] + b[0]; r[1] = a[1] + b[1]; r[2] = a[2] + b[2];
}
static inline void cross3(double r[3], const double a[3], const double b[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void scale3(double r[3], const double a[3], double s) {
  r[0] = a[0] * s; r[1] = a[1] * s; r[2] = a[2] * s;
}
/* r = a + b*s  (VpVxS) */
static inline void axpy3(double r[3], const double a[3], const double b[3], double s) {
  r[0] = a[0] + b[0] * s; r[1] = a[1] + b[1] * s; r[2] = a[2] + b[2] * s;
}
static inline double len3(const double a[3]) { return sqrt(dot3(a, a)); }
static inline double dist2_3(const double a[3], const double b[3]) {
  double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  return d0 * d0 + d1 * d1 + d2 * d2;
}
/* Vnormalize: multiply by the reciprocal length (no zero guard). */
static inline void normalize3(double v[3]) {
  double d = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] *= d; v[1] *= d; v[2] *= d;
}
/* r = M v */
static inline void mxv(double r[3], const m3 M, const double v[3]) {
  double t0 = M[0][0] * v[0] + M[0][1] * v[1] + M[0][2] * v[2];
  double t1 = M[1][0] * v[0] + M[1][1] * v[1] + M[1][2] * v[2];
  double t2 = M[2][0] * v[0] + M[2][1] * v[1] + M[2][2] * v[2];
  r[0] = t0; r[1] = t1; r[2] = t2;
}
/* r = M^T v */
static inline void mtxv(double r[3], const m3 M, const double v[3]) {
  double t0 = M[0][0] * v[0] + M[1][0] * v[1] + M[2][0] * v[2];
  double t1 = M[0][1] * v[0] + M[1][1] * v[1] + M[2][1] * v[2];
  double t2 = M[0][2] * v[0] + M[1][2] * v[1] + M[2][2] * v[2];
  r[0] = t0; r[1] = t1; r[2] = t2;
}
/* r = M v + w */
static inline void mxvpv(double r[3], const m3 M, const double v[3], const double w[3]) {
  double t0 = M[0][0] * v[0] + M[0][1] * v[1] + M[0][2] * v[2] + w[0];
  double t1 = M[1][0] * v[0] + M[1][1] * v[1] + M[1][2] * v[2] + w[1];
  double t2 = M[2][0] * v[0] + M[2][1] * v[1] + M[2][2] * v[2] + w[2];
  r[0] = t0; r[1] = t1; r[2] = t2;
}
/* R = A B */
static inline void mxm(m3 R, const m3 A, const m3 B) {
  m3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
  memcpy(R, t, sizeof(m3));
}
/* R = A^T B */
static inline void mtxm(m3 R, const m3 A, const m3 B) {
  m3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[i][j] = A[0][i] * B[0][j] + A[1][i] * B[1][j] + A[2][i] * B[2][j];
  memcpy(R, t, sizeof(m3));
}

/* ------------------------------------------------------------------------ */
/* sin/cos.  integrate() needs sin/cos of 0.5*angVel*t; the GPU evaluates the
 * same reduction and polynomials with the same rounding so the two sides
 * agree bit for bit.  Method: Cody-Waite reduction by pi/2 split in three
 * parts (each part carries 33 significant bits so k*part is exact for
 * |k| < 2^20), then Taylor series of sin and cos on |r| <= pi/4 truncated
 * after the r^17 / r^18 terms (truncation < 1e-19).                         */
/* ------------------------------------------------------------------------ */
static int g_libm_trig = 0;
void c2o_set_libm_trig(int use_libm) { g_libm_trig = use_libm; }

#define C2O_TWO_OVER_PI 0.6366197723675814
#define C2O_PIO2_A 1.5707963267341256
#define C2O_PIO2_B 6.077100506303966e-11
#define C2O_PIO2_C 2.0222662487959506e-21

static void sincos_poly(double r, double *s, double *c) {
  double r2 = r * r;
  /* sin r = r + r^3 * (-1/3! + r^2/5! - r^4/7! + ... + r^14/17!) */
  double ps = 2.8114572543455207632e-15;          /*  1/17! */
  ps = ps * r2 + -7.6471637318198164759e-13;      /* -1/15! */
  ps = ps * r2 + 1.6059043836821614599e-10;       /*  1/13! */
  ps = ps * r2 + -2.5052108385441718775e-08;      /* -1/11! */
  ps = ps * r2 + 2.7557319223985890653e-06;       /*  1/9!  */
  ps = ps * r2 + -1.9841269841269841270e-04;      /* -1/7!  */
  ps = ps * r2 + 8.3333333333333333333e-03;       /*  1/5!  */
  ps = ps * r2 + -1.6666666666666666667e-01;      /* -1/3!  */
  *s = r + (r * r2) * ps;
  /* cos r = 1 - r^2/2 + r^4 * (1/4! - r^2/6! + ... + r^14/18!) */
  double pc = 1.5619206968586225796e-16;          /*  1/18! */
  pc = pc * r2 + -4.7794773323873852974e-14;      /* -1/16! */
  pc = pc * r2 + 1.1470745597729724714e-11;       /*  1/14! */
  pc = pc * r2 + -2.0876756987868098979e-09;      /* -1/12! */
  pc = pc * r2 + 2.7557319223985890653e-07;       /*  1/10! */
  pc = pc * r2 + -2.4801587301587301587e-05;      /* -1/8!  */
  pc = pc * r2 + 1.3888888888888888889e-03;       /*  1/6!  */
  pc = pc * r2 + -4.1666666666666666667e-02;      /* -1/4!  */
  *c = (1.0 - 0.5 * r2) - (r2 * r2) * pc;
}

void c2o_sincos(double x, double *s, double *c) {
  if (g_libm_trig) { *s = sin(x); *c = cos(x); return; }
  double k = nearbyint(x * C2O_TWO_OVER_PI);
  double r = x - k * C2O_PIO2_A;
  r = r - k * C2O_PIO2_B;
  r = r - k * C2O_PIO2_C;
  double ps, pc;
  sincos_poly(r, &ps, &pc);
  long q = ((long)k) & 3;
  switch (q) {
    case 0: *s = ps; *c = pc; break;
    case 1: *s = pc; *c = -ps; break;
    case 2: *s = -ps; *c = -pc; break;
    default: *s = -pc; *c = ps; break;
  }
}

/* ------------------------------------------------------------------------ */
/* Quaternion / rotation conventions (LinearMath.h).  Quaternions are stored
 * (x, y, z, w) like LinearMath.h:151-154.                                   */
/* ------------------------------------------------------------------------ */

/* Matrix3x3::Quaternion_ (LinearMath.h:759-793).  The reference exits when
 * s == 0 in the trace>0 branch, which cannot happen (trace > 0). */
void c2o_quat_from_matrix(const double R[3][3], double q[4]) {
  const double *v = &R[0][0];
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

/* Matrix3x3::Set_Value(const Quaternion&) (LinearMath.h:809-831). */
void c2o_matrix_from_quat(const double q[4], double R[3][3]) {
  double d = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  double s = 2.0 / d;
  double xs = q[0] * s, ys = q[1] * s, zs = q[2] * s;
  double wx = q[3] * xs, wy = q[3] * ys, wz = q[3] * zs;
  double xx = q[0] * xs, xy = q[0] * ys, xz = q[0] * zs;
  double yy = q[1] * ys, yz = q[1] * zs, zz = q[2] * zs;
  R[0][0] = 1.0 - (yy + zz); R[0][1] = xy - wz; R[0][2] = xz + wy;
  R[1][0] = xy + wz; R[1][1] = 1.0 - (xx + zz); R[1][2] = yz - wx;
  R[2][0] = xz - wy; R[2][1] = yz + wx; R[2][2] = 1.0 - (xx + yy);
}

/* Hamilton product q1 % q2 (LinearMath.h:1018-1025). */
static void quat_mul(const double a[4], const double b[4], double r[4]) {
  double x = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
  double y = a[3] * b[1] + a[1] * b[3] + a[2] * b[0] - a[0] * b[2];
  double z = a[3] * b[2] + a[2] * b[3] + a[0] * b[1] - a[1] * b[0];
  double w = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
  r[0] = x; r[1] = y; r[2] = z; r[3] = w;
}

/* ------------------------------------------------------------------------ */
/* Linear motion (InterpMotion.cpp)                                          */
/* ------------------------------------------------------------------------ */

/* CInterpMotion_Linear(R0,T0,R1,T1): base ctor (InterpMotion.cpp:148-168),
 * velocity() (:486-491), LinearAngularVelocity (:228-270). */
void c2o_motion_init(c2o_motion *m, const double R0[3][3], const double T0[3],
                     const double R1[3][3], const double T1[3]) {
  memcpy(m->R0, R0, sizeof(m3));
  memcpy(m->R1, R1, sizeof(m3));
  memcpy(m->T0, T0, sizeof(double) * 3);
  memcpy(m->T1, T1, sizeof(double) * 3);
  sub3(m->cv, T1, T0);
  double qs[4], qt[4], qs_conj[4], qr[4];
  c2o_quat_from_matrix(R0, qs);
  c2o_quat_from_matrix(R1, qt);
  memcpy(m->q0, qs, sizeof(qs));
  qs_conj[0] = -qs[0]; qs_conj[1] = -qs[1]; qs_conj[2] = -qs[2]; qs_conj[3] = qs[3];
  quat_mul(qs_conj, qt, qr);
  double s = 1 < qr[3] ? 1 : qr[3];
  double sign = s < 0 ? -1 : 1;
  double a = (fabs(s - 1) <= 1e-40 || fabs(s + 1) <= 1e-40)
                 ? (2 * sign)
                 : (sign * acos(2 * s * s - 1) / sqrt(1 - s * s));
  double tan3[3] = {a * qr[0], a * qr[1], a * qr[2]};
  m->angVel = len3(tan3);
  /* the reference compares against the float literal 0.00000001f */
  double len = tan3[0] * tan3[0] + tan3[1] * tan3[1] + tan3[2] * tan3[2];
  if (len < (double)0.00000001f) {
    m->axis[0] = 1.0; m->axis[1] = 0.0; m->axis[2] = 0.0;
  } else {
    double inv = 1.0 / sqrt(len);
    m->axis[0] = tan3[0] * inv; m->axis[1] = tan3[1] * inv; m->axis[2] = tan3[2] * inv;
  }
  /* The rotation is q0 * q_rel(t): q_rel turns about `axis` in the body
   * frame, so in world coordinates the object turns about R0 * axis at every
   * t (R(t) * axis = R0 * axis).  The strict bounds cross this world axis
   * with the world-frame directions they are given. */
  mxv(m->axw, m->R0, m->axis);
}

/* Strict (conservative) mode: the motion bounds use the world-frame axis,
 * every tested leaf pair bounds mint (not only the improving ones), and a
 * touching leaf pair (distance 0) bounds it by 0.  Off = the reference. */
static int g_strict = 0;
void c2o_set_strict(int on) { g_strict = on; }
int c2o_get_strict(void) { return g_strict; }

/* CInterpMotion_Linear::integrate (InterpMotion.cpp:516-568) with
 * AbsoluteRt/DeltaRt (:273-287).  qua7 (optional) = (w,x,y,z,tx,ty,tz). */
void c2o_motion_integrate(const c2o_motion *m, double t, double R[3][3],
                          double T[3], double qua7[7]) {
  double dt = t;
  if (dt > 1) dt = 1;
  double Tn[3];
  Tn[0] = m->T0[0] + dt * m->cv[0];
  Tn[1] = m->T0[1] + dt * m->cv[1];
  Tn[2] = m->T0[2] + dt * m->cv[2];
  double ang = 0.5 * m->angVel * dt;
  double sa, ca;
  c2o_sincos(ang, &sa, &ca);
  double drt[4] = {sa * m->axis[0], sa * m->axis[1], sa * m->axis[2], ca};
  double q[4];
  quat_mul(m->q0, drt, q);
  if (R) c2o_matrix_from_quat(q, R);
  if (T) { T[0] = Tn[0]; T[1] = Tn[1]; T[2] = Tn[2]; }
  if (qua7) {
    qua7[0] = q[3]; qua7[1] = q[0]; qua7[2] = q[1]; qua7[3] = q[2];
    qua7[4] = Tn[0]; qua7[5] = Tn[1]; qua7[6] = Tn[2];
  }
}

/* computeTOC_MotionBound (InterpMotion.cpp:831-881): directional bound of a
 * BV with angular radius angR along N (normalised in place in the reference;
 * here on a copy).  The unused T and d arguments are dropped. */
double c2o_motion_bound(const c2o_motion *m, double angR, const double N[3]) {
  double n[3] = {N[0], N[1], N[2]}, cr[3];
  normalize3(n);
  cross3(cr, g_strict ? m->axw : m->axis, n);
  double w = angR * len3(cr) * m->angVel;
  double v = dot3(m->cv, n);
  if (v < 0) v = 0;
  double path = v + w;
  if (path <= 0) path = 1e-30;
  return path;
}

/* computeTOC(d, r1, S) (InterpMotion.cpp:746-828): leaf bound. */
double c2o_motion_bound_leaf(const c2o_motion *m, double angR, const double S[3]) {
  double n[3] = {S[0], S[1], S[2]};
  normalize3(n);
  double w, v;
  if (m->angVel == 0) {
    w = 0;
    v = dot3(m->cv, n);
    if (v < 0) v = 0;
  } else {
    const double *ax = g_strict ? m->axw : m->axis;
    double cwc[3] = {ax[0] * m->angVel, ax[1] * m->angVel, ax[2] * m->angVel};
    double cr[3];
    cross3(cr, cwc, n);
    w = angR * len3(cr);
    v = dot3(m->cv, n);
    if (v < 0) v = 0;
  }
  double path = w + v;
  if (path == 0) path = 1e-30;
  return path;
}

/* ------------------------------------------------------------------------ */
/* Rectangle-rectangle distance (C2A_RectDist.h:44-934).                     */
/* Rectangle A spans [0,a0]x[0,a1] in its own frame; B spans Tab + Rab *      */
/* [0,b0]x[0,b1].  Returns the distance; S = Q - P is the closest direction   */
/* in A's frame (left untouched when the rectangles overlap, like the        */
/* reference).  The edge-pair cases are tested in the reference's order      */
/* (A1-B1, A1-B0, A0-B1, A0-B0 families, upper/lower edges), then the face-   */
/* normal separation.                                                        */
/* ------------------------------------------------------------------------ */
static inline void clampd(double *v, double lo, double hi) {
  if (*v < lo) *v = lo;
  else if (*v > hi) *v = hi;
}

/* Lumelsky segment parameters (C2A_RectDist.h:81-111). */
static inline void seg_params(double *t, double *u, double a, double b,
                              double AB, double AT, double BT) {
  double denom = 1 - AB * AB;
  if (denom == 0) *t = 0;
  else {
    *t = (AT - BT * AB) / denom;
    clampd(t, 0, a);
  }
  *u = *t * AB - BT;
  if (*u < 0) {
    *u = 0;
    *t = AT;
    clampd(t, 0, a);
  } else if (*u > b) {
    *u = b;
    *t = *u * AB + AT;
    clampd(t, 0, a);
  }
}

/* InsideVoronoi (C2A_RectDist.h:123-154). */
static inline int voronoi(double a, double b, double AnB, double AnT, double AB,
                          double AT, double BT) {
  if (fabs(AnB) < 1e-7) return 0;
  double u = -AnT / AnB;
  clampd(&u, 0, b);
  double t = u * AB + AT;
  clampd(&t, 0, a);
  double v = t * AB - BT;
  if (AnB > 0) {
    if (v > (u + 1e-7)) return 1;
  } else {
    if (v < (u - 1e-7)) return 1;
  }
  return 0;
}

/* Closing step of an edge-edge case: P is on A's edge, Q = Tab + Bf*lf +
 * Bv*u on B's edge (Bf/Bv are columns of Rab; lf==NULL drops that term). */
static inline double edge_finish(double P0, double P1, const double Rab[3][3],
                                 const double Tab[3], int cf, const double *lf,
                                 int cv, double u, double P[3], double Q[3],
                                 double S[3]) {
  P[0] = P0; P[1] = P1; P[2] = 0;
  for (int i = 0; i < 3; i++) {
    if (lf) Q[i] = Tab[i] + Rab[i][cf] * *lf + Rab[i][cv] * u;
    else Q[i] = Tab[i] + Rab[i][cv] * u;
  }
  S[0] = Q[0] - P[0]; S[1] = Q[1] - P[1]; S[2] = Q[2] - P[2];
  return sqrt(dot3(S, S));
}

double c2o_rect_dist(const double Rab[3][3], const double Tab[3],
                     const double a[2], const double b[2], double P[3],
                     double Q[3], double S[3]) {
  const double AB00 = Rab[0][0], AB01 = Rab[0][1], AB10 = Rab[1][0], AB11 = Rab[1][1];
  const double aAB00 = a[0] * AB00, aAB01 = a[0] * AB01;
  const double aAB10 = a[1] * AB10, aAB11 = a[1] * AB11;
  const double bAB00 = b[0] * AB00, bAB10 = b[0] * AB10;
  const double bAB01 = b[1] * AB01, bAB11 = b[1] * AB11;
  double Tba[3];
  mtxv(Tba, Rab, Tab);
  double t, u;

  /* x extents (in B's frame) of A's corners, and of B's corners along A0 */
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

  /* A edge x=a0 (along A1)  x  B edge at b0 (along B1) */
  if ((UA1ux > b[0]) && (UB1ux > a[0]) &&
      ((UA1lx > b[0]) || voronoi(b[1], a[1], AB10, aAB00 - b[0] - Tba[0], AB11,
                                 aAB01 - Tba[1], -Tab[1] - bAB10)) &&
      ((UB1lx > a[0]) || voronoi(a[1], b[1], AB01, Tab[0] + bAB00 - a[0], AB11,
                                 Tab[1] + bAB10, Tba[1] - aAB01))) {
    seg_params(&t, &u, a[1], b[1], AB11, Tab[1] + bAB10, Tba[1] - aAB01);
    return edge_finish(a[0], t, Rab, Tab, 0, &b[0], 1, u, P, Q, S);
  }
  /* A edge x=a0  x  B edge at 0 (along B1) */
  if ((UA1lx < 0) && (LB1ux > a[0]) &&
      ((UA1ux < 0) || voronoi(b[1], a[1], -AB10, Tba[0] - aAB00, AB11,
                              aAB01 - Tba[1], -Tab[1])) &&
      ((LB1lx > a[0]) || voronoi(a[1], b[1], AB01, Tab[0] - a[0], AB11, Tab[1],
                                 Tba[1] - aAB01))) {
    seg_params(&t, &u, a[1], b[1], AB11, Tab[1], Tba[1] - aAB01);
    return edge_finish(a[0], t, Rab, Tab, 0, NULL, 1, u, P, Q, S);
  }
  /* A edge x=0  x  B edge at b0 (along B1) */
  if ((LA1ux > b[0]) && (UB1lx < 0) &&
      ((LA1lx > b[0]) || voronoi(b[1], a[1], AB10, -Tba[0] - b[0], AB11, -Tba[1],
                                 -Tab[1] - bAB10)) &&
      ((UB1ux < 0) || voronoi(a[1], b[1], -AB01, -Tab[0] - bAB00, AB11,
                              Tab[1] + bAB10, Tba[1]))) {
    seg_params(&t, &u, a[1], b[1], AB11, Tab[1] + bAB10, Tba[1]);
    return edge_finish(0, t, Rab, Tab, 0, &b[0], 1, u, P, Q, S);
  }
  /* A edge x=0  x  B edge at 0 (along B1) */
  if ((LA1lx < 0) && (LB1lx < 0) &&
      ((LA1ux < 0) || voronoi(b[1], a[1], -AB10, Tba[0], AB11, -Tba[1], -Tab[1])) &&
      ((LB1ux < 0) || voronoi(a[1], b[1], -AB01, -Tab[0], AB11, Tab[1], Tba[1]))) {
    seg_params(&t, &u, a[1], b[1], AB11, Tab[1], Tba[1]);
    return edge_finish(0, t, Rab, Tab, 0, NULL, 1, u, P, Q, S);
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

  /* A edge x=a0 (along A1)  x  B edge at b1 (along B0) */
  if ((UA1uy > b[1]) && (UB0ux > a[0]) &&
      ((UA1ly > b[1]) || voronoi(b[0], a[1], AB11, aAB01 - Tba[1] - b[1], AB10,
                                 aAB00 - Tba[0], -Tab[1] - bAB11)) &&
      ((UB0lx > a[0]) || voronoi(a[1], b[0], AB00, Tab[0] - a[0] + bAB01, AB10,
                                 Tab[1] + bAB11, Tba[0] - aAB00))) {
    seg_params(&t, &u, a[1], b[0], AB10, Tab[1] + bAB11, Tba[0] - aAB00);
    return edge_finish(a[0], t, Rab, Tab, 1, &b[1], 0, u, P, Q, S);
  }
  /* A edge x=a0  x  B edge at 0 (along B0) */
  if ((UA1ly < 0) && (LB0ux > a[0]) &&
      ((UA1uy < 0) || voronoi(b[0], a[1], -AB11, Tba[1] - aAB01, AB10,
                              aAB00 - Tba[0], -Tab[1])) &&
      ((LB0lx > a[0]) || voronoi(a[1], b[0], AB00, Tab[0] - a[0], AB10, Tab[1],
                                 Tba[0] - aAB00))) {
    seg_params(&t, &u, a[1], b[0], AB10, Tab[1], Tba[0] - aAB00);
    return edge_finish(a[0], t, Rab, Tab, 1, NULL, 0, u, P, Q, S);
  }
  /* A edge x=0  x  B edge at b1 (along B0) */
  if ((LA1uy > b[1]) && (UB0lx < 0) &&
      ((LA1ly > b[1]) || voronoi(b[0], a[1], AB11, -Tba[1] - b[1], AB10, -Tba[0],
                                 -Tab[1] - bAB11)) &&
      ((UB0ux < 0) || voronoi(a[1], b[0], -AB00, -Tab[0] - bAB01, AB10,
                              Tab[1] + bAB11, Tba[0]))) {
    seg_params(&t, &u, a[1], b[0], AB10, Tab[1] + bAB11, Tba[0]);
    return edge_finish(0, t, Rab, Tab, 1, &b[1], 0, u, P, Q, S);
  }
  /* A edge x=0  x  B edge at 0 (along B0) */
  if ((LA1ly < 0) && (LB0lx < 0) &&
      ((LA1uy < 0) || voronoi(b[0], a[1], -AB11, Tba[1], AB10, -Tba[0], -Tab[1])) &&
      ((LB0ux < 0) || voronoi(a[1], b[0], -AB00, -Tab[0], AB10, Tab[1], Tba[0]))) {
    seg_params(&t, &u, a[1], b[0], AB10, Tab[1], Tba[0]);
    return edge_finish(0, t, Rab, Tab, 1, NULL, 0, u, P, Q, S);
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

  /* A edge y=a1 (along A0)  x  B edge at b0 (along B1) */
  if ((UA0ux > b[0]) && (UB1uy > a[1]) &&
      ((UA0lx > b[0]) || voronoi(b[1], a[0], AB00, aAB10 - Tba[0] - b[0], AB01,
                                 aAB11 - Tba[1], -Tab[0] - bAB00)) &&
      ((UB1ly > a[1]) || voronoi(a[0], b[1], AB11, Tab[1] - a[1] + bAB10, AB01,
                                 Tab[0] + bAB00, Tba[1] - aAB11))) {
    seg_params(&t, &u, a[0], b[1], AB01, Tab[0] + bAB00, Tba[1] - aAB11);
    return edge_finish(t, a[1], Rab, Tab, 0, &b[0], 1, u, P, Q, S);
  }
  /* A edge y=a1  x  B edge at 0 (along B1) */
  if ((UA0lx < 0) && (LB1uy > a[1]) &&
      ((UA0ux < 0) || voronoi(b[1], a[0], -AB00, Tba[0] - aAB10, AB01,
                              aAB11 - Tba[1], -Tab[0])) &&
      ((LB1ly > a[1]) || voronoi(a[0], b[1], AB11, Tab[1] - a[1], AB01, Tab[0],
                                 Tba[1] - aAB11))) {
    seg_params(&t, &u, a[0], b[1], AB01, Tab[0], Tba[1] - aAB11);
    return edge_finish(t, a[1], Rab, Tab, 0, NULL, 1, u, P, Q, S);
  }
  /* A edge y=0  x  B edge at b0 (along B1) */
  if ((LA0ux > b[0]) && (UB1ly < 0) &&
      ((LA0lx > b[0]) || voronoi(b[1], a[0], AB00, -b[0] - Tba[0], AB01, -Tba[1],
                                 -bAB00 - Tab[0])) &&
      ((UB1uy < 0) || voronoi(a[0], b[1], -AB11, -Tab[1] - bAB10, AB01,
                              Tab[0] + bAB00, Tba[1]))) {
    seg_params(&t, &u, a[0], b[1], AB01, Tab[0] + bAB00, Tba[1]);
    return edge_finish(t, 0, Rab, Tab, 0, &b[0], 1, u, P, Q, S);
  }
  /* A edge y=0  x  B edge at 0 (along B1) */
  if ((LA0lx < 0) && (LB1ly < 0) &&
      ((LA0ux < 0) || voronoi(b[1], a[0], -AB00, Tba[0], AB01, -Tba[1], -Tab[0])) &&
      ((LB1uy < 0) || voronoi(a[0], b[1], -AB11, -Tab[1], AB01, Tab[0], Tba[1]))) {
    seg_params(&t, &u, a[0], b[1], AB01, Tab[0], Tba[1]);
    return edge_finish(t, 0, Rab, Tab, 0, NULL, 1, u, P, Q, S);
  }

  double LA0ly, LA0uy, UA0ly, UA0uy, LB0ly, LB0uy, UB0ly, UB0uy;
  if (ALLy < AULy) { LA0ly = ALLy; LA0uy = AULy; UA0ly = ALUy; UA0uy = AUUy; }
  else { LA0ly = AULy; LA0uy = ALLy; UA0ly = AUUy; UA0uy = ALUy; }
  if (BLLy < BULy) { LB0ly = BLLy; LB0uy = BULy; UB0ly = BLUy; UB0uy = BUUy; }
  else { LB0ly = BULy; LB0uy = BLLy; UB0ly = BUUy; UB0uy = BLUy; }

  /* A edge y=a1 (along A0)  x  B edge at b1 (along B0) */
  if ((UA0uy > b[1]) && (UB0uy > a[1]) &&
      ((UA0ly > b[1]) || voronoi(b[0], a[0], AB01, aAB11 - Tba[1] - b[1], AB00,
                                 aAB10 - Tba[0], -Tab[0] - bAB01)) &&
      ((UB0ly > a[1]) || voronoi(a[0], b[0], AB10, Tab[1] - a[1] + bAB11, AB00,
                                 Tab[0] + bAB01, Tba[0] - aAB10))) {
    seg_params(&t, &u, a[0], b[0], AB00, Tab[0] + bAB01, Tba[0] - aAB10);
    return edge_finish(t, a[1], Rab, Tab, 1, &b[1], 0, u, P, Q, S);
  }
  /* A edge y=a1  x  B edge at 0 (along B0) */
  if ((UA0ly < 0) && (LB0uy > a[1]) &&
      ((UA0uy < 0) || voronoi(b[0], a[0], -AB01, Tba[1] - aAB11, AB00,
                              aAB10 - Tba[0], -Tab[0])) &&
      ((LB0ly > a[1]) || voronoi(a[0], b[0], AB10, Tab[1] - a[1], AB00, Tab[0],
                                 Tba[0] - aAB10))) {
    seg_params(&t, &u, a[0], b[0], AB00, Tab[0], Tba[0] - aAB10);
    return edge_finish(t, a[1], Rab, Tab, 1, NULL, 0, u, P, Q, S);
  }
  /* A edge y=0  x  B edge at b1 (along B0) */
  if ((LA0uy > b[1]) && (UB0ly < 0) &&
      ((LA0ly > b[1]) || voronoi(b[0], a[0], AB01, -Tba[1] - b[1], AB00, -Tba[0],
                                 -Tab[0] - bAB01)) &&
      ((UB0uy < 0) || voronoi(a[0], b[0], -AB10, -Tab[1] - bAB11, AB00,
                              Tab[0] + bAB01, Tba[0]))) {
    seg_params(&t, &u, a[0], b[0], AB00, Tab[0] + bAB01, Tba[0]);
    return edge_finish(t, 0, Rab, Tab, 1, &b[1], 0, u, P, Q, S);
  }
  /* A edge y=0  x  B edge at 0 (along B0) */
  if ((LA0ly < 0) && (LB0ly < 0) &&
      ((LA0uy < 0) || voronoi(b[0], a[0], -AB01, Tba[1], AB00, -Tba[0], -Tab[0])) &&
      ((LB0uy < 0) || voronoi(a[0], b[0], -AB10, -Tab[1], AB00, Tab[0], Tba[0]))) {
    seg_params(&t, &u, a[0], b[0], AB00, Tab[0], Tba[0]);
    return edge_finish(t, 0, Rab, Tab, 1, NULL, 0, u, P, Q, S);
  }

  /* No edge pair holds the closest points: separation along the two face
   * normals (C2A_RectDist.h:855-933). */
  double sep1, sep2;
  if (Tab[2] > 0.0) {
    sep1 = Tab[2];
    if (Rab[2][0] < 0.0) sep1 += b[0] * Rab[2][0];
    if (Rab[2][1] < 0.0) sep1 += b[1] * Rab[2][1];
  } else {
    sep1 = -Tab[2];
    if (Rab[2][0] > 0.0) sep1 -= b[0] * Rab[2][0];
    if (Rab[2][1] > 0.0) sep1 -= b[1] * Rab[2][1];
  }
  if (Tba[2] < 0) {
    sep2 = -Tba[2];
    if (Rab[0][2] < 0.0) sep2 += a[0] * Rab[0][2];
    if (Rab[1][2] < 0.0) sep2 += a[1] * Rab[1][2];
  } else {
    sep2 = Tba[2];
    if (Rab[0][2] > 0.0) sep2 -= a[0] * Rab[0][2];
    if (Rab[1][2] > 0.0) sep2 -= a[1] * Rab[1][2];
  }
  if (sep1 >= sep2 && sep1 >= 0) {
    P[0] = P[1] = P[2] = 0.0;
    Q[0] = 0.0; Q[1] = 0.0;
    Q[2] = (Tab[2] > 0.0) ? sep1 : -sep1;
    S[0] = Q[0] - P[0]; S[1] = Q[1] - P[1]; S[2] = Q[2] - P[2];
  }
  if (sep2 >= sep1 && sep2 >= 0) {
    Q[0] = Tab[0]; Q[1] = Tab[1]; Q[2] = Tab[2];
    if (Tba[2] < 0) {
      P[0] = Rab[0][2] * sep2 + Tab[0];
      P[1] = Rab[1][2] * sep2 + Tab[1];
      P[2] = Rab[2][2] * sep2 + Tab[2];
    } else {
      P[0] = -Rab[0][2] * sep2 + Tab[0];
      P[1] = -Rab[1][2] * sep2 + Tab[1];
      P[2] = -Rab[2][2] * sep2 + Tab[2];
    }
    S[0] = Q[0] - P[0]; S[1] = Q[1] - P[1]; S[2] = Q[2] - P[2];
  }
  double sep = (sep1 > sep2 ? sep1 : sep2);
  return (sep > 0 ? sep : 0);
}

/* C2A_BV_Distance (C2A_BV.cpp:666-675). */
static double bv_distance(const m3 R, const double T[3], const c2o_bv *b1,
                          const c2o_bv *b2, double S[3]) {
  double P[3], Q[3];
  double d = c2o_rect_dist(R, T, b1->l, b2->l, P, Q, S);
  d -= (b1->r + b2->r);
  return (d < 0.0) ? 0.0 : d;
}

/* ------------------------------------------------------------------------ */
/* Triangle-triangle distance (C2A.cpp:59-405, the feature-tracking copy of  */
/* PQP's TriDist; the hot path's PQP TriDistance computes the same value).   */
/* ------------------------------------------------------------------------ */

/* Segment closest points + separating vector (C2A.cpp:59-163). */
static void seg_points(double VEC[3], double X[3], double Y[3], const double P[3],
                       const double A[3], const double Q[3], const double B[3]) {
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
    memcpy(Y, Q, sizeof(double) * 3);
    t = AT / AA;
    if ((t <= 0) || isnan(t)) {
      memcpy(X, P, sizeof(double) * 3);
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
      memcpy(X, P, sizeof(double) * 3);
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
      memcpy(X, P, sizeof(double) * 3);
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
      if (dot3(VEC, T) < 0) scale3(VEC, VEC, -1);
    }
  }
}

/* One vertex-face test: is a vertex of Tt closest to the face of Ts?
 * (C2A.cpp:255-327 for S's face, 329-386 mirrored).  Returns the vertex
 * index or -1; sets *disjoint when the face normal separates. */
static int vertex_face(const double Ts[3][3], const double Sv[3][3],
                       const double Tt[3][3], double Nrm[3], double *Nl,
                       double proj[3], int *disjoint) {
  double V[3], Z[3];
  cross3(Nrm, Sv[0], Sv[1]);
  *Nl = dot3(Nrm, Nrm);
  if (!(*Nl > 1e-15)) return -1;
  sub3(V, Ts[0], Tt[0]); proj[0] = dot3(V, Nrm);
  sub3(V, Ts[0], Tt[1]); proj[1] = dot3(V, Nrm);
  sub3(V, Ts[0], Tt[2]); proj[2] = dot3(V, Nrm);
  int point = -1;
  if ((proj[0] > 0) && (proj[1] > 0) && (proj[2] > 0)) {
    point = (proj[0] < proj[1]) ? 0 : 1;
    if (proj[2] < proj[point]) point = 2;
  } else if ((proj[0] < 0) && (proj[1] < 0) && (proj[2] < 0)) {
    point = (proj[0] > proj[1]) ? 0 : 1;
    if (proj[2] > proj[point]) point = 2;
  }
  if (point < 0) return -1;
  *disjoint = 1;
  for (int e = 0; e < 3; e++) {
    sub3(V, Tt[point], Ts[e]);
    cross3(Z, Nrm, Sv[e]);
    if (!(dot3(V, Z) > 0)) return -1;
  }
  return point;
}

double c2o_tri_dist(double P[3], double Q[3], const double S[3][3],
                    const double T[3][3], int *typeA, int *fidA, int *typeB,
                    int *fidB) {
  double Sv[3][3], Tv[3][3], VEC[3], V[3], Z[3], minP[3], minQ[3];
  int ta = -1, fa = -1, tb = -1, fb = -1;
  sub3(Sv[0], S[1], S[0]); sub3(Sv[1], S[2], S[1]); sub3(Sv[2], S[0], S[2]);
  sub3(Tv[0], T[1], T[0]); sub3(Tv[1], T[2], T[1]); sub3(Tv[2], T[0], T[2]);
  int shown_disjoint = 0;
  double mindd = dist2_3(S[0], T[0]) + 1;
  double result = -1;
  for (int i = 0; i < 3 && result < 0; i++) {
    for (int j = 0; j < 3; j++) {
      seg_points(VEC, P, Q, S[i], Sv[i], T[j], Tv[j]);
      sub3(V, Q, P);
      double dd = dot3(V, V);
      if (dd <= mindd) {
        memcpy(minP, P, sizeof(minP));
        memcpy(minQ, Q, sizeof(minQ));
        mindd = dd;
        ta = 1; fa = i; tb = 1; fb = j;
        sub3(Z, S[(i + 2) % 3], P);
        double a = dot3(Z, VEC);
        sub3(Z, T[(j + 2) % 3], Q);
        double b = dot3(Z, VEC);
        if ((a <= 0) && (b >= 0)) { result = sqrt(dd); break; }
        double p = dot3(V, VEC);
        if (a < 0) a = 0;
        if (b > 0) b = 0;
        if ((p - a + b) > 0) shown_disjoint = 1;
      }
    }
  }
  if (result < 0) {
    double Nrm[3], Nl, proj[3];
    int pt = vertex_face(S, Sv, T, Nrm, &Nl, proj, &shown_disjoint);
    if (pt >= 0) {
      ta = 2; fa = -1; tb = 0; fb = pt;
      axpy3(P, T[pt], Nrm, proj[pt] / Nl);
      memcpy(Q, T[pt], sizeof(double) * 3);
      result = sqrt(dist2_3(P, Q));
    } else {
      pt = vertex_face(T, Tv, S, Nrm, &Nl, proj, &shown_disjoint);
      if (pt >= 0) {
        ta = 0; fa = pt; tb = 2; fb = -1;
        memcpy(P, S[pt], sizeof(double) * 3);
        axpy3(Q, S[pt], Nrm, proj[pt] / Nl);
        result = sqrt(dist2_3(P, Q));
      } else if (shown_disjoint) {
        memcpy(P, minP, sizeof(minP));
        memcpy(Q, minQ, sizeof(minQ));
        result = sqrt(mindd);
      } else {
        result = 0;   /* overlapping triangles (bCollided) */
      }
    }
  }
  if (typeA) *typeA = ta;
  if (fidA) *fidA = fa;
  if (typeB) *typeB = tb;
  if (fidB) *fidB = fb;
  return result;
}

/* TriDistance / C2A_TriDistance (C2A.cpp:408-424): triangle 2 mapped into
 * model 1's frame by (R, T). */
static double tri_distance(const m3 R, const double T[3], const c2o_tri *t1,
                           const c2o_tri *t2, double p[3], double q[3],
                           int *typeA, int *fidA, int *typeB, int *fidB) {
  double a[3][3], b[3][3];
  memcpy(a, t1->p, sizeof(a));
  mxvpv(b[0], R, t2->p[0], T);
  mxvpv(b[1], R, t2->p[1], T);
  mxvpv(b[2], R, t2->p[2], T);
  return c2o_tri_dist(p, q, a, b, typeA, fidA, typeB, fidB);
}

/* ------------------------------------------------------------------------ */
/* BVH build (C2A_PQP.cpp:331-417, C2A_Build.cpp:253-574, C2A_BV.cpp).       */
/* ------------------------------------------------------------------------ */

/* Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix with the
 * threshold strategy of Numerical Recipes' jacobi (sec. 11.1): threshold
 * 0.2*S/n^2 in the first three sweeps, underflow zeroing after the fourth,
 * at most 50 sweeps.  This restates PQP's Meigen (called at
 * C2A_Build.cpp:416).  Columns of V are eigenvectors, d the eigenvalues. */
void c2o_jacobi_eigen(const double A_in[3][3], double V[3][3], double d[3]) {
  double a[3][3], v[3][3], b[3], z[3];
  memcpy(a, A_in, sizeof(a));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) v[i][j] = (i == j) ? 1.0 : 0.0;
  for (int p = 0; p < 3; p++) { b[p] = d[p] = a[p][p]; z[p] = 0.0; }
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
          z[p] -= h; z[q] += h; d[p] -= h; d[q] += h;
          a[p][q] = 0.0;
#define C2O_ROT(M, i, j, k, l) do { double gg = M[i][j], hh = M[k][l]; \
    M[i][j] = gg - s * (hh + gg * tau); M[k][l] = hh + s * (gg - hh * tau); } while (0)
          for (int j = 0; j < p; j++) C2O_ROT(a, j, p, j, q);
          for (int j = p + 1; j < q; j++) C2O_ROT(a, p, j, j, q);
          for (int j = q + 1; j < 3; j++) C2O_ROT(a, p, j, q, j);
          for (int j = 0; j < 3; j++) C2O_ROT(v, j, p, j, q);
#undef C2O_ROT
        }
      }
    }
    for (int p = 0; p < 3; p++) { b[p] += z[p]; d[p] = b[p]; z[p] = 0.0; }
  }
  memcpy(V, v, sizeof(v));
}

/* get_covariance_triverts (C2A_Build.cpp:279-335). */
static void covariance(m3 M, const c2o_tri *tris, int n) {
  double S1[3] = {0, 0, 0}, S2[3][3] = {{0}};
  for (int i = 0; i < n; i++) {
    const double *p1 = tris[i].p[0], *p2 = tris[i].p[1], *p3 = tris[i].p[2];
    S1[0] += p1[0] + p2[0] + p3[0];
    S1[1] += p1[1] + p2[1] + p3[1];
    S1[2] += p1[2] + p2[2] + p3[2];
    S2[0][0] += (p1[0] * p1[0] + p2[0] * p2[0] + p3[0] * p3[0]);
    S2[1][1] += (p1[1] * p1[1] + p2[1] * p2[1] + p3[1] * p3[1]);
    S2[2][2] += (p1[2] * p1[2] + p2[2] * p2[2] + p3[2] * p3[2]);
    S2[0][1] += (p1[0] * p1[1] + p2[0] * p2[1] + p3[0] * p3[1]);
    S2[0][2] += (p1[0] * p1[2] + p2[0] * p2[2] + p3[0] * p3[2]);
    S2[1][2] += (p1[1] * p1[2] + p2[1] * p2[2] + p3[1] * p3[2]);
  }
  double cnt = (double)(3 * n);
  M[0][0] = S2[0][0] - S1[0] * S1[0] / cnt;
  M[1][1] = S2[1][1] - S1[1] * S1[1] / cnt;
  M[2][2] = S2[2][2] - S1[2] * S1[2] / cnt;
  M[0][1] = S2[0][1] - S1[0] * S1[1] / cnt;
  M[1][2] = S2[1][2] - S1[1] * S1[2] / cnt;
  M[0][2] = S2[0][2] - S1[0] * S1[2] / cnt;
  M[1][0] = M[0][1]; M[2][0] = M[0][2]; M[2][1] = M[1][2];
}

/* get_centroid_triverts (C2A_Build.cpp:253-277). */
static void centroid(double c[3], const c2o_tri *tris, int n) {
  c[0] = c[1] = c[2] = 0.0;
  for (int i = 0; i < n; i++) {
    c[0] += tris[i].p[0][0] + tris[i].p[1][0] + tris[i].p[2][0];
    c[1] += tris[i].p[0][1] + tris[i].p[1][1] + tris[i].p[2][1];
    c[2] += tris[i].p[0][2] + tris[i].p[1][2] + tris[i].p[2][2];
  }
  double cnt = (double)(3 * n);
  c[0] /= cnt; c[1] /= cnt; c[2] /= cnt;
}

/* split_tris (C2A_Build.cpp:344-385): x <= c goes to the first group by
 * swapping into place; an empty side falls back to a half split. */
static int split_tris(c2o_tri *tris, int n, const double a[3], double c) {
  int c1 = 0;
  for (int i = 0; i < n; i++) {
    double p[3];
    add3(p, tris[i].p[0], tris[i].p[1]);
    add3(p, p, tris[i].p[2]);
    double x = dot3(p, a);
    x /= 3.0;
    if (x <= c) {
      c2o_tri tmp = tris[i];
      tris[i] = tris[c1];
      tris[c1] = tmp;
      c1++;
    }
  }
  if ((c1 == 0) || (c1 == n)) c1 = n / 2;
  return c1;
}

static inline double max2(double a, double b) { return a > b ? a : b; }

/* FitToTris_Corner, RSS part (C2A_BV.cpp:354-635) and ComputeAngularRadius
 * (C2A_BV.cpp:71-111; the model centre is zeroed, so the radius is measured
 * from the model origin). */
static void fit_rss(c2o_bv *bv, const m3 O, const c2o_tri *tris, int n) {
  double angR = 0.0;
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      double pr[3] = {tris[i].p[k][0] - 0.0, tris[i].p[k][1] - 0.0, tris[i].p[k][2] - 0.0};
      double d = len3(pr);
      if (d > angR) angR = d;
    }
  bv->angR = angR;
  memcpy(bv->R, O, sizeof(m3));
  int np = 3 * n;
  double (*P)[3] = (double (*)[3])malloc(sizeof(double) * 3 * (size_t)np);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) mtxv(P[3 * i + k], bv->R, tris[i].p[k]);

  double minz = P[0][2], maxz = P[0][2];
  for (int i = 1; i < np; i++) {
    if (P[i][2] < minz) minz = P[i][2];
    else if (P[i][2] > maxz) maxz = P[i][2];
  }
  double r = 0.5 * (maxz - minz);
  double radsqr = r * r;
  double cz = 0.5 * (maxz + minz);

  double minx, maxx, miny, maxy, dz, x;
  int mn = 0, mx = 0;
  for (int i = 1; i < np; i++) {
    if (P[i][0] < P[mn][0]) mn = i;
    else if (P[i][0] > P[mx][0]) mx = i;
  }
  dz = P[mn][2] - cz;
  minx = P[mn][0] + sqrt(max2(radsqr - dz * dz, 0));
  dz = P[mx][2] - cz;
  maxx = P[mx][0] - sqrt(max2(radsqr - dz * dz, 0));
  for (int i = 0; i < np; i++)
    if (P[i][0] < minx) {
      dz = P[i][2] - cz;
      x = P[i][0] + sqrt(max2(radsqr - dz * dz, 0));
      if (x < minx) minx = x;
    }
  for (int i = 0; i < np; i++)
    if (P[i][0] > maxx) {
      dz = P[i][2] - cz;
      x = P[i][0] - sqrt(max2(radsqr - dz * dz, 0));
      if (x > maxx) maxx = x;
    }
  mn = mx = 0;
  for (int i = 1; i < np; i++) {
    if (P[i][1] < P[mn][1]) mn = i;
    else if (P[i][1] > P[mx][1]) mx = i;
  }
  dz = P[mn][2] - cz;
  miny = P[mn][1] + sqrt(max2(radsqr - dz * dz, 0));
  dz = P[mx][2] - cz;
  maxy = P[mx][1] - sqrt(max2(radsqr - dz * dz, 0));
  for (int i = 0; i < np; i++)
    if (P[i][1] < miny) {
      dz = P[i][2] - cz;
      x = P[i][1] + sqrt(max2(radsqr - dz * dz, 0));
      if (x < miny) miny = x;
    }
  for (int i = 0; i < np; i++)
    if (P[i][1] > maxy) {
      dz = P[i][2] - cz;
      x = P[i][1] - sqrt(max2(radsqr - dz * dz, 0));
      if (x > maxy) maxy = x;
    }

  /* grow the rectangle where a corner's quarter-cylinder misses a point */
  const double a = sqrt(0.5);
  for (int i = 0; i < np; i++) {
    double dx, dy, u, t;
    if (P[i][0] > maxx) {
      if (P[i][1] > maxy) {
        dx = P[i][0] - maxx; dy = P[i][1] - maxy;
        u = dx * a + dy * a;
        t = (a * u - dx) * (a * u - dx) + (a * u - dy) * (a * u - dy) +
            (cz - P[i][2]) * (cz - P[i][2]);
        u = u - sqrt(max2(radsqr - t, 0));
        if (u > 0) { maxx += u * a; maxy += u * a; }
      } else if (P[i][1] < miny) {
        dx = P[i][0] - maxx; dy = P[i][1] - miny;
        u = dx * a - dy * a;
        t = (a * u - dx) * (a * u - dx) + (-a * u - dy) * (-a * u - dy) +
            (cz - P[i][2]) * (cz - P[i][2]);
        u = u - sqrt(max2(radsqr - t, 0));
        if (u > 0) { maxx += u * a; miny -= u * a; }
      }
    } else if (P[i][0] < minx) {
      if (P[i][1] > maxy) {
        dx = P[i][0] - minx; dy = P[i][1] - maxy;
        u = dy * a - dx * a;
        t = (-a * u - dx) * (-a * u - dx) + (a * u - dy) * (a * u - dy) +
            (cz - P[i][2]) * (cz - P[i][2]);
        u = u - sqrt(max2(radsqr - t, 0));
        if (u > 0) { minx -= u * a; maxy += u * a; }
      } else if (P[i][1] < miny) {
        dx = P[i][0] - minx; dy = P[i][1] - miny;
        u = -dx * a - dy * a;
        t = (-a * u - dx) * (-a * u - dx) + (-a * u - dy) * (-a * u - dy) +
            (cz - P[i][2]) * (cz - P[i][2]);
        u = u - sqrt(max2(radsqr - t, 0));
        if (u > 0) { minx -= u * a; miny -= u * a; }
      }
    }
  }
  double c0[3] = {minx, miny, cz};
  mxv(bv->Tr, bv->R, c0);
  bv->l[0] = maxx - minx;
  if (bv->l[0] < 0) bv->l[0] = 0;
  bv->l[1] = maxy - miny;
  if (bv->l[1] < 0) bv->l[1] = 0;
  bv->r = r;
  free(P);
}

/* build_recurse (C2A_Build.cpp:393-488).  Explicit stack-free recursion
 * keeps the reference's allocation order: a node's children pair is
 * allocated when the node is processed, first child subtree built first. */
static void build_recurse(c2o_model *m, int bn, int first, int num, int depth) {
  c2o_bv *b = &m->bvs[bn];
  m3 C, E, R;
  double s[3];
  covariance(C, &m->tris[first], num);
  c2o_jacobi_eigen(C, E, s);
  int mn, mid, mx;
  if (s[0] > s[1]) { mx = 0; mn = 1; }
  else { mn = 0; mx = 1; }
  if (s[2] < s[mn]) { mid = mn; mn = 2; }
  else if (s[2] > s[mx]) { mid = mx; mx = 2; }
  else { mid = 2; }
  for (int i = 0; i < 3; i++) { R[i][0] = E[i][mx]; R[i][1] = E[i][mid]; }
  R[0][2] = E[1][mx] * E[2][mid] - E[1][mid] * E[2][mx];
  R[1][2] = E[0][mid] * E[2][mx] - E[0][mx] * E[2][mid];
  R[2][2] = E[0][mx] * E[1][mid] - E[0][mid] * E[1][mx];
  fit_rss(b, R, &m->tris[first], num);
  memcpy(b->R_loc, b->R, sizeof(m3));
  b->depth = depth;
  if (num == 1) {
    b->first_child = -(first + 1);
  } else {
    b->first_child = m->nbv;
    m->nbv += 2;
    double axis[3] = {R[0][0], R[1][0], R[2][0]}, mean[3];
    centroid(mean, &m->tris[first], num);
    double coord = dot3(axis, mean);
    int nfirst = split_tris(&m->tris[first], num, axis, coord);
    int fc = b->first_child;
    build_recurse(m, fc, first, nfirst, depth + 1);
    build_recurse(m, fc + 1, first + nfirst, num - nfirst, depth + 1);
  }
}

/* make_parent_relative (C2A_Build.cpp:493-544): children first, with the
 * parent's model-frame R/Tr, then the node itself. */
static void make_parent_relative(c2o_model *m, int bn, const m3 pR, const double pT[3]) {
  c2o_bv *b = &m->bvs[bn];
  if (b->first_child >= 0) {
    make_parent_relative(m, b->first_child, b->R, b->Tr);
    make_parent_relative(m, b->first_child + 1, b->R, b->Tr);
  }
  m3 Rpc;
  double Tpc[3];
  mtxm(Rpc, pR, b->R);
  memcpy(b->R, Rpc, sizeof(m3));
  sub3(Tpc, b->Tr, pT);
  mtxv(b->Tr, pR, Tpc);
}

int c2o_model_build(c2o_model *m, int n, const double *xyz9, const int *ids,
                    const int *idx3) {
  memset(m, 0, sizeof(*m));
  if (n <= 0) return -1;   /* PQP_ERR_BUILD_EMPTY_MODEL (C2A_PQP.cpp:352-357) */
  m->ntri = n;
  m->tris = (c2o_tri *)calloc((size_t)n, sizeof(c2o_tri));
  m->bvs = (c2o_bv *)calloc((size_t)(2 * n - 1), sizeof(c2o_bv));
  if (!m->tris || !m->bvs) return -2;
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) m->tris[i].p[k][c] = xyz9[9 * i + 3 * k + c];
    m->tris[i].id = ids ? ids[i] : i;
    for (int k = 0; k < 3; k++) m->tris[i].idx[k] = idx3 ? idx3[3 * i + k] : 3 * i + k;
  }
  m->nbv = 1;
  build_recurse(m, 0, 0, n, 0);
  m3 I = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double Z[3] = {0, 0, 0};
  make_parent_relative(m, 0, I, Z);
  for (int i = 0; i < m->nbv; i++) {
    c2o_bv *b = &m->bvs[i];
    b->size = sqrt(b->l[0] * b->l[0] + b->l[1] * b->l[1]) + 2 * b->r;
  }
  m->last_tri = 0;   /* last_tri = tris after the build (C2A_PQP.cpp:401) */
  return 0;
}

void c2o_model_free(c2o_model *m) {
  free(m->tris);
  free(m->bvs);
  memset(m, 0, sizeof(*m));
}

/* ------------------------------------------------------------------------ */
/* CA step: TOCStepRecurse_Dis (C2A.cpp:1113-1354)                           */
/* ------------------------------------------------------------------------ */
typedef struct {
  const c2o_model *m1, *m2;
  const c2o_motion *mo1, *mo2;
  const double (*r1)[3];     /* current rotation of object 1 */
  const double *tt1;         /* current translation of object 1 */
  c2o_result *res;
  m3 Rrel;                   /* res->R */
  double Trel[3];            /* res->T */
  int last1, last2;          /* model last_tri updates */
} step_ctx;

static inline int desc_pred(const c2o_result *res, double d, double mt) {
  return mt < res->UpboundTOC &&
         ((d < (res->distance - res->abs_err)) || (d * (1 + res->rel_err) < res->distance));
}

static void toc_recurse(step_ctx *cx, const m3 R, const double T[3], int b1, int b2) {
  const c2o_bv *B1 = &cx->m1->bvs[b1], *B2 = &cx->m2->bvs[b2];
  c2o_result *res = cx->res;
  int l1 = B1->first_child < 0, l2 = B2->first_child < 0;
  if (l1 && l2) {
    double p[3], q[3];
    int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
    double dTri = tri_distance(cx->Rrel, cx->Trel, &cx->m1->tris[t1],
                               &cx->m2->tris[t2], p, q, NULL, NULL, NULL, NULL);
    const int improving = dTri <= res->distance;
    if (improving || g_strict) {
      double tmp1[3], tmp2[3], w1[3], w2[3], S1[3], S2[3];
      mxv(tmp1, cx->r1, p);
      add3(w1, tmp1, cx->tt1);
      mxv(tmp2, cx->r1, q);
      add3(w2, tmp2, cx->tt1);
      sub3(S1, w2, w1);
      scale3(S2, S1, -1);
      double mb1 = c2o_motion_bound_leaf(cx->mo1, B1->angR, S1);
      double mb2 = c2o_motion_bound_leaf(cx->mo2, B2->angR, S2);
      double mt = dTri / (mb1 + mb2);
      if (mt < 0.0) mt = 0.0;
      /* strict: a touching pair has no direction (0/NaN); it bounds mint by 0 */
      if (g_strict && dTri == 0.0) mt = 0.0;
      if (mt <= res->mint) res->mint = mt;
      if (improving) {
        res->distance = dTri;
        memcpy(res->p1, p, sizeof(p));
        memcpy(res->p2, q, sizeof(q));
        cx->last1 = t1;
        cx->last2 = t2;
      }
    }
    res->num_tri_tests++;
    return;
  }
  int a1, a2, c1, c2;
  m3 R1, R2;
  double T1[3], T2[3], Tt[3];
  if (l2 || (!l1 && (B1->size > B2->size))) {
    a1 = B1->first_child; c1 = a1 + 1; a2 = c2 = b2;
    const c2o_bv *A = &cx->m1->bvs[a1], *C = &cx->m1->bvs[c1];
    mtxm(R1, A->R, R);
    sub3(Tt, T, A->Tr);
    mtxv(T1, A->R, Tt);
    mtxm(R2, C->R, R);
    sub3(Tt, T, C->Tr);
    mtxv(T2, C->R, Tt);
  } else {
    a1 = c1 = b1; a2 = B2->first_child; c2 = a2 + 1;
    const c2o_bv *A = &cx->m2->bvs[a2], *C = &cx->m2->bvs[c2];
    mxm(R1, R, A->R);
    mxvpv(T1, R, A->Tr, T);
    mxm(R2, R, C->R);
    mxvpv(T2, R, C->Tr, T);
  }
  const c2o_bv *A1 = &cx->m1->bvs[a1], *A2 = &cx->m2->bvs[a2];
  const c2o_bv *C1 = &cx->m1->bvs[c1], *C2 = &cx->m2->bvs[c2];
  double S[3], tmp[3], Sw[3], Sn[3];
  double minta, mintb;
  double d1 = bv_distance(R1, T1, A1, A2, S);
  if (d1 != 0.0) {
    mxv(tmp, A1->R_loc, S);
    mxv(Sw, cx->r1, tmp);
    scale3(Sn, Sw, -1);
    double mb1 = c2o_motion_bound(cx->mo1, A1->angR, Sw);
    double mb2 = c2o_motion_bound(cx->mo2, A2->angR, Sn);
    mintb = d1 / (mb1 + mb2);
    if (mintb <= 0) mintb = 0.0;
  } else {
    mintb = 0.0;
  }
  double d2 = bv_distance(R2, T2, C1, C2, S);
  if (d2 != 0.0) {
    mxv(tmp, C1->R_loc, S);
    mxv(Sw, cx->r1, tmp);
    scale3(Sn, Sw, -1);
    double mb1 = c2o_motion_bound(cx->mo1, C1->angR, Sw);
    double mb2 = c2o_motion_bound(cx->mo2, C2->angR, Sn);
    minta = d2 / (mb1 + mb2);
    if (minta <= 0) minta = 0.0;
  } else {
    minta = 0.0;
  }
  res->num_bv_tests += 2;
  if (d2 < d1) {
    if (desc_pred(res, d2, minta)) toc_recurse(cx, R2, T2, c1, c2);
    else if (minta < res->mint) res->mint = minta;
    if (desc_pred(res, d1, mintb)) toc_recurse(cx, R1, T1, a1, a2);
    else if (mintb < res->mint) res->mint = mintb;
  } else {
    if (desc_pred(res, d1, mintb)) toc_recurse(cx, R1, T1, a1, a2);
    else if (mintb < res->mint) res->mint = mintb;
    if (desc_pred(res, d2, minta)) toc_recurse(cx, R2, T2, c1, c2);
    else if (minta < res->mint) res->mint = minta;
  }
}

/* Relative transform of model 2 in model 1's frame and the root-pair
 * transform (C2A.cpp:1793-1806). */
static void root_transform(const c2o_model *m1, const c2o_model *m2,
                           const m3 R1, const double T1[3], const m3 R2,
                           const double T2[3], m3 Rrel, double Trel[3], m3 R,
                           double T[3]) {
  double Tt[3];
  mtxm(Rrel, R1, R2);
  sub3(Tt, T2, T1);
  mtxv(Trel, R1, Tt);
  m3 Rt;
  mxm(Rt, Rrel, m2->bvs[0].R);
  mtxm(R, m1->bvs[0].R, Rt);
  mxvpv(Tt, Rrel, m2->bvs[0].Tr, Trel);
  sub3(Tt, Tt, m1->bvs[0].Tr);
  mtxv(T, m1->bvs[0].R, Tt);
}

/* C2A_TimeOfContactStep, rotation path (C2A.cpp:1778-1931). */
void c2o_toc_step(const c2o_model *m1, const c2o_model *m2,
                  const c2o_motion *mo1, const c2o_motion *mo2,
                  const double R1[3][3], const double T1[3],
                  const double R2[3][3], const double T2[3],
                  c2o_model *upd1, c2o_model *upd2, c2o_result *res) {
  step_ctx cx;
  m3 R;
  double T[3], p[3], q[3];
  cx.m1 = m1; cx.m2 = m2; cx.mo1 = mo1; cx.mo2 = mo2;
  cx.r1 = R1; cx.tt1 = T1; cx.res = res;
  root_transform(m1, m2, R1, T1, R2, T2, cx.Rrel, cx.Trel, R, T);
  cx.last1 = upd1 ? upd1->last_tri : -1;
  cx.last2 = upd2 ? upd2->last_tri : -1;
  res->distance = tri_distance(cx.Rrel, cx.Trel, &m1->tris[res->last_triA],
                               &m2->tris[res->last_triB], p, q, NULL, NULL, NULL, NULL);
  if (res->numCA == 0) res->mint = 1;
  if (res->mint <= 0.005 || res->distance <= 0.5 || res->numCA > 5) {
    res->abs_err = 0;
    res->rel_err = 0;
  } else {
    res->abs_err = 1e+30;
    res->rel_err = (res->numCA <= 2) ? 3 : 0.5;
  }
  res->mint = 1;
  memcpy(res->R, cx.Rrel, sizeof(m3));      /* res->R, res->T (C2A.cpp:1793-1796) */
  memcpy(res->T, cx.Trel, sizeof(cx.Trel));
  toc_recurse(&cx, R, T, 0, 0);
  if (upd1) upd1->last_tri = cx.last1;
  if (upd2) upd2->last_tri = cx.last2;
}

/* ------------------------------------------------------------------------ */
/* Translation-only CA (C2A_Solve when both rotations are below 1e-8, the    */
/* global b_TanslationCCD): C2A.cpp:1362-1521 (traversal), 1818-1852 (seed   */
/* bound), 1900-1917 (final pose), 2032-2049 (toc = mint, no CA loop);       */
/* InterpMotion.cpp:289-298 (ConservD), 570-655 (CAonRSS), 657-743           */
/* (CAonNonAdjacentTriangles).  Only object 1's velocity enters: object 2 is */
/* treated as static.  m_toc_delta is C2A_Solve's d_delta (the query's       */
/* tolerance_d, C2A.cpp:2384-2388).                                          */
/* ------------------------------------------------------------------------ */
static int g_solve_translation = 1;
void c2o_set_solve_translation(int on) { g_solve_translation = on; }

/* ConservD (InterpMotion.cpp:289-298): GMP_CCD_SECURITY_DISTANCE_RATIO 0.1. */
static double conserv_d(double d, double delta) { return d - 0.1 * delta; }

/* CAonRSS (InterpMotion.cpp:570-655) for BV pair (b1, b2) at transform
 * (R, T): local conservative advancement of the pair under object 1's
 * translation, capped at *mint.  The reference reads an uninitialised
 * direction when the two rectangles overlap at the start (C2ARectDist leaves
 * S unset); here such a pair is touching now: time 0, distance 0.  The
 * bValid_C_clo == false branch (:607-617) is dead: C2ARectDist always
 * reports valid closest points. */
static int ca_on_rss(const c2o_motion *mo1, const m3 r1, const m3 R, const double T[3],
                     const c2o_bv *b1, const c2o_bv *b2, double delta, double *mint,
                     double *distance) {
  double S[3] = {0.0, 0.0, 0.0}, tmp[3], Sw[3], Vel[3], T2[3];
  double d = bv_distance(R, T, b1, b2, S);
  mtxv(tmp, r1, mo1->cv);
  mtxv(Vel, b1->R_loc, tmp);
  if (d == 0.0) {
    *mint = 0.0;
    *distance = 0.0;
    return 1;
  }
  mxv(tmp, b1->R_loc, S);
  mxv(Sw, r1, tmp);
  double tocf = conserv_d(d, delta) / c2o_motion_bound_leaf(mo1, b1->angR, Sw);
  double total = tocf;
  int it = 1;
  while ((d >= delta) && (total <= mint[0]) && (it < 50)) {
    for (int i = 0; i < 3; i++) T2[i] = T[i] + Vel[i] * (-total);
    d = bv_distance(R, T2, b1, b2, S);
    if (d == 0) break;
    mxv(tmp, b1->R_loc, S);
    mxv(Sw, r1, tmp);
    tocf = conserv_d(d, delta) / c2o_motion_bound_leaf(mo1, b1->angR, Sw);
    it++;
    if (tocf < delta) break;
    total += tocf;
  }
  if (total < 1.0) {
    *mint = total;
    double dd = (total - delta) * len3(mo1->cv);
    *distance = dd < 0 ? 0 : dd;
    return 1;
  }
  return 0;
}

/* CAonNonAdjacentTriangles (InterpMotion.cpp:657-743): triangle A (model-1
 * frame) advancing with object 1's velocity against triangle B (already in
 * model-1 frame). */
static int ca_on_tris(const c2o_motion *mo1, const m3 r1, const double A[3][3],
                      const double B[3][3], double delta, double *mint, double *distance) {
  double Vel[3], p[3], q[3], n[3], At[3][3];
  mtxv(Vel, r1, mo1->cv);
  double d = c2o_tri_dist(p, q, A, B, NULL, NULL, NULL, NULL);
  sub3(n, q, p);
  normalize3(n);
  double u = dot3(Vel, n);
  if (u <= 0) u = 1e-30;
  if (d == 0) {
    *mint = 0.0;
    *distance = 0.0;
    return 1;
  }
  double dt = conserv_d(d, delta) / u;
  int it = 1;
  if (dt >= mint[0]) return 0;
  double total = dt;
  while ((d > delta) && (total <= mint[0]) && (it < 50)) {
    for (int v = 0; v < 3; v++)
      for (int i = 0; i < 3; i++) At[v][i] = A[v][i] + Vel[i] * total;
    d = c2o_tri_dist(p, q, At, B, NULL, NULL, NULL, NULL);
    if (d == 0.0) break;
    sub3(n, q, p);
    normalize3(n);
    u = dot3(Vel, n);
    if (u <= 0) u = 1e-30;
    double tofc = conserv_d(d, delta) / u;
    it++;
    if (tofc < delta) break;
    total += tofc;
  }
  if (total <= mint[0] && total >= 0.0) {
    *mint = total;
    *distance = total * len3(mo1->cv) + delta;
    return 1;
  }
  return 0;
}

typedef struct {
  const c2o_model *m1, *m2;
  const c2o_motion *mo1;
  const double (*r1)[3];
  m3 Rrel;
  double Trel[3];
  double delta;
  c2o_result *res;
  int la, lb;                /* res->last_triA/B */
} trans_ctx;

/* Triangle t of model 2 in model 1's frame (MxVpV, C2A.cpp:1403-1405). */
static void tri_in_frame1(double out[3][3], const m3 R, const double T[3], const c2o_tri *t) {
  mxvpv(out[0], R, t->p[0], T);
  mxvpv(out[1], R, t->p[1], T);
  mxvpv(out[2], R, t->p[2], T);
}

static void tri_copy(double out[3][3], const c2o_tri *t) { memcpy(out, t->p, sizeof(t->p)); }

/* TOCStepRecurse_Dis_Translation (C2A.cpp:1362-1521). */
static void trans_recurse(trans_ctx *cx, const m3 R, const double T[3], int b1, int b2) {
  const c2o_bv *B1 = &cx->m1->bvs[b1], *B2 = &cx->m2->bvs[b2];
  c2o_result *res = cx->res;
  int l1 = B1->first_child < 0, l2 = B2->first_child < 0;
  if (l1 && l2) {
    int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
    double A[3][3], Bt[3][3];
    tri_copy(A, &cx->m1->tris[t1]);
    tri_in_frame1(Bt, cx->Rrel, cx->Trel, &cx->m2->tris[t2]);
    if (ca_on_tris(cx->mo1, cx->r1, A, Bt, cx->delta, &res->mint, &res->distance)) {
      cx->la = t1;
      cx->lb = t2;
    }
    res->num_tri_tests++;
    return;
  }
  int a1, a2, c1, c2;
  m3 R1, R2;
  double T1[3], T2[3], Tt[3];
  if (l2 || (!l1 && (B1->size > B2->size))) {
    a1 = B1->first_child; c1 = a1 + 1; a2 = c2 = b2;
    const c2o_bv *A = &cx->m1->bvs[a1], *C = &cx->m1->bvs[c1];
    mtxm(R1, A->R, R);
    sub3(Tt, T, A->Tr);
    mtxv(T1, A->R, Tt);
    mtxm(R2, C->R, R);
    sub3(Tt, T, C->Tr);
    mtxv(T2, C->R, Tt);
  } else {
    a1 = c1 = b1; a2 = B2->first_child; c2 = a2 + 1;
    const c2o_bv *A = &cx->m2->bvs[a2], *C = &cx->m2->bvs[c2];
    mxm(R1, R, A->R);
    mxvpv(T1, R, A->Tr, T);
    mxm(R2, R, C->R);
    mxvpv(T2, R, C->Tr, T);
  }
  double d1 = 1e+30, d2 = 1e+30, minta = res->mint, mintc = res->mint;
  ca_on_rss(cx->mo1, cx->r1, R1, T1, &cx->m1->bvs[a1], &cx->m2->bvs[a2], cx->delta, &minta, &d1);
  ca_on_rss(cx->mo1, cx->r1, R2, T2, &cx->m1->bvs[c1], &cx->m2->bvs[c2], cx->delta, &mintc, &d2);
  res->num_bv_tests += 2;
  if (d2 < d1) {
    if (mintc < res->mint &&
        ((d2 < (res->distance - res->abs_err)) || (d2 * (1 + res->rel_err) < res->distance)))
      trans_recurse(cx, R2, T2, c1, c2);
    if (minta < res->mint &&
        ((d1 < (res->distance - res->abs_err)) || (d1 * (1 + res->rel_err) < res->distance)))
      trans_recurse(cx, R1, T1, a1, a2);
  } else {
    if (minta < res->mint &&
        ((d1 < (res->distance - res->abs_err)) || (d1 * (1 + res->rel_err) < res->distance)))
      trans_recurse(cx, R1, T1, a1, a2);
    if (mintc < res->mint &&
        ((d2 < (res->distance - res->abs_err)) || (d2 * (1 + res->rel_err) < res->distance)))
      trans_recurse(cx, R2, T2, c1, c2);
  }
}

/* The translation-only query: one step from the start poses, then
 * toc = mint (C2A.cpp:2032-2049).  The reference copies uninitialised
 * closest points into res->p1/p2 (C2A.cpp:1411-1412); here they are the
 * closest points of the final triangle distance.  The models' last_tri are
 * not updated (the branch writes res->last_triA/B only). */
static double query_toc_translation(c2o_model *m1, c2o_model *m2, c2o_motion *mo1,
                                    c2o_motion *mo2, c2o_result *res, double delta) {
  trans_ctx cx;
  m3 R, R1, R2;
  double T[3], T1[3], T2[3], Ttmp[3];
  memcpy(R1, mo1->R0, sizeof(m3));
  memcpy(T1, mo1->T0, sizeof(T1));
  memcpy(R2, mo2->R0, sizeof(m3));
  memcpy(T2, mo2->T0, sizeof(T2));
  cx.m1 = m1; cx.m2 = m2; cx.mo1 = mo1; cx.r1 = R1; cx.delta = delta; cx.res = res;
  root_transform(m1, m2, R1, T1, R2, T2, cx.Rrel, cx.Trel, R, T);
  cx.la = res->last_triA;
  cx.lb = res->last_triB;
  /* seed bound from the caller's triangle pair (C2A.cpp:1818-1852) */
  double A[3][3], B[3][3], mint = 1.0, dTri = 0.0;
  tri_copy(A, &m1->tris[cx.la]);
  tri_in_frame1(B, cx.Rrel, cx.Trel, &m2->tris[cx.lb]);
  if (ca_on_tris(mo1, R1, A, B, delta, &mint, &dTri)) {
    res->mint = mint;
    res->distance = dTri;
  } else {
    res->mint = 1.0;
    res->distance = 1e+30;
  }
  res->abs_err = 0;
  res->rel_err = 0;
  trans_recurse(&cx, R, T, 0, 0);
  /* distance at the advanced pose (C2A.cpp:1906-1916) */
  c2o_motion_integrate(mo1, res->mint, R1, T1, NULL);
  c2o_motion_integrate(mo2, res->mint, R2, T2, NULL);
  m3 Rr;
  double Tr[3];
  mtxm(Rr, R1, R2);
  sub3(Ttmp, T2, T1);
  mtxv(Tr, R1, Ttmp);
  res->distance = tri_distance(Rr, Tr, &m1->tris[cx.la], &m2->tris[cx.lb], res->p1, res->p2,
                               NULL, NULL, NULL, NULL);
  memcpy(res->R, Rr, sizeof(m3));           /* res->R, res->T (C2A.cpp:1902-1904) */
  memcpy(res->T, Tr, sizeof(Tr));
  res->toc = res->mint;
  res->collisionfree = res->toc >= 1.0 ? 1 : 0;
  res->numCA = 0;
  if (!res->collisionfree) c2o_motion_integrate(mo1, res->toc, res->R_toc, res->T_toc, NULL);
  /* the models' last_tri stay as they were; res->last_triA/B become the last
   * improving pair (C2A.cpp:1413-1414), which out_last_tri* report here */
  res->out_last_triA = cx.la;
  res->out_last_triB = cx.lb;
  return res->toc;
}

/* The CA loop of C2A_QueryTimeOfContact, rotation path (C2A.cpp:2029-2143),
 * over an abstract step: step(ctx, lamda, res) runs C2A_TimeOfContactStep at
 * time lamda (numCA, mint and UpboundTOC as the loop leaves them in res) and
 * sets res->distance and res->mint.  The first step runs at lamda 0.  Leaves
 * toc, collisionfree, numCA and the iteration-cap flag in res; returns the
 * TOC lamda at which R_toc/T_toc must be evaluated, or -1 when collision
 * free. */
typedef void (*ca_step_fn)(void *ctx, double lamda, c2o_result *res);

static double ca_loop(ca_step_fn step, void *ctx, c2o_result *res, double tol_d,
                      double tol_t, int max_iter) {
  double lamda = 0.0, lastLamda, dist, mint;
  int nItrs = 0;
  step(ctx, 0.0, res);
  dist = res->distance;
  mint = res->mint;
  res->numCA = 1;
  lastLamda = mint;                          /* C2A.cpp:2072 */
  while (dist > tol_d) {
    nItrs++;
    if (nItrs > max_iter) { res->flags |= C2O_FLAG_ITER_CAP; break; }
    if (mint >= 1.0) {                       /* C2A.cpp:2086-2093 */
      res->collisionfree = 1;
      res->toc = 0;
      return -1.0;
    }
    double dlamda = mint;
    if (dlamda < tol_t) break;               /* C2A.cpp:2098-2101 */
    lamda += dlamda;
    if (lamda >= 1.0) {                      /* C2A.cpp:2103-2110 */
      res->collisionfree = 1;
      res->toc = 0;
      return -1.0;
    }
    lastLamda = lamda;
    res->numCA++;
    res->UpboundTOC = 1.0 - lamda;
    step(ctx, lamda, res);
    dist = res->distance;
    mint = res->mint;
  }
  /* the dist == 0 && mint < 0 branch (C2A.cpp:2125-2128) cannot run: mint is
   * clamped at 0 */
  double toc = lastLamda;
  res->collisionfree = 0;
  if (toc >= 1 - tol_t) toc = 0;             /* C2A.cpp:2136-2137 */
  res->toc = toc;
  return toc;
}

typedef struct {
  c2o_model *m1, *m2;
  c2o_motion *mo1, *mo2;
  int calls;
} query_ctx;

static void query_step(void *vctx, double lamda, c2o_result *res) {
  query_ctx *cx = (query_ctx *)vctx;
  m3 R1, R2;
  double T1[3], T2[3];
  if (cx->calls++ == 0) {                    /* the first step: the start poses as given */
    memcpy(R1, cx->mo1->R0, sizeof(m3));
    memcpy(T1, cx->mo1->T0, sizeof(T1));
    memcpy(R2, cx->mo2->R0, sizeof(m3));
    memcpy(T2, cx->mo2->T0, sizeof(T2));
  } else {
    c2o_motion_integrate(cx->mo1, lamda, R1, T1, NULL);
    c2o_motion_integrate(cx->mo2, lamda, R2, T2, NULL);
  }
  c2o_toc_step(cx->m1, cx->m2, cx->mo1, cx->mo2, R1, T1, R2, T2, cx->m1, cx->m2, res);
}

/* C2A_QueryTimeOfContact (C2A.cpp:1987-2146); with both rotations below
 * 1e-8 and the translation switch on (C2A_Solve's b_TanslationCCD), the
 * translation-only branch. */
double c2o_query_toc(c2o_model *m1, c2o_model *m2, c2o_motion *mo1,
                     c2o_motion *mo2, c2o_result *res, double tol_d,
                     double tol_t, int max_iter) {
  res->num_bv_tests = 0;
  res->num_tri_tests = 0;
  res->num_contact = 0;
  res->UpboundTOC = 1;
  res->numCA = 0;
  res->flags = 0;
  if (mo1->angVel < 1e-8 && mo2->angVel < 1e-8) res->flags |= C2O_FLAG_TRANSLATION_ONLY;
  if (g_solve_translation && !g_strict && (res->flags & C2O_FLAG_TRANSLATION_ONLY))
    return query_toc_translation(m1, m2, mo1, mo2, res, tol_d);
  query_ctx cx = {m1, m2, mo1, mo2, 0};
  const double toc = ca_loop(query_step, &cx, res, tol_d, tol_t, max_iter);
  if (toc >= 0.0) c2o_motion_integrate(mo1, toc, res->R_toc, res->T_toc, NULL);
  res->out_last_triA = m1->last_tri;
  res->out_last_triB = m2->last_tri;
  return res->toc;
}

/* The CA loop over a scripted step (tests of the loop's branches): call k
 * (k = 0 first) sets distance = dist[k] and mint = mint[k] and records the
 * lamda, numCA and UpboundTOC it was called with.  Calls beyond n reuse the
 * last entry.  Returns the number of step calls. */
typedef struct {
  const double *dist, *mint;
  int n, calls, cap;
  double *lamda_out, *up_out;
  int *numca_out;
} scripted_ctx;

static void scripted_step(void *vctx, double lamda, c2o_result *res) {
  scripted_ctx *cx = (scripted_ctx *)vctx;
  const int k = cx->calls < cx->n ? cx->calls : cx->n - 1;
  if (cx->calls < cx->cap) {
    cx->lamda_out[cx->calls] = lamda;
    cx->up_out[cx->calls] = res->UpboundTOC;
    cx->numca_out[cx->calls] = res->numCA;
  }
  cx->calls++;
  res->distance = cx->dist[k];
  res->mint = cx->mint[k];
}

int c2o_ca_loop_scripted(const double *dist, const double *mint, int n, double tol_d,
                         double tol_t, int max_iter, c2o_result *res, double *lamda_out,
                         double *up_out, int *numca_out, int cap) {
  scripted_ctx cx = {dist, mint, n, 0, cap, lamda_out, up_out, numca_out};
  memset(res, 0, sizeof(*res));
  res->UpboundTOC = 1;
  (void)ca_loop(scripted_step, &cx, res, tol_d, tol_t, max_iter);
  return cx.calls;
}

/* ------------------------------------------------------------------------ */
/* Contact enumeration (C2A.cpp:1525-1775)                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
  const c2o_model *m1, *m2;
  m3 Rrel;
  double Trel[3];
  double thr;
  int count, cap;
  c2o_contact *out;
  long long bv_tests;
} contact_ctx;

static void contact_recurse(contact_ctx *cx, const m3 R, const double T[3], int b1, int b2) {
  const c2o_bv *B1 = &cx->m1->bvs[b1], *B2 = &cx->m2->bvs[b2];
  int l1 = B1->first_child < 0, l2 = B2->first_child < 0;
  if (l1 && l2) {
    double p[3], q[3];
    int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
    int ta, fa, tb, fb;
    const c2o_tri *T1 = &cx->m1->tris[t1], *T2 = &cx->m2->tris[t2];
    double d = tri_distance(cx->Rrel, cx->Trel, T1, T2, p, q, &ta, &fa, &tb, &fb);
    if (ta == -1 || tb == -1) return;
    if (d <= cx->thr) {
      /* records fill from the end so the list reads like the reference's
       * push_front order */
      if (cx->count < cx->cap && cx->out) {
        c2o_contact *c = &cx->out[cx->count];
        c->triA = t1; c->triB = t2;
        c->typeA = ta + 1; c->typeB = tb + 1;
        for (int k = 0; k < 3; k++) { c->fidA[k] = -1; c->fidB[k] = -1; }
        if (ta == 0) c->fidA[0] = T1->idx[fa];
        if (ta == 1) { c->fidA[0] = T1->idx[fa]; c->fidA[1] = T1->idx[(fa + 1) % 3]; }
        if (ta == 2) { c->fidA[0] = T1->idx[0]; c->fidA[1] = T1->idx[1]; c->fidA[2] = T1->idx[2]; }
        if (tb == 0) c->fidB[0] = T2->idx[fb];
        if (tb == 1) { c->fidB[0] = T2->idx[fb]; c->fidB[1] = T2->idx[(fb + 1) % 3]; }
        if (tb == 2) { c->fidB[0] = T2->idx[0]; c->fidB[1] = T2->idx[1]; c->fidB[2] = T2->idx[2]; }
        memcpy(c->PA, p, sizeof(p));
        double tmp[3];
        sub3(tmp, q, cx->Trel);
        mtxv(c->PB, cx->Rrel, tmp);
        c->distance = d;
      }
      cx->count++;
    }
    return;
  }
  int a1, a2, c1, c2;
  m3 R1, R2;
  double T1[3], T2[3], Tt[3], S[3];
  if (l2 || (!l1 && (B1->size > B2->size))) {
    a1 = B1->first_child; c1 = a1 + 1; a2 = c2 = b2;
    const c2o_bv *A = &cx->m1->bvs[a1], *C = &cx->m1->bvs[c1];
    mtxm(R1, A->R, R); sub3(Tt, T, A->Tr); mtxv(T1, A->R, Tt);
    mtxm(R2, C->R, R); sub3(Tt, T, C->Tr); mtxv(T2, C->R, Tt);
  } else {
    a1 = c1 = b1; a2 = B2->first_child; c2 = a2 + 1;
    const c2o_bv *A = &cx->m2->bvs[a2], *C = &cx->m2->bvs[c2];
    mxm(R1, R, A->R); mxvpv(T1, R, A->Tr, T);
    mxm(R2, R, C->R); mxvpv(T2, R, C->Tr, T);
  }
  double d1 = bv_distance(R1, T1, &cx->m1->bvs[a1], &cx->m2->bvs[a2], S);
  double d2 = bv_distance(R2, T2, &cx->m1->bvs[c1], &cx->m2->bvs[c2], S);
  cx->bv_tests += 2;
  if (d2 < d1) {
    if (d2 < cx->thr) contact_recurse(cx, R2, T2, c1, c2);
    if (d1 < cx->thr) contact_recurse(cx, R1, T1, a1, a2);
  } else {
    if (d1 < cx->thr) contact_recurse(cx, R1, T1, a1, a2);
    if (d2 < cx->thr) contact_recurse(cx, R2, T2, c1, c2);
  }
}

int c2o_query_contact(const c2o_model *m1, const c2o_model *m2,
                      const double R1[3][3], const double T1[3],
                      const double R2[3][3], const double T2[3],
                      double threshold, c2o_contact *out, int cap,
                      long long *bv_tests) {
  contact_ctx cx;
  m3 R;
  double T[3];
  cx.m1 = m1; cx.m2 = m2; cx.thr = threshold; cx.count = 0; cx.cap = cap;
  cx.out = out; cx.bv_tests = 0;
  root_transform(m1, m2, R1, T1, R2, T2, cx.Rrel, cx.Trel, R, T);
  contact_recurse(&cx, R, T, 0, 0);
  /* reverse to the reference's push_front list order */
  int n = cx.count < cap ? cx.count : cap;
  if (out)
    for (int i = 0, j = n - 1; i < j; i++, j--) {
      c2o_contact t = out[i]; out[i] = out[j]; out[j] = t;
    }
  if (bv_tests) *bv_tests = cx.bv_tests;
  return cx.count;
}

/* ------------------------------------------------------------------------ */
/* Static distance query (C2A_PQP.cpp:970-1057 -> 481-620, qsize <= 2)       */
/* ------------------------------------------------------------------------ */
typedef struct {
  c2o_model *m1, *m2;
  m3 Rrel;
  double Trel[3];
  double rel_err, abs_err;
  int overlap_only;
  c2o_distance_result *r;
} distance_ctx;

/* C2ADistanceRecurse (C2A_PQP.cpp:481-620): closer child pair first, a pair
 * is entered while d < D - abs_err or d (1 + rel_err) < D; leaves replace the
 * bound only when strictly closer. */
static void distance_recurse(distance_ctx *cx, const m3 R, const double T[3], int b1, int b2) {
  const c2o_bv *B1 = &cx->m1->bvs[b1], *B2 = &cx->m2->bvs[b2];
  int l1 = B1->first_child < 0, l2 = B2->first_child < 0;
  c2o_distance_result *r = cx->r;
  /* C2ACollideRecurse enters a pair only if C2A_BV_Overlap holds
   * (C2A_PQP.cpp:1068-1070 -> C2A_BV.cpp:641-650: RectDist <= r1 + r2) */
  if (cx->overlap_only) {
    double P[3], Q[3], S[3];
    if (!(c2o_rect_dist(R, T, B1->l, B2->l, P, Q, S) <= B1->r + B2->r)) return;
  }
  if (l1 && l2) {
    double p[3], q[3];
    int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
    int ta, fa, tb, fb;
    r->tri_tests++;
    double d = tri_distance(cx->Rrel, cx->Trel, &cx->m1->tris[t1], &cx->m2->tris[t2], p, q,
                            &ta, &fa, &tb, &fb);
    if (d < r->distance) {
      r->distance = d;
      memcpy(r->p1, p, sizeof(p));
      memcpy(r->p2, q, sizeof(q));   /* model 1's frame until the end */
      cx->m1->last_tri = t1;
      cx->m2->last_tri = t2;
    }
    return;
  }
  int a1, a2, c1, c2;
  m3 R1, R2;
  double T1[3], T2[3], Tt[3], S[3];
  if (l2 || (!l1 && (B1->size > B2->size))) {
    a1 = B1->first_child; c1 = a1 + 1; a2 = c2 = b2;
    const c2o_bv *A = &cx->m1->bvs[a1], *C = &cx->m1->bvs[c1];
    mtxm(R1, A->R, R); sub3(Tt, T, A->Tr); mtxv(T1, A->R, Tt);
    mtxm(R2, C->R, R); sub3(Tt, T, C->Tr); mtxv(T2, C->R, Tt);
  } else {
    a1 = c1 = b1; a2 = B2->first_child; c2 = a2 + 1;
    const c2o_bv *A = &cx->m2->bvs[a2], *C = &cx->m2->bvs[c2];
    mxm(R1, R, A->R); mxvpv(T1, R, A->Tr, T);
    mxm(R2, R, C->R); mxvpv(T2, R, C->Tr, T);
  }
  r->bv_tests += 2;
  double d1 = bv_distance(R1, T1, &cx->m1->bvs[a1], &cx->m2->bvs[a2], S);
  double d2 = bv_distance(R2, T2, &cx->m1->bvs[c1], &cx->m2->bvs[c2], S);
#define ENTER(d) (((d) < (r->distance - cx->abs_err)) || ((d) * (1 + cx->rel_err) < r->distance))
  if (d2 < d1) {
    if (ENTER(d2)) distance_recurse(cx, R2, T2, c1, c2);
    if (ENTER(d1)) distance_recurse(cx, R1, T1, a1, a2);
  } else {
    if (ENTER(d1)) distance_recurse(cx, R1, T1, a1, a2);
    if (ENTER(d2)) distance_recurse(cx, R2, T2, c1, c2);
  }
#undef ENTER
}

static void distance_query(c2o_model *m1, c2o_model *m2, const double R1[3][3],
                           const double T1[3], const double R2[3][3], const double T2[3],
                           double rel_err, double abs_err, int overlap_only,
                           c2o_distance_result *r) {
  distance_ctx cx;
  cx.overlap_only = overlap_only;
  m3 R;
  double T[3], p[3], q[3];
  int ta, fa, tb, fb;
  cx.m1 = m1; cx.m2 = m2; cx.rel_err = rel_err; cx.abs_err = abs_err; cx.r = r;
  root_transform(m1, m2, R1, T1, R2, T2, cx.Rrel, cx.Trel, R, T);
  /* first bound: the models' last_tri (C2A_PQP.cpp:998-1005) */
  r->distance = tri_distance(cx.Rrel, cx.Trel, &m1->tris[m1->last_tri], &m2->tris[m2->last_tri],
                             p, q, &ta, &fa, &tb, &fb);
  memcpy(r->p1, p, sizeof(p));
  memcpy(r->p2, q, sizeof(q));
  r->bv_tests = 0;
  r->tri_tests = 0;
  distance_recurse(&cx, R, T, 0, 0);
  /* p2 into model 2's frame (C2A_PQP.cpp:1043-1047) */
  double u[3];
  sub3(u, r->p2, cx.Trel);
  mtxv(r->p2, cx.Rrel, u);
  r->last_tri[0] = m1->last_tri;
  r->last_tri[1] = m2->last_tri;
}

void c2o_distance(c2o_model *m1, c2o_model *m2, const double R1[3][3], const double T1[3],
                  const double R2[3][3], const double T2[3], double rel_err, double abs_err,
                  c2o_distance_result *r) {
  distance_query(m1, m2, R1, T1, R2, T2, rel_err, abs_err, 0, r);
}

/* C2A_Collide, C2A_DistanceResult form (C2A_PQP.cpp:1199-1279): the distance
 * search entering only overlapping RSS pairs, the root pair included. */
void c2o_collide_distance(c2o_model *m1, c2o_model *m2, const double R1[3][3],
                          const double T1[3], const double R2[3][3], const double T2[3],
                          double rel_err, double abs_err, c2o_distance_result *r) {
  distance_query(m1, m2, R1, T1, R2, T2, rel_err, abs_err, 1, r);
}

/* ------------------------------------------------------------------------ */
/* C2A_Solve (C2A.cpp:2315-2444)                                             */
/* ------------------------------------------------------------------------ */
int c2o_solve(const double R00[3][3], const double T00[3],
              const double R01[3][3], const double T01[3], c2o_model *m1,
              const double R10[3][3], const double T10[3],
              const double R11[3][3], const double T11[3], c2o_model *m2,
              double R0out[3][3], double T0out[3],
              double R1out[3][3], double T1out[3], c2o_result *res) {
  c2o_motion mo1, mo2;
  c2o_motion_init(&mo1, R00, T00, R01, T01);
  c2o_motion_init(&mo2, R10, T10, R11, T11);
  res->last_triA = m1->last_tri;
  res->last_triB = m2->last_tri;
  c2o_query_toc(m1, m2, &mo1, &mo2, res, 1e-4, 1e-4, 150);
  if (!res->collisionfree) {
    double q7[7];
    m3 Ra, Rb;
    double Ta[3], Tb[3];
    c2o_motion_integrate(&mo1, res->toc, Ra, Ta, q7);
    if (R0out) c2o_matrix_from_quat((double[4]){q7[1], q7[2], q7[3], q7[0]}, R0out);
    if (T0out) memcpy(T0out, &q7[4], sizeof(double) * 3);
    c2o_motion_integrate(&mo2, res->toc, Rb, Tb, q7);
    if (R1out) c2o_matrix_from_quat((double[4]){q7[1], q7[2], q7[3], q7[0]}, R1out);
    if (T1out) memcpy(T1out, &q7[4], sizeof(double) * 3);
    double thr = 2 * res->distance + 0.001;
    res->num_contact = c2o_query_contact(m1, m2, Ra, Ta, Rb, Tb, thr, NULL, 0, NULL);
  }
  return 1;   /* TOCFound, always (C2A.cpp:2442) */
}

/* ------------------------------------------------------------------------ */
/* Batch helper                                                              */
/* ------------------------------------------------------------------------ */
static void pose_unpack(const double *p12, m3 R, double T[3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R[i][j] = p12[3 * i + j];
  T[0] = p12[9]; T[1] = p12[10]; T[2] = p12[11];
}

typedef struct {
  const c2o_model *m1, *m2;
  int nq;
  const double *p00, *p01, *p10, *p11;
  const int *seeds;
  double tol_d, tol_t;
  int max_iter, with_contacts, chain;
  c2o_result *out;
  int next;
  pthread_mutex_t mu;
} batch_job;

static void run_one(batch_job *job, int i, c2o_model *w1, c2o_model *w2) {
  m3 R00, R01, R10, R11;
  double T00[3], T01[3], T10[3], T11[3];
  pose_unpack(job->p00 + 12 * (size_t)i, R00, T00);
  pose_unpack(job->p01 + 12 * (size_t)i, R01, T01);
  pose_unpack(job->p10 + 12 * (size_t)i, R10, T10);
  pose_unpack(job->p11 + 12 * (size_t)i, R11, T11);
  c2o_motion mo1, mo2;
  c2o_motion_init(&mo1, R00, T00, R01, T01);
  c2o_motion_init(&mo2, R10, T10, R11, T11);
  c2o_result *res = &job->out[i];
  memset(res, 0, sizeof(*res));
  int sA = job->seeds ? job->seeds[2 * i] : -1;
  int sB = job->seeds ? job->seeds[2 * i + 1] : -1;
  /* the demo protocol seeds from the models' last_tri
   * (mainTorusknot.cpp:314-315), so an explicit seed is the model state */
  if (sA >= 0) w1->last_tri = sA;
  if (sB >= 0) w2->last_tri = sB;
  res->last_triA = w1->last_tri;
  res->last_triB = w2->last_tri;
  c2o_query_toc(w1, w2, &mo1, &mo2, res, job->tol_d, job->tol_t, job->max_iter);
  if (job->with_contacts && !res->collisionfree) {
    m3 Ra, Rb;
    double Ta[3], Tb[3];
    c2o_motion_integrate(&mo1, res->toc, Ra, Ta, NULL);
    c2o_motion_integrate(&mo2, res->toc, Rb, Tb, NULL);
    double thr = 2 * res->distance + 0.001;
    res->num_contact = c2o_query_contact(w1, w2, Ra, Ta, Rb, Tb, thr, NULL, 0, NULL);
  }
}

static void *batch_worker(void *arg) {
  batch_job *job = (batch_job *)arg;
  /* shallow per-thread copies: only last_tri is mutated */
  c2o_model w1 = *job->m1, w2 = *job->m2;
  for (;;) {
    pthread_mutex_lock(&job->mu);
    int i = job->next++;
    pthread_mutex_unlock(&job->mu);
    if (i >= job->nq) break;
    w1.last_tri = job->m1->last_tri;
    w2.last_tri = job->m2->last_tri;
    run_one(job, i, &w1, &w2);
  }
  return NULL;
}

int c2o_batch_toc(const c2o_model *m1, const c2o_model *m2, int nq,
                  const double *pose00, const double *pose01,
                  const double *pose10, const double *pose11,
                  const int *seeds, double tol_d, double tol_t, int max_iter,
                  int with_contacts, int chain, int nthreads, c2o_result *out) {
  batch_job job;
  job.m1 = m1; job.m2 = m2; job.nq = nq;
  job.p00 = pose00; job.p01 = pose01; job.p10 = pose10; job.p11 = pose11;
  job.seeds = seeds; job.tol_d = tol_d; job.tol_t = tol_t;
  job.max_iter = max_iter; job.with_contacts = with_contacts; job.chain = chain;
  job.out = out; job.next = 0;
  if (chain || nthreads <= 1) {
    c2o_model w1 = *m1, w2 = *m2;
    for (int i = 0; i < nq; i++) {
      if (!chain) { w1.last_tri = m1->last_tri; w2.last_tri = m2->last_tri; }
      run_one(&job, i, &w1, &w2);
    }
    return 0;
  }
  pthread_mutex_init(&job.mu, NULL);
  pthread_t th[256];
  if (nthreads > 256) nthreads = 256;
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, batch_worker, &job);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&job.mu);
  return 0;
}

/* Independent distance queries over threads (the CPU baseline of the GPU
 * batch): query i seeds the models' last_tri with seeds[2i], seeds[2i+1]. */
typedef struct {
  const c2o_model *m1, *m2;
  const double *poses;
  const int *seeds;
  double rel_err, abs_err;
  c2o_distance_result *out;
  int nq, next;
  pthread_mutex_t mu;
} distance_job;

static void distance_one(distance_job *job, int i, c2o_model *w1, c2o_model *w2) {
  m3 R1, R2;
  double T1[3], T2[3];
  pose_unpack(job->poses + 24 * (size_t)i, R1, T1);
  pose_unpack(job->poses + 24 * (size_t)i + 12, R2, T2);
  w1->last_tri = job->seeds ? job->seeds[2 * i] : 0;
  w2->last_tri = job->seeds ? job->seeds[2 * i + 1] : 0;
  c2o_distance(w1, w2, R1, T1, R2, T2, job->rel_err, job->abs_err, &job->out[i]);
}

static void *distance_worker(void *arg) {
  distance_job *job = (distance_job *)arg;
  c2o_model w1 = *job->m1, w2 = *job->m2;   /* only last_tri is written */
  for (;;) {
    pthread_mutex_lock(&job->mu);
    int i = job->next++;
    pthread_mutex_unlock(&job->mu);
    if (i >= job->nq) break;
    distance_one(job, i, &w1, &w2);
  }
  return NULL;
}

void c2o_distance_batch(const c2o_model *m1, const c2o_model *m2, int nq, const double *poses24,
                        const int *seeds, double rel_err, double abs_err, int nthreads,
                        c2o_distance_result *out) {
  distance_job job;
  job.m1 = m1; job.m2 = m2; job.poses = poses24; job.seeds = seeds;
  job.rel_err = rel_err; job.abs_err = abs_err; job.out = out; job.nq = nq; job.next = 0;
  if (nthreads <= 1) {
    c2o_model w1 = *m1, w2 = *m2;
    for (int i = 0; i < nq; i++) distance_one(&job, i, &w1, &w2);
    return;
  }
  pthread_mutex_init(&job.mu, NULL);
  pthread_t th[256];
  if (nthreads > 256) nthreads = 256;
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, distance_worker, &job);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&job.mu);
}

/* ------------------------------------------------------------------------ */
/* Static queries beyond C2A_Distance's depth-first form                    */
/* ------------------------------------------------------------------------ */

/* Child pair transforms of an expansion (the descend rule and transforms
 * shared by every recursion of C2A_PQP.cpp, e.g. :527-576 and :1343-1392):
 * returns the two child pairs (a1, a2), (c1, c2) and their transforms. */
static void expand_pair(const c2o_model *m1, const c2o_model *m2, const m3 R, const double T[3],
                        int b1, int b2, int *a1, int *a2, int *c1, int *c2, m3 R1, double T1[3],
                        m3 R2, double T2[3]) {
  const c2o_bv *B1 = &m1->bvs[b1], *B2 = &m2->bvs[b2];
  int l1 = B1->first_child < 0, l2 = B2->first_child < 0;
  double Tt[3];
  if (l2 || (!l1 && (B1->size > B2->size))) {
    *a1 = B1->first_child; *c1 = *a1 + 1; *a2 = *c2 = b2;
    const c2o_bv *A = &m1->bvs[*a1], *C = &m1->bvs[*c1];
    mtxm(R1, A->R, R); sub3(Tt, T, A->Tr); mtxv(T1, A->R, Tt);
    mtxm(R2, C->R, R); sub3(Tt, T, C->Tr); mtxv(T2, C->R, Tt);
  } else {
    *a1 = *c1 = b1; *a2 = B2->first_child; *c2 = *a2 + 1;
    const c2o_bv *A = &m2->bvs[*a2], *C = &m2->bvs[*c2];
    mxm(R1, R, A->R); mxvpv(T1, R, A->Tr, T);
    mxm(R2, R, C->R); mxvpv(T2, R, C->Tr, T);
  }
}

/* PQP's BVTQ (the bounded priority queue of BV tests behind the qsize > 2
 * searches, C2A_PQP.cpp:624-786 and :1606-1759; PQP 1.3's BVTQ.h is not
 * vendored, version unpinned).  Restated as a binary min-heap on d over
 * `size` slots: AddTest sifts up past parents with a larger d, ExtractMinTest
 * moves the last test to the root and sifts it down past the smaller child
 * while that child's d is smaller.  Tests of equal d may leave in another
 * order than PQP's heap gives them. */
typedef struct {
  m3 R;
  double T[3];
  double d;
  int b1, b2;
} c2o_bvt;
typedef struct {
  int size, n;
  c2o_bvt *h;   /* h[1..n] */
} c2o_bvtq;

static void bvtq_init(c2o_bvtq *q, int size) {
  q->size = size;
  q->n = 0;
  q->h = (c2o_bvt *)malloc(sizeof(c2o_bvt) * (size_t)(size + 1));
}
static void bvtq_free(c2o_bvtq *q) { free(q->h); }
static void bvtq_add(c2o_bvtq *q, const c2o_bvt *t) {
  int i = ++q->n;
  while (i > 1 && q->h[i / 2].d > t->d) {
    q->h[i] = q->h[i / 2];
    i /= 2;
  }
  q->h[i] = *t;
}
static c2o_bvt bvtq_extract_min(c2o_bvtq *q) {
  c2o_bvt top = q->h[1];
  c2o_bvt last = q->h[q->n--];
  int i = 1;
  for (;;) {
    int c = 2 * i;
    if (c > q->n) break;
    if (c + 1 <= q->n && q->h[c + 1].d < q->h[c].d) c++;
    if (!(q->h[c].d < last.d)) break;
    q->h[i] = q->h[c];
    i = c;
  }
  if (q->n >= 1) q->h[i] = last;
  return top;
}

/* The two child tests of an expansion with their RSS distances
 * (C2A_PQP.cpp:693-771 / :1661-1746). */
static void child_tests(const c2o_model *m1, const c2o_model *m2, const c2o_bvt *p, c2o_bvt *t1,
                        c2o_bvt *t2) {
  double S[3];
  expand_pair(m1, m2, p->R, p->T, p->b1, p->b2, &t1->b1, &t1->b2, &t2->b1, &t2->b2, t1->R,
              t1->T, t2->R, t2->T);
  t1->d = bv_distance(t1->R, t1->T, &m1->bvs[t1->b1], &m2->bvs[t1->b2], S);
  t2->d = bv_distance(t2->R, t2->T, &m1->bvs[t2->b1], &m2->bvs[t2->b2], S);
}

typedef struct {
  c2o_model *m1, *m2;
  m3 Rrel;
  double Trel[3];
  double rel_err, abs_err;
  int qsize;
  c2o_distance_result *r;
} dqueue_ctx;

/* C2ADistanceQueueRecurse (C2A_PQP.cpp:624-786). */
static void distance_queue_recurse(dqueue_ctx *cx, const m3 R, const double T[3], int b1,
                                   int b2) {
  c2o_bvtq q;
  bvtq_init(&q, cx->qsize);
  c2o_bvt mt;
  memcpy(mt.R, R, sizeof(m3));
  memcpy(mt.T, T, sizeof(mt.T));
  mt.b1 = b1;
  mt.b2 = b2;
  mt.d = 0.0;
  c2o_distance_result *r = cx->r;
  for (;;) {
    const c2o_bv *B1 = &cx->m1->bvs[mt.b1], *B2 = &cx->m2->bvs[mt.b2];
    int l1 = B1->first_child < 0, l2 = B2->first_child < 0;
    if (l1 && l2) {   /* :645-669 */
      double p[3], qq[3];
      int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1, ta, fa, tb, fb;
      r->tri_tests++;
      double d = tri_distance(cx->Rrel, cx->Trel, &cx->m1->tris[t1], &cx->m2->tris[t2], p, qq,
                              &ta, &fa, &tb, &fb);
      if (d < r->distance) {
        r->distance = d;
        memcpy(r->p1, p, sizeof(p));
        memcpy(r->p2, qq, sizeof(qq));
        cx->m1->last_tri = t1;
        cx->m2->last_tri = t2;
      }
    } else if (q.n == q.size - 1) {   /* :670-676: no room for two tests */
      distance_queue_recurse(cx, mt.R, mt.T, mt.b1, mt.b2);
    } else {   /* :677-772 */
      c2o_bvt t1, t2;
      r->bv_tests += 2;
      child_tests(cx->m1, cx->m2, &mt, &t1, &t2);
      bvtq_add(&q, &t1);
      bvtq_add(&q, &t2);
    }
    if (q.n == 0) break;   /* :774-785 */
    mt = bvtq_extract_min(&q);
    if ((mt.d + cx->abs_err >= r->distance) && ((mt.d * (1 + cx->rel_err)) >= r->distance))
      break;
  }
  bvtq_free(&q);
}

/* C2A_Distance with qsize > 2 (C2A_PQP.cpp:970-1057, :1031-1037). */
void c2o_distance_queue(c2o_model *m1, c2o_model *m2, const double R1[3][3], const double T1[3],
                        const double R2[3][3], const double T2[3], double rel_err,
                        double abs_err, int qsize, c2o_distance_result *r) {
  if (qsize <= 2) {
    c2o_distance(m1, m2, R1, T1, R2, T2, rel_err, abs_err, r);
    return;
  }
  dqueue_ctx cx;
  m3 R;
  double T[3], p[3], q[3];
  int ta, fa, tb, fb;
  cx.m1 = m1; cx.m2 = m2; cx.rel_err = rel_err; cx.abs_err = abs_err; cx.qsize = qsize;
  cx.r = r;
  root_transform(m1, m2, R1, T1, R2, T2, cx.Rrel, cx.Trel, R, T);
  r->distance = tri_distance(cx.Rrel, cx.Trel, &m1->tris[m1->last_tri], &m2->tris[m2->last_tri],
                             p, q, &ta, &fa, &tb, &fb);
  memcpy(r->p1, p, sizeof(p));
  memcpy(r->p2, q, sizeof(q));
  r->bv_tests = 0;
  r->tri_tests = 0;
  distance_queue_recurse(&cx, R, T, 0, 0);
  double u[3];
  sub3(u, r->p2, cx.Trel);
  mtxv(r->p2, cx.Rrel, u);
  r->last_tri[0] = m1->last_tri;
  r->last_tri[1] = m2->last_tri;
}

/* C2A_Distance_At_ContactSpace (C2A_PQP.cpp:1282-1496): the depth-first
 * search at zero error bounds from distance = upper_bound, no seed pair;
 * every call returns at once after the first leaf closer than the bound
 * (:1290).  found = 1 when that happened (the closest-pair fields are then
 * set; otherwise distance = upper_bound and p1/p2 are not written). */
typedef struct {
  c2o_model *m1, *m2;
  m3 Rrel;
  double Trel[3];
  double upper;
  c2o_distance_result *r;
} dcs_ctx;

static void distance_cs_recurse(dcs_ctx *cx, const m3 R, const double T[3], int b1, int b2) {
  c2o_distance_result *r = cx->r;
  if (r->distance < cx->upper) return;   /* :1290 */
  const c2o_bv *B1 = &cx->m1->bvs[b1], *B2 = &cx->m2->bvs[b2];
  int l1 = B1->first_child < 0, l2 = B2->first_child < 0;
  if (l1 && l2) {   /* :1299-1328 */
    double p[3], q[3];
    int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1, ta, fa, tb, fb;
    r->tri_tests++;
    double d = tri_distance(cx->Rrel, cx->Trel, &cx->m1->tris[t1], &cx->m2->tris[t2], p, q,
                            &ta, &fa, &tb, &fb);
    if (d < r->distance) {
      r->distance = d;
      memcpy(r->p1, p, sizeof(p));
      memcpy(r->p2, q, sizeof(q));
      cx->m1->last_tri = t1;
      cx->m2->last_tri = t2;
    }
    return;
  }
  int a1, a2, c1, c2;
  m3 R1, R2;
  double T1[3], T2[3], S[3];
  expand_pair(cx->m1, cx->m2, R, T, b1, b2, &a1, &a2, &c1, &c2, R1, T1, R2, T2);
  r->bv_tests += 2;
  double d1 = bv_distance(R1, T1, &cx->m1->bvs[a1], &cx->m2->bvs[a2], S);
  double d2 = bv_distance(R2, T2, &cx->m1->bvs[c1], &cx->m2->bvs[c2], S);
  if (d2 < d1) {   /* :1391-1414 */
    if (d2 < r->distance) distance_cs_recurse(cx, R2, T2, c1, c2);
    if (d1 < r->distance) distance_cs_recurse(cx, R1, T1, a1, a2);
  } else {
    if (d1 < r->distance) distance_cs_recurse(cx, R1, T1, a1, a2);
    if (d2 < r->distance) distance_cs_recurse(cx, R2, T2, c1, c2);
  }
}

int c2o_distance_cs(c2o_model *m1, c2o_model *m2, const double R1[3][3], const double T1[3],
                    const double R2[3][3], const double T2[3], double upper_bound,
                    c2o_distance_result *r) {
  dcs_ctx cx;
  m3 R;
  double T[3];
  cx.m1 = m1; cx.m2 = m2; cx.upper = upper_bound; cx.r = r;
  root_transform(m1, m2, R1, T1, R2, T2, cx.Rrel, cx.Trel, R, T);
  r->distance = upper_bound;
  r->bv_tests = 0;
  r->tri_tests = 0;
  r->last_tri[0] = m1->last_tri;
  r->last_tri[1] = m2->last_tri;
  distance_cs_recurse(&cx, R, T, 0, 0);
  int found = r->distance < upper_bound;
  if (found) {   /* :1484-1488 */
    double u[3];
    sub3(u, r->p2, cx.Trel);
    mtxv(r->p2, cx.Rrel, u);
  }
  r->last_tri[0] = m1->last_tri;
  r->last_tri[1] = m2->last_tri;
  return found;
}

/* C2A_Tolerance (C2A_PQP.cpp:1762-1845): is any triangle pair within
 * `tolerance`?  Depth first, closer child pair first, a pair entered while
 * its RSS distance is <= tolerance, stopping at the first such triangle
 * pair (:1498-1604); qsize > 2 searches best first (:1606-1759). */
typedef struct {
  const c2o_model *m1, *m2;
  m3 Rrel;
  double Trel[3];
  double tol;
  int qsize;
  c2o_tolerance_result *r;
} tol_ctx;

static int tol_leaf(tol_ctx *cx, int b1, int b2) {
  const c2o_bv *B1 = &cx->m1->bvs[b1], *B2 = &cx->m2->bvs[b2];
  double p[3], q[3];
  int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1, ta, fa, tb, fb;
  c2o_tolerance_result *r = cx->r;
  r->tri_tests++;
  double d = tri_distance(cx->Rrel, cx->Trel, &cx->m1->tris[t1], &cx->m2->tris[t2], p, q, &ta,
                          &fa, &tb, &fb);
  if (d <= r->tolerance) {   /* :1518-1529 */
    r->closer = 1;
    r->distance = d;
    memcpy(r->p1, p, sizeof(p));
    memcpy(r->p2, q, sizeof(q));
    r->tri[0] = t1;
    r->tri[1] = t2;
    return 1;
  }
  return 0;
}

static void tolerance_recurse(tol_ctx *cx, const m3 R, const double T[3], int b1, int b2) {
  const c2o_bv *B1 = &cx->m1->bvs[b1], *B2 = &cx->m2->bvs[b2];
  c2o_tolerance_result *r = cx->r;
  if (B1->first_child < 0 && B2->first_child < 0) {
    tol_leaf(cx, b1, b2);
    return;
  }
  int a1, a2, c1, c2;
  m3 R1, R2;
  double T1[3], T2[3], S[3];
  expand_pair(cx->m1, cx->m2, R, T, b1, b2, &a1, &a2, &c1, &c2, R1, T1, R2, T2);
  r->bv_tests += 2;
  double d1 = bv_distance(R1, T1, &cx->m1->bvs[a1], &cx->m2->bvs[a2], S);
  double d2 = bv_distance(R2, T2, &cx->m1->bvs[c1], &cx->m2->bvs[c2], S);
  if (d2 < d1) {   /* :1590-1603 */
    if (d2 <= r->tolerance) tolerance_recurse(cx, R2, T2, c1, c2);
    if (r->closer) return;
    if (d1 <= r->tolerance) tolerance_recurse(cx, R1, T1, a1, a2);
  } else {
    if (d1 <= r->tolerance) tolerance_recurse(cx, R1, T1, a1, a2);
    if (r->closer) return;
    if (d2 <= r->tolerance) tolerance_recurse(cx, R2, T2, c1, c2);
  }
}

/* C2AToleranceQueueRecurse (C2A_PQP.cpp:1606-1759). */
static void tolerance_queue_recurse(tol_ctx *cx, const m3 R, const double T[3], int b1,
                                    int b2) {
  c2o_bvtq q;
  bvtq_init(&q, cx->qsize);
  c2o_bvt mt;
  memcpy(mt.R, R, sizeof(m3));
  memcpy(mt.T, T, sizeof(mt.T));
  mt.b1 = b1;
  mt.b2 = b2;
  mt.d = 0.0;
  c2o_tolerance_result *r = cx->r;
  for (;;) {
    const c2o_bv *B1 = &cx->m1->bvs[mt.b1], *B2 = &cx->m2->bvs[mt.b2];
    int l1 = B1->first_child < 0, l2 = B2->first_child < 0;
    if (l1 && l2) {
      if (tol_leaf(cx, mt.b1, mt.b2)) break;   /* :1644 */
    } else if (q.n == q.size - 1) {   /* :1647-1654 */
      tolerance_queue_recurse(cx, mt.R, mt.T, mt.b1, mt.b2);
      if (r->closer == 1) break;
    } else {
      c2o_bvt t1, t2;
      r->bv_tests += 2;
      child_tests(cx->m1, cx->m2, &mt, &t1, &t2);
      if (t1.d <= r->tolerance) bvtq_add(&q, &t1);   /* :1747-1750 */
      if (t2.d <= r->tolerance) bvtq_add(&q, &t2);
    }
    if (q.n == 0 || q.h[1].d > r->tolerance) {   /* :1752-1756 */
      r->closer = 0;
      break;
    }
    mt = bvtq_extract_min(&q);
  }
  bvtq_free(&q);
}

void c2o_tolerance(const c2o_model *m1, const c2o_model *m2, const double R1[3][3],
                   const double T1[3], const double R2[3][3], const double T2[3],
                   double tolerance, int qsize, c2o_tolerance_result *r) {
  tol_ctx cx;
  m3 R;
  double T[3], S[3];
  cx.m1 = m1; cx.m2 = m2; cx.qsize = qsize; cx.r = r;
  root_transform(m1, m2, R1, T1, R2, T2, cx.Rrel, cx.Trel, R, T);
  if (tolerance < 0.0) tolerance = 0.0;   /* :1786 */
  r->tolerance = tolerance;
  r->bv_tests = 0;
  r->tri_tests = 0;
  r->closer = 0;
  r->distance = 0.0;
  r->tri[0] = r->tri[1] = -1;
  memset(r->p1, 0, sizeof(r->p1));
  memset(r->p2, 0, sizeof(r->p2));
  /* the root pair's distance for a trivial reject (:1817-1819) */
  double d = bv_distance(R, T, &m1->bvs[0], &m2->bvs[0], S);
  if (d <= r->tolerance) {
    if (qsize <= 2) tolerance_recurse(&cx, R, T, 0, 0);
    else tolerance_queue_recurse(&cx, R, T, 0, 0);
  }
  if (r->closer) {   /* p2 into model 2's frame (:1834-1838) */
    double u[3];
    sub3(u, r->p2, cx.Trel);
    mtxv(r->p2, cx.Rrel, u);
  }
}

/* PQP's TriContact (declared at C2A_PQP.cpp:441 and C2A.cpp:37, defined in
 * PQP 1.3's TriDist/Tri-overlap code, not vendored): the separating-axis
 * test of two triangles over 17 candidate axes -- the two face normals,
 * the nine edge-edge cross products and the six in-plane edge normals --
 * with the vertices taken relative to P1.  Projections that overlap on
 * every axis mean contact; touching counts as contact. */
static int project6(const double ax[3], const double p1[3], const double p2[3],
                    const double p3[3], const double q1[3], const double q2[3],
                    const double q3[3]) {
  double P1 = dot3(ax, p1), P2 = dot3(ax, p2), P3 = dot3(ax, p3);
  double Q1 = dot3(ax, q1), Q2 = dot3(ax, q2), Q3 = dot3(ax, q3);
  double mx1 = P1 > P2 ? P1 : P2, mn1 = P1 < P2 ? P1 : P2;
  mx1 = mx1 > P3 ? mx1 : P3;
  mn1 = mn1 < P3 ? mn1 : P3;
  double mx2 = Q1 > Q2 ? Q1 : Q2, mn2 = Q1 < Q2 ? Q1 : Q2;
  mx2 = mx2 > Q3 ? mx2 : Q3;
  mn2 = mn2 < Q3 ? mn2 : Q3;
  if (mn1 > mx2) return 0;
  if (mn2 > mx1) return 0;
  return 1;
}

int c2o_tri_contact(const double P1[3], const double P2[3], const double P3[3],
                    const double Q1[3], const double Q2[3], const double Q3[3]) {
  double p1[3] = {0, 0, 0}, p2[3], p3[3], q1[3], q2[3], q3[3];
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
  if (!project6(n1, p1, p2, p3, q1, q2, q3)) return 0;
  if (!project6(m1, p1, p2, p3, q1, q2, q3)) return 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cross3(ef, e[i], f[j]);
      if (!project6(ef, p1, p2, p3, q1, q2, q3)) return 0;
    }
  for (int i = 0; i < 3; i++) {
    if (!project6(g[i], p1, p2, p3, q1, q2, q3)) return 0;
    if (!project6(h[i], p1, p2, p3, q1, q2, q3)) return 0;
  }
  return 1;
}

/* C2A_Collide, PQP_CollideResult form (C2A_PQP.cpp:911-968 ->
 * CollideRecurse :798-909): every triangle pair in contact under pairs of
 * overlapping RSS volumes, in visiting order (first child pair first, no
 * distance order); flag == 2 (C2A_FIRST_CONTACT) stops at the first. */
typedef struct {
  const c2o_model *m1, *m2;
  m3 Rrel;
  double Trel[3];
  int flag, cap;
  int *pairs;
  c2o_collide_result *r;
} coll_ctx;

static void collide_recurse(coll_ctx *cx, const m3 R, const double T[3], int b1, int b2) {
  const c2o_bv *B1 = &cx->m1->bvs[b1], *B2 = &cx->m2->bvs[b2];
  c2o_collide_result *r = cx->r;
  r->bv_tests++;   /* :804 */
  {   /* BV_Overlap for RSS: RectDist <= r1 + r2 (:807) */
    double P[3], Q[3], S[3];
    if (!(c2o_rect_dist(R, T, B1->l, B2->l, P, Q, S) <= B1->r + B2->r)) return;
  }
  int l1 = B1->first_child < 0, l2 = B2->first_child < 0;
  if (l1 && l2) {   /* :813-854 */
    int t1 = -B1->first_child - 1, t2 = -B2->first_child - 1;
    const c2o_tri *A = &cx->m1->tris[t1], *B = &cx->m2->tris[t2];
    double q1[3], q2[3], q3[3];
    r->tri_tests++;
    mxvpv(q1, cx->Rrel, B->p[0], cx->Trel);
    mxvpv(q2, cx->Rrel, B->p[1], cx->Trel);
    mxvpv(q3, cx->Rrel, B->p[2], cx->Trel);
    if (c2o_tri_contact(A->p[0], A->p[1], A->p[2], q1, q2, q3)) {
      if (r->num_pairs < cx->cap) {
        cx->pairs[2 * r->num_pairs] = t1;
        cx->pairs[2 * r->num_pairs + 1] = t2;
      }
      r->num_pairs++;
    }
    return;
  }
  int a1, a2, c1, c2;
  m3 R1, R2;
  double T1[3], T2[3];
  expand_pair(cx->m1, cx->m2, R, T, b1, b2, &a1, &a2, &c1, &c2, R1, T1, R2, T2);
  collide_recurse(cx, R1, T1, a1, a2);   /* :869-907 */
  if (cx->flag == 2 && r->num_pairs > 0) return;
  collide_recurse(cx, R2, T2, c1, c2);
}

void c2o_collide_pairs(const c2o_model *m1, const c2o_model *m2, const double R1[3][3],
                       const double T1[3], const double R2[3][3], const double T2[3], int flag,
                       int cap, int *pairs, c2o_collide_result *r) {
  coll_ctx cx;
  m3 R;
  double T[3];
  cx.m1 = m1; cx.m2 = m2; cx.flag = flag; cx.cap = cap; cx.pairs = pairs; cx.r = r;
  root_transform(m1, m2, R1, T1, R2, T2, cx.Rrel, cx.Trel, R, T);
  r->bv_tests = 0;
  r->tri_tests = 0;
  r->num_pairs = 0;
  collide_recurse(&cx, R, T, 0, 0);
}

/* Batches of independent static queries for the tests (poses24 as in
 * c2o_distance_batch; every query starts from the models as given). */
void c2o_distance_queue_batch(const c2o_model *m1, const c2o_model *m2, int nq,
                              const double *poses24, const int *seeds, double rel_err,
                              double abs_err, int qsize, c2o_distance_result *out) {
  c2o_model w1 = *m1, w2 = *m2;
  for (int i = 0; i < nq; i++) {
    m3 R1, R2;
    double T1[3], T2[3];
    pose_unpack(poses24 + 24 * (size_t)i, R1, T1);
    pose_unpack(poses24 + 24 * (size_t)i + 12, R2, T2);
    w1.last_tri = seeds ? seeds[2 * i] : 0;
    w2.last_tri = seeds ? seeds[2 * i + 1] : 0;
    c2o_distance_queue(&w1, &w2, R1, T1, R2, T2, rel_err, abs_err, qsize, &out[i]);
  }
}

void c2o_distance_cs_batch(const c2o_model *m1, const c2o_model *m2, int nq,
                           const double *poses24, const int *seeds, const double *upper,
                           c2o_distance_result *out, int *found) {
  c2o_model w1 = *m1, w2 = *m2;
  for (int i = 0; i < nq; i++) {
    m3 R1, R2;
    double T1[3], T2[3];
    pose_unpack(poses24 + 24 * (size_t)i, R1, T1);
    pose_unpack(poses24 + 24 * (size_t)i + 12, R2, T2);
    w1.last_tri = seeds ? seeds[2 * i] : 0;
    w2.last_tri = seeds ? seeds[2 * i + 1] : 0;
    memset(&out[i], 0, sizeof(out[i]));
    found[i] = c2o_distance_cs(&w1, &w2, R1, T1, R2, T2, upper[i], &out[i]);
  }
}

void c2o_tolerance_batch(const c2o_model *m1, const c2o_model *m2, int nq, const double *poses24,
                         const double *tolerance, int qsize, c2o_tolerance_result *out) {
  for (int i = 0; i < nq; i++) {
    m3 R1, R2;
    double T1[3], T2[3];
    pose_unpack(poses24 + 24 * (size_t)i, R1, T1);
    pose_unpack(poses24 + 24 * (size_t)i + 12, R2, T2);
    c2o_tolerance(m1, m2, R1, T1, R2, T2, tolerance[i], qsize, &out[i]);
  }
}

void c2o_collide_pairs_batch(const c2o_model *m1, const c2o_model *m2, int nq,
                             const double *poses24, int flag, int cap, int *pairs,
                             c2o_collide_result *out) {
  for (int i = 0; i < nq; i++) {
    m3 R1, R2;
    double T1[3], T2[3];
    pose_unpack(poses24 + 24 * (size_t)i, R1, T1);
    pose_unpack(poses24 + 24 * (size_t)i + 12, R2, T2);
    c2o_collide_pairs(m1, m2, R1, T1, R2, T2, flag, cap, pairs + 2 * (size_t)cap * i, &out[i]);
  }
}
