// LinearMath.h -- the reference's linear-algebra types (C2A/LinearMath.h:
// 25-452, definitions :500-1183), source compatible: Coord3D, Quaternion
// (x, y, z, w storage), Matrix3x3 (row-major, m[i][j]) and Transform
// (rotation + translation), with the reference's operator meanings:
//
//   Coord3D    a * b   dot product          a % b   cross product
//   Matrix3x3  M * v   M v                  M % v   M^T v
//              A * B   A B     A / B  A B^T    A % B  A^T B
//   Transform  X * p   R p + T (point)      X & v   R v (vector)
//   Quaternion p * q   4-D dot product      p % q   Hamilton product
//   v ^= M : v = M v    v ^= X : v = X * v    v &= X : v = X & v
//
// Sums are evaluated left to right, as the reference writes them, so the
// results are the reference's bit for bit.  Where the reference calls
// exit(0) (a near-singular Inverse()), this returns a zero matrix instead.
// Members the reference declares without defining (Quaternion::Normalize,
// Identity, !=, <=, >=, the array setter; Coord3D * Matrix3x3) are defined
// with their evident meaning.
#ifndef C2A_COMPAT_LINEARMATH_H
#define C2A_COMPAT_LINEARMATH_H

#include <math.h>

typedef double Real;

class Quaternion;
class Matrix3x3;
class Transform;

namespace c2a_lm {
inline Real dot3(Real a0, Real a1, Real a2, Real b0, Real b1, Real b2) {
  return a0 * b0 + a1 * b1 + a2 * b2;
}
}  // namespace c2a_lm

class Coord3D {
 public:
  Coord3D() {}
  Coord3D(Real x, Real y, Real z) { Set_Value(x, y, z); }
  Coord3D(const Real t[]) { Set_Value(t); }
  inline Coord3D(const Quaternion& q);   // the vector part

  int Size() const { return 3; }
  const Real& X() const { return v_[0]; }
  const Real& Y() const { return v_[1]; }
  const Real& Z() const { return v_[2]; }
  Real& X() { return v_[0]; }
  Real& Y() { return v_[1]; }
  Real& Z() { return v_[2]; }
  const Real& operator[](int i) const { return v_[i % 3]; }
  Real& operator[](int i) { return v_[i % 3]; }
  const Real* Value() const { return v_; }
  void Get_Value(Real v[]) const { v[0] = v_[0]; v[1] = v_[1]; v[2] = v_[2]; }
  void Get_Value(Real& x, Real& y, Real& z) const { x = v_[0]; y = v_[1]; z = v_[2]; }
  void Set_X(Real x) { v_[0] = x; }
  void Set_Y(Real y) { v_[1] = y; }
  void Set_Z(Real z) { v_[2] = z; }
  void Set_Value(const Real v[]) { v_[0] = v[0]; v_[1] = v[1]; v_[2] = v[2]; }
  void Set_Value(Real x, Real y, Real z) { v_[0] = x; v_[1] = y; v_[2] = z; }

  Real Dist_Sq(const Coord3D& t) const {
    const Real x = t.v_[0] - v_[0], y = t.v_[1] - v_[1], z = t.v_[2] - v_[2];
    return x * x + y * y + z * z;
  }
  Real Dist(const Coord3D& t) const { return sqrt(Dist_Sq(t)); }
  Real Length_Sq() const { return v_[0] * v_[0] + v_[1] * v_[1] + v_[2] * v_[2]; }
  Real Length() const { return sqrt(Length_Sq()); }

  // merge this triple into mn / mx component by component
  void Min(Coord3D& mn) const {
    for (int i = 0; i < 3; i++) mn.v_[i] = mn.v_[i] > v_[i] ? v_[i] : mn.v_[i];
  }
  void Max(Coord3D& mx) const {
    for (int i = 0; i < 3; i++) mx.v_[i] = mx.v_[i] > v_[i] ? mx.v_[i] : v_[i];
  }
  void Min_Max(Coord3D& mn, Coord3D& mx) const { Min(mn); Max(mx); }

  // leaves vectors of length <= 1e-20 unchanged
  void Normalize() {
    if (Length() > 1.0e-20) *this /= Length();
  }
  void Zero() { v_[0] = v_[1] = v_[2] = 0.0; }
  void Identity() { Zero(); }
  void Negate() { v_[0] = -v_[0]; v_[1] = -v_[1]; v_[2] = -v_[2]; }
  void Add(Real a, Real b, Real c) { v_[0] += a; v_[1] += b; v_[2] += c; }
  void Sub(Real a, Real b, Real c) { v_[0] -= a; v_[1] -= b; v_[2] -= c; }

  void operator+=(const Coord3D& t) { v_[0] += t.v_[0]; v_[1] += t.v_[1]; v_[2] += t.v_[2]; }
  void operator-=(const Coord3D& t) { v_[0] -= t.v_[0]; v_[1] -= t.v_[1]; v_[2] -= t.v_[2]; }
  void operator*=(Real s) { v_[0] *= s; v_[1] *= s; v_[2] *= s; }
  void operator/=(Real s) { *this *= 1.0 / s; }   // multiplies by the reciprocal
  inline void operator^=(const Matrix3x3& m);   // v = M v
  inline void operator^=(const Transform& m);   // v = R v + T
  inline void operator&=(const Transform& m);   // v = R v

  bool operator==(const Coord3D& t) const {
    return v_[0] == t.v_[0] && v_[1] == t.v_[1] && v_[2] == t.v_[2];
  }
  bool operator!=(const Coord3D& t) const { return !(*this == t); }
  bool operator<=(const Coord3D& t) const {
    return v_[0] <= t.v_[0] && v_[1] <= t.v_[1] && v_[2] <= t.v_[2];
  }
  bool operator>=(const Coord3D& t) const {
    return v_[0] >= t.v_[0] && v_[1] >= t.v_[1] && v_[2] >= t.v_[2];
  }

 private:
  Real v_[3];
};

inline Coord3D operator-(const Coord3D& t) { return Coord3D(-t[0], -t[1], -t[2]); }
inline Coord3D operator+(const Coord3D& a, const Coord3D& b) {
  return Coord3D(a[0] + b[0], a[1] + b[1], a[2] + b[2]);
}
inline Coord3D operator-(const Coord3D& a, const Coord3D& b) {
  return Coord3D(a[0] - b[0], a[1] - b[1], a[2] - b[2]);
}
inline Coord3D operator*(Real s, const Coord3D& t) { return Coord3D(s * t[0], s * t[1], s * t[2]); }
inline Coord3D operator*(const Coord3D& t, Real s) { return s * t; }
inline Coord3D operator/(const Coord3D& t, Real s) {
  const Real r = 1.0 / s;
  return Coord3D(t[0] * r, t[1] * r, t[2] * r);
}
inline Real operator*(const Coord3D& a, const Coord3D& b) {
  return c2a_lm::dot3(a[0], a[1], a[2], b[0], b[1], b[2]);
}
inline Coord3D operator%(const Coord3D& a, const Coord3D& b) {
  return Coord3D(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                 a[0] * b[1] - a[1] * b[0]);
}

class Quaternion {
 public:
  Quaternion() {}
  Quaternion(Real x, Real y, Real z, Real w) { Set_Value(x, y, z, w); }
  Quaternion(const Real t[]) { Set_Value(t); }
  Quaternion(const Coord3D& t) { Set_Value(t); }            // [x, y, z, 0]
  Quaternion(const Coord3D& axis, Real theta) { Set_Value(axis, theta); }
  Quaternion(Real ax, Real ay, Real az) { Set_Value(ax, ay, az); }   // Euler x, y', z''

  int Size() const { return 4; }
  const Real& X() const { return q_[0]; }
  const Real& Y() const { return q_[1]; }
  const Real& Z() const { return q_[2]; }
  const Real& W() const { return q_[3]; }
  Real& X() { return q_[0]; }
  Real& Y() { return q_[1]; }
  Real& Z() { return q_[2]; }
  Real& W() { return q_[3]; }
  void Set_X(Real x) { q_[0] = x; }
  void Set_Y(Real y) { q_[1] = y; }
  void Set_Z(Real z) { q_[2] = z; }
  void Set_W(Real w) { q_[3] = w; }
  // out-of-range indices read element 0, as the reference's do
  Real& operator[](int i) { return q_[(i >= 0 && i < 4) ? i : 0]; }
  Real operator[](int i) const { return q_[(i >= 0 && i < 4) ? i : 0]; }
  const Real* Value() const { return q_; }
  void Get_Value(Real v[]) const { for (int i = 0; i < 4; i++) v[i] = q_[i]; }
  void Get_Value(Real& x, Real& y, Real& z, Real& w) const {
    x = q_[0]; y = q_[1]; z = q_[2]; w = q_[3];
  }
  void Set_Value(const Real v[]) { for (int i = 0; i < 4; i++) q_[i] = v[i]; }
  void Set_Value(Real x, Real y, Real z, Real w) { q_[0] = x; q_[1] = y; q_[2] = z; q_[3] = w; }
  void Set_Value(const Coord3D& t) { Set_Value(t[0], t[1], t[2], 0.0); }
  // rotation by theta about axis (normalised first)
  void Set_Value(const Coord3D& axis, Real theta) {
    Coord3D u = axis;
    u.Normalize();
    u *= sin(0.5 * theta);
    Set_Value(u[0], u[1], u[2], cos(0.5 * theta));
  }
  // about x, then the new y, then the new z
  void Set_Value(Real ax, Real ay, Real az) {
    Set_Value(sin(0.5 * ax), 0.0, 0.0, cos(0.5 * ax));
    *this *= Quaternion(0.0, sin(0.5 * ay), 0.0, cos(0.5 * ay));
    *this *= Quaternion(0.0, 0.0, sin(0.5 * az), cos(0.5 * az));
  }

  Real Length_Sq() const { return q_[0] * q_[0] + q_[1] * q_[1] + q_[2] * q_[2] + q_[3] * q_[3]; }
  Real Length() const { return sqrt(Length_Sq()); }
  void Normalize() { *this /= Length(); }
  void Invert() { q_[0] = -q_[0]; q_[1] = -q_[1]; q_[2] = -q_[2]; }   // the conjugate
  void Negate() { Invert(); }
  void Zero() { q_[0] = q_[1] = q_[2] = q_[3] = 0.0; }
  void Identity() { Set_Value(0.0, 0.0, 0.0, 1.0); }
  // (x, y, z, -w): the same rotation's inverse up to sign
  Quaternion Inverse() { return Quaternion(q_[0], q_[1], q_[2], -q_[3]); }
  Real Angle() { return (q_[3] < -1.0 || q_[3] > 1.0) ? 0.0 : 2.0 * acos(q_[3]); }

  void operator+=(const Quaternion& q) { for (int i = 0; i < 4; i++) q_[i] += q.q_[i]; }
  void operator-=(const Quaternion& q) { for (int i = 0; i < 4; i++) q_[i] -= q.q_[i]; }
  void operator*=(Real s) { for (int i = 0; i < 4; i++) q_[i] *= s; }
  void operator/=(Real s) { *this *= 1.0 / s; }
  // Hamilton product in place: vector part a x b + w_a b + w_b a
  void operator*=(const Quaternion& q) {
    const Coord3D a(q_[0], q_[1], q_[2]), b(q.q_[0], q.q_[1], q.q_[2]);
    const Coord3D v = (a % b) + q_[3] * b + q.q_[3] * a;
    const Real w = q_[3] * q.q_[3] - (a * b);
    Set_Value(v[0], v[1], v[2], w);
  }

  bool operator==(const Quaternion& q) const {
    return q_[0] == q.q_[0] && q_[1] == q.q_[1] && q_[2] == q.q_[2] && q_[3] == q.q_[3];
  }
  bool operator!=(const Quaternion& q) const { return !(*this == q); }
  bool operator<=(const Quaternion& q) const {
    return q_[0] <= q.q_[0] && q_[1] <= q.q_[1] && q_[2] <= q.q_[2] && q_[3] <= q.q_[3];
  }
  bool operator>=(const Quaternion& q) const {
    return q_[0] >= q.q_[0] && q_[1] >= q.q_[1] && q_[2] >= q.q_[2] && q_[3] >= q.q_[3];
  }
  // of q and -q, the one farther from this quaternion
  inline Quaternion Farthest(const Quaternion& q) const;

 private:
  Real q_[4];   // x, y, z, w (LinearMath.h:134-135, 151-154)
};

inline Coord3D::Coord3D(const Quaternion& q) { Set_Value(q[0], q[1], q[2]); }

inline Quaternion operator-(const Quaternion& q) { return Quaternion(-q[0], -q[1], -q[2], -q[3]); }
inline Quaternion operator+(const Quaternion& a, const Quaternion& b) {
  return Quaternion(a[0] + b[0], a[1] + b[1], a[2] + b[2], a[3] + b[3]);
}
inline Quaternion operator-(const Quaternion& a, const Quaternion& b) {
  return Quaternion(a[0] - b[0], a[1] - b[1], a[2] - b[2], a[3] - b[3]);
}
inline Quaternion operator*(Real s, const Quaternion& q) {
  return Quaternion(s * q[0], s * q[1], s * q[2], s * q[3]);
}
inline Quaternion operator*(const Quaternion& q, Real s) { return s * q; }
inline Quaternion operator/(const Quaternion& q, Real s) {
  const Real r = 1.0 / s;
  return Quaternion(q[0] * r, q[1] * r, q[2] * r, q[3] * r);
}
inline Real operator*(const Quaternion& a, const Quaternion& b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}
inline Quaternion operator%(const Quaternion& a, const Quaternion& b) {
  return Quaternion(a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1],
                    a[3] * b[1] + a[1] * b[3] + a[2] * b[0] - a[0] * b[2],
                    a[3] * b[2] + a[2] * b[3] + a[0] * b[1] - a[1] * b[0],
                    a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2]);
}
inline Quaternion Quaternion::Farthest(const Quaternion& q) const {
  const Quaternion d = *this - q, s = *this + q;
  return (d * d > s * s) ? q : -q;
}

class Matrix3x3 {
 public:
  Matrix3x3() {}
  Matrix3x3(const Real v[3][3]) { Set_Value(v); }
  Matrix3x3(const Real v[]) { Set_Value(v); }
  Matrix3x3(Real v0, Real v1, Real v2, Real v3, Real v4, Real v5, Real v6, Real v7, Real v8) {
    const Real v[9] = {v0, v1, v2, v3, v4, v5, v6, v7, v8};
    Set_Value(v);
  }

  int Size() const { return 9; }
  const Real* operator[](int i) const { return &m_[(i % 3) * 3]; }
  Real* operator[](int i) { return &m_[(i % 3) * 3]; }
  const Real* Value() const { return m_; }
  void Get_Value(Real v[]) const { for (int i = 0; i < 9; i++) v[i] = m_[i]; }
  void Get_Value(Real v[3][3]) const {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) v[i][j] = m_[3 * i + j];
  }
  void Set_Value(const Real v[]) { for (int i = 0; i < 9; i++) m_[i] = v[i]; }
  void Set_Value(const Real v[3][3]) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) m_[3 * i + j] = v[i][j];
  }
  // Rotation of a unit quaternion (LinearMath.h:809-831; the same expression
  // the CCD engine uses, so TOC poses match bit for bit).
  void Set_Value(const Quaternion& q);
  // Rz (Ry Rx): about x, then y, then z (fixed axes)
  inline void Set_Value(Real rx, Real ry, Real rz);
  // The unit quaternion of this rotation (LinearMath.h:759-790).
  Quaternion Quaternion_();

  bool IsZero() const {
    for (int i = 0; i < 9; i++)
      if (m_[i] != 0.0) return false;
    return true;
  }
  // cofactor a(r1, c1) a(r2, c2) - a(r1, c2) a(r2, c1)
  Real Cofac(int r1, int c1, int r2, int c2) const {
    return m_[r1 * 3 + c1] * m_[r2 * 3 + c2] - m_[r1 * 3 + c2] * m_[r2 * 3 + c1];
  }
  // adjugate over the determinant; a zero matrix when det < 1e-15
  inline Matrix3x3 Inverse() const;

  void Zero() { for (int i = 0; i < 9; i++) m_[i] = 0.0; }
  void Identity() { for (int i = 0; i < 9; i++) m_[i] = (i % 4 == 0) ? 1.0 : 0.0; }
  void Transpose() {
    Real t = m_[1]; m_[1] = m_[3]; m_[3] = t;
    t = m_[2]; m_[2] = m_[6]; m_[6] = t;
    t = m_[5]; m_[5] = m_[7]; m_[7] = t;
  }

  void operator+=(const Matrix3x3& m) { for (int i = 0; i < 9; i++) m_[i] += m.m_[i]; }
  void operator-=(const Matrix3x3& m) { for (int i = 0; i < 9; i++) m_[i] -= m.m_[i]; }
  void operator*=(Real s) { for (int i = 0; i < 9; i++) m_[i] *= s; }
  void operator/=(Real s) { *this *= 1.0 / s; }

 private:
  Real m_[9];
};

inline Matrix3x3 operator-(const Matrix3x3& m) {
  Real v[9];
  for (int i = 0; i < 9; i++) v[i] = -m.Value()[i];
  return Matrix3x3(v);
}
inline Matrix3x3 operator+(const Matrix3x3& a, const Matrix3x3& b) {
  Real v[9];
  for (int i = 0; i < 9; i++) v[i] = a.Value()[i] + b.Value()[i];
  return Matrix3x3(v);
}
inline Matrix3x3 operator-(const Matrix3x3& a, const Matrix3x3& b) {
  Real v[9];
  for (int i = 0; i < 9; i++) v[i] = a.Value()[i] - b.Value()[i];
  return Matrix3x3(v);
}
inline Matrix3x3 operator*(Real s, const Matrix3x3& m) {
  Real v[9];
  for (int i = 0; i < 9; i++) v[i] = s * m.Value()[i];
  return Matrix3x3(v);
}
inline Matrix3x3 operator*(const Matrix3x3& m, Real s) { return s * m; }
inline Matrix3x3 operator/(const Matrix3x3& m, Real s) { return (1.0 / s) * m; }
// M v
inline Coord3D operator*(const Matrix3x3& m, const Coord3D& t) {
  return Coord3D(c2a_lm::dot3(t[0], t[1], t[2], m[0][0], m[0][1], m[0][2]),
                 c2a_lm::dot3(t[0], t[1], t[2], m[1][0], m[1][1], m[1][2]),
                 c2a_lm::dot3(t[0], t[1], t[2], m[2][0], m[2][1], m[2][2]));
}
// v^T M (declared by the reference without a definition)
inline Coord3D operator*(const Coord3D& t, const Matrix3x3& m) {
  return Coord3D(c2a_lm::dot3(t[0], t[1], t[2], m[0][0], m[1][0], m[2][0]),
                 c2a_lm::dot3(t[0], t[1], t[2], m[0][1], m[1][1], m[2][1]),
                 c2a_lm::dot3(t[0], t[1], t[2], m[0][2], m[1][2], m[2][2]));
}
// M^T v
inline Coord3D operator%(const Matrix3x3& m, const Coord3D& t) { return t * m; }
// A B
inline Matrix3x3 operator*(const Matrix3x3& a, const Matrix3x3& b) {
  Real v[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      v[3 * i + j] = c2a_lm::dot3(a[i][0], a[i][1], a[i][2], b[0][j], b[1][j], b[2][j]);
  return Matrix3x3(v);
}
// A B^T
inline Matrix3x3 operator/(const Matrix3x3& a, const Matrix3x3& b) {
  Real v[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      v[3 * i + j] = c2a_lm::dot3(a[i][0], a[i][1], a[i][2], b[j][0], b[j][1], b[j][2]);
  return Matrix3x3(v);
}
// A^T B
inline Matrix3x3 operator%(const Matrix3x3& a, const Matrix3x3& b) {
  Real v[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      v[3 * i + j] = c2a_lm::dot3(a[0][i], a[1][i], a[2][i], b[0][j], b[1][j], b[2][j]);
  return Matrix3x3(v);
}

inline void Matrix3x3::Set_Value(Real rx, Real ry, Real rz) {
  Real c = cos(rx), s = sin(rx);
  const Matrix3x3 X(1.0, 0.0, 0.0, 0.0, c, -s, 0.0, s, c);
  c = cos(ry);
  s = sin(ry);
  const Matrix3x3 Y(c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c);
  c = cos(rz);
  s = sin(rz);
  const Matrix3x3 Z(c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0);
  *this = Z * (Y * X);
}

inline Matrix3x3 Matrix3x3::Inverse() const {
  const Coord3D co(Cofac(1, 1, 2, 2), Cofac(1, 2, 2, 0), Cofac(1, 0, 2, 1));
  const Real det = m_[0] * co[0] + m_[1] * co[1] + m_[2] * co[2];
  if (det < 1.0e-15) {
    Matrix3x3 z;
    z.Zero();
    return z;
  }
  const Real s = Real(1.0f) / det;
  return Matrix3x3(co[0] * s, Cofac(0, 2, 2, 1) * s, Cofac(0, 1, 1, 2) * s,
                   co[1] * s, Cofac(0, 0, 2, 2) * s, Cofac(0, 2, 1, 0) * s,
                   co[2] * s, Cofac(0, 1, 2, 0) * s, Cofac(0, 0, 1, 1) * s);
}

class Transform {
 public:
  Transform() {}
  // a row-major 3x4 matrix whose fourth column is the translation
  Transform(const Real v[]) { Set_Value(v); }
  Transform(Real v0, Real v1, Real v2, Real v3, Real v4, Real v5, Real v6, Real v7, Real v8,
            Real v9, Real v10, Real v11)
      : R_(v0, v1, v2, v4, v5, v6, v8, v9, v10), T_(v3, v7, v11) {}
  Transform(const Matrix3x3& r) : R_(r), T_(0.0, 0.0, 0.0) {}
  Transform(const Matrix3x3& r, const Coord3D& t) : R_(r), T_(t) {}

  int Size() const { return 12; }
  const Coord3D& Translation() const { return T_; }
  Coord3D& Translation() { return T_; }
  const Matrix3x3& Rotation() const { return R_; }
  Matrix3x3& Rotation() { return R_; }
  const Quaternion Quaternion_() { return R_.Quaternion_(); }
  void Set_Value(const Transform& m) { R_ = m.R_; T_ = m.T_; }
  void Set_Value(const Real v[]) {
    const Real r[9] = {v[0], v[1], v[2], v[4], v[5], v[6], v[8], v[9], v[10]};
    R_.Set_Value(r);
    T_.Set_Value(v[3], v[7], v[11]);
  }
  void Set_Value(const Real r[], const Real t[]) { R_.Set_Value(r); T_.Set_Value(t); }
  void Set_Value(const Matrix3x3& r, const Coord3D& t) { R_ = r; T_ = t; }
  void Set_Rotation(const Matrix3x3& r) { R_ = r; }
  void Set_Rotation(const Quaternion& q) { R_.Set_Value(q); }
  void Set_Translation(const Coord3D& t) { T_ = t; }
  // this = m^-1 of a rigid m: R^T, -R^T T
  void Invert(const Transform& m) {
    R_ = m.R_;
    R_.Transpose();
    const Coord3D t = m.R_ % m.T_;
    T_.Set_Value(-t[0], -t[1], -t[2]);
  }
  void Identity() { R_.Identity(); T_.Identity(); }

 private:
  Matrix3x3 R_;
  Coord3D T_;
};

// R p + T
inline Coord3D operator*(const Transform& m, const Coord3D& t) {
  const Matrix3x3& R = m.Rotation();
  const Coord3D& T = m.Translation();
  return Coord3D(c2a_lm::dot3(t[0], t[1], t[2], R[0][0], R[0][1], R[0][2]) + T[0],
                 c2a_lm::dot3(t[0], t[1], t[2], R[1][0], R[1][1], R[1][2]) + T[1],
                 c2a_lm::dot3(t[0], t[1], t[2], R[2][0], R[2][1], R[2][2]) + T[2]);
}
// R v
inline Coord3D operator&(const Transform& m, const Coord3D& t) { return m.Rotation() * t; }

inline void Coord3D::operator^=(const Matrix3x3& m) { *this = m * *this; }
inline void Coord3D::operator^=(const Transform& m) { *this = m * *this; }
inline void Coord3D::operator&=(const Transform& m) { *this = m & *this; }

#endif
