// LinearMath.h -- the subset of the reference's linear-algebra types that its
// CCD interface passes around (C2A/LinearMath.h:25-447): Coord3D, Quaternion
// (x, y, z, w storage), Matrix3x3 (row-major, m[i][j]) and Transform
// (rotation + translation).  Source compatible for callers of C2A_Solve.
#ifndef C2A_COMPAT_LINEARMATH_H
#define C2A_COMPAT_LINEARMATH_H

#include <math.h>

typedef double Real;

class Quaternion;

class Coord3D {
 public:
  Coord3D() {}
  Coord3D(Real x, Real y, Real z) { Set_Value(x, y, z); }
  explicit Coord3D(const Real t[]) { Set_Value(t); }

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
  void Identity() { v_[0] = v_[1] = v_[2] = 0.0; }
  Real Length_Sq() const { return v_[0] * v_[0] + v_[1] * v_[1] + v_[2] * v_[2]; }
  Real Length() const { return sqrt(Length_Sq()); }

 private:
  Real v_[3];
};

class Quaternion {
 public:
  Quaternion() {}
  Quaternion(Real x, Real y, Real z, Real w) { Set_Value(x, y, z, w); }
  const Real& X() const { return q_[0]; }
  const Real& Y() const { return q_[1]; }
  const Real& Z() const { return q_[2]; }
  const Real& W() const { return q_[3]; }
  Real& X() { return q_[0]; }
  Real& Y() { return q_[1]; }
  Real& Z() { return q_[2]; }
  Real& W() { return q_[3]; }
  const Real* Value() const { return q_; }
  void Set_Value(Real x, Real y, Real z, Real w) { q_[0] = x; q_[1] = y; q_[2] = z; q_[3] = w; }

 private:
  Real q_[4];   // x, y, z, w (LinearMath.h:134-135, 151-154)
};

class Matrix3x3 {
 public:
  Matrix3x3() {}
  explicit Matrix3x3(const Real v[3][3]) { Set_Value(v); }
  const Real* operator[](int i) const { return &m_[(i % 3) * 3]; }
  Real* operator[](int i) { return &m_[(i % 3) * 3]; }
  void Get_Value(Real v[3][3]) const {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) v[i][j] = m_[3 * i + j];
  }
  void Set_Value(const Real v[3][3]) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) m_[3 * i + j] = v[i][j];
  }
  // Rotation of a unit quaternion (LinearMath.h:809-831; the same expression
  // the CCD engine uses, so TOC poses match bit for bit).
  void Set_Value(const Quaternion& q);
  void Identity() {
    for (int i = 0; i < 9; i++) m_[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }

 private:
  Real m_[9];
};

class Transform {
 public:
  Transform() {}
  Transform(const Matrix3x3& r, const Coord3D& t) : R_(r), T_(t) {}
  const Coord3D& Translation() const { return T_; }
  Coord3D& Translation() { return T_; }
  const Matrix3x3& Rotation() const { return R_; }
  Matrix3x3& Rotation() { return R_; }
  void Set_Value(const Matrix3x3& r, const Coord3D& t) { R_ = r; T_ = t; }
  void Set_Rotation(const Matrix3x3& r) { R_ = r; }
  void Set_Rotation(const Quaternion& q) { R_.Set_Value(q); }
  void Set_Translation(const Coord3D& t) { T_ = t; }
  void Identity() { R_.Identity(); T_.Identity(); }

 private:
  Matrix3x3 R_;
  Coord3D T_;
};

#endif
