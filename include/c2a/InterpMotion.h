// InterpMotion.h -- the linear motion of the reference's CCD interface
// (C2A/InterpMotion.h:53-226, InterpMotion.cpp:148-287, 486-568):
// translation T0 + t (T1 - T0), rotation q0 * [axis sin(theta t / 2),
// cos(theta t / 2)] about the body-frame axis of q0* q1, t clamped to 1.
#ifndef C2A_COMPAT_INTERPMOTION_H
#define C2A_COMPAT_INTERPMOTION_H

#include "LinearMath.h"
#include "C2A_Internal.h"

class CInterpMotion {
 public:
  virtual ~CInterpMotion() {}
  // Pose at time dt in [0, 1]; qua = (w, x, y, z, tx, ty, tz)
  // (InterpMotion.cpp:206-221, 557-565).  Also stores it in `transform`.
  virtual bool integrate(const double dt, PQP_REAL qua[7]) = 0;
  bool integrate(const double dt, PQP_REAL R[3][3], PQP_REAL T[3]);

  Transform transform;     // pose of the last integrate()
  Transform transform_s;   // start pose
  Transform transform_t;   // end pose
  Coord3D m_axis;          // body-frame rotation axis
  Real m_angVel;           // rotation angle over the motion
  PQP_REAL m_toc_delta;    // kept for source compatibility

  // Engine motion record (R0, T0, q0, cv, axis, angVel).
  const double* engine_motion() const { return m_; }

 protected:
  double m_[24];
};

class CInterpMotion_Linear : public CInterpMotion {
 public:
  CInterpMotion_Linear(const PQP_REAL R0[3][3], const PQP_REAL T0[3], const PQP_REAL R1[3][3],
                       const PQP_REAL T1[3]);
  // q = (w, x, y, z, tx, ty, tz) (InterpMotion.cpp:206-221); the rotations
  // are converted to matrices first, so this form can differ from the
  // reference in the last bits (the matrix form above is exact).
  CInterpMotion_Linear(const PQP_REAL q0[7], const PQP_REAL q1[7]);
  virtual bool integrate(const double dt, PQP_REAL qua[7]);
  using CInterpMotion::integrate;

 private:
  void init(const double* pose_s, const double* pose_e);
};

#endif
