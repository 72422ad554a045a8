// scorer.h -- the learned front-ordering model: a pair scorer
// 32 -> 64 -> 64 -> 1 (ReLU) that predicts log(closest triangle distance in
// the pair's subtree + 1e-4).  Weights are BF16 for the matrix cores, with FP32
// accumulation and FP32 biases and master weights.  Shared by the GPU
// kernels (scorer.hip, the traversal) and the host test harness.
#pragma once

#include <math.h>
#include <stdint.h>

#include "c2a_math.h"

namespace c2a {

constexpr int NF = 32;   // features
constexpr int NH = 64;   // hidden width
// FP32 parameter vector: W1[NF][NH], b1[NH], W2[NH][NH], b2[NH], w3[NH], b3
constexpr int P_W1 = 0;
constexpr int P_B1 = P_W1 + NF * NH;
constexpr int P_W2 = P_B1 + NH;
constexpr int P_B2 = P_W2 + NH * NH;
constexpr int P_W3 = P_B2 + NH;
constexpr int P_B3 = P_W3 + NH;
constexpr int NPARAM = P_B3 + 1;   // 6337

// Device copy of a scorer: BF16 weights (row-major [in][out], and transposed
// [out][in] for the traversal's register-resident pass) and FP32 biases.
struct ScorerDev {
  const uint16_t* w1;   // NF x NH
  const uint16_t* w2;   // NH x NH
  const uint16_t* w1t;  // NH x NF
  const uint16_t* w2t;  // NH x NH
  const float* b1;
  const float* b2;
  const float* w3;
  float b3;
  int32_t on;           // order the traversal by score
};

// Round-to-nearest-even FP32 -> BF16 (NaN kept quiet).
C2A_HD uint16_t f32_to_bf16(float f) {
  union { float f; uint32_t u; } x;
  x.f = f;
  if ((x.u & 0x7f800000u) == 0x7f800000u && (x.u & 0x007fffffu)) return (uint16_t)((x.u >> 16) | 0x40);
  const uint32_t r = 0x7fffu + ((x.u >> 16) & 1u);
  return (uint16_t)((x.u + r) >> 16);
}
C2A_HD float bf16_to_f32(uint16_t h) {
  union { float f; uint32_t u; } x;
  x.u = (uint32_t)h << 16;
  return x.f;
}

C2A_HD float slog(double v) { return (float)log(v + 1e-4); }

// The 32 features of a pair (frozen list).  Lengths are in units of the
// query's scale s (the angular radius of mesh 1's root: the largest vertex
// distance from its origin), so that meshes of any size look alike:
//  0-1 side lengths of node 1's rectangle, 2-3 of node 2's, 4-5 radii,
//  6-8 relative translation, 9-17 relative rotation (row-major),
//  18 log RSS distance, 19 log admissible time, 20-21 log motion bounds,
//  22-23 depths of the two nodes, 24-25 angular radii, 26 |cv| of object 1,
//  27 rotation angle of object 1, 28 |cv| of object 2, 29 rotation angle of
//  object 2, 30 remaining time (UpboundTOC), 31 leaf flags (node1 + 2 node2) / 3.
// inv_s = 1 / s.  The training label is slog(closest distance * inv_s).
C2A_HD void pair_features(float* f, const double* l1, const double* l2, double r1, double r2,
                          const double* R, const double* T, double d, double mt, double mb1,
                          double mb2, int depth1, int depth2, double angR1, double angR2,
                          const Motion& mo1, const Motion& mo2, double up, int leafbits,
                          double inv_s) {
  f[0] = (float)(l1[0] * inv_s);
  f[1] = (float)(l1[1] * inv_s);
  f[2] = (float)(l2[0] * inv_s);
  f[3] = (float)(l2[1] * inv_s);
  f[4] = (float)(r1 * inv_s);
  f[5] = (float)(r2 * inv_s);
  for (int i = 0; i < 3; i++) f[6 + i] = (float)(T[i] * inv_s);
  for (int i = 0; i < 9; i++) f[9 + i] = (float)R[i];
  f[18] = slog(d * inv_s);
  f[19] = slog(mt);
  f[20] = slog(mb1 * inv_s);
  f[21] = slog(mb2 * inv_s);
  f[22] = (float)depth1 * 0.0625f;
  f[23] = (float)depth2 * 0.0625f;
  f[24] = (float)(angR1 * inv_s);
  f[25] = (float)(angR2 * inv_s);
  f[26] = (float)(sqrt(dot3(mo1.cv, mo1.cv)) * inv_s);
  f[27] = (float)mo1.angVel;
  f[28] = (float)(sqrt(dot3(mo2.cv, mo2.cv)) * inv_s);
  f[29] = (float)mo2.angVel;
  f[30] = (float)up;
  f[31] = (float)leafbits * (1.0f / 3.0f);
}

// Scalar forward pass with the same BF16 rounding as the matrix-core path
// (inputs, weights and the first hidden layer rounded to BF16; FP32 sums).
// Host harness only; the GPU uses the matrix cores.
inline float score_scalar(const float* x, const float* p) {
  float h1[NH], h2[NH];
  for (int j = 0; j < NH; j++) {
    float s = 0.0f;
    for (int k = 0; k < NF; k++)
      s += bf16_to_f32(f32_to_bf16(x[k])) * bf16_to_f32(f32_to_bf16(p[P_W1 + k * NH + j]));
    s += p[P_B1 + j];
    h1[j] = bf16_to_f32(f32_to_bf16(s > 0.0f ? s : 0.0f));
  }
  for (int j = 0; j < NH; j++) {
    float s = 0.0f;
    for (int k = 0; k < NH; k++) s += h1[k] * bf16_to_f32(f32_to_bf16(p[P_W2 + k * NH + j]));
    s += p[P_B2 + j];
    h2[j] = s > 0.0f ? s : 0.0f;
  }
  float y = p[P_B3];
  for (int j = 0; j < NH; j++) y += h2[j] * p[P_W3 + j];
  return y;
}

// Scalar forward pass over a BF16 feature row and the BF16 transposed weights
// of a ScorerDev whose arrays are host memory (the host test harness).
inline float score_dev_scalar(const ScorerDev& S, const uint16_t* xb) {
  float h1[NH];
  for (int j = 0; j < NH; j++) {
    float s = 0.0f;
    for (int k = 0; k < NF; k++) s += bf16_to_f32(xb[k]) * bf16_to_f32(S.w1t[j * NF + k]);
    s += S.b1[j];
    h1[j] = bf16_to_f32(f32_to_bf16(s > 0.0f ? s : 0.0f));
  }
  float y = S.b3;
  for (int j = 0; j < NH; j++) {
    float s = 0.0f;
    for (int k = 0; k < NH; k++) s += h1[k] * bf16_to_f32(S.w2t[j * NH + k]);
    s += S.b2[j];
    y += (s > 0.0f ? s : 0.0f) * S.w3[j];
  }
  return y;
}

}  // namespace c2a
