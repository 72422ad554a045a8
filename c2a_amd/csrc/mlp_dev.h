// mlp_dev.h -- matrix-core (v_mfma_f32_32x32x16_bf16) forward and backward
// passes of the pair scorer over 64 rows per wavefront.  Operand layout of
// the 32x32x16 BF16 instruction of gfx950: lane l supplies A[l % 32][8(l/32)+i]
// and B[8(l/32)+i][l % 32], i = 0..7; accumulator r of lane l is
// C[8(r/4) + 4(l/32) + r%4][l % 32].
//
// Where a product's k runs over the rows of an accumulator (the weight
// gradients, the second layer of score64), lane l's eight k-slots of step s
// are its own accumulator registers 8s..8s+7, i.e. rows
// 16s + 8(i/4) + 4(l/32) + i%4, and the other operand is gathered at those
// same rows: the sum over k does not depend on the order of k.
#pragma once

#include <hip/hip_runtime.h>

#include "scorer.h"

namespace c2a {

typedef short bf16x8 __attribute__((ext_vector_type(8)));   // BF16 bit patterns
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct MlpLds {                 // per-wavefront scratch (20.5 KB)
  uint16_t x[64][NF];           // input rows, BF16
  uint16_t h[64][NH];           // first hidden layer after ReLU, BF16
  uint16_t g[64][NH];           // gradient w.r.t. the second hidden layer, BF16
  float s[64];                  // per-row scores
  float dy[64];                 // per-row loss gradient
};

// Gradient accumulators of one wavefront (accumulator layout).
struct MlpGrad {
  f32x16 w1[2];      // W1[feature = crow][32 nt + l%32]
  f32x16 w2[2][2];   // W2[32 mt + crow][32 nt + l%32]
  float b1[2], b2[2], w3[2];   // column 32 c + l%32 (summed over both half-waves)
  float b3;
  float loss;
};

__device__ inline int crow(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// the row of accumulator register 8 s + i of this lane (k-slot i of step s)
__device__ inline int krow(int s, int i, int lane) { return crow(8 * s + i, lane); }

__device__ inline bf16x8 a_frag(const uint16_t* base, int ld, int lane, int kofs) {
  // A[m = lane % 32][k = kofs + 8 (lane / 32) + i] from a row-major BF16 tile
  const uint16_t* p = base + (lane & 31) * ld + kofs + 8 * (lane >> 5);
  bf16x8 v;
  for (int i = 0; i < 8; i++) v[i] = (short)p[i];
  return v;
}

__device__ inline bf16x8 b_frag(const uint16_t* w, int ld, int lane, int kofs, int nofs) {
  // B[k = kofs + 8 (lane / 32) + i][n = nofs + lane % 32] from a row-major BF16 matrix
  const uint16_t* p = w + (kofs + 8 * (lane >> 5)) * ld + nofs + (lane & 31);
  bf16x8 v;
  for (int i = 0; i < 8; i++) v[i] = (short)p[i * ld];
  return v;
}

__device__ inline bf16x8 bt_frag(const uint16_t* w, int ld, int lane, int kofs, int nofs) {
  // B[k][n] = W[n][k] (transposed): W row-major with leading dimension ld
  const uint16_t* p = w + (nofs + (lane & 31)) * ld + kofs + 8 * (lane >> 5);
  bf16x8 v;
  for (int i = 0; i < 8; i++) v[i] = (short)p[i];
  return v;
}

__device__ inline f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16v8, a),
                                                 __builtin_bit_cast(bf16v8, b), c, 0, 0, 0);
}

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward over the 64 rows in m.x: leaves relu(layer 1) as BF16 in m.h,
// relu(layer 2) in h2[rt][ct] (accumulator layout) and the scores in m.s.
// All 64 lanes must be active.
__device__ inline void mlp_forward64(const ScorerDev& S, MlpLds& m, int lane, f32x16 h2[2][2]) {
  for (int rt = 0; rt < 2; rt++)
    for (int ct = 0; ct < 2; ct++) {
      f32x16 acc = {0};
      for (int ks = 0; ks < NF / 16; ks++)
        acc = mfma(a_frag(&m.x[32 * rt][0], NF, lane, 16 * ks),
                   b_frag(S.w1, NH, lane, 16 * ks, 32 * ct), acc);
      const int col = 32 * ct + (lane & 31);
      const float bias = S.b1[col];
      wave_lds_sync();
      for (int r = 0; r < 16; r++) {
        const float v = acc[r] + bias;
        m.h[32 * rt + crow(r, lane)][col] = f32_to_bf16(v > 0.0f ? v : 0.0f);
      }
    }
  wave_lds_sync();
  for (int rt = 0; rt < 2; rt++)
    for (int ct = 0; ct < 2; ct++) {
      f32x16 acc = {0};
      for (int ks = 0; ks < NH / 16; ks++)
        acc = mfma(a_frag(&m.h[32 * rt][0], NH, lane, 16 * ks),
                   b_frag(S.w2, NH, lane, 16 * ks, 32 * ct), acc);
      const float bias = S.b2[32 * ct + (lane & 31)];
      for (int r = 0; r < 16; r++) {
        const float v = acc[r] + bias;
        acc[r] = v > 0.0f ? v : 0.0f;
      }
      h2[rt][ct] = acc;
    }
  // layer 3: each lane's column contribution, summed over the 32 lanes of
  // its half-wavefront (which share the rows)
  const float w3a = S.w3[lane & 31], w3b = S.w3[32 + (lane & 31)];
  for (int rt = 0; rt < 2; rt++) {
    float part[16];
    for (int r = 0; r < 16; r++) part[r] = h2[rt][0][r] * w3a + h2[rt][1][r] * w3b;
    for (int o = 1; o < 32; o <<= 1)
      for (int r = 0; r < 16; r++) part[r] += __shfl_xor(part[r], o);
    if ((lane & 31) == 0)
      for (int r = 0; r < 16; r++) m.s[32 * rt + crow(r, lane)] = part[r] + S.b3;
  }
  wave_lds_sync();
}

// two runs of four consecutive BF16 values (each 8-byte aligned)
__device__ inline bf16x8 ld_frag2(const uint16_t* p0, const uint16_t* p1) {
  const uint64_t u0 = *reinterpret_cast<const uint64_t*>(p0);
  const uint64_t u1 = *reinterpret_cast<const uint64_t*>(p1);
  bf16x8 v;
  for (int i = 0; i < 4; i++) {
    v[i] = (short)(u0 >> (16 * i));
    v[4 + i] = (short)(u1 >> (16 * i));
  }
  return v;
}

// Scores of 64 BF16 feature rows X[64][NF] (row-major, any address space),
// computed transposed so that nothing leaves the registers:
// H1^T = W1^T X^T and H2^T = W2^T H1^T.  The accumulator of H1^T already
// holds, lane by lane, the B operand of the second product: step ks takes
// registers 8 (ks % 2) .. +7 of row tile ks / 2, hidden units
// 16 ks + 4 (lane / 32) + i and 16 ks + 8 + 4 (lane / 32) + i (i = 0..3),
// and W2^T is read at the same units.  Returns the score of row `lane`.
// All 64 lanes must be active.
__device__ inline float score64(const ScorerDev& S, const uint16_t* X, int lane) {
  const int hi = lane >> 5, lo = lane & 31;
  float y[2];
  for (int nt = 0; nt < 2; nt++) {   // feature rows 32 nt .. 32 nt + 31
    f32x16 h1[2];
    for (int mt = 0; mt < 2; mt++) {
      f32x16 acc = {0};
      for (int ks = 0; ks < NF / 16; ks++) {
        const uint16_t* pa = S.w1t + (32 * mt + lo) * NF + 16 * ks + 8 * hi;
        const uint16_t* pb = X + (32 * nt + lo) * NF + 16 * ks + 8 * hi;
        acc = mfma(ld_frag2(pa, pa + 4), ld_frag2(pb, pb + 4), acc);
      }
      h1[mt] = acc;
    }
    bf16x8 hb[NH / 16];
    for (int ks = 0; ks < NH / 16; ks++)
      for (int i = 0; i < 8; i++) {
        const int unit = 32 * (ks >> 1) + krow(ks & 1, i, lane);
        const float v = h1[ks >> 1][8 * (ks & 1) + i] + S.b1[unit];
        hb[ks][i] = (short)f32_to_bf16(v > 0.0f ? v : 0.0f);
      }
    float part = 0.0f;
    for (int mt = 0; mt < 2; mt++) {
      f32x16 acc = {0};
      for (int ks = 0; ks < NH / 16; ks++) {
        const uint16_t* pw = S.w2t + (32 * mt + lo) * NH + 16 * ks + 4 * hi;
        acc = mfma(ld_frag2(pw, pw + 8), hb[ks], acc);
      }
      for (int r = 0; r < 16; r++) {
        const int unit = 32 * mt + crow(r, lane);
        const float v = acc[r] + S.b2[unit];
        part += (v > 0.0f ? v : 0.0f) * S.w3[unit];
      }
    }
    part += __shfl_xor(part, 32);
    y[nt] = part + S.b3;
  }
  return hi ? y[1] : y[0];
}

__device__ inline void grad_zero(MlpGrad& G) {
  const f32x16 z = {0};
  G.w1[0] = G.w1[1] = z;
  G.w2[0][0] = G.w2[0][1] = G.w2[1][0] = G.w2[1][1] = z;
  for (int c = 0; c < 2; c++) G.b1[c] = G.b2[c] = G.w3[c] = 0.0f;
  G.b3 = 0.0f;
  G.loss = 0.0f;
}

// Backward of the mean squared error (score - y)^2 over the rows [0, nvalid)
// of the current 64, scaled by `scale` (2 / batch size); adds to G.
// m.x, m.h and the targets (in m.dy on entry) come from the caller;
// h2 is the forward's second hidden layer.
__device__ inline void mlp_backward64(const ScorerDev& S, MlpLds& m, int lane, f32x16 h2[2][2],
                                      int nvalid, float scale, MlpGrad& G) {
  // dL/ds per row
  {
    const int row = lane;
    const float e = row < nvalid ? m.s[row] - m.dy[row] : 0.0f;
    G.loss += e * e;
    wave_lds_sync();
    m.dy[row] = e * scale;
    wave_lds_sync();
  }
  const float w3c[2] = {S.w3[lane & 31], S.w3[32 + (lane & 31)]};
  // second hidden layer gradient (accumulator layout) and w3, b2, b3
  f32x16 dh2[2][2];
  float gb3 = 0.0f;
  for (int rt = 0; rt < 2; rt++)
    for (int ct = 0; ct < 2; ct++) {
      f32x16 d;
      for (int r = 0; r < 16; r++) {
        const float dy = m.dy[32 * rt + crow(r, lane)];
        const float hv = h2[rt][ct][r];
        G.w3[ct] += hv * dy;
        d[r] = hv > 0.0f ? dy * w3c[ct] : 0.0f;
        G.b2[ct] += d[r];
        if (ct == 0 && (lane & 31) == 0) gb3 += dy;
      }
      dh2[rt][ct] = d;
    }
  G.b3 += gb3;
  // W2 gradient: H1^T (64 x rows) times dH2 (rows x 64); the accumulator
  // layout of dH2 is already the B operand for k = rows of its row tile
  for (int mt = 0; mt < 2; mt++)
    for (int nt = 0; nt < 2; nt++)
      for (int rt = 0; rt < 2; rt++)
        for (int ks = 0; ks < 2; ks++) {
          bf16x8 a, b;
          for (int i = 0; i < 8; i++) {
            a[i] = (short)m.h[32 * rt + krow(ks, i, lane)][32 * mt + (lane & 31)];
            b[i] = (short)f32_to_bf16(dh2[rt][nt][8 * ks + i]);
          }
          G.w2[mt][nt] = mfma(a, b, G.w2[mt][nt]);
        }
  // dH1 = dH2 W2^T through LDS (A operand needs row-major dH2)
  wave_lds_sync();
  for (int rt = 0; rt < 2; rt++)
    for (int ct = 0; ct < 2; ct++)
      for (int r = 0; r < 16; r++)
        m.g[32 * rt + crow(r, lane)][32 * ct + (lane & 31)] = f32_to_bf16(dh2[rt][ct][r]);
  wave_lds_sync();
  f32x16 dh1[2][2];
  for (int rt = 0; rt < 2; rt++)
    for (int nt = 0; nt < 2; nt++) {
      f32x16 acc = {0};
      for (int ks = 0; ks < NH / 16; ks++)
        acc = mfma(a_frag(&m.g[32 * rt][0], NH, lane, 16 * ks),
                   bt_frag(S.w2, NH, lane, 16 * ks, 32 * nt), acc);
      for (int r = 0; r < 16; r++) {
        const float h1 = bf16_to_f32(m.h[32 * rt + crow(r, lane)][32 * nt + (lane & 31)]);
        acc[r] = h1 > 0.0f ? acc[r] : 0.0f;
        G.b1[nt] += acc[r];
      }
      dh1[rt][nt] = acc;
    }
  // W1 gradient: X^T (32 x rows) times dH1 (rows x 64)
  for (int nt = 0; nt < 2; nt++)
    for (int rt = 0; rt < 2; rt++)
      for (int ks = 0; ks < 2; ks++) {
        bf16x8 a, b;
        for (int i = 0; i < 8; i++) {
          a[i] = (short)m.x[32 * rt + krow(ks, i, lane)][lane & 31];
          b[i] = (short)f32_to_bf16(dh1[rt][nt][8 * ks + i]);
        }
        G.w1[nt] = mfma(a, b, G.w1[nt]);
      }
  wave_lds_sync();
}

}  // namespace c2a
