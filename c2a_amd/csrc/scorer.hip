// scorer.hip -- training and inference of the learned front-ordering scorer
// on the matrix cores (mlp_dev.h), and its C ABI.  Training is plain
// mini-batch Adam on the mean squared error between the score and
// log(closest subtree triangle distance + 1e-4); gradients are reduced in a
// fixed order, so a training run is reproducible.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <stdio.h>

#include <cmath>
#include <new>
#include <vector>

#include "../../include/c2a_amd.h"
#include "mlp_dev.h"
#include "scorer.h"

namespace c2a {

// Scores of n feature rows as the traversal computes them (score64 over BF16
// rows staged in LDS, one wavefront per 64 rows).
__global__ void __launch_bounds__(64) infer_kernel(ScorerDev S, const float* X, int64_t n,
                                                   float* out) {
  __shared__ __align__(16) uint16_t x[64][NF];
  const int lane = threadIdx.x;
  for (int64_t base = (int64_t)blockIdx.x * 64; base < n; base += (int64_t)gridDim.x * 64) {
    const int64_t row = base + lane;
    for (int k = 0; k < NF; k++) x[lane][k] = f32_to_bf16(row < n ? X[row * NF + k] : 0.0f);
    wave_lds_sync();
    const float y = score64(S, &x[0][0], lane);
    if (row < n) out[row] = y;
    wave_lds_sync();
  }
}

// Scores of n feature rows through the training forward pass (FP32 in, one
// wavefront per 64 rows).
__global__ void __launch_bounds__(64) score_rows_kernel(ScorerDev S, const float* X, int64_t n,
                                                        float* out) {
  __shared__ MlpLds m;
  const int lane = threadIdx.x;
  for (int64_t base = (int64_t)blockIdx.x * 64; base < n; base += (int64_t)gridDim.x * 64) {
    const int64_t row = base + lane;
    for (int k = 0; k < NF; k++) m.x[lane][k] = f32_to_bf16(row < n ? X[row * NF + k] : 0.0f);
    wave_lds_sync();
    f32x16 h2[2][2];
    mlp_forward64(S, m, lane, h2);
    if (row < n) out[row] = m.s[lane];
    wave_lds_sync();
  }
}

// Partial gradients of rows [blockIdx.x * per_block, ...) written to
// partial[blockIdx.x * (NPARAM + 1) ...] (the last entry is the squared error).
__global__ void __launch_bounds__(64) grad_kernel(ScorerDev S, const float* X, const float* Y,
                                                  const int32_t* idx, int64_t n, int per_block,
                                                  float scale, float* partial) {
  __shared__ MlpLds m;
  const int lane = threadIdx.x;
  MlpGrad G;
  grad_zero(G);
  const int64_t b0 = (int64_t)blockIdx.x * per_block;
  const int64_t b1 = b0 + per_block < n ? b0 + per_block : n;
  for (int64_t base = b0; base < b1; base += 64) {
    const int64_t row = base + lane;
    const bool v = row < b1;
    const int64_t src = v ? (idx ? (int64_t)idx[row] : row) : 0;
    for (int k = 0; k < NF; k++) m.x[lane][k] = f32_to_bf16(v ? X[src * NF + k] : 0.0f);
    wave_lds_sync();
    f32x16 h2[2][2];
    mlp_forward64(S, m, lane, h2);
    m.dy[lane] = v ? Y[src] : 0.0f;
    wave_lds_sync();
    mlp_backward64(S, m, lane, h2, (int)(b1 - base), scale, G);
  }
  float* o = partial + (int64_t)blockIdx.x * (NPARAM + 1);
  for (int nt = 0; nt < 2; nt++)
    for (int r = 0; r < 16; r++) o[P_W1 + crow(r, lane) * NH + 32 * nt + (lane & 31)] = G.w1[nt][r];
  for (int mt = 0; mt < 2; mt++)
    for (int nt = 0; nt < 2; nt++)
      for (int r = 0; r < 16; r++)
        o[P_W2 + (32 * mt + crow(r, lane)) * NH + 32 * nt + (lane & 31)] = G.w2[mt][nt][r];
  for (int c = 0; c < 2; c++) {
    const float b1v = G.b1[c] + __shfl_xor(G.b1[c], 32);
    const float b2v = G.b2[c] + __shfl_xor(G.b2[c], 32);
    const float w3v = G.w3[c] + __shfl_xor(G.w3[c], 32);
    if (lane < 32) {
      o[P_B1 + 32 * c + lane] = b1v;
      o[P_B2 + 32 * c + lane] = b2v;
      o[P_W3 + 32 * c + lane] = w3v;
    }
  }
  const float b3v = G.b3 + __shfl(G.b3, 32);
  float loss = G.loss;
  for (int o2 = 32; o2 >= 1; o2 >>= 1) loss += __shfl_xor(loss, o2);
  if (lane == 0) {
    o[P_B3] = b3v;
    o[NPARAM] = loss;
  }
}

// Sums the per-block partials in block order (reproducible).
__global__ void reduce_kernel(const float* partial, int nblocks, float* grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > NPARAM) return;
  float s = 0.0f;
  for (int b = 0; b < nblocks; b++) s += partial[(int64_t)b * (NPARAM + 1) + i];
  grad[i] = s;
}

// Adam step on the FP32 master parameters; refreshes the BF16 weights.
__global__ void adam_kernel(float* p, float* mom, float* vel, const float* grad, float lr,
                            float b1, float b2, float eps, int step, uint16_t* w1,
                            uint16_t* w2, uint16_t* w1t, uint16_t* w2t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  const float g = grad[i];
  const float mi = b1 * mom[i] + (1.0f - b1) * g;
  const float vi = b2 * vel[i] + (1.0f - b2) * g * g;
  mom[i] = mi;
  vel[i] = vi;
  const float mh = mi / (1.0f - powf(b1, (float)step));
  const float vh = vi / (1.0f - powf(b2, (float)step));
  const float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
  p[i] = pi;
  if (i >= P_W1 && i < P_B1) {
    const int k = i - P_W1;
    w1[k] = w1t[(k % NH) * NF + k / NH] = f32_to_bf16(pi);
  }
  if (i >= P_W2 && i < P_B2) {
    const int k = i - P_W2;
    w2[k] = w2t[(k % NH) * NH + k / NH] = f32_to_bf16(pi);
  }
}

__global__ void to_bf16_kernel(const float* p, uint16_t* w1, uint16_t* w2, uint16_t* w1t,
                               uint16_t* w2t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NF * NH) w1[i] = w1t[(i % NH) * NF + i / NH] = f32_to_bf16(p[P_W1 + i]);
  if (i < NH * NH) w2[i] = w2t[(i % NH) * NH + i / NH] = f32_to_bf16(p[P_W2 + i]);
}

}  // namespace c2a

using namespace c2a;

// Device state of a scorer: FP32 master parameters (+ Adam moments) and the
// BF16 weight copies the matrix cores read.
struct c2a_scorer {
  float* p = nullptr;       // NPARAM + 1 (the last float is b3 mirrored for the kernels)
  float* mom = nullptr;
  float* vel = nullptr;
  uint16_t* w1 = nullptr;
  uint16_t* w2 = nullptr;
  uint16_t* w1t = nullptr;
  uint16_t* w2t = nullptr;
  int step = 0;
  ~c2a_scorer() {
    if (p) (void)hipFree(p);
    if (mom) (void)hipFree(mom);
    if (vel) (void)hipFree(vel);
    if (w1) (void)hipFree(w1);
    if (w2) (void)hipFree(w2);
    if (w1t) (void)hipFree(w1t);
    if (w2t) (void)hipFree(w2t);
  }
};

namespace c2a {
int set_error(int code, const char* msg);   // capi.cpp: sets c2a_last_error
}

namespace {

int hip_err(hipError_t e) {
  return e == hipSuccess ? C2A_OK : c2a::set_error(C2A_ERR_HIP, hipGetErrorString(e));
}
int bad_arg(const char* what) { return c2a::set_error(C2A_ERR_BAD_ARG, what); }
int oom(const char* what) { return c2a::set_error(C2A_ERR_OOM, what); }

// Device scratch of one call, freed on every return path.
template <class T>
struct DevArray {
  T* p = nullptr;
  ~DevArray() {
    if (p) (void)hipFree(p);
  }
  bool alloc(size_t n) { return hipMalloc(&p, (n ? n : 1) * sizeof(T)) == hipSuccess; }
};

ScorerDev dev_view(const c2a_scorer* s, float b3) {
  ScorerDev d;
  d.w1 = s->w1;
  d.w2 = s->w2;
  d.w1t = s->w1t;
  d.w2t = s->w2t;
  d.b1 = s->p + P_B1;
  d.b2 = s->p + P_B2;
  d.w3 = s->p + P_W3;
  d.b3 = b3;
  d.on = 1;
  return d;
}

float read_b3(const c2a_scorer* s) {
  float b3 = 0.0f;
  (void)hipMemcpy(&b3, s->p + P_B3, sizeof(float), hipMemcpyDeviceToHost);
  return b3;
}

// splitmix64 stream for the initial weights
uint64_t mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

namespace c2a {

ScorerDev scorer_view(const c2a_scorer* s) { return dev_view(s, read_b3(s)); }

}  // namespace c2a

extern "C" {

int c2a_scorer_create(uint64_t seed, c2a_scorer** out) {
  if (!out) return bad_arg("null output handle");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return c2a::set_error(C2A_ERR_NO_DEVICE, "no HIP device available");
  c2a_scorer* s = new (std::nothrow) c2a_scorer();
  if (!s) return oom("host allocation failed");
  const size_t pb = (NPARAM + 1) * sizeof(float);
  if (hipMalloc(&s->p, pb) != hipSuccess || hipMalloc(&s->mom, pb) != hipSuccess ||
      hipMalloc(&s->vel, pb) != hipSuccess ||
      hipMalloc(&s->w1, NF * NH * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&s->w2, NH * NH * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&s->w1t, NF * NH * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&s->w2t, NH * NH * sizeof(uint16_t)) != hipSuccess) {
    delete s;
    return oom("scorer parameter buffers");
  }
  // He-uniform initial weights, zero biases
  std::vector<float> h(NPARAM + 1, 0.0f);
  uint64_t st = mix(seed ^ 0x5C0E5C0Eull);
  auto uni = [&]() {
    st = mix(st);
    return (double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  for (int i = 0; i < NF * NH; i++) h[P_W1 + i] = (float)(uni() * sqrt(6.0 / NF));
  for (int i = 0; i < NH * NH; i++) h[P_W2 + i] = (float)(uni() * sqrt(6.0 / NH));
  for (int i = 0; i < NH; i++) h[P_W3 + i] = (float)(uni() * sqrt(6.0 / NH) * 0.5);
  int rc = c2a_scorer_set_params(s, h.data());
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return C2A_OK;
}

int c2a_scorer_destroy(c2a_scorer* s) {
  delete s;
  return C2A_OK;
}

int c2a_scorer_set_params(c2a_scorer* s, const float* params) {
  if (!s || !params) return bad_arg("null scorer or parameters");
  const hipError_t e0 = hipMemcpy(s->p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice);
  if (e0 != hipSuccess) return hip_err(e0);
  (void)hipMemset(s->mom, 0, (NPARAM + 1) * sizeof(float));
  (void)hipMemset(s->vel, 0, (NPARAM + 1) * sizeof(float));
  s->step = 0;
  hipLaunchKernelGGL(to_bf16_kernel, dim3((NH * NH + 255) / 256), dim3(256), 0, 0, s->p, s->w1,
                     s->w2, s->w1t, s->w2t);
  return hip_err(hipDeviceSynchronize());
}

int c2a_scorer_get_params(const c2a_scorer* s, float* params) {
  if (!s || !params) return bad_arg("null scorer or parameters");
  return hip_err(hipMemcpy(params, s->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
}

// File format: the 8 bytes "C2ASCR01", the parameter count as a
// little-endian uint32, then the FP32 parameters (little-endian).
int c2a_scorer_save(const c2a_scorer* s, const char* path) {
  if (!s || !path) return bad_arg("null scorer or path");
  std::vector<float> h(NPARAM);
  int rc = c2a_scorer_get_params(s, h.data());
  if (rc) return rc;
  FILE* f = fopen(path, "wb");
  if (!f) return bad_arg("cannot open the scorer file for writing");
  const uint32_t n = NPARAM;
  const bool ok = fwrite("C2ASCR01", 1, 8, f) == 8 && fwrite(&n, 4, 1, f) == 1 &&
                  fwrite(h.data(), sizeof(float), NPARAM, f) == (size_t)NPARAM;
  return (fclose(f) == 0 && ok) ? C2A_OK : C2A_ERR_BAD_ARG;
}

int c2a_scorer_load(const char* path, c2a_scorer** out) {
  if (!path || !out) return bad_arg("null path or output handle");
  *out = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return bad_arg("cannot open the scorer file");
  char magic[8];
  uint32_t n = 0;
  std::vector<float> h(NPARAM);
  const bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "C2ASCR01", 8) == 0 &&
                  fread(&n, 4, 1, f) == 1 && n == (uint32_t)NPARAM &&
                  fread(h.data(), sizeof(float), NPARAM, f) == (size_t)NPARAM;
  fclose(f);
  if (!ok) return bad_arg("not a scorer file (header, count or size)");
  for (float v : h)
    if (!std::isfinite(v)) return bad_arg("scorer file holds a non-finite parameter");
  c2a_scorer* s = nullptr;
  int rc = c2a_scorer_create(0, &s);
  if (rc) return rc;
  rc = c2a_scorer_set_params(s, h.data());
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return C2A_OK;
}

int c2a_scorer_infer(const c2a_scorer* s, const float* feats, int64_t n, float* out) {
  if (!s || n < 0 || (n > 0 && (!feats || !out))) return bad_arg("bad scorer arguments");
  if (n == 0) return C2A_OK;
  DevArray<float> dx, dy;
  if (!dx.alloc((size_t)n * NF) || !dy.alloc((size_t)n)) return oom("scorer device buffers");
  hipError_t e = hipMemcpy(dx.p, feats, n * NF * sizeof(float), hipMemcpyHostToDevice);
  const int grid = (int)((n + 63) / 64 < 4096 ? (n + 63) / 64 : 4096);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(infer_kernel, dim3(grid), dim3(64), 0, 0, dev_view(s, read_b3(s)), dx.p,
                       n, dy.p);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dy.p, n * sizeof(float), hipMemcpyDeviceToHost);
  return hip_err(e);
}

// Gradient of the mean squared error over all n rows (for checking), and the
// mean squared error itself.
int c2a_scorer_grad(const c2a_scorer* s, const float* feats, const float* labels, int64_t n,
                    float* grad_out, float* loss_out) {
  if (!s || n <= 0 || !feats || !labels || !grad_out) return bad_arg("bad scorer arguments");
  const int per_block = 256;
  const int nb = (int)((n + per_block - 1) / per_block);
  DevArray<float> dx, dy, part, g;
  if (!dx.alloc((size_t)n * NF) || !dy.alloc((size_t)n) ||
      !part.alloc((size_t)nb * (NPARAM + 1)) || !g.alloc(NPARAM + 1))
    return oom("scorer gradient buffers");
  hipError_t e = hipMemcpy(dx.p, feats, n * NF * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dy.p, labels, n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(grad_kernel, dim3(nb), dim3(64), 0, 0, dev_view(s, read_b3(s)), dx.p,
                       dy.p, (const int32_t*)nullptr, n, per_block, (float)(2.0 / (double)n),
                       part.p);
    hipLaunchKernelGGL(reduce_kernel, dim3((NPARAM + 256) / 256), dim3(256), 0, 0, part.p, nb,
                       g.p);
    e = hipGetLastError();
  }
  std::vector<float> h(NPARAM + 1);
  if (e == hipSuccess)
    e = hipMemcpy(h.data(), g.p, (NPARAM + 1) * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) {
    memcpy(grad_out, h.data(), NPARAM * sizeof(float));
    if (loss_out) *loss_out = (float)(h[NPARAM] / (double)n);
  }
  return hip_err(e);
}

int c2a_scorer_train(c2a_scorer* s, const float* feats, const float* labels, int64_t n,
                     int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss_out) {
  if (!s || n <= 0 || !feats || !labels || epochs <= 0 || batch <= 0)
    return bad_arg("bad scorer training arguments");
  const int per_block = 256;
  const int nbmax = (batch + per_block - 1) / per_block;
  DevArray<float> dx_, dy_, part_, g_;
  DevArray<int32_t> didx_;
  if (!dx_.alloc((size_t)n * NF) || !dy_.alloc((size_t)n) || !didx_.alloc((size_t)n) ||
      !part_.alloc((size_t)nbmax * (NPARAM + 1)) || !g_.alloc(NPARAM + 1))
    return oom("scorer training buffers");
  float *dx = dx_.p, *dy = dy_.p, *part = part_.p, *g = g_.p;
  int32_t* didx = didx_.p;
  hipError_t e = hipMemcpy(dx, feats, n * NF * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dy, labels, n * sizeof(float), hipMemcpyHostToDevice);
  std::vector<int32_t> perm((size_t)n);
  uint64_t st = mix(seed ^ 0xA11CEull);
  std::vector<float> gh(NPARAM + 1);
  for (int ep = 0; ep < epochs && e == hipSuccess; ep++) {
    for (int64_t i = 0; i < n; i++) perm[(size_t)i] = (int32_t)i;
    for (int64_t i = n - 1; i > 0; i--) {   // seeded Fisher-Yates shuffle
      st = mix(st);
      const int64_t j = (int64_t)(st % (uint64_t)(i + 1));
      const int32_t t = perm[(size_t)i];
      perm[(size_t)i] = perm[(size_t)j];
      perm[(size_t)j] = t;
    }
    e = hipMemcpy(didx, perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice);
    double eloss = 0.0;
    for (int64_t b0 = 0; b0 < n && e == hipSuccess; b0 += batch) {
      const int64_t nb_rows = b0 + batch < n ? batch : n - b0;
      const int nb = (int)((nb_rows + per_block - 1) / per_block);
      hipLaunchKernelGGL(grad_kernel, dim3(nb), dim3(64), 0, 0, dev_view(s, read_b3(s)), dx, dy,
                         didx + b0, nb_rows, per_block, (float)(2.0 / (double)nb_rows), part);
      hipLaunchKernelGGL(reduce_kernel, dim3((NPARAM + 256) / 256), dim3(256), 0, 0, part, nb, g);
      s->step++;
      hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, s->p, s->mom,
                         s->vel, g, lr, 0.9f, 0.999f, 1e-8f, s->step, s->w1, s->w2, s->w1t, s->w2t);
      e = hipGetLastError();
      float l = 0.0f;
      if (e == hipSuccess) e = hipMemcpy(&l, g + NPARAM, sizeof(float), hipMemcpyDeviceToHost);
      eloss += l;
    }
    if (loss_out) loss_out[ep] = (float)(eloss / (double)n);
  }
  return hip_err(e);
}

}  // extern "C"
