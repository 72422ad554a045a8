// bvh_gpu.hip -- the RSS hierarchy build on gfx950, bit for bit the host
// builder's (bvh.cpp), which follows the reference's top-down build
// (C2A_Build.cpp:393-574, C2A_BV.cpp:71-111 and 354-635).
//
// Level by level, every node of a level at once, in three sizes: a
// 256-thread workgroup per node above kWaveMax triangles, one wavefront per
// node down to kThreadMax, one thread per node below.  A workgroup (or
// wavefront) stages the node's per-triangle covariance terms through LDS and
// nine lanes add them to the nine sums in triangle order (sums are
// order-dependent: one dependent add per triangle and sum is the only
// sequential part); every minimum, maximum and arg-extreme of the fit is
// reduced across the lanes (exact in any order); the corner pass evaluates 64
// points at a time against the current rectangle and advances to the first
// one that grows it; and the in-place partition of the split is reproduced in
// parallel from prefix counts (see the split below).  A one-thread node runs
// the host's sequential code.
// Node numbers follow the host's depth-first order through dfs_first_child.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "bvh.h"
#include "bvh_fit.h"
#include "bvh_gpu.h"

namespace c2a {

namespace {

constexpr int kWaveMax = 2048;   // nodes above this many triangles: 256 threads
constexpr int kThreadMax = 24;   // nodes at most this large: one thread

// node classes: 0 workgroup, 1 wavefront, 2 thread
__host__ __device__ inline int node_class(int32_t num) {
  return num > kWaveMax ? 0 : (num > kThreadMax ? 1 : 2);
}

struct BuildJob {
  int32_t bn, first, num, depth, rights, pad[3];
};

struct BuildArgs {
  const double* src;       // n x 9, caller order
  int32_t* perm;           // position -> caller triangle
  int32_t* perm_tmp;       // the split's output arrangement
  int32_t* tape;           // 2n: the partition's queue tape
  int32_t* cand;           // 3n: corner candidates
  GpuNode* nodes;          // 2n - 1, zeroed
  const BuildJob* jobs[3];  // this level's nodes of each class
  BuildJob* next[3];        // the next level's nodes of each class
  int32_t* counts;          // [0..2] next level's node counts, [3] max depth
  int32_t refit;            // 1: fit the listed nodes only (topology kept)
};

__device__ inline void push_child(const BuildArgs& a, const BuildJob& j) {
  const int c = node_class(j.num);
  a.next[c][atomicAdd(&a.counts[c], 1)] = j;
}

__device__ inline const double* tri_at(const BuildArgs& a, int32_t pos) {
  return a.src + 9 * (int64_t)a.perm[pos];
}

// point i (3 * triangle + vertex) of a node, projected into orientation R
__device__ inline void proj(const BuildArgs& a, const double* R, int32_t first, int i, double* P) {
  const double* v = tri_at(a, first + i / 3) + 3 * (i % 3);
  mtxv(P, R, v);
}

// ---- workgroup reductions over B threads (B / 64 wavefronts) --------------

template <int B>
__device__ double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sh[0];
  for (int k = 1; k < B / 64; k++) r = sh[k] > r ? sh[k] : r;
  __syncthreads();
  return r;
}

template <int B>
__device__ double block_min(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sh[0];
  for (int k = 1; k < B / 64; k++) r = sh[k] < r ? sh[k] : r;
  __syncthreads();
  return r;
}

// (value, index) with the smallest value, first index among ties; sign = -1
// turns it into the largest value
template <int B>
__device__ int block_argext(double v, int idx, double sign, double* shv, int* shi) {
  double key = sign * v;
  for (int o = 32; o > 0; o >>= 1) {
    const double kw = __shfl_xor(key, o);
    const int iw = __shfl_xor(idx, o);
    if (kw < key || (kw == key && iw < idx)) {
      key = kw;
      idx = iw;
    }
  }
  if ((threadIdx.x & 63) == 0) {
    shv[threadIdx.x >> 6] = key;
    shi[threadIdx.x >> 6] = idx;
  }
  __syncthreads();
  double bk = shv[0];
  int bi = shi[0];
  for (int k = 1; k < B / 64; k++)
    if (shv[k] < bk || (shv[k] == bk && shi[k] < bi)) {
      bk = shv[k];
      bi = shi[k];
    }
  __syncthreads();
  return bi;
}

// exclusive prefix sum of v over the workgroup; total gets the sum
template <int B>
__device__ int block_scan(int v, int* sh, int& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  int before = 0;
  total = 0;
  for (int k = 0; k < B / 64; k++) {
    if (k < w) before += sh[k];
    total += sh[k];
  }
  __syncthreads();
  return before + x - v;
}

// ---- one large node per workgroup -----------------------------------------

template <int B>
__global__ void __launch_bounds__(B) build_node(BuildArgs a, int cls) {
  __shared__ double chunk[B * 9];
  __shared__ double shv[B / 64];
  __shared__ int shi[B / 64];
  __shared__ int shs[B / 64];
  __shared__ double acc[9], R[9], box[4];
  __shared__ int flag;
  const BuildJob job = a.jobs[cls][blockIdx.x];
  const int tid = threadIdx.x;
  const int32_t first = job.first, num = job.num;
  const int np = 3 * num;

  // covariance sums, in triangle order (C2A_Build.cpp:279-335): every
  // thread computes its triangle's nine terms into LDS, then lane k < 9 adds
  // term k of each triangle, in order, to sum k (one dependent add per
  // triangle and sum)
  double run = 0.0;
  for (int base = 0; base < num; base += B) {
    const int i = base + tid;
    if (i < num) cov_terms(tri_at(a, first + i), chunk + 9 * tid);
    __syncthreads();
    if (tid < 9) {
      const int m = num - base < B ? num - base : B;
      double sk = run;
      int j = 0;
      for (; j + 8 <= m; j += 8) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t[u] = chunk[9 * (j + u) + tid];
#pragma unroll
        for (int u = 0; u < 8; u++) sk += t[u];
      }
      for (; j < m; j++) sk += chunk[9 * j + tid];
      run = sk;
    }
    __syncthreads();
  }
  if (tid < 9) acc[tid] = run;
  __syncthreads();
  if (tid == 0) {
    double C[9], E[9], sv[3], Rl[9], s9[9];
    for (int k = 0; k < 9; k++) s9[k] = acc[k];
    cov_finish(s9, num, C);
    jacobi3(C, E, sv);
    orient_from_eigen(E, sv, Rl);
    for (int k = 0; k < 9; k++) R[k] = Rl[k];
  }
  __syncthreads();
  double Rr[9];
  for (int k = 0; k < 9; k++) Rr[k] = R[k];

  // angular radius, slab (z) extent, then the in-plane arg-extremes
  double amax = 0.0, zmin = INFINITY, zmax = -INFINITY;
  double xmn = INFINITY, xmx = -INFINITY, ymn = INFINITY, ymx = -INFINITY;
  int ixmn = 0x7fffffff, ixmx = 0x7fffffff, iymn = 0x7fffffff, iymx = 0x7fffffff;
  for (int i = tid; i < np; i += B) {
    const double* v = tri_at(a, first + i / 3) + 3 * (i % 3);
    const double d = vertex_radius(v);
    amax = d > amax ? d : amax;
    double P[3];
    mtxv(P, Rr, v);
    zmin = P[2] < zmin ? P[2] : zmin;
    zmax = P[2] > zmax ? P[2] : zmax;
    if (P[0] < xmn) { xmn = P[0]; ixmn = i; }
    if (P[0] > xmx) { xmx = P[0]; ixmx = i; }
    if (P[1] < ymn) { ymn = P[1]; iymn = i; }
    if (P[1] > ymx) { ymx = P[1]; iymx = i; }
  }
  const double angR = block_max<B>(amax, shv);
  const double minz = block_min<B>(zmin, shv);
  const double maxz = block_max<B>(zmax, shv);
  ixmn = block_argext<B>(xmn, ixmn, 1.0, shv, shi);
  ixmx = block_argext<B>(xmx, ixmx, -1.0, shv, shi);
  iymn = block_argext<B>(ymn, iymn, 1.0, shv, shi);
  iymx = block_argext<B>(ymx, iymx, -1.0, shv, shi);
  const double r = 0.5 * (maxz - minz);
  const double radsqr = r * r;
  const double cz = 0.5 * (maxz + minz);
  double P[3];
  proj(a, Rr, first, ixmn, P);
  const double lox0 = extent_lo(P[0], P[2], cz, radsqr);
  proj(a, Rr, first, ixmx, P);
  const double hix0 = extent_hi(P[0], P[2], cz, radsqr);
  proj(a, Rr, first, iymn, P);
  const double loy0 = extent_lo(P[1], P[2], cz, radsqr);
  proj(a, Rr, first, iymx, P);
  const double hiy0 = extent_hi(P[1], P[2], cz, radsqr);
  double lox = lox0, hix = hix0, loy = loy0, hiy = hiy0;
  for (int i = tid; i < np; i += B) {
    proj(a, Rr, first, i, P);
    if (P[0] < lox0) { const double x = extent_lo(P[0], P[2], cz, radsqr); lox = x < lox ? x : lox; }
    if (P[0] > hix0) { const double x = extent_hi(P[0], P[2], cz, radsqr); hix = x > hix ? x : hix; }
    if (P[1] < loy0) { const double x = extent_lo(P[1], P[2], cz, radsqr); loy = x < loy ? x : loy; }
    if (P[1] > hiy0) { const double x = extent_hi(P[1], P[2], cz, radsqr); hiy = x > hiy ? x : hiy; }
  }
  double minx = block_min<B>(lox, shv), maxx = block_max<B>(hix, shv);
  double miny = block_min<B>(loy, shv), maxy = block_max<B>(hiy, shv);

  // corner pass: the candidates in point order, then one lane in sequence
  int ncand = 0;
  int32_t* cand = a.cand + 3 * (int64_t)first;
  for (int base = 0; base < np; base += B) {
    const int i = base + tid;
    int f = 0;
    if (i < np) {
      proj(a, Rr, first, i, P);
      f = corner_candidate(P[0], P[1], minx, maxx, miny, maxy) ? 1 : 0;
    }
    int tot;
    const int at = block_scan<B>(f, shs, tot);
    if (f) cand[ncand + at] = i;
    ncand += tot;
  }
  // The host grows the rectangle point by point (C2A_BV.cpp:542-606); a
  // point changes it only when it lies outside the current rounded corner.
  // Wavefront 0 evaluates 64 candidates at once against the current
  // rectangle: the points before the first one that changes it are no-ops
  // in sequence too, and that one's result is the next rectangle.  Each
  // chunk's projected points are staged in LDS by the whole workgroup.
  __syncthreads();
  double* cp = chunk;
  for (int base = 0; base < ncand; base += B) {
    const int m = ncand - base < B ? ncand - base : B;
    if (tid < m) proj(a, Rr, first, cand[base + tid], cp + 3 * tid);
    __syncthreads();
    if (tid < 64) {
      int pos = 0;
      while (pos < m) {
        const int j = pos + tid;
        double x0 = minx, x1 = maxx, y0 = miny, y1 = maxy;
        bool changed = false;
        if (j < m) {
          corner_grow(cp[3 * j], cp[3 * j + 1], cp[3 * j + 2], cz, radsqr, x0, x1, y0, y1);
          changed = __double_as_longlong(x0) != __double_as_longlong(minx) ||
                    __double_as_longlong(x1) != __double_as_longlong(maxx) ||
                    __double_as_longlong(y0) != __double_as_longlong(miny) ||
                    __double_as_longlong(y1) != __double_as_longlong(maxy);
        }
        const unsigned long long hit = __ballot(changed);
        if (hit == 0) {
          pos += 64;
          continue;
        }
        const int f = __ffsll(hit) - 1;
        minx = __shfl(x0, f);
        maxx = __shfl(x1, f);
        miny = __shfl(y0, f);
        maxy = __shfl(y1, f);
        pos += f + 1;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    GpuNode& nd = a.nodes[job.bn];
    for (int k = 0; k < 9; k++) nd.R[k] = Rr[k];
    nd.angR = angR;
    fit_finish(Rr, minx, maxx, miny, maxy, cz, r, nd.Tr, nd.l, &nd.r);
    for (int k = 0; k < 9; k++) nd.Rloc[k] = Rr[k];
    if (!a.refit) {
      nd.depth = job.depth;
      nd.first_child = dfs_first_child(job.depth, first, job.rights);
      atomicMax(&a.counts[3], job.depth);
      const double ax[3] = {Rr[0], Rr[3], Rr[6]};
      double s9[9];
      for (int k = 0; k < 9; k++) s9[k] = acc[k];
      box[0] = split_plane(s9, num, ax);
    }
  }
  __syncthreads();
  if (a.refit) return;   // the tree stays: no split

  // split (split_tris, C2A_Build.cpp:344-385): the in-place partition swaps
  // each triangle at or below the plane with the first one above it, so the
  // triangles above it form a queue that a "below" triangle rotates by one
  // (front to back) and an "above" triangle extends.  The arrangement is
  // rebuilt from prefix counts: the queue's pushes are written to a tape,
  // a rotation's push refers back to the tape entry it pops, and pointer
  // jumping resolves the references.
  const double axis[3] = {Rr[0], Rr[3], Rr[6]};
  const double coord = box[0];
  int32_t* tape = a.tape + 2 * (int64_t)first;
  int nsmall = 0, nbig = 0, ncyc = 0;
  for (int base = 0; base < num; base += B) {
    const int i = base + tid;
    int s = 0, b = 0;
    if (i < num) {
      s = split_coord(tri_at(a, first + i), axis) <= coord ? 1 : 0;
      b = 1 - s;
    }
    int tot;
    const int ex = block_scan<B>(s | (b << 16), shs, tot);
    const int sb = nsmall + (ex & 0xffff), bb = nbig + (ex >> 16);
    const int cyc = (s && bb > 0) ? 1 : 0;
    int ctot;
    const int cb = ncyc + block_scan<B>(cyc, shs, ctot);
    if (i < num) {
      const int32_t elem = a.perm[first + i];
      if (s) a.perm_tmp[first + sb] = elem;
      if (b) tape[bb + cb] = elem;                 // push index = pushes before
      else if (cyc) tape[bb + cb] = -(cb + 1);     // refers to the popped entry
    }
    nsmall += tot & 0xffff;
    nbig += tot >> 16;
    ncyc += ctot;
  }
  __syncthreads();
  const int T = nbig + ncyc;
  for (;;) {
    if (tid == 0) flag = 0;
    __syncthreads();
    for (int p = tid; p < T; p += B) {
      const int32_t v = tape[p];
      if (v < 0) {
        tape[p] = tape[-v - 1];
        flag = 1;
      }
    }
    __syncthreads();
    if (!flag) break;
    __syncthreads();
  }
  for (int j = tid; j < nbig; j += B) a.perm_tmp[first + nsmall + j] = tape[ncyc + j];
  __syncthreads();
  for (int i = tid; i < num; i += B) a.perm[first + i] = a.perm_tmp[first + i];
  if (tid == 0) {
    int c1 = nsmall;
    if (c1 == 0 || c1 == num) c1 = num / 2;
    const int fc = dfs_first_child(job.depth, first, job.rights);
    BuildJob kids[2];
    kids[0] = BuildJob{fc, first, c1, job.depth + 1, job.rights, {0, 0, 0}};
    kids[1] = BuildJob{fc + 1, first + c1, num - c1, job.depth + 1, job.rights + 1, {0, 0, 0}};
    push_child(a, kids[0]);
    push_child(a, kids[1]);
  }
}

// ---- small subtrees: one thread each, the host's sequential code -----------

__device__ void fit_small(const BuildArgs& a, const BuildJob& j, double* S1) {
  double acc[9];
  for (int k = 0; k < 9; k++) acc[k] = 0.0;
  for (int i = 0; i < j.num; i++) cov_add(acc, tri_at(a, j.first + i));
  double C[9], E[9], sv[3], R[9];
  cov_finish(acc, j.num, C);
  jacobi3(C, E, sv);
  orient_from_eigen(E, sv, R);
  for (int k = 0; k < 3; k++) S1[k] = acc[k];
  const int np = 3 * j.num;
  double angR = 0.0;
  for (int i = 0; i < np; i++) {
    const double d = vertex_radius(tri_at(a, j.first + i / 3) + 3 * (i % 3));
    if (d > angR) angR = d;
  }
  double P[3], Q[3];
  proj(a, R, j.first, 0, P);
  double minz = P[2], maxz = P[2];
  for (int i = 1; i < np; i++) {
    proj(a, R, j.first, i, P);
    if (P[2] < minz) minz = P[2];
    else if (P[2] > maxz) maxz = P[2];
  }
  const double r = 0.5 * (maxz - minz);
  const double radsqr = r * r;
  const double cz = 0.5 * (maxz + minz);
  double lo[2], hi[2];
  for (int c = 0; c < 2; c++) {
    int mn = 0, mx = 0;
    proj(a, R, j.first, 0, P);
    double vmn = P[c], vmx = P[c];
    for (int i = 1; i < np; i++) {
      proj(a, R, j.first, i, P);
      if (P[c] < vmn) { mn = i; vmn = P[c]; }
      else if (P[c] > vmx) { mx = i; vmx = P[c]; }
    }
    proj(a, R, j.first, mn, P);
    lo[c] = extent_lo(P[c], P[2], cz, radsqr);
    proj(a, R, j.first, mx, Q);
    hi[c] = extent_hi(Q[c], Q[2], cz, radsqr);
    for (int i = 0; i < np; i++) {
      proj(a, R, j.first, i, P);
      if (P[c] < lo[c]) {
        const double x = extent_lo(P[c], P[2], cz, radsqr);
        if (x < lo[c]) lo[c] = x;
      }
    }
    for (int i = 0; i < np; i++) {
      proj(a, R, j.first, i, P);
      if (P[c] > hi[c]) {
        const double x = extent_hi(P[c], P[2], cz, radsqr);
        if (x > hi[c]) hi[c] = x;
      }
    }
  }
  double minx = lo[0], maxx = hi[0], miny = lo[1], maxy = hi[1];
  for (int i = 0; i < np; i++) {
    proj(a, R, j.first, i, P);
    corner_grow(P[0], P[1], P[2], cz, radsqr, minx, maxx, miny, maxy);
  }
  GpuNode& nd = a.nodes[j.bn];
  for (int k = 0; k < 9; k++) nd.R[k] = R[k];
  nd.angR = angR;
  fit_finish(R, minx, maxx, miny, maxy, cz, r, nd.Tr, nd.l, &nd.r);
  for (int k = 0; k < 9; k++) nd.Rloc[k] = R[k];
  nd.depth = j.depth;
}

__global__ void __launch_bounds__(64) build_thread(BuildArgs a, int32_t njobs) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= njobs) return;
  const BuildJob j = a.jobs[2][t];
  double S1[3];
  fit_small(a, j, S1);
  if (a.refit) return;   // the tree stays: no split
  GpuNode& nd = a.nodes[j.bn];
  atomicMax(&a.counts[3], j.depth);
  if (j.num == 1) {
    nd.first_child = -(j.first + 1);
    return;
  }
  const int fc = dfs_first_child(j.depth, j.first, j.rights);
  nd.first_child = fc;
  const double axis[3] = {nd.R[0], nd.R[3], nd.R[6]};
  const double coord = split_plane(S1, j.num, axis);
  int c1 = 0;
  for (int i = 0; i < j.num; i++) {
    if (split_coord(tri_at(a, j.first + i), axis) <= coord) {
      const int32_t x = a.perm[j.first + i];
      a.perm[j.first + i] = a.perm[j.first + c1];
      a.perm[j.first + c1] = x;
      c1++;
    }
  }
  if (c1 == 0 || c1 == j.num) c1 = j.num / 2;
  push_child(a, BuildJob{fc, j.first, c1, j.depth + 1, j.rights, {0, 0, 0}});
  push_child(a, BuildJob{fc + 1, j.first + c1, j.num - c1, j.depth + 1, j.rights + 1, {0, 0, 0}});
}

// ---- parent-relative transforms, sizes and the reordered triangles --------

__global__ void model_frames(GpuNode* nodes, int64_t nn, double* mRT, int32_t* parent) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= nn) return;
  const GpuNode& nd = nodes[v];
  for (int k = 0; k < 9; k++) mRT[12 * v + k] = nd.R[k];
  for (int k = 0; k < 3; k++) mRT[12 * v + 9 + k] = nd.Tr[k];
  if (nd.first_child >= 0) {
    parent[nd.first_child] = (int32_t)v;
    parent[nd.first_child + 1] = (int32_t)v;
  }
}

__global__ void relative_frames(GpuNode* nodes, int64_t nn, const double* mRT,
                                const int32_t* parent) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= nn) return;
  GpuNode& nd = nodes[v];
  double pR[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, pT[3] = {0, 0, 0};
  if (v > 0) {
    const int32_t p = parent[v];
    for (int k = 0; k < 9; k++) pR[k] = mRT[12 * (int64_t)p + k];
    for (int k = 0; k < 3; k++) pT[k] = mRT[12 * (int64_t)p + 9 + k];
  }
  parent_relative(pR, pT, nd.R, nd.Tr);
  nd.size = node_size(nd.l, nd.r);
}

__global__ void gather_tris(const double* src, const int32_t* ids, const int32_t* vidx,
                            const int32_t* perm, int64_t n, double* tris, int32_t* ids_out,
                            int32_t* vidx_out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t s = perm[i];
  for (int k = 0; k < 9; k++) tris[9 * i + k] = src[9 * (int64_t)s + k];
  ids_out[i] = ids ? ids[s] : s;
  for (int k = 0; k < 3; k++) vidx_out[3 * i + k] = vidx ? vidx[3 * (int64_t)s + k] : 3 * s + k;
}

__global__ void gather_xyz(const double* src, const int32_t* perm, int64_t n, double* tris) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t s = perm[i];
  for (int k = 0; k < 9; k++) tris[9 * i + k] = src[9 * (int64_t)s + k];
}

__global__ void iota(int32_t* p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int32_t)i;
}

struct Dev {
  void* p = nullptr;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 1); }
};

#define BVH_TRY(expr)                    \
  do {                                   \
    hipError_t e_ = (expr);              \
    if (e_ != hipSuccess) return e_;     \
  } while (0)

inline unsigned grid_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t build_bvh_device(const double* xyz9, const int32_t* ids, const int32_t* vidx3,
                            int64_t n, Bvh* out, void** nodes_dev, void** tris_dev,
                            hipStream_t s) {
  *nodes_dev = nullptr;
  *tris_dev = nullptr;
  const int64_t nn = 2 * n - 1;
  Dev src, dids, dvidx, perm, perm_tmp, tape, cand, counts, mRT, parent, tris, ids_out, vidx_out;
  Dev lists[2][3];   // ping-pong job lists of the three node classes
  GpuNode* nodes = nullptr;
  BVH_TRY(src.alloc((size_t)n * 9 * sizeof(double)));
  BVH_TRY(perm.alloc((size_t)n * 4));
  BVH_TRY(perm_tmp.alloc((size_t)n * 4));
  BVH_TRY(tape.alloc((size_t)n * 8));
  BVH_TRY(cand.alloc((size_t)n * 12));
  for (int l = 0; l < 2; l++)
    for (int c = 0; c < 3; c++) BVH_TRY(lists[l][c].alloc((size_t)n * sizeof(BuildJob)));
  BVH_TRY(counts.alloc(4 * sizeof(int32_t)));
  BVH_TRY(hipMalloc((void**)&nodes, (size_t)nn * sizeof(GpuNode)));
  Dev nodes_guard;
  nodes_guard.p = nodes;
  BVH_TRY(hipMemcpyAsync(src.p, xyz9, (size_t)n * 9 * sizeof(double), hipMemcpyHostToDevice, s));
  if (ids) {
    BVH_TRY(dids.alloc((size_t)n * 4));
    BVH_TRY(hipMemcpyAsync(dids.p, ids, (size_t)n * 4, hipMemcpyHostToDevice, s));
  }
  if (vidx3) {
    BVH_TRY(dvidx.alloc((size_t)n * 12));
    BVH_TRY(hipMemcpyAsync(dvidx.p, vidx3, (size_t)n * 12, hipMemcpyHostToDevice, s));
  }
  BVH_TRY(hipMemsetAsync(nodes, 0, (size_t)nn * sizeof(GpuNode), s));
  BVH_TRY(hipMemsetAsync(counts.p, 0, 4 * sizeof(int32_t), s));
  hipLaunchKernelGGL(iota, dim3(grid_of(n, 256)), dim3(256), 0, s, (int32_t*)perm.p, n);
  BVH_TRY(hipGetLastError());

  BuildArgs a;
  memset(&a, 0, sizeof(a));   // refit = 0: a build
  a.src = (const double*)src.p;
  a.perm = (int32_t*)perm.p;
  a.perm_tmp = (int32_t*)perm_tmp.p;
  a.tape = (int32_t*)tape.p;
  a.cand = (int32_t*)cand.p;
  a.nodes = nodes;
  a.counts = (int32_t*)counts.p;
  const BuildJob root{0, 0, (int32_t)n, 0, 0, {0, 0, 0}};
  int32_t host_counts[4] = {0, 0, 0, 0};
  BVH_TRY(hipStreamSynchronize(s));
  const int rc = node_class((int32_t)n);
  BVH_TRY(hipMemcpy(lists[0][rc].p, &root, sizeof(root), hipMemcpyHostToDevice));
  int64_t cnt[3] = {0, 0, 0};
  cnt[rc] = 1;
  for (int level = 0; cnt[0] + cnt[1] + cnt[2] > 0; level++) {
    const int cur = level & 1, nxt = cur ^ 1;
    for (int c = 0; c < 3; c++) {
      a.jobs[c] = (const BuildJob*)lists[cur][c].p;
      a.next[c] = (BuildJob*)lists[nxt][c].p;
    }
    BVH_TRY(hipMemsetAsync(a.counts, 0, 3 * sizeof(int32_t), s));
    if (cnt[0] > 0)
      hipLaunchKernelGGL(build_node<256>, dim3((unsigned)cnt[0]), dim3(256), 0, s, a, 0);
    if (cnt[1] > 0)
      hipLaunchKernelGGL(build_node<64>, dim3((unsigned)cnt[1]), dim3(64), 0, s, a, 1);
    if (cnt[2] > 0)
      hipLaunchKernelGGL(build_thread, dim3(grid_of(cnt[2], 64)), dim3(64), 0, s, a,
                         (int32_t)cnt[2]);
    BVH_TRY(hipGetLastError());
    BVH_TRY(hipMemcpyAsync(host_counts, a.counts, sizeof(host_counts), hipMemcpyDeviceToHost, s));
    BVH_TRY(hipStreamSynchronize(s));
    for (int c = 0; c < 3; c++) cnt[c] = host_counts[c];
  }
  BVH_TRY(mRT.alloc((size_t)nn * 12 * sizeof(double)));
  BVH_TRY(parent.alloc((size_t)nn * 4));
  hipLaunchKernelGGL(model_frames, dim3(grid_of(nn, 256)), dim3(256), 0, s, nodes, nn,
                     (double*)mRT.p, (int32_t*)parent.p);
  BVH_TRY(hipGetLastError());
  hipLaunchKernelGGL(relative_frames, dim3(grid_of(nn, 256)), dim3(256), 0, s, nodes, nn,
                     (const double*)mRT.p, (const int32_t*)parent.p);
  BVH_TRY(hipGetLastError());
  BVH_TRY(tris.alloc((size_t)n * 9 * sizeof(double)));
  BVH_TRY(ids_out.alloc((size_t)n * 4));
  BVH_TRY(vidx_out.alloc((size_t)n * 12));
  hipLaunchKernelGGL(gather_tris, dim3(grid_of(n, 256)), dim3(256), 0, s, a.src,
                     (const int32_t*)dids.p, (const int32_t*)dvidx.p, a.perm, n,
                     (double*)tris.p, (int32_t*)ids_out.p, (int32_t*)vidx_out.p);
  BVH_TRY(hipGetLastError());
  out->nodes.resize((size_t)nn);
  out->tris.resize((size_t)n * 9);
  out->tri_id.resize((size_t)n);
  out->tri_vidx.resize((size_t)n * 3);
  out->src_index.resize((size_t)n);
  BVH_TRY(hipMemcpyAsync(out->src_index.data(), a.perm, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  BVH_TRY(hipMemcpyAsync(out->nodes.data(), nodes, (size_t)nn * sizeof(GpuNode),
                         hipMemcpyDeviceToHost, s));
  BVH_TRY(hipMemcpyAsync(out->tris.data(), tris.p, (size_t)n * 9 * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  BVH_TRY(hipMemcpyAsync(out->tri_id.data(), ids_out.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  BVH_TRY(hipMemcpyAsync(out->tri_vidx.data(), vidx_out.p, (size_t)n * 12, hipMemcpyDeviceToHost,
                         s));
  BVH_TRY(hipMemcpyAsync(host_counts, a.counts, sizeof(host_counts), hipMemcpyDeviceToHost, s));
  BVH_TRY(hipStreamSynchronize(s));
  out->max_depth = host_counts[3];
  // the mesh keeps the device copies
  *nodes_dev = nodes;
  nodes_guard.p = nullptr;
  *tris_dev = tris.p;
  tris.p = nullptr;
  return hipSuccess;
}

// Refit on the device: the tree is known, so each depth's nodes go in one
// launch per size class (no host round trip between levels), then the
// parent-relative frames and the triangles in build order.  Copies the
// result back into *b.
hipError_t refit_bvh_device(const double* xyz9, Bvh* b, void* nodes_dev, void* tris_dev,
                            hipStream_t s) {
  const int64_t n = (int64_t)b->tri_id.size();
  const int64_t nn = 2 * n - 1;
  if (n <= 0 || (int64_t)b->src_index.size() != n || (int64_t)b->nodes.size() != nn)
    return hipErrorInvalidValue;
  std::vector<int32_t> first, num, depth;
  node_ranges(b->nodes, first, num, depth);
  // level by level: a node's corner candidates live in its triangle range of
  // a.cand, which its descendants' ranges overlap, so only nodes of one
  // depth (disjoint ranges) may run together
  int max_depth = 0;
  for (int64_t v = 0; v < nn; v++) max_depth = std::max(max_depth, (int)depth[(size_t)v]);
  std::vector<BuildJob> jobs[3];
  // jobs of each class ordered by depth, with per-depth offsets
  std::vector<std::vector<int32_t>> cnt(3, std::vector<int32_t>((size_t)max_depth + 1, 0));
  for (int64_t v = 0; v < nn; v++) cnt[node_class(num[(size_t)v])][(size_t)depth[(size_t)v]]++;
  std::vector<std::vector<int32_t>> off(3, std::vector<int32_t>((size_t)max_depth + 2, 0));
  for (int c = 0; c < 3; c++) {
    for (int d = 0; d <= max_depth; d++) off[c][(size_t)d + 1] = off[c][(size_t)d] + cnt[c][(size_t)d];
    jobs[c].resize((size_t)off[c][(size_t)max_depth + 1]);
  }
  {
    std::vector<std::vector<int32_t>> fill = off;
    for (int64_t v = 0; v < nn; v++) {
      const int c = node_class(num[(size_t)v]);
      jobs[c][(size_t)fill[c][(size_t)depth[(size_t)v]]++] =
          BuildJob{(int32_t)v, first[(size_t)v], num[(size_t)v], depth[(size_t)v], 0, {0, 0, 0}};
    }
  }
  Dev src, perm, cand, mRT, parent, lists[3];
  BVH_TRY(src.alloc((size_t)n * 9 * sizeof(double)));
  BVH_TRY(perm.alloc((size_t)n * 4));
  BVH_TRY(cand.alloc((size_t)n * 12));
  BVH_TRY(hipMemcpyAsync(src.p, xyz9, (size_t)n * 9 * sizeof(double), hipMemcpyHostToDevice, s));
  BVH_TRY(hipMemcpyAsync(perm.p, b->src_index.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
  BuildArgs a;
  memset(&a, 0, sizeof(a));
  a.src = (const double*)src.p;
  a.perm = (int32_t*)perm.p;
  a.cand = (int32_t*)cand.p;
  a.nodes = (GpuNode*)nodes_dev;
  a.refit = 1;
  for (int c = 0; c < 3; c++) {
    if (jobs[c].empty()) continue;
    BVH_TRY(lists[c].alloc(jobs[c].size() * sizeof(BuildJob)));
    BVH_TRY(hipMemcpyAsync(lists[c].p, jobs[c].data(), jobs[c].size() * sizeof(BuildJob),
                           hipMemcpyHostToDevice, s));
  }
  for (int d = 0; d <= max_depth; d++) {
    for (int c = 0; c < 3; c++)
      a.jobs[c] = (const BuildJob*)lists[c].p + off[c][(size_t)d];
    const int32_t n0 = cnt[0][(size_t)d], n1 = cnt[1][(size_t)d], n2 = cnt[2][(size_t)d];
    if (n0 > 0) hipLaunchKernelGGL(build_node<256>, dim3((unsigned)n0), dim3(256), 0, s, a, 0);
    if (n1 > 0) hipLaunchKernelGGL(build_node<64>, dim3((unsigned)n1), dim3(64), 0, s, a, 1);
    if (n2 > 0)
      hipLaunchKernelGGL(build_thread, dim3(grid_of((int64_t)n2, 64)), dim3(64), 0, s, a, n2);
    BVH_TRY(hipGetLastError());
  }
  BVH_TRY(mRT.alloc((size_t)nn * 12 * sizeof(double)));
  BVH_TRY(parent.alloc((size_t)nn * 4));
  hipLaunchKernelGGL(model_frames, dim3(grid_of(nn, 256)), dim3(256), 0, s, a.nodes, nn,
                     (double*)mRT.p, (int32_t*)parent.p);
  BVH_TRY(hipGetLastError());
  hipLaunchKernelGGL(relative_frames, dim3(grid_of(nn, 256)), dim3(256), 0, s, a.nodes, nn,
                     (const double*)mRT.p, (const int32_t*)parent.p);
  BVH_TRY(hipGetLastError());
  hipLaunchKernelGGL(gather_xyz, dim3(grid_of(n, 256)), dim3(256), 0, s, a.src, a.perm, n,
                     (double*)tris_dev);
  BVH_TRY(hipGetLastError());
  BVH_TRY(hipMemcpyAsync(b->nodes.data(), nodes_dev, (size_t)nn * sizeof(GpuNode),
                         hipMemcpyDeviceToHost, s));
  BVH_TRY(hipMemcpyAsync(b->tris.data(), tris_dev, (size_t)n * 9 * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  BVH_TRY(hipStreamSynchronize(s));
  return hipSuccess;
}

}  // namespace c2a
