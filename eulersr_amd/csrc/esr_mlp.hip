// Learned warm start on gfx950: kNN/IDW feature assembly, a small MLP
// (BF16 MFMA with FP32 accumulation) for inference and training, AdamW, and
// the assembly of the warm-start state with its fallback test.
//
// The GEMM kernel computes C = op(A) op(B) (+ bias) with A(i,k) at
// A[i*sa_i + k*sa_k] and B(k,j) at B[k*sb_k + j*sb_j], so the forward
// product X W^T, the data gradient dZ W and the weight gradient dZ^T X are
// the same kernel with different strides. Block tile 64x64, four waves each
// owning 32x32 = 2x2 tiles of v_mfma_f32_16x16x32_bf16, K step 32 staged
// through LDS. Each output tile is produced by one block over the full K
// range in a fixed order, so every result (weight gradients included) is
// bit-reproducible.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "esr_mlp.h"

namespace esr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float32 -> bfloat16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int TM = 64, TN = 64, TK = 32, TKP = TK + 8;

// Stage a 64 x 32 tile T[r][k] of the operand X(r, k) = X[r*s_r + k*s_k]
// for rows [r0, rmax) and k in [k0, kmax). mode 1: k contiguous (one 16-byte
// load of 8 k per thread); mode 2: r contiguous (8 r per thread); mode 0:
// element-wise. Out-of-range entries are zero.
__device__ __forceinline__ void stage(uint16_t (*T)[TKP], const uint16_t* __restrict__ X, int64_t s_r, int64_t s_k,
                                      int r0, int rmax, int64_t k0, int64_t kmax, int mode, int tid) {
  if (mode == 1) {
    const int r = tid >> 2, kk = (tid & 3) * 8;
    const int gr = r0 + r;
    const int64_t gk = k0 + kk;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (gr < rmax && gk + 8 <= kmax) {
      v = *reinterpret_cast<const uint4*>(X + (int64_t)gr * s_r + gk);
      *reinterpret_cast<uint4*>(&T[r][kk]) = v;
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) T[r][kk + q] = (gr < rmax && gk + q < kmax) ? X[(int64_t)gr * s_r + gk + q] : (uint16_t)0;
    }
  } else if (mode == 2) {
    const int k = tid >> 3, rr = (tid & 7) * 8;
    const int64_t gk = k0 + k;
    const int gr = r0 + rr;
    if (gk < kmax && gr + 8 <= rmax) {
      const uint4 v = *reinterpret_cast<const uint4*>(X + gk * s_k + gr);
      const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
      for (int q = 0; q < 8; ++q) T[rr + q][k] = h[q];
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) T[rr + q][k] = (gk < kmax && gr + q < rmax) ? X[gk * s_k + gr + q] : (uint16_t)0;
    }
  } else {
    for (int e = tid; e < TM * TK; e += 256) {
      const int r = e / TK, k = e % TK;
      const int gr = r0 + r;
      const int64_t gk = k0 + k;
      T[r][k] = (gr < rmax && gk < kmax) ? X[(int64_t)gr * s_r + gk * s_k] : (uint16_t)0;
    }
  }
}

// blockIdx.z selects a K chunk [z*kchunk, (z+1)*kchunk); its fp32 partial
// goes to C32 + z*c_split (bias and bf16 output only when gridDim.z == 1).
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[TM][TKP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][TKP];  // Bs[j][k]
  const int bm = blockIdx.y * TM, bn = blockIdx.x * TN;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const int64_t kbeg = (int64_t)blockIdx.z * g.kchunk;
  const int64_t kend = kbeg + g.kchunk < g.K ? kbeg + g.kchunk : g.K;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int64_t k0 = kbeg; k0 < kend; k0 += TK) {
    stage(As, g.A, g.sa_i, g.sa_k, bm, g.M, k0, kend, g.mode_a, tid);
    stage(Bs, g.B, g.sb_j, g.sb_k, bn, g.N, k0, kend, g.mode_b, tid);
    __syncthreads();
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      af[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&As[wm + t * 16 + lr][lk]));
      bfr[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&Bs[wn + t * 16 + lr][lk]));
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
  float* C32 = g.C32 ? g.C32 + (int64_t)blockIdx.z * g.c_split : nullptr;
  // D(row, col): row = 4*(lane/16) + r, col = lane % 16 within each 16x16 tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = bm + wm + a * 16 + (lane >> 4) * 4 + r;
        const int j = bn + wn + b * 16 + lr;
        if (i < g.M && j < g.N) {
          float v = acc[a][b][r];
          if (g.bias) v += g.bias[j];
          if (C32) C32[(int64_t)i * g.ldc32 + j] = v;
          if (g.C16) g.C16[(int64_t)i * g.ldc16 + j] = f2bf(g.relu ? fmaxf(v, 0.f) : v);
        }
      }
}

// out[e] = sum over s of part[s*n + e]: four interleaved partial sums
// (s mod 4) combined in a fixed order
__global__ void k_sum_splits(const float* __restrict__ part, int32_t ns, int64_t n, float* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
  int32_t q = 0;
  for (; q + 4 <= ns; q += 4) {
    a0 += part[(int64_t)q * n + e];
    a1 += part[(int64_t)(q + 1) * n + e];
    a2 += part[(int64_t)(q + 2) * n + e];
    a3 += part[(int64_t)(q + 3) * n + e];
  }
  for (; q < ns; ++q) a0 += part[(int64_t)q * n + e];
  out[e] = (a0 + a1) + (a2 + a3);
}

// partial column sums of an fp32 (rows x n) matrix: block (x, y) of 256
// threads covers 64 columns and rows [y*chunk, (y+1)*chunk); thread row-lane
// t = tid/64 sums rows r0 + t, r0 + t + 4, ... with two accumulators; the
// four lanes are combined in a fixed order into part[y*n + j]
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ X, int64_t rows, int32_t n,
                                                     int64_t ld, int64_t chunk, float* __restrict__ part) {
  __shared__ float sh[4][64];
  const int c = threadIdx.x & 63, t = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + c;
  const int64_t r0 = (int64_t)blockIdx.y * chunk;
  const int64_t r1 = r0 + chunk < rows ? r0 + chunk : rows;
  float a0 = 0.f, a1 = 0.f;
  if (j < n) {
    int64_t i = r0 + t;
    for (; i + 4 < r1; i += 8) {
      a0 += X[i * ld + j];
      a1 += X[(i + 4) * ld + j];
    }
    for (; i < r1; i += 4) a0 += X[i * ld + j];
  }
  sh[t][c] = a0 + a1;
  __syncthreads();
  if (t == 0 && j < n) part[(int64_t)blockIdx.y * n + j] = (sh[0][c] + sh[1][c]) + (sh[2][c] + sh[3][c]);
}

// MSE of (rows x n) predictions against targets; dY = 2 (Y - T) / (rows n)
// written as fp32 and bf16; per-block partial sums (double) for the loss.
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, int64_t total, float scale,
                      float* __restrict__ dY32, uint16_t* __restrict__ dY16, double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const float d = Y[e] - T[e];
    acc += (double)d * (double)d;
    const float g = scale * d;
    dY32[e] = g;
    dY16[e] = f2bf(g);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// dZ = dH where the forward pre-activation is positive (ReLU)
__global__ void k_relu_bwd(const float* __restrict__ Z, const float* __restrict__ dH, int64_t total,
                           float* __restrict__ dZ32, uint16_t* __restrict__ dZ16) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const float g = Z[e] > 0.f ? dH[e] : 0.f;
  dZ32[e] = g;
  dZ16[e] = f2bf(g);
}

// decoupled-weight-decay Adam on fp32 master weights; refreshes the bf16 copy
__global__ void k_adamw(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                        float* __restrict__ v, uint16_t* __restrict__ w16, int64_t n, float lr, float b1, float b2,
                        float eps, float wd, float c1, float c2) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.f - b1) * gi;
  const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  const float mh = mi / c1, vh = vi / c2;
  float wi = w[e];
  wi = wi - lr * (mh / (sqrtf(vh) + eps) + wd * wi);
  w[e] = wi;
  if (w16) w16[e] = f2bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = f2bf(x[e]);
}

// ---------------------------------------------------------------------------
// warm-start features and state assembly
// ---------------------------------------------------------------------------

// One row per (case, HR cell). Columns: for each of the k LR neighbours its
// normalised primitives (4), offset / h_LR (2) and IDW weight (1); then the
// wall features (4), the normalised Mach and angle of attack (2); zero
// padding up to in_pad. Also the IDW prediction sum_j w_j prim_j
// (eval/core.py:135-142).
// Row assembly with the neighbour count K known at compile time: the row is
// built in registers and written as 16-byte stores (in_pad % 8 == 0, the
// row start is then 16-byte aligned). K = 0: any k, element stores.
template <int K>
__global__ void k_warm_features(int32_t n_hr, int32_t n_lr, int32_t k, const int32_t* __restrict__ idx,
                                const float* __restrict__ rel, const float* __restrict__ w,
                                const float* __restrict__ wall, int32_t n_cases, const float* __restrict__ lrW,
                                const float* __restrict__ scal, float4 mu, float4 isig, float gm1, int32_t in_pad,
                                uint16_t* __restrict__ X, float* __restrict__ idw) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= (int64_t)n_hr * n_cases) return;
  const int32_t c = (int32_t)(row / n_hr), i = (int32_t)(row - (int64_t)c * n_hr);
  if (K > 0) k = K;
  uint16_t* x = X + row * in_pad;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  constexpr int NV = K > 0 ? ((7 * K + 6 + 7) / 8) * 8 : 8;
  float v[NV];
  int col = 0;
  for (int j = 0; j < k; ++j) {
    const int32_t l = idx[(int64_t)i * k + j];
    const float4 W = reinterpret_cast<const float4*>(lrW)[(int64_t)c * n_lr + l];
    const float r = W.x, u = W.y / r, vv = W.z / r;
    const float P = (W.w - 0.5f * r * (u * u + vv * vv)) * gm1;
    const float wj = w[(int64_t)i * k + j];
    acc.x += wj * r;
    acc.y += wj * u;
    acc.z += wj * vv;
    acc.w += wj * P;
    const float2 rl = reinterpret_cast<const float2*>(rel)[(int64_t)i * k + j];
    const float f[7] = {(r - mu.x) * isig.x, (u - mu.y) * isig.y, (vv - mu.z) * isig.z, (P - mu.w) * isig.w,
                        rl.x, rl.y, wj};
    if (K > 0) {
#pragma unroll
      for (int q = 0; q < 7; ++q) v[7 * j + q] = f[q];
    } else {
      for (int q = 0; q < 7; ++q) x[col + q] = f2bf(f[q]);
      col += 7;
    }
  }
  const float4 wf = reinterpret_cast<const float4*>(wall)[i];
  const float tail[6] = {wf.x, wf.y, wf.z, wf.w, scal[2 * c], scal[2 * c + 1]};
  if (K > 0) {
#pragma unroll
    for (int q = 0; q < 6; ++q) v[7 * K + q] = tail[q];
#pragma unroll
    for (int q = 7 * K + 6; q < NV; ++q) v[q] = 0.f;
    if ((in_pad & 7) == 0 && in_pad >= NV) {
#pragma unroll
      for (int b = 0; b < NV / 8; ++b) {
        uint4 u4;
        u4.x = f2bf(v[8 * b + 0]) | ((uint32_t)f2bf(v[8 * b + 1]) << 16);
        u4.y = f2bf(v[8 * b + 2]) | ((uint32_t)f2bf(v[8 * b + 3]) << 16);
        u4.z = f2bf(v[8 * b + 4]) | ((uint32_t)f2bf(v[8 * b + 5]) << 16);
        u4.w = f2bf(v[8 * b + 6]) | ((uint32_t)f2bf(v[8 * b + 7]) << 16);
        reinterpret_cast<uint4*>(x)[b] = u4;
      }
      for (int q = NV; q < in_pad; ++q) x[q] = 0;
    } else {
#pragma unroll
      for (int q = 0; q < 7 * K + 6; ++q) x[q] = f2bf(v[q]);
      for (int q = 7 * K + 6; q < in_pad; ++q) x[q] = 0;
    }
  } else {
    for (int q = 0; q < 6; ++q) x[col + q] = f2bf(tail[q]);
    for (col += 6; col < in_pad; ++col) x[col] = 0;
  }
  reinterpret_cast<float4*>(idw)[row] = acc;
}

// prim = IDW + sig * Y (Y null: IDW only) -> conserved; flags a case whose
// prediction has rho <= 0, p <= 0 or a non-finite value.
__global__ void k_warm_apply(int64_t rows, int32_t n_hr, const float* __restrict__ idw, const float* __restrict__ Y,
                             float4 sig, float gm1, float* __restrict__ Wout, int32_t* __restrict__ bad) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  float4 p = reinterpret_cast<const float4*>(idw)[row];
  if (Y) {
    const float4 d = reinterpret_cast<const float4*>(Y)[row];
    p.x += sig.x * d.x;
    p.y += sig.y * d.y;
    p.z += sig.z * d.z;
    p.w += sig.w * d.w;
  }
  const bool ok = (p.x > 0.f) && (p.w > 0.f) && isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(p.w);
  if (!ok) atomicOr(&bad[row / n_hr], 1);
  const float E = p.w / gm1 + 0.5f * p.x * (p.y * p.y + p.z * p.z);
  reinterpret_cast<float4*>(Wout)[row] = make_float4(p.x, p.x * p.y, p.x * p.z, E);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

namespace {
int stage_mode(const uint16_t* X, int64_t s_r, int64_t s_k) {
  const bool al = (reinterpret_cast<uintptr_t>(X) & 15u) == 0;
  if (s_k == 1 && s_r % 8 == 0 && al) return 1;
  if (s_r == 1 && s_k % 8 == 0 && al) return 2;
  return 0;
}
}  // namespace

void gemm_bf16(const GemmArgs& g0, hipStream_t st) {
  if (g0.M <= 0 || g0.N <= 0) return;
  GemmArgs g = g0;
  if (g.kchunk <= 0 || g.kchunk >= g.K) g.kchunk = g.K > 0 ? g.K : 1;
  const int32_t nz = (int32_t)((g.K + g.kchunk - 1) / g.kchunk);
  if (nz > 1 && (g.bias || g.C16 || g.kchunk % TK != 0))
    throw Error(ESR_ERR_VALUE, "split-K GEMM takes fp32 output only and a K chunk multiple of 32");
  g.mode_a = stage_mode(g.A, g.sa_i, g.sa_k);
  g.mode_b = stage_mode(g.B, g.sb_j, g.sb_k);
  dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, nz > 0 ? nz : 1);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, g);
  hip_check(hipGetLastError(), "k_gemm_bf16");
}

void sum_splits(const float* part, int32_t ns, int64_t n, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, part, ns, n, out);
  hip_check(hipGetLastError(), "k_sum_splits");
}

int32_t colsum_splits(int64_t rows) {
  const int64_t s = (rows + 255) / 256;
  return (int32_t)(s < 1 ? 1 : (s > 256 ? 256 : s));
}

void colsum(const float* X, int64_t rows, int32_t n, int64_t ld, float* part, float* out, hipStream_t st) {
  const int32_t ns = colsum_splits(rows);
  const int64_t chunk = (rows + ns - 1) / ns;
  hipLaunchKernelGGL(k_colsum_part, dim3((n + 63) / 64, ns), dim3(256), 0, st, X, rows, n, ld, chunk, part);
  hip_check(hipGetLastError(), "k_colsum_part");
  sum_splits(part, ns, n, out, st);
}

int32_t mse_blocks(int64_t total) {
  int64_t b = (total + 255) / 256;
  return (int32_t)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

void mse(const float* Y, const float* T, int64_t total, float scale, float* dY32, uint16_t* dY16, double* part,
         hipStream_t st) {
  hipLaunchKernelGGL(k_mse, dim3(mse_blocks(total)), dim3(256), 0, st, Y, T, total, scale, dY32, dY16, part);
  hip_check(hipGetLastError(), "k_mse");
}

void relu_bwd(const float* Z, const float* dH, int64_t total, float* dZ32, uint16_t* dZ16, hipStream_t st) {
  hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, Z, dH, total, dZ32, dZ16);
  hip_check(hipGetLastError(), "k_relu_bwd");
}

void adamw(float* w, const float* g, float* m, float* v, uint16_t* w16, int64_t n, float lr, float b1, float b2,
           float eps, float wd, float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(k_adamw, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, g, m, v, w16, n, lr, b1, b2,
                     eps, wd, c1, c2);
  hip_check(hipGetLastError(), "k_adamw");
}

void to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y, n);
  hip_check(hipGetLastError(), "k_to_bf16");
}

void warm_features(int32_t n_hr, int32_t n_lr, int32_t k, const int32_t* idx, const float* rel, const float* w,
                   const float* wall, int32_t n_cases, const float* lrW, const float* scal, float4 mu, float4 isig,
                   float gm1, int32_t in_pad, uint16_t* X, float* idw, hipStream_t st) {
  const int64_t rows = (int64_t)n_hr * n_cases;
  const dim3 grid((unsigned)((rows + 255) / 256));
  if (k == 6)
    hipLaunchKernelGGL(k_warm_features<6>, grid, dim3(256), 0, st, n_hr, n_lr, k, idx, rel, w, wall, n_cases, lrW,
                       scal, mu, isig, gm1, in_pad, X, idw);
  else
    hipLaunchKernelGGL(k_warm_features<0>, grid, dim3(256), 0, st, n_hr, n_lr, k, idx, rel, w, wall, n_cases, lrW,
                       scal, mu, isig, gm1, in_pad, X, idw);
  hip_check(hipGetLastError(), "k_warm_features");
}

void warm_apply(int64_t rows, int32_t n_hr, const float* idw, const float* Y, float4 sig, float gm1, float* Wout,
                int32_t* bad, hipStream_t st) {
  hipLaunchKernelGGL(k_warm_apply, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, rows, n_hr, idw, Y, sig,
                     gm1, Wout, bad);
  hip_check(hipGetLastError(), "k_warm_apply");
}

}  // namespace esr
