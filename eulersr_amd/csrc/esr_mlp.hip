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

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[TM][TKP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][TKP];  // Bs[j][k]
  const int bm = blockIdx.y * TM, bn = blockIdx.x * TN;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.K; k0 += TK) {
    for (int e = tid; e < TM * TK; e += 256) {
      const int i = e / TK, k = e % TK;
      const int gi = bm + i, gk = k0 + k;
      As[i][k] = (gi < g.M && gk < g.K) ? g.A[(int64_t)gi * g.sa_i + (int64_t)gk * g.sa_k] : (uint16_t)0;
    }
    for (int e = tid; e < TN * TK; e += 256) {
      const int j = e / TK, k = e % TK;
      const int gj = bn + j, gk = k0 + k;
      Bs[j][k] = (gj < g.N && gk < g.K) ? g.B[(int64_t)gk * g.sb_k + (int64_t)gj * g.sb_j] : (uint16_t)0;
    }
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
          if (g.C32) g.C32[(int64_t)i * g.ldc32 + j] = v;
          if (g.C16) g.C16[(int64_t)i * g.ldc16 + j] = f2bf(g.relu ? fmaxf(v, 0.f) : v);
        }
      }
}

// column sums of an fp32 (rows x n) matrix, one thread per column, fixed order
__global__ void k_colsum(const float* __restrict__ X, int64_t rows, int32_t n, int64_t ld, float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s = 0.f;
  for (int64_t i = 0; i < rows; ++i) s += X[i * ld + j];
  out[j] = s;
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
__global__ void k_warm_features(int32_t n_hr, int32_t n_lr, int32_t k, const int32_t* __restrict__ idx,
                                const float* __restrict__ rel, const float* __restrict__ w,
                                const float* __restrict__ wall, int32_t n_cases, const float* __restrict__ lrW,
                                const float* __restrict__ scal, float4 mu, float4 isig, float gm1, int32_t in_pad,
                                uint16_t* __restrict__ X, float* __restrict__ idw) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= (int64_t)n_hr * n_cases) return;
  const int32_t c = (int32_t)(row / n_hr), i = (int32_t)(row - (int64_t)c * n_hr);
  uint16_t* x = X + row * in_pad;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  int col = 0;
  for (int j = 0; j < k; ++j) {
    const int32_t l = idx[(int64_t)i * k + j];
    const float4 W = reinterpret_cast<const float4*>(lrW)[(int64_t)c * n_lr + l];
    const float r = W.x, u = W.y / r, v = W.z / r;
    const float P = (W.w - 0.5f * r * (u * u + v * v)) * gm1;
    const float wj = w[(int64_t)i * k + j];
    acc.x += wj * r;
    acc.y += wj * u;
    acc.z += wj * v;
    acc.w += wj * P;
    x[col++] = f2bf((r - mu.x) * isig.x);
    x[col++] = f2bf((u - mu.y) * isig.y);
    x[col++] = f2bf((v - mu.z) * isig.z);
    x[col++] = f2bf((P - mu.w) * isig.w);
    x[col++] = f2bf(rel[((int64_t)i * k + j) * 2]);
    x[col++] = f2bf(rel[((int64_t)i * k + j) * 2 + 1]);
    x[col++] = f2bf(wj);
  }
  for (int q = 0; q < 4; ++q) x[col++] = f2bf(wall[(int64_t)i * 4 + q]);
  x[col++] = f2bf(scal[2 * c]);
  x[col++] = f2bf(scal[2 * c + 1]);
  for (; col < in_pad; ++col) x[col] = 0;
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

void gemm_bf16(const GemmArgs& g, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return;
  dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, g);
  hip_check(hipGetLastError(), "k_gemm_bf16");
}

void colsum(const float* X, int64_t rows, int32_t n, int64_t ld, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3((n + 63) / 64), dim3(64), 0, st, X, rows, n, ld, out);
  hip_check(hipGetLastError(), "k_colsum");
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
  hipLaunchKernelGGL(k_warm_features, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, n_hr, n_lr, k, idx, rel,
                     w, wall, n_cases, lrW, scal, mu, isig, gm1, in_pad, X, idw);
  hip_check(hipGetLastError(), "k_warm_features");
}

void warm_apply(int64_t rows, int32_t n_hr, const float* idw, const float* Y, float4 sig, float gm1, float* Wout,
                int32_t* bad, hipStream_t st) {
  hipLaunchKernelGGL(k_warm_apply, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, rows, n_hr, idw, Y, sig,
                     gm1, Wout, bad);
  hip_check(hipGetLastError(), "k_warm_apply");
}

}  // namespace esr
