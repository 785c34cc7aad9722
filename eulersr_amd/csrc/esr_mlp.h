// Launch interface of the warm-start / MLP kernels (esr_mlp.hip). Internal.
#pragma once

#include "esr_internal.h"

namespace esr {

struct GemmArgs {
  int32_t M, N, K;
  const uint16_t* A;
  int64_t sa_i, sa_k;
  const uint16_t* B;
  int64_t sb_k, sb_j;
  const float* bias;  // N entries or null
  float* C32;         // fp32 output (row stride ldc32) or null
  int64_t ldc32;
  uint16_t* C16;      // bf16 output (row stride ldc16) or null
  int64_t ldc16;
  int32_t relu;       // ReLU on the bf16 output
  int64_t kchunk;     // split-K chunk (0: no split); partial z at C32 + z*c_split
  int64_t c_split;
  int32_t mode_a, mode_b;  // operand staging modes (set by gemm_bf16)
};

constexpr int32_t kMaxSplits = 64;

void gemm_bf16(const GemmArgs& g, hipStream_t st);
void sum_splits(const float* part, int32_t ns, int64_t n, float* out, hipStream_t st);
int32_t colsum_splits(int64_t rows);
void colsum(const float* X, int64_t rows, int32_t n, int64_t ld, float* part, float* out, hipStream_t st);
int32_t mse_blocks(int64_t total);
void mse(const float* Y, const float* T, int64_t total, float scale, float* dY32, uint16_t* dY16, double* part,
         hipStream_t st);
void relu_bwd(const float* Z, const float* dH, int64_t total, float* dZ32, uint16_t* dZ16, hipStream_t st);
void adamw(float* w, const float* g, float* m, float* v, uint16_t* w16, int64_t n, float lr, float b1, float b2,
           float eps, float wd, float c1, float c2, hipStream_t st);
void to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t st);
void warm_features(int32_t n_hr, int32_t n_lr, int32_t k, const int32_t* idx, const float* rel, const float* w,
                   const float* wall, int32_t n_cases, const float* lrW, const float* scal, float4 mu, float4 isig,
                   float gm1, int32_t in_pad, uint16_t* X, float* idw, hipStream_t st);
void warm_apply(int64_t rows, int32_t n_hr, const float* idw, const float* Y, float4 sig, float gm1, float* Wout,
                int32_t* bad, hipStream_t st);

}  // namespace esr
