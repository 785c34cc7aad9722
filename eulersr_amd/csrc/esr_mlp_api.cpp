// C ABI of the warm-start network and feature kernels (include/eulersr.h).
//
// esr_mlp holds a multilayer perceptron (ReLU hidden layers, linear output)
// with float32 master parameters, their bfloat16 copies used by the MFMA
// GEMMs, AdamW moments and the activation workspace for up to max_rows rows.
// Parameter layout (host and device): for each layer l, W_l (out x in,
// row-major) then b_l (out).
#include <cmath>
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "esr_mlp.h"

using namespace esr;

struct esr_mlp {
  int32_t L = 0;
  std::vector<int32_t> w;           // widths, w[0] = input, w[L] = output
  int64_t max_rows = 0;
  std::vector<int64_t> woff, boff;  // parameter offsets
  int64_t nparam = 0;
  int32_t maxw = 0;
  DevBuf P, G, M1, M2, P16;
  std::vector<std::unique_ptr<DevBuf>> Z, H;  // per layer pre-activation (fp32), hidden activation (bf16)
  DevBuf dA32, dA16, dB32, part;
  DevBuf split;  // split-K partials of the weight gradients, kMaxSplits x max(out*in)
  DevBuf cpart;  // column-sum partials of the bias gradients
};

namespace {

template <class F>
esr_status guard(F&& f) {
  try {
    f();
    return ESR_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return ESR_ERR_MEMORY;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return ESR_ERR_DEVICE;
  }
}

hipStream_t S(void* p) { return reinterpret_cast<hipStream_t>(p); }

void check_rows(const esr_mlp* m, int64_t rows) {
  if (rows <= 0 || rows > m->max_rows)
    throw Error(ESR_ERR_VALUE, "rows must be in [1, max_rows] (" + std::to_string(m->max_rows) + ")");
}

// forward pass over `rows` rows of X (bf16, rows x w[0]); output in Z[L-1],
// or in Y (rows x w[L], fp32) when given. Training keeps every layer's fp32
// pre-activation Z[l] for the backward pass; inference (Y given) writes only
// the hidden layers' bf16 activations, which the next layer reads.
void forward(esr_mlp* m, int64_t rows, const uint16_t* X, hipStream_t st, float* Y = nullptr) {
  const uint16_t* in = X;
  for (int l = 0; l < m->L; ++l) {
    GemmArgs g{};
    g.M = (int32_t)rows;
    g.N = m->w[l + 1];
    g.K = m->w[l];
    g.A = in;
    g.sa_i = m->w[l];
    g.sa_k = 1;
    g.B = m->P16.as<uint16_t>() + m->woff[l];
    g.sb_k = 1;
    g.sb_j = m->w[l];
    g.bias = m->P.as<float>() + m->boff[l];
    const bool hidden = l < m->L - 1;
    g.C32 = Y ? (hidden ? nullptr : Y) : m->Z[l]->as<float>();
    g.ldc32 = m->w[l + 1];
    g.C16 = hidden ? m->H[l]->as<uint16_t>() : nullptr;
    g.ldc16 = m->w[l + 1];
    g.relu = hidden ? 1 : 0;
    gemm_bf16(g, st);
    if (hidden) in = m->H[l]->as<uint16_t>();
  }
}

// forward + MSE + backward; gradients into G, loss returned
double loss_grad(esr_mlp* m, int64_t rows, const uint16_t* X, const float* T, hipStream_t st) {
  forward(m, rows, X, st);
  const int32_t nout = m->w[m->L];
  const int64_t total = rows * nout;
  const int32_t nb = mse_blocks(total);
  mse(m->Z[m->L - 1]->as<float>(), T, total, 2.0f / (float)total, m->dA32.as<float>(), m->dA16.as<uint16_t>(),
      m->part.as<double>(), st);
  for (int l = m->L - 1; l >= 0; --l) {
    const uint16_t* in = (l == 0) ? X : m->H[l - 1]->as<uint16_t>();
    const int32_t no = m->w[l + 1], ni = m->w[l];
    // weight gradient dZ^T in   (no x ni), K = rows, split over K so the
    // few output tiles still fill the chip; partials summed in fixed order
    const int64_t tiles = (int64_t)((no + 63) / 64) * ((ni + 63) / 64);
    int64_t ns = std::min<int64_t>({(int64_t)kMaxSplits, (512 + tiles - 1) / tiles, (rows + 255) / 256});
    if (ns < 1) ns = 1;
    int64_t kchunk = (rows + ns - 1) / ns;
    kchunk = (kchunk + 31) / 32 * 32;
    ns = (rows + kchunk - 1) / kchunk;
    GemmArgs g{};
    g.M = no;
    g.N = ni;
    g.K = (int32_t)rows;
    g.A = m->dA16.as<uint16_t>();
    g.sa_i = 1;
    g.sa_k = no;
    g.B = in;
    g.sb_k = ni;
    g.sb_j = 1;
    g.ldc32 = ni;
    if (ns > 1) {
      g.C32 = m->split.as<float>();
      g.kchunk = kchunk;
      g.c_split = (int64_t)no * ni;
      gemm_bf16(g, st);
      sum_splits(m->split.as<float>(), (int32_t)ns, (int64_t)no * ni, m->G.as<float>() + m->woff[l], st);
    } else {
      g.C32 = m->G.as<float>() + m->woff[l];
      gemm_bf16(g, st);
    }
    colsum(m->dA32.as<float>(), rows, no, no, m->cpart.as<float>(), m->G.as<float>() + m->boff[l], st);
    if (l > 0) {
      // data gradient dZ W   (rows x ni), K = no
      GemmArgs d{};
      d.M = (int32_t)rows;
      d.N = ni;
      d.K = no;
      d.A = m->dA16.as<uint16_t>();
      d.sa_i = no;
      d.sa_k = 1;
      d.B = m->P16.as<uint16_t>() + m->woff[l];
      d.sb_k = ni;
      d.sb_j = 1;
      d.C32 = m->dB32.as<float>();
      d.ldc32 = ni;
      gemm_bf16(d, st);
      relu_bwd(m->Z[l - 1]->as<float>(), m->dB32.as<float>(), rows * ni, m->dA32.as<float>(),
               m->dA16.as<uint16_t>(), st);
    }
  }
  std::vector<double> part(nb);
  hip_check(hipMemcpyAsync(part.data(), m->part.p, nb * sizeof(double), hipMemcpyDeviceToHost, st), "loss D2H");
  hip_check(hipStreamSynchronize(st), "loss sync");
  double s = 0.0;
  for (double v : part) s += v;
  return s / (double)total;
}

}  // namespace

extern "C" {

esr_status esr_mlp_create(int32_t n_layers, const int32_t* widths, int64_t max_rows, esr_mlp** out) {
  return guard([&] {
    if (!out || !widths || n_layers < 1 || max_rows < 1) throw Error(ESR_ERR_VALUE, "bad MLP arguments");
    std::unique_ptr<esr_mlp> m(new esr_mlp());
    m->L = n_layers;
    m->w.assign(widths, widths + n_layers + 1);
    for (int32_t v : m->w)
      if (v < 1 || v > 4096) throw Error(ESR_ERR_VALUE, "layer widths must be in [1, 4096]");
    m->max_rows = max_rows;
    int64_t off = 0;
    for (int l = 0; l < n_layers; ++l) {
      m->woff.push_back(off);
      off += (int64_t)m->w[l + 1] * m->w[l];
      m->boff.push_back(off);
      off += m->w[l + 1];
    }
    m->nparam = off;
    for (int32_t v : m->w) m->maxw = std::max(m->maxw, v);
    m->P.alloc(off * sizeof(float));
    m->G.alloc(off * sizeof(float));
    m->M1.alloc(off * sizeof(float));
    m->M2.alloc(off * sizeof(float));
    m->P16.alloc(off * sizeof(uint16_t));
    for (int l = 0; l < n_layers; ++l) {
      auto z = std::make_unique<DevBuf>();
      z->alloc((size_t)max_rows * m->w[l + 1] * sizeof(float));
      m->Z.push_back(std::move(z));
      auto h = std::make_unique<DevBuf>();
      h->alloc((size_t)max_rows * m->w[l + 1] * sizeof(uint16_t));
      m->H.push_back(std::move(h));
    }
    m->dA32.alloc((size_t)max_rows * m->maxw * sizeof(float));
    m->dA16.alloc((size_t)max_rows * m->maxw * sizeof(uint16_t));
    m->dB32.alloc((size_t)max_rows * m->maxw * sizeof(float));
    m->part.alloc(1024 * sizeof(double));
    int64_t maxprod = 1;
    for (int l = 0; l < n_layers; ++l) maxprod = std::max<int64_t>(maxprod, (int64_t)m->w[l] * m->w[l + 1]);
    m->split.alloc((size_t)kMaxSplits * maxprod * sizeof(float));
    m->cpart.alloc((size_t)256 * m->maxw * sizeof(float));
    hip_check(hipMemset(m->P.p, 0, off * sizeof(float)), "memset");
    hip_check(hipMemset(m->P16.p, 0, off * sizeof(uint16_t)), "memset");
    hip_check(hipMemset(m->M1.p, 0, off * sizeof(float)), "memset");
    hip_check(hipMemset(m->M2.p, 0, off * sizeof(float)), "memset");
    *out = m.release();
  });
}

esr_status esr_mlp_destroy(esr_mlp* m) {
  return guard([&] { delete m; });
}

esr_status esr_mlp_param_count(const esr_mlp* m, int64_t* n) {
  return guard([&] {
    if (!m || !n) throw Error(ESR_ERR_VALUE, "null argument");
    *n = m->nparam;
  });
}

esr_status esr_mlp_set_params(esr_mlp* m, const float* host) {
  return guard([&] {
    if (!m || !host) throw Error(ESR_ERR_VALUE, "null argument");
    hip_check(hipMemcpy(m->P.p, host, m->nparam * sizeof(float), hipMemcpyHostToDevice), "params H2D");
    to_bf16(m->P.as<float>(), m->P16.as<uint16_t>(), m->nparam, nullptr);
    hip_check(hipMemset(m->M1.p, 0, m->nparam * sizeof(float)), "memset");
    hip_check(hipMemset(m->M2.p, 0, m->nparam * sizeof(float)), "memset");
    hip_check(hipDeviceSynchronize(), "params sync");
  });
}

esr_status esr_mlp_get_params(const esr_mlp* m, float* host) {
  return guard([&] {
    if (!m || !host) throw Error(ESR_ERR_VALUE, "null argument");
    hip_check(hipDeviceSynchronize(), "params sync");
    hip_check(hipMemcpy(host, m->P.p, m->nparam * sizeof(float), hipMemcpyDeviceToHost), "params D2H");
  });
}

esr_status esr_mlp_forward(esr_mlp* m, int64_t rows, const uint16_t* X, float* Y, void* stream) {
  return guard([&] {
    if (!m || !X || !Y) throw Error(ESR_ERR_VALUE, "null argument");
    hipStream_t st = S(stream);
    const int32_t nin = m->w[0], nout = m->w[m->L];
    for (int64_t r0 = 0; r0 < rows; r0 += m->max_rows) {
      const int64_t r = std::min(m->max_rows, rows - r0);
      forward(m, r, X + r0 * nin, st, Y + r0 * nout);
    }
  });
}

esr_status esr_mlp_loss_grad(esr_mlp* m, int64_t rows, const uint16_t* X, const float* T, double* loss,
                             float* grads_host, void* stream) {
  return guard([&] {
    if (!m || !X || !T) throw Error(ESR_ERR_VALUE, "null argument");
    check_rows(m, rows);
    hipStream_t st = S(stream);
    const double l = loss_grad(m, rows, X, T, st);
    if (loss) *loss = l;
    if (grads_host)
      hip_check(hipMemcpy(grads_host, m->G.p, m->nparam * sizeof(float), hipMemcpyDeviceToHost), "grads D2H");
  });
}

esr_status esr_mlp_adamw(esr_mlp* m, float lr, float beta1, float beta2, float eps, float weight_decay, int32_t t,
                         void* stream) {
  return guard([&] {
    if (!m || t < 1) throw Error(ESR_ERR_VALUE, "bad AdamW arguments");
    hipStream_t st = S(stream);
    const float c1 = 1.0f - std::pow(beta1, (float)t), c2 = 1.0f - std::pow(beta2, (float)t);
    for (int l = 0; l < m->L; ++l) {
      const int64_t nw = (int64_t)m->w[l + 1] * m->w[l];
      adamw(m->P.as<float>() + m->woff[l], m->G.as<float>() + m->woff[l], m->M1.as<float>() + m->woff[l],
            m->M2.as<float>() + m->woff[l], m->P16.as<uint16_t>() + m->woff[l], nw, lr, beta1, beta2, eps,
            weight_decay, c1, c2, st);
      adamw(m->P.as<float>() + m->boff[l], m->G.as<float>() + m->boff[l], m->M1.as<float>() + m->boff[l],
            m->M2.as<float>() + m->boff[l], nullptr, m->w[l + 1], lr, beta1, beta2, eps, 0.0f, c1, c2, st);
    }
  });
}

esr_status esr_warm_features(const esr_knn_map* map, int32_t n_cases, const float* lr_states, const float* case_scal,
                             const float* mu, const float* sig, double gamma, int32_t in_pad, uint16_t* X,
                             float* idw, void* stream) {
  return guard([&] {
    if (!map || !lr_states || !case_scal || !mu || !sig || !X || !idw || n_cases < 1)
      throw Error(ESR_ERR_VALUE, "null argument");
    if (in_pad < 7 * map->k + 6) throw Error(ESR_ERR_VALUE, "in_pad smaller than the feature count");
    const float4 m4 = make_float4(mu[0], mu[1], mu[2], mu[3]);
    const float4 is = make_float4(1.0f / sig[0], 1.0f / sig[1], 1.0f / sig[2], 1.0f / sig[3]);
    warm_features(map->n_hr, map->n_lr, map->k, map->knn_idx, map->knn_rel, map->knn_w, map->wall, n_cases,
                  lr_states, case_scal, m4, is, (float)(gamma - 1.0), in_pad, X, idw, S(stream));
  });
}

esr_status esr_warm_apply(int64_t rows, int32_t n_hr, const float* idw, const float* delta, const float* sig,
                          double gamma, float* W_out, int32_t* bad, void* stream) {
  return guard([&] {
    if (!idw || !W_out || !bad || !sig || rows < 1 || n_hr < 1) throw Error(ESR_ERR_VALUE, "null argument");
    warm_apply(rows, n_hr, idw, delta, make_float4(sig[0], sig[1], sig[2], sig[3]), (float)(gamma - 1.0), W_out,
               bad, S(stream));
  });
}

}  // extern "C"
