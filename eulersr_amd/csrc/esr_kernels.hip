// gfx950 kernels of the explicit 2D unstructured finite-volume Euler step.
//
// One work-item per (cell, case). Cells are in reverse Cuthill-McKee order;
// the state of a group of C cases that share a mesh is stored
// case-innermost (element (cell, c) at cell*C + c), so neighbour gathers of
// one wave hit contiguous memory when C >= 64 and geometry loads are
// wave-uniform. Wave64 everywhere; no float atomics (the only atomic is an
// integer max on non-negative float bits, which is order independent).
//
// Per stage of a Runge-Kutta step:
//   k_recon  (MUSCL only, per cell): primitives, boundary ghosts, weighted
//            LSQ gradient, Venkatakrishnan limiter and the three limited face
//            states of each cell (helper.py:136-153, 233-256;
//            euler_solver.py:36-122).
//   k_face   (per logical face): numerical flux times face length, once per
//            face in the owner's orientation. The reference evaluates every
//            interior face once from each side (euler_solver.py:107-109);
//            the partner's flux is the exact negation here, which equals the
//            reference up to rounding (F(L,R,n) = -F(R,L,-n)).
//   k_update (per cell): deterministic gather of the three face fluxes
//            (no atomics), division by the area, sponge source and the
//            Runge-Kutta stage update (euler_solver.py:141-143, 602-667;
//            helper.py:262-295).
// Reductions (FP64 norms, FP64 CD/CL wall sums, dt minimum) use fixed-order
// wave/LDS trees and per-range partials, so every run is bit-reproducible.
#include <hip/hip_runtime.h>

#include <cmath>

#include "esr_launch.h"

namespace esr {

// ---------------------------------------------------------------------------
// arithmetic helpers
// ---------------------------------------------------------------------------

// a / b through the hardware reciprocal (1 ulp) instead of the IEEE division
// sequence: the kernels are issue-bound and the reference tolerance is far
// above the difference.
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float frcp(float b) { return __builtin_amdgcn_rcpf(b); }

// Correctly rounded square root, the value of sqrtf: the hardware v_sqrt_f32
// result s is within one ulp, and the sign of the residual x - s' s of each
// neighbour s' = s -/+ 1 ulp (one fma each) tells whether the neighbour is the
// rounded root. sqrtf's generic lowering adds a rescaling of inputs below
// 2^-96 (to keep denormals exact) and zero/infinity fix-ups around the same
// test; inputs here (c^2 = gamma P / rho, Roe densities clamped to >= 1e-6,
// rho / P) are normal, and +0, +inf and NaN come out right without them.
__device__ __forceinline__ float fsqrt(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sd = __uint_as_float(__float_as_uint(s) - 1u);
  const float su = __uint_as_float(__float_as_uint(s) + 1u);
  const float rd = __builtin_fmaf(-sd, s, x);
  const float ru = __builtin_fmaf(-su, s, x);
  const float r = (rd <= 0.0f) ? sd : s;
  return (ru > 0.0f) ? su : r;
}

// Natural logarithm, the value of logf for inputs of at least 2^-96: log2 x
// from the hardware v_log_f32, times ln 2 split as LN2_HI + LN2_LO (LN2_HI =
// ln 2 to 24 bits rounded toward zero, LN2_LO the float nearest the rest), the
// product y ln 2 formed with its rounding error (an fma) and rounded once.
// Zero and infinities pass through as log2 gives them (-inf, +inf), NaN stays
// NaN; logf's rescaling of inputs below 2^-96 is left out (the inputs here are
// pressures, densities and their ratios).
__device__ __forceinline__ float flog(float x) {
  constexpr float LN2_HI = 0x1.62e42ep-1f, LN2_LO = 0x1.efa39ep-25f;
  const float y = __builtin_amdgcn_logf(x);
  const float r = y * LN2_HI;
  float e = __builtin_fmaf(y, LN2_HI, -r);
  e = __builtin_fmaf(y, LN2_LO, e);
  const float l = __builtin_fmaf(y, LN2_HI, e);
  return (fabsf(y) < __builtin_inff()) ? l : y;
}

// ---------------------------------------------------------------------------
// state conversions (helper.py:25-46)
// ---------------------------------------------------------------------------

struct Prim {
  float r, u, v, P;
};

// getPrimitive: P = (E - 0.5 rho (u^2+v^2)) * M^2 * (gamma-1)
__device__ __forceinline__ Prim prim_of(const float4 W, float M2, float gm1) {
  Prim q;
  const float ir = frcp(W.x);
  q.r = W.x;
  q.u = W.y * ir;
  q.v = W.z * ir;
  q.P = (W.w - 0.5f * q.r * (q.u * q.u + q.v * q.v)) * M2 * gm1;
  return q;
}

// getConserved: E = P / ((gamma-1) M^2) + 0.5 rho (u^2+v^2)
__device__ __forceinline__ float4 cons_of(const Prim q, float inv_gm1M2) {
  return make_float4(q.r, q.r * q.u, q.r * q.v, q.P * inv_gm1M2 + 0.5f * q.r * (q.u * q.u + q.v * q.v));
}

__device__ __forceinline__ Prim as_prim(const float4 p) { return Prim{p.x, p.y, p.z, p.w}; }

// Boundary ghost state, helper.py:173-204 (walls and the subsonic inlet use
// the default gamma = 1.4 and M = 1 of getPrimitive/getConserved).
// Loads and stores at a 32-bit element offset from a wave-uniform base (a
// group's slab or a mesh array) in global memory (every buffer here comes
// from hipMalloc): the access is a global one (not flat, which a pointer read
// from a struct would otherwise get) addressed as a scalar base plus a
// 32-bit vector byte offset, with no 64-bit address arithmetic. Every slab of
// a group and every mesh array stays below 4 GiB (checked when groups are
// built).
typedef __attribute__((address_space(1))) const char gconst_char;
typedef __attribute__((address_space(1))) char gchar;
// the access goes through a builtin type (HIP's float4 / float2 are classes
// whose copy operations take generic references)
typedef float raw_float4 __attribute__((ext_vector_type(4)));
typedef float raw_float2 __attribute__((ext_vector_type(2)));
template <class T>
struct Raw {
  typedef T type;
  static __device__ __forceinline__ T of(const T v) { return v; }
  static __device__ __forceinline__ T to(const T v) { return v; }
};
template <>
struct Raw<float4> {
  typedef raw_float4 type;
  static __device__ __forceinline__ float4 of(const raw_float4 v) { return make_float4(v.x, v.y, v.z, v.w); }
  static __device__ __forceinline__ raw_float4 to(const float4 v) { return raw_float4{v.x, v.y, v.z, v.w}; }
};
template <>
struct Raw<float2> {
  typedef raw_float2 type;
  static __device__ __forceinline__ float2 of(const raw_float2 v) { return make_float2(v.x, v.y); }
  static __device__ __forceinline__ raw_float2 to(const float2 v) { return raw_float2{v.x, v.y}; }
};
template <class T>
__device__ __forceinline__ T ld(const T* base, uint32_t e) {
  typedef __attribute__((address_space(1))) const typename Raw<T>::type gR;
  return Raw<T>::of(*(gR*)((gconst_char*)(const char*)base + e * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ void st(T* base, uint32_t e, const T v) {
  typedef __attribute__((address_space(1))) typename Raw<T>::type gR;
  *(gR*)((gchar*)(char*)base + e * (uint32_t)sizeof(T)) = Raw<T>::to(v);
}

__device__ __forceinline__ float4 ghost_state(uint32_t kind, const float4 Wi, const float2 n, const float4 value,
                                              const float4 insub, const float4* __restrict__ Wg, uint32_t wrap_e) {
  if (kind == G_WALL) {
    const Prim p = prim_of(Wi, 1.0f, 0.4f);
    const float vn = p.u * n.x + p.v * n.y;
    const float bx = p.u - 2.0f * vn * n.x;
    const float by = p.v - 2.0f * vn * n.y;
    return cons_of(Prim{p.r, bx, by, p.P}, 2.5f);
  }
  if (kind == G_INLET) return value;
  if (kind == G_OUTLET) return Wi;
  if (kind == G_SUBSONIC) {
    const Prim p = prim_of(Wi, 1.0f, 0.4f);
    const float r = insub.x, u = insub.y, v = insub.z;
    return make_float4(r, r * u, r * v, p.P * 2.5f + 0.5f * r * (u * u + v * v));
  }
  return ld(Wg, wrap_e);
}

// ---------------------------------------------------------------------------
// work-item mapping
// ---------------------------------------------------------------------------

struct Item {
  const GroupDev* G;
  int32_t idx;  // cell (cell kernels) or logical face (face kernels)
  int32_t c;
  uint32_t e;   // element offset in the group's state slab (cell kernels)
  bool active;  // the case still runs (LAZY mapping: stores only when true)
};

// XCD-aware block order (a tuning choice): the hardware deals workgroup p to
// XCD p % 8, so the logical block (p % 8) * (grid / 8) + p / 8 gives every
// XCD (and its L2) a contiguous run of logical blocks -- in RCM order, a
// contiguous run of cells whose neighbours are mostly in the same run. Block
// tables hold a multiple of 8 entries (padding: group -1); `xcd` (a kernel
// argument) selects the order.
constexpr int32_t N_XCD = 8;
__device__ __forceinline__ int32_t logical_block(int xcd) {
  if (!xcd) return (int32_t)blockIdx.x;
  const int32_t per = gridDim.x / N_XCD;
  return (int32_t)(blockIdx.x % N_XCD) * per + (int32_t)(blockIdx.x / N_XCD);
}

// LAZY: a work-item of a stopped case is mapped too, with it.active = false,
// so the kernel issues its loads without first waiting for the case's flag
// and only skips its stores; otherwise such an item returns false at once.
template <bool FACES, bool LAZY = false>
__device__ __forceinline__ bool map_item(const GroupDev* __restrict__ groups, const BlockRef* __restrict__ blk_group,
                                         const CaseDev* __restrict__ cases, int xcd, Item& it) {
  const BlockRef e = blk_group[logical_block(xcd)];
  if (e.group < 0) return false;
  const GroupDev* G = groups + e.group;
  const int32_t C = G->C;
  // t / C by the group's magic number (in-group indices are 32-bit)
  const uint32_t t = (uint32_t)e.block * blockDim.x + threadIdx.x;
  uint32_t q = t;
  if (C > 1) {
    const uint32_t h = __umulhi(G->cmagic, t);
    q = (h + ((t - h) >> 1)) >> (G->cshift - 1);
  }
  const int32_t idx = (int32_t)q;
  const int32_t c = (int32_t)(t - q * (uint32_t)C);
  if (idx >= (FACES ? G->m.nf : G->m.n)) return false;
  if (LAZY) {
    it.active = cases[G->case0 + c].active != 0;
  } else {
    if (!cases[G->case0 + c].active) return false;
    it.active = true;
  }
  it.G = G;
  it.idx = idx;
  it.c = c;
  it.e = FACES ? 0u : t;
  return true;
}

// element offsets inside a group's slabs (state [n][C], reconstruction
// [3][n][C], face flux [nf][C]); 32-bit, checked when the groups are built
__device__ __forceinline__ uint32_t sel(const GroupDev* G, int32_t cell, int32_t c) {
  return (uint32_t)cell * (uint32_t)G->C + (uint32_t)c;
}
__device__ __forceinline__ uint32_t rel(const GroupDev* G, int k, int32_t cell, int32_t c) {
  return ((uint32_t)k * (uint32_t)G->m.n + (uint32_t)cell) * (uint32_t)G->C + (uint32_t)c;
}
__device__ __forceinline__ uint32_t fel(const GroupDev* G, int32_t face, int32_t c) {
  return (uint32_t)face * (uint32_t)G->C + (uint32_t)c;
}
// mesh [k][n] arrays
__device__ __forceinline__ uint32_t kn(int k, int32_t n, int32_t cell) {
  return (uint32_t)k * (uint32_t)n + (uint32_t)cell;
}

// area / sum_k (c + |u . n_k|) len_k of one cell from its state W: the time
// step of the cell at CFL 1 (helper.py:8-18 before the minimum over cells
// and the CFL factor). k_dt takes its minimum (the reference's global dt);
// local time stepping uses it per cell.
__device__ __forceinline__ float cell_dt1(const MeshDev& m, const Phys& ph, int32_t cell, const float4 W) {
  const Prim q = prim_of(W, ph.M2, ph.gm1);
  const float cel = fsqrt(fabsf(ph.g * q.P / q.r)) * ph.invM;
  float tk[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float2 nk = ld(m.nrm, kn(k, m.n, cell));
    const float un = q.u * nk.x + q.v * nk.y;
    tk[k] = (cel + fabsf(un)) * ld(m.len, kn(k, m.n, cell));
  }
  return ld(m.area, (uint32_t)cell) / ((tk[0] + tk[1]) + tk[2]);
}

// ---------------------------------------------------------------------------
// reconstruction: LSQ gradient + Venkatakrishnan + MUSCL face states
// ---------------------------------------------------------------------------

// venkatakrishnan(a, b) with K = 0 (euler_solver.py:36-39)
__device__ __forceinline__ float venkat(float a, float b) {
  return fdiv(a * a + 2.0f * a * b, a * a + 2.0f * (b * b) + a * b + 1e-7f);
}

__global__ __launch_bounds__(BLOCK) void k_recon(const GroupDev* __restrict__ groups,
                                                 const BlockRef* __restrict__ blk_group,
                                                 const CaseDev* __restrict__ cases,
                                                 const float4* __restrict__ Ws, float4* __restrict__ rec, int xcd) {
  Item it;
  if (!map_item<false, true>(groups, blk_group, cases, xcd, it)) return;
  const GroupDev* G = it.G;
  const MeshDev& m = G->m;
  const Phys& ph = G->ph;
  const int32_t n = m.n, cell = it.idx;
  const CaseDev& cs = cases[G->case0 + it.c];
  const float4* Wg = Ws + G->st_off;
  float4* rg = rec + G->rec_off;

  // Every load is issued before the arithmetic: the face records, the
  // neighbours' states (a boundary face reads the cell's own state, in
  // bounds, and replaces it by the ghost state below) and the LSQ weights.
  // Loaded face by face, each neighbour load would wait behind the previous
  // face's branch. (The case's inlet state, which only inlet ghosts read,
  // stays in their branch: loaded for every cell it cost 4 % at 512 cases.)
  const float4 Wi = ld(Wg, it.e);
  uint32_t f[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) f[k] = ld(m.fref, kn(k, n, cell));
  float4 Wn[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) Wn[k] = ld(Wg, sel(G, fref_is_bnd(f[k]) ? cell : (int32_t)(f[k] & FREF_IDX), it.c));
  // weighted least squares (helper.py:136-153): grad = A^-1 sum_k w_k dW_k dl_k
  const float2 w0 = ld(m.wdl, kn(0, n, cell)), w1 = ld(m.wdl, kn(1, n, cell)), w2 = ld(m.wdl, kn(2, n, cell));
  const float4 ai = ld(m.ainv, (uint32_t)cell);
  float2 dm[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) dm[k] = ld(m.dxm, kn(k, n, cell));
  const Prim pi = prim_of(Wi, ph.M2, ph.gm1);
  Prim pr[3];
  bool bnd[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float4 WR = Wn[k];
    bnd[k] = fref_is_bnd(f[k]);
    if (bnd[k]) {
      const uint32_t kind = (f[k] >> FREF_KIND_SHIFT) & 7u;
      const float2 nk = (kind == G_WALL) ? ld(m.nrm, kn(k, n, cell)) : make_float2(0.f, 0.f);
      WR = ghost_state(kind, Wi, nk, cs.value, m.inlet_subsonic, Wg, sel(G, (int32_t)(f[k] & FREF_BIDX), it.c));
    }
    pr[k] = prim_of(WR, ph.M2, ph.gm1);
  }

  const float qi[4] = {pi.r, pi.u, pi.v, pi.P};
  float D[3][4];
  float phi[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    float q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) q[k] = (v == 0) ? pr[k].r : (v == 1) ? pr[k].u : (v == 2) ? pr[k].v : pr[k].P;
    const float d0 = q[0] - qi[v], d1 = q[1] - qi[v], d2 = q[2] - qi[v];
    const float bx = (d0 * w0.x + d1 * w1.x) + d2 * w2.x;
    const float by = (d0 * w0.y + d1 * w1.y) + d2 * w2.y;
    const float gx = ai.x * bx + ai.y * by;
    const float gy = ai.z * bx + ai.w * by;
    const float mn = fminf(fminf(qi[v], q[0]), fminf(q[1], q[2]));
    const float mx = fmaxf(fmaxf(qi[v], q[0]), fmaxf(q[1], q[2]));
    // Venkatakrishnan limiter, min over the three faces (euler_solver.py:45-67)
    float p3[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float d = dm[k].x * gx + dm[k].y * gy;
      D[k][v] = d;
      // phi = venkat(max - q, d) for d > 1e-7, venkat(min - q, d) for
      // d < -1e-7, else 1 (branch-free: one evaluation per face)
      const float pv = venkat((d > 0.f ? mx : mn) - qi[v], d);
      p3[k] = (fabsf(d) > 1e-7f) ? pv : 1.0f;
    }
    phi[v] = fminf(fminf(p3[0], p3[1]), p3[2]);
  }

  if (!it.active) return;
  // limited face states (primitives); no slope on marker > 1 faces
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float4 o;
    o.x = bnd[k] ? qi[0] : qi[0] + phi[0] * D[k][0];
    o.y = bnd[k] ? qi[1] : qi[1] + phi[1] * D[k][1];
    o.z = bnd[k] ? qi[2] : qi[2] + phi[2] * D[k][2];
    o.w = bnd[k] ? qi[3] : qi[3] + phi[3] * D[k][3];
    st(rg, rel(G, k, cell, it.c), o);
  }
}

// ---------------------------------------------------------------------------
// numerical fluxes (euler_solver.py:145-591); each returns the total flux
// F = central - D of one face for the owner's normal.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void central_flux(const float4 WL, const Prim& L, const float4 WR, const Prim& R,
                                             float nx, float ny, float i2, float F[4]) {
  // getFlux central part (euler_solver.py:382-417)
  const float unL = L.u * nx + L.v * ny;
  const float unR = R.u * nx + R.v * ny;
  F[0] = 0.5f * (L.r * unL + R.r * unR);
  F[1] = 0.5f * ((L.r * L.u * unL + i2 * L.P * nx) + (R.r * R.u * unR + i2 * R.P * nx));
  F[2] = 0.5f * ((L.r * L.v * unL + i2 * L.P * ny) + (R.r * R.v * unR + i2 * R.P * ny));
  F[3] = 0.5f * ((WL.w + i2 * L.P) * unL + (WR.w + i2 * R.P) * unR);
}

// physical flux in the HLLC/AUSM form: P n / M^2
__device__ __forceinline__ void phys_flux(const float4 W, const Prim& q, float un, float nx, float ny, float invM2,
                                          float F[4]) {
  F[0] = q.r * un;
  F[1] = q.r * q.u * un + q.P * nx * invM2;
  F[2] = q.r * q.v * un + q.P * ny * invM2;
  F[3] = (W.w + q.P * invM2) * un;
}

__device__ __forceinline__ float sound(const Prim& q, const Phys& ph) {
  return fsqrt(fabsf(ph.g * q.P * frcp(q.r))) * ph.invM;
}

// HLLC flux F_HLLC (euler_solver.py:273-372). Every branch of the scheme
// takes the physical flux of one side K (left for SL >= 0 or the left star
// state, right otherwise), so only that side's physical flux is formed.
__device__ __forceinline__ void hllc_flux(const float4 WL, const Prim& L, const float4 WR, const Prim& R, float nx,
                                          float ny, const Phys& ph, float F[4]) {
  const float unL = L.u * nx + L.v * ny;
  const float unR = R.u * nx + R.v * ny;
  const float tx = -ny, ty = nx;
  const float utL = L.u * tx + L.v * ty;
  const float utR = R.u * tx + R.v * ty;
  const float aL = sound(L, ph);
  const float aR = sound(R, ph);
  float SL = fminf(unL - aL, unR - aR);
  float SR = fmaxf(unL + aL, unR + aR);
  SL = (SL == SR) ? SL - 1e-6f : SL;
  SR = (SL == SR) ? SR + 1e-6f : SR;
  float den = L.r * (SL - unL) - R.r * (SR - unR);
  den = (den == 0.0f) ? 1e-12f : den;
  const float Ss = fdiv(R.P - L.P + L.r * unL * (SL - unL) - R.r * unR * (SR - unR), den);
  // SL >= 0: F_L; left star state (SL <= 0 <= S*): F_L + SL (U*_L - U_L);
  // right star state (S* <= 0 <= SR): F_R + SR (U*_R - U_R); otherwise F_R
  const bool left = (SL <= 0.0f && Ss >= 0.0f);
  const bool takeL = (SL >= 0.0f) || left;
  const bool star = !(SL >= 0.0f) && (left || (Ss <= 0.0f && SR >= 0.0f));
  const Prim& K = takeL ? L : R;
  const float4 WK = takeL ? WL : WR;
  const float unK = takeL ? unL : unR;
  phys_flux(WK, K, unK, nx, ny, ph.invM2, F);
  if (star) {
    const float PsL = L.P + L.r * (SL - unL) * (Ss - unL);
    const float PsR = R.P + R.r * (SR - unR) * (Ss - unR);
    const float Ps = 0.5f * (PsL + PsR);
    const float S = left ? SL : SR;
    const float utK = left ? utL : utR;
    const float idn = frcp(S - Ss);
    const float rs = K.r * (S - unK) * idn;
    const float ux = Ss * nx + utK * tx, uy = Ss * ny + utK * ty;
    const float Es = ((S - unK) * WK.w - K.P * unK + Ps * Ss) * idn;
    F[0] = F[0] + S * (rs - WK.x);
    F[1] = F[1] + S * (rs * ux - WK.y);
    F[2] = F[2] + S * (rs * uy - WK.z);
    F[3] = F[3] + S * (Es - WK.w);
  }
}

// HLLC dissipation D = central - F_HLLC (euler_solver.py:373-374)
__device__ __forceinline__ void diss_hllc(const float4 WL, const Prim& L, const float4 WR, const Prim& R, float nx,
                                          float ny, const Phys& ph, float D[4]) {
  float F[4], FL[4], FR[4];
  hllc_flux(WL, L, WR, R, nx, ny, ph, F);
  phys_flux(WL, L, L.u * nx + L.v * ny, nx, ny, ph.invM2, FL);
  phys_flux(WR, R, R.u * nx + R.v * ny, nx, ny, ph.invM2, FR);
#pragma unroll
  for (int v = 0; v < 4; ++v) D[v] = 0.5f * (FL[v] + FR[v]) - F[v];
}

__device__ __forceinline__ float wave_speed(const Prim& q, float nx, float ny, const Phys& ph) {
  return sound(q, ph) + fabsf(q.u * nx + q.v * ny);
}

// entropy variables (helper.py:48-70)
__device__ __forceinline__ float4 entropy_vars(const float4 W, const Phys& ph) {
  const float r = W.x, ir = frcp(r), u = W.y * ir, v = W.z * ir;
  const float P = (W.w - 0.5f * r * (u * u + v * v)) * ph.gm1;
  const float iP = frcp(P);
  const float s = flog(P) - ph.g * flog(r);  // ln(P / rho^gamma)
  float4 V;
  V.x = (ph.g - s) * ph.inv_gm1 - 0.5f * r * (u * u + v * v) * iP;
  V.y = W.y * iP;
  V.z = W.z * iP;
  V.w = -r * iP;
  return V;
}

// d/dt getConserved_from_Entropy(eta + t deta) at t = 0 (the jax.jvp of
// euler_solver.py:234-238), chain rule over helper.py:72-86.
__device__ __forceinline__ void jvp_u_of_v(const float4 e, const float4 d, const Phys& ph, float out[4]) {
  const float ie = frcp(e.w);
  const float u = -e.y * ie;
  const float du = -(d.y * ie - e.y * d.w * (ie * ie));
  const float v = -e.z * ie;
  const float dv = -(d.z * ie - e.z * d.w * (ie * ie));
  const float q = 0.5f * (u * u + v * v);
  const float dq = u * du + v * dv;
  const float s = ph.g - ph.gm1 * (e.x - q * e.w);
  const float ds = -ph.gm1 * (d.x - (dq * e.w + q * d.w));
  const float a = powf(-e.w, ph.k_exp);
  const float da = ph.k_exp * powf(-e.w, ph.k_exp - 1.0f) * (-d.w);
  const float b = expf(s * ph.k_exp);
  const float db = b * ds * ph.k_exp;
  const float rho = a * b;
  const float drho = da * b + a * db;
  const float dP = -(drho * ie - rho * d.w * (ie * ie));
  const float kin = u * u + v * v;
  const float dkin = 2.0f * u * du + 2.0f * v * dv;
  out[0] = drho;
  out[1] = drho * u + rho * du;
  out[2] = drho * v + rho * dv;
  out[3] = dP * ph.inv_gm1 + 0.5f * (drho * kin + rho * dkin);
}

__device__ __forceinline__ float log_mean(float aL, float aR) {
  const float xi = fdiv(aL, aR);
  const float f = fdiv(xi - 1.0f, xi + 1.0f);
  const float u = f * f;
  const float F = (u < 1e-3f) ? 1.0f + u * (1.0f / 3.0f) + (u * u) * 0.2f + (u * (u * u)) * (1.0f / 7.0f)
                              : fdiv(flog(xi), 2.0f * f);
  return fdiv(aL + aR, 2.0f * F);
}

// Ismail-Roe entropy-conservative flux (euler_solver.py:532-588)
__device__ __forceinline__ void ismail_roe_flux(const float4 WL, const float4 WR, float nx, float ny, const Phys& ph,
                                                float F[4]) {
  float ZL[4], ZR[4];
  {
    const float r = WL.x, ir = frcp(r), u = WL.y * ir, v = WL.z * ir;
    const float P = (WL.w - 0.5f * r * (u * u + v * v)) * ph.gm1;
    const float s = fsqrt(fdiv(r, P));
    ZL[0] = s;
    ZL[1] = s * u;
    ZL[2] = s * v;
    ZL[3] = s * P;
  }
  {
    const float r = WR.x, ir = frcp(r), u = WR.y * ir, v = WR.z * ir;
    const float P = (WR.w - 0.5f * r * (u * u + v * v)) * ph.gm1;
    const float s = fsqrt(fdiv(r, P));
    ZR[0] = s;
    ZR[1] = s * u;
    ZR[2] = s * v;
    ZR[3] = s * P;
  }
  float Zb[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) Zb[v] = 0.5f * (ZL[v] + ZR[v]);
  const float Z1 = log_mean(ZL[0], ZR[0]);
  const float Z4 = log_mean(ZL[3], ZR[3]);
  const float iZ0 = frcp(Zb[0]);
  const float rh = Zb[0] * Z4;
  const float uh = Zb[1] * iZ0;
  const float vh = Zb[2] * iZ0;
  const float P1 = Zb[3] * iZ0;
  const float P2 = fdiv(ph.c_ir1 * Z4, Z1) + ph.c_ir2 * Zb[3] * iZ0;
  const float ah2 = fdiv(ph.g * P2, rh);
  const float Hh = ah2 * ph.inv_gm1 + (uh * uh + vh * vh) * 0.5f;
  const float un = uh * nx + vh * ny;
  F[0] = rh * un;
  F[1] = rh * uh * un + P1 * nx;
  F[2] = rh * vh * un + P1 * ny;
  F[3] = rh * un * Hh;
}

// Roe dissipation 0.5 R |Lambda S| R^T dv (euler_solver.py:145-199) from the
// face states' conserved W and primitive q forms.
__device__ __forceinline__ void diss_roe(const float4 WL, const Prim& L, const float4 WR, const Prim& R, float nx,
                                         float ny, const Phys& ph, float D[4]) {
  // Roe average (helper.py:101-127): pressure without the M^2 factor of the
  // primitives, P = q.P / M^2; rho <= 0 is clamped to 1e-6 (then u, v, P
  // come from the conserved state), P <= 0 to 1e-6
  const float i0 = frcp(L.r), i1 = frcp(R.r);
  float rL = L.r, uL = L.u, vL = L.v, PL = L.P * ph.invM2, irL = i0;
  float rR = R.r, uR = R.u, vR = R.v, PR = R.P * ph.invM2, irR = i1;
  if (!(rL > 0.0f)) {
    rL = (rL <= 0.0f) ? 1e-6f : rL;
    irL = frcp(rL);
    uL = WL.y * irL;
    vL = WL.z * irL;
    PL = (WL.w - 0.5f * rL * (uL * uL + vL * vL)) * ph.gm1;
  }
  if (!(rR > 0.0f)) {
    rR = (rR <= 0.0f) ? 1e-6f : rR;
    irR = frcp(rR);
    uR = WR.y * irR;
    vR = WR.z * irR;
    PR = (WR.w - 0.5f * rR * (uR * uR + vR * vR)) * ph.gm1;
  }
  PL = (PL <= 0.0f) ? 1e-6f : PL;
  PR = (PR <= 0.0f) ? 1e-6f : PR;
  const float sL = fsqrt(rL), sR = fsqrt(rR);
  const float isum = frcp(sL + sR);
  const float r = sL * sR;
  const float u = (sL * uL + sR * uR) * isum;
  const float v = (sL * vL + sR * vR) * isum;
  const float HL = (WL.w + PL) * irL;
  const float HR = (WR.w + PR) * irR;
  const float H = (sL * HL + sR * HR) * isum;
  const float a = fsqrt(fabsf(ph.gm1 * (H - 0.5f * (u * u + v * v))));
  // eigenvectors R (columns), euler_solver.py:156-175
  float Rm[4][4];
  Rm[0][0] = 1.0f;
  Rm[1][0] = u - a * nx;
  Rm[2][0] = v - a * ny;
  Rm[3][0] = H - a * (u * nx + v * ny);
  Rm[0][1] = 1.0f;
  Rm[1][1] = u;
  Rm[2][1] = v;
  Rm[3][1] = 0.5f * (u * u + v * v);
  Rm[0][2] = 0.0f;
  Rm[1][2] = -ny;
  Rm[2][2] = nx;
  Rm[3][2] = -u * ny + v * nx;
  Rm[0][3] = 1.0f;
  Rm[1][3] = u + a * nx;
  Rm[2][3] = v + a * ny;
  Rm[3][3] = H + a * (u * nx + v * ny);
  const float un = u * nx + v * ny;
  const float lam[4] = {un - a, un, un, un + a};
  const float S[4] = {r * ph.inv_2g, ph.gm1 * r * ph.inv_g, r * ph.inv_g, r * ph.inv_2g};
  // difference of the entropy variables V(W_R) - V(W_L) (helper.py:58-69)
  // of the unclamped states: V = [(g - s)/(g - 1) - rho |u|^2 / (2P),
  // rho u / P, rho v / P, -rho / P], s = ln(P / rho^g); the entropy jump
  // taken as ln(P_R/P_L) - g ln(rho_R/rho_L) (two logarithms, no power)
  float dv[4];
  {
    const float P0 = L.P * ph.invM2, P1 = R.P * ph.invM2;
    const float iP0 = frcp(P0), iP1 = frcp(P1);
    const float ds = flog(P1 * iP0) - ph.g * flog(R.r * i0);
    const float k0 = L.r * (L.u * L.u + L.v * L.v), k1 = R.r * (R.u * R.u + R.v * R.v);  // rho |u|^2
    dv[0] = -ds * ph.inv_gm1 - 0.5f * (k1 * iP1 - k0 * iP0);
    dv[1] = WR.y * iP1 - WL.y * iP0;
    dv[2] = WR.z * iP1 - WL.z * iP0;
    dv[3] = L.r * iP0 - R.r * iP1;
  }
  float DRH[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float RH = ((Rm[0][i] * dv[0] + Rm[1][i] * dv[1]) + Rm[2][i] * dv[2]) + Rm[3][i] * dv[3];
    DRH[i] = fabsf(lam[i] * S[i]) * RH;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
    D[i] = 0.5f * (((Rm[i][0] * DRH[0] + Rm[i][1] * DRH[1]) + Rm[i][2] * DRH[2]) + Rm[i][3] * DRH[3]);
}

__device__ __forceinline__ void ausm_split(float Mn, float& Mp, float& Mm, float& Pp, float& Pm) {
  const float absm = fabsf(Mn);
  const bool sub = absm < 1.0f;
  const float q = (Mn * Mn - 1.0f) * (Mn * Mn - 1.0f);
  const float beta = 0.125f, alpha = 0.1875f;
  Mp = sub ? 0.25f * ((Mn + 1.0f) * (Mn + 1.0f)) + beta * q : 0.5f * (Mn + absm);
  Mm = sub ? -0.25f * ((Mn - 1.0f) * (Mn - 1.0f)) - beta * q : 0.5f * (Mn - absm);
  const float sg = (Mn > 0.0f) ? 1.0f : ((Mn < 0.0f) ? -1.0f : 0.0f);
  Pp = sub ? 0.25f * ((Mn + 1.0f) * (Mn + 1.0f)) * (2.0f - Mn) + alpha * Mn * q : 0.5f * (1.0f + sg);
  Pm = sub ? 0.25f * ((Mn - 1.0f) * (Mn - 1.0f)) * (2.0f + Mn) - alpha * Mn * q : 0.5f * (1.0f - sg);
}

// AUSM+ flux F_AUSM (euler_solver.py:490-529)
__device__ __forceinline__ void ausm_flux(const float4 WL, const Prim& L, const float4 WR, const Prim& R, float nx,
                                          float ny, const Phys& ph, float FA[4]) {
  const float unL = L.u * nx + L.v * ny;
  const float unR = R.u * nx + R.v * ny;
  const float aL = sound(L, ph);
  const float aR = sound(R, ph);
  const float af = 0.5f * (aL + aR);
  const float MnL = fdiv(unL, aL + 1e-12f);
  const float MnR = fdiv(unR, aR + 1e-12f);
  float MpL, MmL_, PpL, PmL_, MpR_, MmR, PpR_, PmR;
  ausm_split(MnL, MpL, MmL_, PpL, PmL_);
  ausm_split(MnR, MpR_, MmR, PpR_, PmR);
  const float mL = af * MpL * L.r;
  const float mR = af * MmR * R.r;
  const float pf = PpL * L.P + PmR * R.P;
  const float HL = fdiv(WL.w + L.P * ph.invM2, L.r);
  const float HR = fdiv(WR.w + R.P * ph.invM2, R.r);
  FA[0] = mL + mR;
  FA[1] = mL * L.u + mR * R.u + pf * nx * ph.invM2;
  FA[2] = mL * L.v + mR * R.v + pf * ny * ph.invM2;
  FA[3] = mL * HL + mR * HR;
}

// AUSM+ dissipation D = central - F_AUSM (euler_solver.py:530)
__device__ __forceinline__ void diss_ausm(const float4 WL, const Prim& L, const float4 WR, const Prim& R, float nx,
                                          float ny, const Phys& ph, float D[4]) {
  float FA[4], FL[4], FR[4];
  ausm_flux(WL, L, WR, R, nx, ny, ph, FA);
  const float unL = L.u * nx + L.v * ny;
  const float unR = R.u * nx + R.v * ny;
  phys_flux(WL, L, unL, nx, ny, ph.invM2, FL);
  phys_flux(WR, R, unR, nx, ny, ph.invM2, FR);
#pragma unroll
  for (int v = 0; v < 4; ++v) D[v] = 0.5f * (FL[v] + FR[v]) - FA[v];
}

// Flux of one face from its left / right states in conserved (W) and
// primitive (q) form (the same state; the primitives are the reconstructed
// values themselves, not a conversion of W back).
template <int FLUX, int ENT>
__device__ __forceinline__ void face_flux(const float4 WL, const Prim& L, const float4 WR, const Prim& R, float nx,
                                          float ny, const Phys& ph, float lf_c, float F[4]) {
  float Fc[4], D[4];
  constexpr bool IR = (FLUX == ESR_FLUX_TADMOR) || (ENT != 0);
  // HLLC and AUSM+ define their dissipation as central - F_scheme, so the
  // total flux central - D is the scheme's own flux (to rounding)
  if (!IR && FLUX == ESR_FLUX_HLLC) {
    hllc_flux(WL, L, WR, R, nx, ny, ph, F);
    return;
  }
  if (!IR && FLUX == ESR_FLUX_AUSM) {
    ausm_flux(WL, L, WR, R, nx, ny, ph, F);
    return;
  }
  if (IR)
    ismail_roe_flux(WL, WR, nx, ny, ph, Fc);
  else
    central_flux(WL, L, WR, R, nx, ny, ph.invM2, Fc);

  if (FLUX == ESR_FLUX_HLLC) {
    diss_hllc(WL, L, WR, R, nx, ny, ph, D);
  } else if (FLUX == ESR_FLUX_ROE) {
    diss_roe(WL, L, WR, R, nx, ny, ph, D);
  } else if (FLUX == ESR_FLUX_AUSM) {
    diss_ausm(WL, L, WR, R, nx, ny, ph, D);
  } else if (FLUX == ESR_FLUX_LF) {
    const float h = 0.5f * lf_c;
    D[0] = h * (WR.x - WL.x);
    D[1] = h * (WR.y - WL.y);
    D[2] = h * (WR.z - WL.z);
    D[3] = h * (WR.w - WL.w);
  } else {  // Rusanov; entropy-stable variant for Tadmor / entropy=True
    const float Cm = fmaxf(wave_speed(R, nx, ny, ph), wave_speed(L, nx, ny, ph));
    const float h = 0.5f * Cm;
    if (FLUX == ESR_FLUX_TADMOR || ENT) {
      const float4 eL = entropy_vars(WL, ph);
      const float4 eR = entropy_vars(WR, ph);
      const float4 eb = make_float4(0.5f * (eL.x + eR.x), 0.5f * (eL.y + eR.y), 0.5f * (eL.z + eR.z),
                                    0.5f * (eL.w + eR.w));
      const float4 de = make_float4(eR.x - eL.x, eR.y - eL.y, eR.z - eL.z, eR.w - eL.w);
      float dU[4];
      jvp_u_of_v(eb, de, ph, dU);
#pragma unroll
      for (int v = 0; v < 4; ++v) D[v] = h * dU[v];
    } else {
      D[0] = h * (WR.x - WL.x);
      D[1] = h * (WR.y - WL.y);
      D[2] = h * (WR.z - WL.z);
      D[3] = h * (WR.w - WL.w);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) F[v] = Fc[v] - D[v];
}

// Conserved left (owner) and right (partner or ghost) states of a logical face.
// Both states are loaded before either is used (a boundary face reads the
// owner's element in place of a partner, in bounds, and replaces it by the
// ghost state), so the partner load does not wait behind the face's branch.
template <int RECON>
__device__ __forceinline__ void face_states(const GroupDev* G, int32_t face, int32_t c, const CaseDev& cs,
                                            const float4* __restrict__ Ws, const float4* __restrict__ rec,
                                            float4& WL, Prim& L, float4& WR, Prim& R) {
  const MeshDev& m = G->m;
  const Phys& ph = G->ph;
  const float4* Wsg = Ws + G->st_off;
  const float4* rg = rec + G->rec_off;
  const uint32_t own = ld(m.f_own, (uint32_t)face);
  const uint32_t f = ld(m.f_ref, (uint32_t)face);
  const int32_t i = (int32_t)(own & FREF_IDX);
  const int k = (int)(own >> 30);
  const bool bnd = fref_is_bnd(f);
  const int32_t j = bnd ? i : (int32_t)(f & FREF_IDX);
  float4 qL, qR;
  if (RECON == ESR_RECON_MUSCL) {
    qL = ld(rg, rel(G, k, i, c));
    qR = ld(rg, rel(G, bnd ? k : (int)(f >> 30), j, c));
  } else {
    qL = ld(Wsg, sel(G, i, c));
    qR = ld(Wsg, sel(G, j, c));
  }
  const float4 value = cs.value;
  if (bnd) {
    const uint32_t kind = (f >> FREF_KIND_SHIFT) & 7u;
    const float4 Wi = (RECON == ESR_RECON_MUSCL) ? ld(Wsg, sel(G, i, c)) : qL;
    const float2 nk = (kind == G_WALL) ? ld(m.f_nrm, (uint32_t)face) : make_float2(0.f, 0.f);
    const float4 Wgh = ghost_state(kind, Wi, nk, value, m.inlet_subsonic, Wsg, sel(G, (int32_t)(f & FREF_BIDX), c));
    if (RECON == ESR_RECON_MUSCL) {
      const Prim g = prim_of(Wgh, ph.M2, ph.gm1);
      qR = make_float4(g.r, g.u, g.v, g.P);
    } else {
      qR = Wgh;
    }
  }
  if (RECON == ESR_RECON_MUSCL) {
    L = as_prim(qL);
    R = as_prim(qR);
    WL = cons_of(L, ph.inv_gm1M2);
    WR = cons_of(R, ph.inv_gm1M2);
  } else {
    WL = qL;
    WR = qR;
    L = prim_of(WL, ph.M2, ph.gm1);
    R = prim_of(WR, ph.M2, ph.gm1);
  }
}

// ---------------------------------------------------------------------------
// per-face flux: F * len of each logical face, owner's orientation
// ---------------------------------------------------------------------------

// face_flux of a launch class: FLUX_HLLC_ROE picks HLLC or Roe by the
// group's flux (uniform over a block)
template <int FLUX, int ENT>
__device__ __forceinline__ void class_flux(int32_t gflux, const float4 WL, const Prim& L, const float4 WR,
                                           const Prim& R, float nx, float ny, const Phys& ph, float lf_c,
                                           float F[4]) {
  if (FLUX == FLUX_HLLC_ROE) {
    if (gflux == ESR_FLUX_ROE)
      face_flux<ESR_FLUX_ROE, ENT>(WL, L, WR, R, nx, ny, ph, lf_c, F);
    else
      face_flux<ESR_FLUX_HLLC, ENT>(WL, L, WR, R, nx, ny, ph, lf_c, F);
  } else {
    face_flux<FLUX, ENT>(WL, L, WR, R, nx, ny, ph, lf_c, F);
  }
}

template <int RECON, int FLUX, int ENT>
__global__ __launch_bounds__(BLOCK) void k_face(const GroupDev* __restrict__ groups,
                                                const BlockRef* __restrict__ blk_group,
                                                const CaseDev* __restrict__ cases, const float4* __restrict__ Ws,
                                                const float4* __restrict__ rec, float4* __restrict__ fx, int xcd) {
  Item it;
  if (!map_item<true, true>(groups, blk_group, cases, xcd, it)) return;
  const GroupDev* G = it.G;
  const MeshDev& m = G->m;
  const CaseDev& cs = cases[G->case0 + it.c];
  const float2 nk = ld(m.f_nrm, (uint32_t)it.idx);
  const float sk = ld(m.f_len, (uint32_t)it.idx);
  float4 WL, WR;
  Prim L, R;
  face_states<RECON>(G, it.idx, it.c, cs, Ws, rec, WL, L, WR, R);
  const float lf_c = (FLUX == ESR_FLUX_LF) ? __uint_as_float(cs.lf_cmax) : 0.0f;
  float F[4];
  class_flux<FLUX, ENT>(G->flux, WL, L, WR, R, nk.x, nk.y, G->ph, lf_c, F);
  if (!it.active) return;
  st(fx + G->fx_off, fel(G, it.idx, it.c), make_float4(sk * F[0], sk * F[1], sk * F[2], sk * F[3]));
}

// Lax-Friedrichs: per-case maximum wave speed over both states of every face
// (euler_solver.py:267-269), integer max on the float bits (values >= 0).
template <int RECON>
__global__ __launch_bounds__(BLOCK) void k_lf_cmax(const GroupDev* __restrict__ groups,
                                                   const BlockRef* __restrict__ blk_group, CaseDev* cases,
                                                   const float4* __restrict__ Ws, const float4* __restrict__ rec, int xcd) {
  Item it;
  if (!map_item<true>(groups, blk_group, cases, xcd, it)) return;
  const GroupDev* G = it.G;
  const Phys& ph = G->ph;
  CaseDev& cs = cases[G->case0 + it.c];
  float4 WL, WR;
  Prim L, R;
  face_states<RECON>(G, it.idx, it.c, cs, Ws, rec, WL, L, WR, R);
  const float2 nk = ld(G->m.f_nrm, (uint32_t)it.idx);
  const float cl = wave_speed(L, nk.x, nk.y, ph), cr = wave_speed(R, nk.x, nk.y, ph);
  const float cm = (cl != cl) ? cl : ((cr != cr) ? cr : fmaxf(cl, cr));
  atomicMax(&cs.lf_cmax, __float_as_uint(cm));
}

// ---------------------------------------------------------------------------
// deterministic face-to-cell gather + residual + stage update
// ---------------------------------------------------------------------------

// residual of one cell from its three face fluxes (F * len, outward), the
// sponge source, and the stage update selected by `mode` (a StageMode)
// (s: the cell's element offset in the group slabs W0, out and acc, which
// are the group's base pointers)
// LOCAL: local time stepping (each cell's own dt, CaseDev.cfl_local); a
// compile-time choice, so the global-step kernels carry none of its code.
// The inputs of the update (1/area, sponge weight, the step's start state W0,
// the case's inlet state, time steps and sponge strength) are loaded by
// stage_loads at the top of a kernel, with its other loads, not after the
// fluxes.
struct StageIn {
  float ia, sg;
  float4 w0;
  float4 value;
  float dt, dt_half, dt_sixth, strength;
};
__device__ __forceinline__ StageIn stage_loads(int mode, const GroupDev* G, const CaseDev& cs, int32_t cell,
                                               uint32_t s, const float4 Wi, const float4* W0) {
  StageIn in;
  in.ia = ld(G->m.inv_area, (uint32_t)cell);
  in.sg = G->sponge ? ld(G->sponge_g, (uint32_t)cell) : 0.0f;
  in.w0 = (mode == MODE_EE || mode == MODE_RESID) ? Wi : ld(W0, s);
  in.value = cs.value;
  in.dt = cs.dt;
  in.dt_half = cs.dt_half;
  in.dt_sixth = cs.dt_sixth;
  in.strength = cs.strength;
  return in;
}

template <bool LOCAL>
__device__ __forceinline__ void stage_epilogue(int mode, const GroupDev* G, const CaseDev& cs, int32_t cell, uint32_t s,
                                               const float4 Wi, const float4 Fk[3], const StageIn& in, float4* out,
                                               float4* acc) {
  const float ia = in.ia;
  float R[4];
  R[0] = ((Fk[0].x + Fk[1].x) + Fk[2].x) * ia;
  R[1] = ((Fk[0].y + Fk[1].y) + Fk[2].y) * ia;
  R[2] = ((Fk[0].z + Fk[1].z) + Fk[2].z) * ia;
  R[3] = ((Fk[0].w + Fk[1].w) + Fk[2].w) * ia;
  if (G->sponge) {
    const float sigma = in.strength * in.sg;
    R[0] = R[0] - sigma * (in.value.x - Wi.x);
    R[1] = R[1] - sigma * (in.value.y - Wi.y);
    R[2] = R[2] - sigma * (in.value.z - Wi.z);
    R[3] = R[3] - sigma * (in.value.w - Wi.w);
  }
  if (mode == MODE_RESID) {
    st(out, s, make_float4(R[0], R[1], R[2], R[3]));
    return;
  }
  const float4 w0 = in.w0;
  // time step: the case's global dt, or with local time stepping the cell's
  // own dt from the step's start state w0 (the same dt in every stage)
  float dt = in.dt, dt_half = in.dt_half, dt_sixth = in.dt_sixth;
  if (LOCAL) {
    dt = cell_dt1(G->m, G->ph, cell, w0) * cs.cfl_local;
    dt_half = dt / 2.0f;
    dt_sixth = dt / 6.0f;
  }
  float4 o;
  if (mode == MODE_EE || mode == MODE_S1) {
    o = make_float4(w0.x - dt * R[0], w0.y - dt * R[1], w0.z - dt * R[2], w0.w - dt * R[3]);
  } else if (mode == MODE_SRK2_2) {
    o = make_float4(0.5f * ((w0.x + Wi.x) - dt * R[0]), 0.5f * ((w0.y + Wi.y) - dt * R[1]),
                    0.5f * ((w0.z + Wi.z) - dt * R[2]), 0.5f * ((w0.w + Wi.w) - dt * R[3]));
  } else if (mode == MODE_RK2_2) {
    o = make_float4(0.5f * ((w0.x - dt * R[0]) + Wi.x), 0.5f * ((w0.y - dt * R[1]) + Wi.y),
                    0.5f * ((w0.z - dt * R[2]) + Wi.z), 0.5f * ((w0.w - dt * R[3]) + Wi.w));
  } else if (mode == MODE_RK4_1) {
    const float h = dt_half;
    o = make_float4(w0.x - h * R[0], w0.y - h * R[1], w0.z - h * R[2], w0.w - h * R[3]);
    st(acc, s, make_float4(R[0], R[1], R[2], R[3]));
  } else if (mode == MODE_RK4_2 || mode == MODE_RK4_3) {
    const float h = (mode == MODE_RK4_2) ? dt_half : dt;
    o = make_float4(w0.x - h * R[0], w0.y - h * R[1], w0.z - h * R[2], w0.w - h * R[3]);
    const float4 a = ld(acc, s);
    st(acc, s, make_float4(a.x + 2.0f * R[0], a.y + 2.0f * R[1], a.z + 2.0f * R[2], a.w + 2.0f * R[3]));
  } else {  // MODE_RK4_4
    const float h = dt_sixth;
    const float4 a = ld(acc, s);
    o = make_float4(w0.x - h * (a.x + R[0]), w0.y - h * (a.y + R[1]), w0.z - h * (a.z + R[2]),
                    w0.w - h * (a.w + R[3]));
  }
  st(out, s, o);
}

template <int MODE, bool LOCAL>
__global__ __launch_bounds__(BLOCK) void k_update(const GroupDev* __restrict__ groups,
                                                  const BlockRef* __restrict__ blk_group,
                                                  const CaseDev* __restrict__ cases,
                                                  const float4* __restrict__ Ws, const float4* W0,
                                                  const float4* __restrict__ fx, float4* out, float4* acc, int xcd) {
  Item it;
  if (!map_item<false, true>(groups, blk_group, cases, xcd, it)) return;
  const GroupDev* G = it.G;
  const MeshDev& m = G->m;
  const int32_t n = m.n, cell = it.idx;
  const CaseDev& cs = cases[G->case0 + it.c];
  const float4* fg = fx + G->fx_off;
  const int64_t so = G->st_off;
  const float4 Wi = ld(Ws + so, it.e);
  const StageIn in = stage_loads(MODE, G, cs, cell, it.e, Wi, W0 + so);
  float4 Fk[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint32_t cf = ld(m.cface, kn(k, n, cell));
    const float4 f = ld(fg, fel(G, (int32_t)(cf & ~CFACE_NEG), it.c));
    Fk[k] = (cf & CFACE_NEG) ? make_float4(-f.x, -f.y, -f.z, -f.w) : f;
  }
  if (!it.active) return;
  stage_epilogue<LOCAL>(MODE, G, cs, cell, it.e, Wi, Fk, in, out + so, acc ? acc + so : acc);
}

// ---------------------------------------------------------------------------
// cell-centric stage: the three face fluxes of a cell from its own and its
// neighbours' face states (each interior face is evaluated by both of its
// cells, as the reference's per-cell flux arrays do), then the residual and
// the stage update -- one launch instead of face + update.
// ---------------------------------------------------------------------------

template <int RECON, int FLUX, int ENT>
__global__ __launch_bounds__(BLOCK) void k_cell(const GroupDev* __restrict__ groups,
                                                const BlockRef* __restrict__ blk_group,
                                                const CaseDev* __restrict__ cases, const float4* __restrict__ Ws,
                                                const float4* W0, const float4* __restrict__ rec, float4* out,
                                                float4* acc, int mode, int xcd) {
  Item it;
  if (!map_item<false, true>(groups, blk_group, cases, xcd, it)) return;
  const GroupDev* G = it.G;
  const MeshDev& m = G->m;
  const Phys& ph = G->ph;
  const int32_t n = m.n, cell = it.idx;
  const CaseDev& cs = cases[G->case0 + it.c];
  const float4* Wg = Ws + G->st_off;
  const float4* rg = rec + G->rec_off;
  const float4 Wi = ld(Wg, it.e);
  const float lf_c = (FLUX == ESR_FLUX_LF) ? __uint_as_float(cs.lf_cmax) : 0.0f;
  // constant reconstruction: the cell's own state on every face
  const Prim Pi = (RECON == ESR_RECON_MUSCL) ? Prim{} : prim_of(Wi, ph.M2, ph.gm1);
  // Every load of the three faces is issued before any flux: the face
  // records, then the cell's own face states and the partners' (a boundary
  // face reads an in-bounds stand-in, its own face state or the cell's state,
  // and replaces it by the ghost state below). Loaded face by face, each
  // face's partner load would wait for the previous face's flux.
  uint32_t f[3];
  float2 nk[3];
  float sk[3];
  float4 qL[3], qR[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    f[k] = ld(m.fref, kn(k, n, cell));
    nk[k] = ld(m.nrm, kn(k, n, cell));
    sk[k] = ld(m.len, kn(k, n, cell));
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const bool bnd = fref_is_bnd(f[k]);
    const int32_t j = bnd ? cell : (int32_t)(f[k] & FREF_IDX);
    if (RECON == ESR_RECON_MUSCL) {
      qL[k] = ld(rg, rel(G, k, cell, it.c));
      qR[k] = ld(rg, rel(G, bnd ? k : (int)(f[k] >> 30), j, it.c));
    } else {
      qR[k] = ld(Wg, sel(G, j, it.c));
    }
  }
  const int64_t so = G->st_off;
  const StageIn in = stage_loads(mode, G, cs, cell, it.e, Wi, W0 + so);
  float4 Fk[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float4 WL, WR;
    Prim L, R;
    if (RECON == ESR_RECON_MUSCL) {
      L = as_prim(qL[k]);
      WL = cons_of(L, ph.inv_gm1M2);
    } else {
      L = Pi;
      WL = Wi;
    }
    // partner state: the loaded one, or for a boundary face the ghost state
    // (in primitive form for MUSCL, whose both paths build WR from R)
    float4 q = qR[k];
    if (fref_is_bnd(f[k])) {
      const uint32_t kind = (f[k] >> FREF_KIND_SHIFT) & 7u;
      const float2 nw = (kind == G_WALL) ? nk[k] : make_float2(0.f, 0.f);
      const float4 Wgh = ghost_state(kind, Wi, nw, in.value, m.inlet_subsonic, Wg,
                                     sel(G, (int32_t)(f[k] & FREF_BIDX), it.c));
      if (RECON == ESR_RECON_MUSCL) {
        const Prim g = prim_of(Wgh, ph.M2, ph.gm1);
        q = make_float4(g.r, g.u, g.v, g.P);
      } else {
        q = Wgh;
      }
    }
    if (RECON == ESR_RECON_MUSCL) {
      R = as_prim(q);
      WR = cons_of(R, ph.inv_gm1M2);
    } else {
      WR = q;
      R = prim_of(WR, ph.M2, ph.gm1);
    }
    float F[4];
    class_flux<FLUX, ENT>(G->flux, WL, L, WR, R, nk[k].x, nk[k].y, ph, lf_c, F);
    Fk[k] = make_float4(sk[k] * F[0], sk[k] * F[1], sk[k] * F[2], sk[k] * F[3]);
  }
  if (!it.active) return;
  stage_epilogue<false>(mode, G, cs, cell, it.e, Wi, Fk, in, out + so, acc ? acc + so : acc);
}

// ---------------------------------------------------------------------------
// per-case reductions
// ---------------------------------------------------------------------------

// Block layout of a reduction over one group: blockIdx.y = case chunk of up
// to BLOCK cases, blockIdx.x = cell range. Thread t handles case c0 + t % Ck
// and cells cs = t / Ck, cs + P, ... of the range (P = BLOCK / Ck lanes per
// case); partials are combined over cs in a fixed order.
struct RedMap {
  int32_t c0, Ck, P, cell_begin, cell_end;
};
constexpr int32_t RED_ITERS = 8;

__host__ __device__ inline int32_t red_cells_per_range(int32_t C) {
  const int32_t Ck = C < BLOCK ? C : BLOCK;
  return (BLOCK / Ck) * RED_ITERS;
}

int32_t reduce_ranges(const GroupDev& g) {
  const int32_t cpr = red_cells_per_range(g.C);
  return (g.m.n + cpr - 1) / cpr;
}

__device__ __forceinline__ RedMap red_map(const GroupDev* G, int32_t range, int32_t chunk) {
  RedMap r;
  r.c0 = chunk * BLOCK;
  r.Ck = min(G->C - r.c0, BLOCK);
  r.P = BLOCK / r.Ck;
  const int32_t cpr = red_cells_per_range(G->C);
  r.cell_begin = range * cpr;
  r.cell_end = min(r.cell_begin + cpr, G->m.n);
  return r;
}

void red_table(const GroupDev& g, int32_t gi, std::vector<RedRef>& out) {
  const int32_t nr = reduce_ranges(g), nc = (g.C + BLOCK - 1) / BLOCK;
  for (int32_t y = 0; y < nc; ++y)
    for (int32_t x = 0; x < nr; ++x) out.push_back({gi, x, y, 0});
}

__device__ void check_cases(int32_t c0, int32_t c1, CaseDev* cases, CaseLoop* loops, const double* part_d,
                            const double* part_r, const int32_t* __restrict__ nranges, int32_t rstride,
                            const double* __restrict__ forces, int64_t step);

// ca.counter != null: the last block of this launch to finish (every earlier
// launch of the stream has completed) runs the stationarity check of every
// case, then resets the counter -- one launch less per check.
__global__ __launch_bounds__(BLOCK) void k_norms(const GroupDev* __restrict__ groups, const RedRef* __restrict__ tbl,
                                                 const CaseDev* __restrict__ cases,
                                                 const float4* __restrict__ Wn, float4* __restrict__ Wref,
                                                 double* __restrict__ part_d, double* __restrict__ part_r,
                                                 int32_t rstride, int copy_ref, CheckArgs ca) {
  const RedRef e = tbl[blockIdx.x];
  const GroupDev* G = groups + e.group;
  const RedMap r = red_map(G, e.range, e.chunk);
  __shared__ double sd[BLOCK], sr[BLOCK];
  const int t = threadIdx.x;
  double accd = 0.0, accr = 0.0;
  int32_t c = -1;
  if (t < r.P * r.Ck) {
    c = r.c0 + t % r.Ck;
    const int32_t cs0 = t / r.Ck;
    if (cases[G->case0 + c].check) {
      for (int32_t cell = r.cell_begin + cs0; cell < r.cell_end; cell += r.P) {
        const int64_t s = G->st_off + (int64_t)cell * G->C + c;
        const float4 a = Wn[s];
        const float4 b = Wref[s];
        const double d0 = (double)a.x - (double)b.x, d1 = (double)a.y - (double)b.y;
        const double d2 = (double)a.z - (double)b.z, d3 = (double)a.w - (double)b.w;
        accd += ((d0 * d0 + d1 * d1) + d2 * d2) + d3 * d3;
        accr += (((double)b.x * b.x + (double)b.y * b.y) + (double)b.z * b.z) + (double)b.w * b.w;
        if (copy_ref) Wref[s] = a;
      }
    }
  }
  sd[t] = accd;
  sr[t] = accr;
  __syncthreads();
  if (t < r.Ck) {
    const int32_t cc = r.c0 + t;
    double a = 0.0, b = 0.0;
    for (int32_t q = 0; q < r.P; ++q) {
      a += sd[q * r.Ck + t];
      b += sr[q * r.Ck + t];
    }
    part_d[(int64_t)(G->case0 + cc) * rstride + e.range] = a;
    part_r[(int64_t)(G->case0 + cc) * rstride + e.range] = b;
  }
  if (!ca.counter) return;
  __shared__ int last;
  __threadfence();
  __syncthreads();
  if (t == 0) last = (atomicAdd(ca.counter, 1u) == gridDim.x - 1) ? 1 : 0;
  __syncthreads();
  if (!last) return;
  __threadfence();  // agent-scope acquire: the other blocks' partials are visible
  int64_t step = ca.step;
  if (ca.dstep) step = ca.dstep[1];
  check_cases(0, ca.n_cases, ca.cases, ca.loops, part_d, part_r, ca.nranges, rstride, ca.forces, step);
  __syncthreads();
  if (t == 0) {
    if (ca.dstep) ca.dstep[0] = step;
    *ca.counter = 0u;
  }
}

// helper.get_dt per cell: area / sum_k (c + |u.n_k|) len_k; min over cells.
// local == 1 stores the per-cell value instead (local time stepping).
__global__ __launch_bounds__(BLOCK) void k_dt(const GroupDev* __restrict__ gdev, const CaseDev* __restrict__ cases,
                                              const float4* __restrict__ W, float* __restrict__ part,
                                              int32_t rstride) {
  const GroupDev* G = gdev;
  const RedMap r = red_map(G, blockIdx.x, blockIdx.y);
  const MeshDev& m = G->m;
  const Phys& ph = G->ph;
  __shared__ float sm[BLOCK];
  const int t = threadIdx.x;
  float acc = INFINITY;
  if (t < r.P * r.Ck) {
    const int32_t c = r.c0 + t % r.Ck;
    const int32_t cs0 = t / r.Ck;
    for (int32_t cell = r.cell_begin + cs0; cell < r.cell_end; cell += r.P) {
      const float d = cell_dt1(m, ph, cell, W[G->st_off + (int64_t)cell * G->C + c]);
      acc = (d != d) ? d : ((acc != acc) ? acc : fminf(acc, d));
    }
  }
  sm[t] = acc;
  __syncthreads();
  if (t < r.Ck) {
    float a = INFINITY;
    for (int32_t q = 0; q < r.P; ++q) {
      const float x = sm[q * r.Ck + t];
      a = (x != x) ? x : ((a != a) ? a : fminf(a, x));
    }
    part[(int64_t)(G->case0 + r.c0 + t) * rstride + blockIdx.x] = a;
  }
}

// CD/CL wall sums in FP64 (helper.py:366-415). Pressure with the default
// gamma = 1.4 and M = 1. One block per case of the group.
__global__ __launch_bounds__(BLOCK) void k_forces(const GroupDev* __restrict__ groups,
                                                  const BlockRef* __restrict__ tbl,
                                                  const CaseDev* __restrict__ cases, const float4* __restrict__ W,
                                                  double* __restrict__ cd_cl) {
  const BlockRef e = tbl[blockIdx.x];
  const GroupDev* G = groups + e.group;
  const MeshDev& m = G->m;
  const int32_t c = e.block;
  __shared__ double sx[BLOCK], sy[BLOCK];
  double ax = 0.0, ay = 0.0;
  if (cases[G->case0 + c].check) {
    for (int32_t w = threadIdx.x; w < m.n_wall; w += BLOCK) {
      const int32_t cell = m.wall_cell[w];
      const Prim q = prim_of(W[G->st_off + (int64_t)cell * G->C + c], 1.0f, 0.4f);
      const double P = (double)q.P;
      ax += P * (double)m.wall_nx[w] * (double)m.wall_ds[w];
      ay += P * (double)m.wall_ny[w] * (double)m.wall_ds[w];
    }
  }
  sx[threadIdx.x] = ax;
  sy[threadIdx.x] = ay;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sx[threadIdx.x] += sx[threadIdx.x + s];
      sy[threadIdx.x] += sy[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cd_cl[2 * (int64_t)(G->case0 + c)] = sx[0];
    cd_cl[2 * (int64_t)(G->case0 + c) + 1] = sy[0];
  }
}

// Sums of case c's partial norms by one wave: lane l adds ranges l, l+64,
// ... in order, then a fixed shuffle tree; every lane returns the totals.
__device__ __forceinline__ void wave_case_sums(int32_t c, const double* part_d, const double* part_r,
                                               const int32_t* __restrict__ nranges, int32_t rstride, double& d,
                                               double& r) {
  const int lane = threadIdx.x & 63;
  double a = 0.0, b = 0.0;
  const int32_t nr = nranges[c];
  for (int32_t q = lane; q < nr; q += 64) {
    a += part_d[(int64_t)c * rstride + q];
    b += part_r[(int64_t)c * rstride + q];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  d = __shfl(a, 0, 64);
  r = __shfl(b, 0, 64);
}

// Stationarity check of main.py:149-177 for case c (check flag set), given
// the FP64 sums d = ||W - W_ref||^2, r = ||W_ref||^2.
__device__ void apply_check(int32_t c, CaseDev* cases, CaseLoop* loops, double d, double r,
                            const double* __restrict__ forces, int64_t step) {
  CaseDev& cs = cases[c];
  CaseLoop& L = loops[c];
  const int64_t block_steps = (step - L.ref_step) > 1 ? (step - L.ref_step) : 1;
  const double rel = sqrt(d) / ((double)block_steps * (sqrt(r) + 1e-16));
  const bool under = rel <= L.threshold;
  const int32_t new_streak = under ? L.streak + 1 : 0;
  const int64_t new_start = (under && L.streak == 0) ? step : L.streak_start;
  const bool confirmed = (new_streak >= L.patience) && !L.stop;
  if (confirmed) L.stopping_step = new_start;
  L.stop = L.stop || confirmed;
  L.streak = new_streak;
  L.streak_start = new_start;
  L.ref_step = step;
  L.last_rel = rel;
  L.recent[L.n_checks % RECENT_CHECKS] = rel;
  L.n_checks += 1;
  if (!(d == d) || isinf(d)) L.diverged = 1;
  if (L.track_eng && forces) {
    const double cd = forces[2 * (int64_t)c] / L.q_ref, cl = forces[2 * (int64_t)c + 1] / L.q_ref;
    L.cd = cd;
    L.cl = cl;
    const bool eu = (fabs(cd - L.target_cd) <= L.cd_tol) && (fabs(cl - L.target_cl) <= L.cl_tol);
    const int32_t ns = eu ? L.eng_streak + 1 : 0;
    const int64_t st = (eu && L.eng_streak == 0) ? step : L.eng_streak_start;
    const bool ec = (ns >= L.eng_patience) && !L.eng_stop;
    if (ec) L.eng_stopping_step = st;
    L.eng_stop = L.eng_stop || ec;
    L.eng_streak = ns;
    L.eng_streak_start = st;
  }
  L.steps_run = step;
  // a non-finite residual ends the case (unconverged) instead of running to N
  if (L.stop || L.diverged || step >= L.N) {
    L.done = 1;
    cs.active = 0;
  }
  cs.check = 0;
}

// checks of cases [c0, c1) with the waves of this block (wave w: cases
// c0 + w, c0 + w + waves, ...)
__device__ void check_cases(int32_t c0, int32_t c1, CaseDev* cases, CaseLoop* loops, const double* part_d,
                            const double* part_r, const int32_t* __restrict__ nranges, int32_t rstride,
                            const double* __restrict__ forces, int64_t step) {
  const int wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  for (int32_t c = c0 + wave; c < c1; c += waves) {
    if (!cases[c].check) continue;  // wave-uniform
    double d, r;
    wave_case_sums(c, part_d, part_r, nranges, rstride, d, r);
    if ((threadIdx.x & 63) == 0) apply_check(c, cases, loops, d, r, forces, step);
  }
}

// one wave per case, four cases per block
__global__ __launch_bounds__(256) void k_check(int32_t n_cases, CaseDev* cases, CaseLoop* loops,
                                               const double* __restrict__ part_d, const double* __restrict__ part_r,
                                               const int32_t* __restrict__ nranges, int32_t rstride,
                                               const double* __restrict__ forces, int64_t step, int32_t K,
                                               int64_t* dstep) {
  if (dstep) step = dstep[1];
  const int32_t c0 = blockIdx.x * 4;
  check_cases(c0, min(c0 + 4, n_cases), cases, loops, part_d, part_r, nranges, rstride, forces, step);
  if (dstep && blockIdx.x == 0 && threadIdx.x == 0) dstep[0] = step;
}

__global__ void k_dt_final(int32_t n_cases, CaseDev* cases, const float* __restrict__ part,
                           const int32_t* __restrict__ nranges, int32_t rstride, const float* __restrict__ cfl,
                           float* __restrict__ dt_out) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cases) return;
  float a = INFINITY;
  for (int32_t q = 0; q < nranges[c]; ++q) {
    const float x = part[(int64_t)c * rstride + q];
    a = (x != x) ? x : ((a != a) ? a : fminf(a, x));
  }
  const float dt = a * cfl[c];
  cases[c].dt = dt;
  cases[c].dt_half = dt / 2.0f;
  cases[c].dt_sixth = dt / 6.0f;
  dt_out[c] = dt;
}

// ---------------------------------------------------------------------------
// layout conversion and small utilities
// ---------------------------------------------------------------------------

__global__ void k_gather_in(const GroupDev* __restrict__ gdev, int32_t c, const int32_t* __restrict__ perm,
                            const float4* __restrict__ user, float4* __restrict__ slab) {
  const GroupDev* G = gdev;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G->m.n) return;
  slab[G->st_off + (int64_t)i * G->C + c] = user[perm[i]];
}

__global__ void k_scatter_out(const GroupDev* __restrict__ gdev, int32_t c, const int32_t* __restrict__ perm,
                              const float4* __restrict__ slab, float4* __restrict__ user) {
  const GroupDev* G = gdev;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G->m.n) return;
  user[perm[i]] = slab[G->st_off + (int64_t)i * G->C + c];
}

// Exhaustive check of fsqrt and flog against sqrtf and logf over the 2^31
// non-negative float bit patterns (work-item t takes t, t + T, t + 2T, ...
// for T work-items in all): counts[0] square-root mismatches at inputs of at
// least 2^-96, counts[1] below; counts[2], counts[3] the same for the
// logarithm. NaN equals NaN.
__global__ __launch_bounds__(BLOCK) void k_math_check(unsigned long long* counts, uint32_t per_item) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t T = gridDim.x * blockDim.x;
  unsigned long long c[4] = {0, 0, 0, 0};
  for (uint32_t i = 0; i < per_item; ++i) {
    const uint32_t b = t + i * T;
    if (b > 0x7FFFFFFFu) break;
    const float x = __uint_as_float(b);
    const int tiny = (b < 0x0F800000u) ? 1 : 0;  // below 2^-96
    const float s0 = fsqrt(x), s1 = sqrtf(x), l0 = flog(x), l1 = logf(x);
    if (__float_as_uint(s0) != __float_as_uint(s1) && !(s0 != s0 && s1 != s1)) c[tiny] += 1;
    if (__float_as_uint(l0) != __float_as_uint(l1) && !(l0 != l0 && l1 != l1)) c[2 + tiny] += 1;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (c[k]) atomicAdd(&counts[k], c[k]);
}

__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_set_active(int32_t n_cases, CaseDev* cases, const CaseLoop* loops, int32_t all) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cases) return;
  cases[c].active = all ? 1 : (loops ? !loops[c].done : 1);
  cases[c].check = 0;
}

__global__ void k_zero_lf(int32_t n_cases, CaseDev* cases) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n_cases) cases[c].lf_cmax = 0u;
}

__global__ void k_count_active(int32_t n_cases, const CaseDev* __restrict__ cases, int32_t* count) {
  __shared__ int32_t s[BLOCK];
  int32_t a = 0;
  for (int32_t c = threadIdx.x; c < n_cases; c += BLOCK) a += cases[c].active ? 1 : 0;
  s[threadIdx.x] = a;
  __syncthreads();
  for (int q = BLOCK / 2; q > 0; q >>= 1) {
    if ((int)threadIdx.x < q) s[threadIdx.x] += s[threadIdx.x + q];
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = s[0];
}

__global__ void k_set_case(CaseDev* dst, CaseDev v) { *dst = v; }
__global__ void k_set_step(int64_t* dstep, int64_t v) {
  dstep[0] = v;
  dstep[1] = v;
}
__global__ void k_set_group(GroupDev* dst, GroupDev v) { *dst = v; }

__global__ void k_mark_check(int32_t n_cases, CaseDev* cases, const CaseLoop* __restrict__ loops, int64_t step,
                             int32_t K, int32_t all, int64_t* dstep) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (dstep) {
    step = dstep[0] + K;
    if (c == 0) dstep[1] = step;
  }
  if (c >= n_cases) return;
  if (all) {
    cases[c].check = 1;
    return;
  }
  cases[c].check = (cases[c].active && ((step % K) == 0 || step == loops[c].N)) ? 1 : 0;
}

__global__ void k_probe_final(int32_t n_cases, CaseDev* cases, const double* __restrict__ part_d,
                              const double* __restrict__ part_r, const int32_t* __restrict__ nranges, int32_t rstride,
                              double* __restrict__ out) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cases) return;
  double d = 0.0, r = 0.0;
  for (int32_t q = 0; q < nranges[c]; ++q) {
    d += part_d[(int64_t)c * rstride + q];
    r += part_r[(int64_t)c * rstride + q];
  }
  out[c] = sqrt(d) / (sqrt(r) + 1e-16);
  cases[c].check = 0;
}

__global__ __launch_bounds__(BLOCK) void k_commit(const GroupDev* __restrict__ groups,
                                                  const BlockRef* __restrict__ blk_group,
                                                  const CaseDev* __restrict__ cases, const float4* __restrict__ src,
                                                  float4* __restrict__ dst, int xcd) {
  Item it;
  if (!map_item<false>(groups, blk_group, cases, xcd, it)) return;
  st(dst + it.G->st_off, it.e, ld(src + it.G->st_off, it.e));
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

static void check_launch(const char* what) { hip_check(hipGetLastError(), what); }

void launch_recon(const LaunchSet& ls, const CaseDev* cases, const float4* Ws, float4* rec, hipStream_t st) {
  if (ls.nblocks <= 0) return;
  hipLaunchKernelGGL(k_recon, dim3(ls.nblocks), dim3(ls.block), 0, st, ls.groups, ls.blk_group, cases, Ws, rec,
                     ls.xcd);
  check_launch("k_recon");
}

void launch_lf_cmax(const LaunchSet& ls, const CaseDev* cases, const float4* Ws, const float4* rec, hipStream_t st) {
  if (ls.fnblocks <= 0) return;
  CaseDev* cm = const_cast<CaseDev*>(cases);
  if (ls.recon == ESR_RECON_MUSCL)
    hipLaunchKernelGGL(k_lf_cmax<ESR_RECON_MUSCL>, dim3(ls.fnblocks), dim3(ls.block), 0, st, ls.groups, ls.fblk_group,
                       cm, Ws, rec, ls.xcd);
  else
    hipLaunchKernelGGL(k_lf_cmax<ESR_RECON_CONSTANT>, dim3(ls.fnblocks), dim3(ls.block), 0, st, ls.groups,
                       ls.fblk_group, cm, Ws, rec, ls.xcd);
  check_launch("k_lf_cmax");
}

template <int RECON, int ENT>
static void launch_face_f(const LaunchSet& ls, const CaseDev* cases, const float4* Ws, const float4* rec, float4* fx,
                          hipStream_t st) {
  const dim3 g(ls.fnblocks), b(ls.block);
  switch (ls.flux) {
    case ESR_FLUX_LF:
      hipLaunchKernelGGL((k_face<RECON, ESR_FLUX_LF, ENT>), g, b, 0, st, ls.groups, ls.fblk_group, cases, Ws, rec, fx, ls.xcd);
      break;
    case ESR_FLUX_RUSANOV:
      hipLaunchKernelGGL((k_face<RECON, ESR_FLUX_RUSANOV, ENT>), g, b, 0, st, ls.groups, ls.fblk_group, cases, Ws,
                         rec, fx, ls.xcd);
      break;
    case ESR_FLUX_TADMOR:
      hipLaunchKernelGGL((k_face<RECON, ESR_FLUX_TADMOR, ENT>), g, b, 0, st, ls.groups, ls.fblk_group, cases, Ws,
                         rec, fx, ls.xcd);
      break;
    case ESR_FLUX_AUSM:
      hipLaunchKernelGGL((k_face<RECON, ESR_FLUX_AUSM, ENT>), g, b, 0, st, ls.groups, ls.fblk_group, cases, Ws, rec,
                         fx, ls.xcd);
      break;
    case ESR_FLUX_HLLC:
      hipLaunchKernelGGL((k_face<RECON, ESR_FLUX_HLLC, ENT>), g, b, 0, st, ls.groups, ls.fblk_group, cases, Ws, rec,
                         fx, ls.xcd);
      break;
    case ESR_FLUX_ROE:
      hipLaunchKernelGGL((k_face<RECON, ESR_FLUX_ROE, ENT>), g, b, 0, st, ls.groups, ls.fblk_group, cases, Ws, rec,
                         fx, ls.xcd);
      break;
    case FLUX_HLLC_ROE:
      hipLaunchKernelGGL((k_face<RECON, FLUX_HLLC_ROE, ENT>), g, b, 0, st, ls.groups, ls.fblk_group, cases, Ws,
                         rec, fx, ls.xcd);
      break;
    default: throw Error(ESR_ERR_VALUE, "unknown flux");
  }
}

void launch_face(const LaunchSet& ls, const CaseDev* cases, const float4* Ws, const float4* rec, float4* fx,
                 hipStream_t st) {
  if (ls.fnblocks <= 0) return;
  if (ls.recon == ESR_RECON_MUSCL) {
    if (ls.entropy)
      launch_face_f<ESR_RECON_MUSCL, 1>(ls, cases, Ws, rec, fx, st);
    else
      launch_face_f<ESR_RECON_MUSCL, 0>(ls, cases, Ws, rec, fx, st);
  } else {
    if (ls.entropy)
      launch_face_f<ESR_RECON_CONSTANT, 1>(ls, cases, Ws, rec, fx, st);
    else
      launch_face_f<ESR_RECON_CONSTANT, 0>(ls, cases, Ws, rec, fx, st);
  }
  check_launch("k_face");
}

template <int RECON, int ENT>
static void launch_cell_f(const LaunchSet& ls, int mode, const CaseDev* cases, const float4* Ws, const float4* W0,
                          const float4* rec, float4* out, float4* acc, hipStream_t st) {
  const dim3 g(ls.nblocks), b(ls.block);
#define ESR_CELL_CASE(FL)                                                                                    \
  case FL:                                                                                                   \
    hipLaunchKernelGGL((k_cell<RECON, FL, ENT>), g, b, 0, st, ls.groups, ls.blk_group, cases, Ws, W0, rec, out, \
                       acc, mode, ls.xcd);                                                                   \
    break;
  switch (ls.flux) {
    ESR_CELL_CASE(ESR_FLUX_LF)
    ESR_CELL_CASE(ESR_FLUX_RUSANOV)
    ESR_CELL_CASE(ESR_FLUX_TADMOR)
    ESR_CELL_CASE(ESR_FLUX_AUSM)
    ESR_CELL_CASE(ESR_FLUX_HLLC)
    ESR_CELL_CASE(ESR_FLUX_ROE)
    ESR_CELL_CASE(FLUX_HLLC_ROE)
    default: throw Error(ESR_ERR_VALUE, "unknown flux");
  }
#undef ESR_CELL_CASE
}

void launch_cell(const LaunchSet& ls, int mode, const CaseDev* cases, const float4* Ws, const float4* W0,
                 const float4* rec, float4* out, float4* acc, hipStream_t st) {
  if (ls.nblocks <= 0) return;
  if (mode < MODE_RESID || mode > MODE_RK4_4) throw Error(ESR_ERR_VALUE, "unknown stage mode");
  if (ls.recon == ESR_RECON_MUSCL) {
    if (ls.entropy)
      launch_cell_f<ESR_RECON_MUSCL, 1>(ls, mode, cases, Ws, W0, rec, out, acc, st);
    else
      launch_cell_f<ESR_RECON_MUSCL, 0>(ls, mode, cases, Ws, W0, rec, out, acc, st);
  } else {
    if (ls.entropy)
      launch_cell_f<ESR_RECON_CONSTANT, 1>(ls, mode, cases, Ws, W0, rec, out, acc, st);
    else
      launch_cell_f<ESR_RECON_CONSTANT, 0>(ls, mode, cases, Ws, W0, rec, out, acc, st);
  }
  check_launch("k_cell");
}

template <bool LOCAL>
static void launch_update_l(const LaunchSet& ls, int mode, const CaseDev* cases, const float4* Ws, const float4* W0,
                            const float4* fx, float4* out, float4* acc, hipStream_t st) {
#define ESR_UPD_CASE(MD)                                                                                     \
  case MD:                                                                                                   \
    hipLaunchKernelGGL((k_update<MD, LOCAL>), dim3(ls.nblocks), dim3(ls.block), 0, st, ls.groups, ls.blk_group, \
                       cases, Ws, W0, fx, out, acc, ls.xcd);                                                 \
    break;
  switch (mode) {
    ESR_UPD_CASE(MODE_RESID)
    ESR_UPD_CASE(MODE_EE)
    ESR_UPD_CASE(MODE_S1)
    ESR_UPD_CASE(MODE_SRK2_2)
    ESR_UPD_CASE(MODE_RK2_2)
    ESR_UPD_CASE(MODE_RK4_1)
    ESR_UPD_CASE(MODE_RK4_2)
    ESR_UPD_CASE(MODE_RK4_3)
    ESR_UPD_CASE(MODE_RK4_4)
    default:
      throw Error(ESR_ERR_VALUE, "unknown stage mode");
  }
#undef ESR_UPD_CASE
}

void launch_update(const LaunchSet& ls, int mode, bool local_dt, const CaseDev* cases, const float4* Ws,
                   const float4* W0, const float4* fx, float4* out, float4* acc, hipStream_t st) {
  if (ls.nblocks <= 0) return;
  if (local_dt)
    launch_update_l<true>(ls, mode, cases, Ws, W0, fx, out, acc, st);
  else
    launch_update_l<false>(ls, mode, cases, Ws, W0, fx, out, acc, st);
  check_launch("k_update");
}

static dim3 red_grid(const GroupDev& g) {
  return dim3((unsigned)reduce_ranges(g), (unsigned)((g.C + BLOCK - 1) / BLOCK));
}

void launch_norms(const GroupDev* groups, const RedRef* tbl, int32_t n, const CaseDev* cases, const float4* Wn,
                  float4* Wref, double* part_d, double* part_r, int32_t rstride, int copy_ref, hipStream_t st,
                  const CheckArgs* fused) {
  if (n <= 0) return;
  CheckArgs ca{};
  if (fused) ca = *fused;
  hipLaunchKernelGGL(k_norms, dim3((unsigned)n), dim3(BLOCK), 0, st, groups, tbl, cases, Wn, Wref, part_d, part_r,
                     rstride, copy_ref, ca);
  check_launch("k_norms");
}

void launch_dt(const GroupDev& g, const GroupDev* gdev, const CaseDev* cases, const float4* W, float* part,
               int32_t rstride, int /*local*/, hipStream_t st) {
  hipLaunchKernelGGL(k_dt, red_grid(g), dim3(BLOCK), 0, st, gdev, cases, W, part, rstride);
  check_launch("k_dt");
}

void launch_forces(const GroupDev* groups, const BlockRef* tbl, int32_t n, const CaseDev* cases, const float4* W,
                   double* cd_cl, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_forces, dim3((unsigned)n), dim3(BLOCK), 0, st, groups, tbl, cases, W, cd_cl);
  check_launch("k_forces");
}

void launch_check(int32_t n_cases, CaseDev* cases, CaseLoop* loops, const double* part_d, const double* part_r,
                  const int32_t* nranges, int32_t rstride, const double* forces, int64_t step, int32_t K,
                  hipStream_t st, int64_t* dstep) {
  hipLaunchKernelGGL(k_check, dim3((n_cases + 3) / 4), dim3(256), 0, st, n_cases, cases, loops, part_d, part_r,
                     nranges, rstride, forces, step, K, dstep);
  check_launch("k_check");
}

void launch_dt_final(int32_t n_cases, CaseDev* cases, const float* part, const int32_t* nranges, int32_t rstride,
                     const float* cfl, float* dt_out, hipStream_t st) {
  hipLaunchKernelGGL(k_dt_final, dim3((n_cases + 127) / 128), dim3(128), 0, st, n_cases, cases, part, nranges,
                     rstride, cfl, dt_out);
  check_launch("k_dt_final");
}

void launch_gather_in_dev(int32_t n, const GroupDev* gdev, int32_t c, const int32_t* perm, const float4* user,
                          float4* slab, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_in, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, gdev, c, perm, user, slab);
  check_launch("k_gather_in");
}

void launch_scatter_out_dev(int32_t n, const GroupDev* gdev, int32_t c, const int32_t* perm, const float4* slab,
                            float4* user, hipStream_t st) {
  hipLaunchKernelGGL(k_scatter_out, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, gdev, c, perm, slab, user);
  check_launch("k_scatter_out");
}

void launch_set_case(CaseDev* dst, const CaseDev& v, hipStream_t st) {
  hipLaunchKernelGGL(k_set_case, dim3(1), dim3(1), 0, st, dst, v);
  check_launch("k_set_case");
}

void launch_set_step(int64_t* dstep, int64_t v, hipStream_t st) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, st, dstep, v);
  check_launch("k_set_step");
}

void launch_set_group(GroupDev* dst, const GroupDev& v, hipStream_t st) {
  hipLaunchKernelGGL(k_set_group, dim3(1), dim3(1), 0, st, dst, v);
  check_launch("k_set_group");
}

void launch_mark_check(int32_t n_cases, CaseDev* cases, const CaseLoop* loops, int64_t step, int32_t K, int32_t all,
                       hipStream_t st, int64_t* dstep) {
  hipLaunchKernelGGL(k_mark_check, dim3((n_cases + 127) / 128), dim3(128), 0, st, n_cases, cases, loops, step, K, all,
                     dstep);
  check_launch("k_mark_check");
}

void launch_probe_final(int32_t n_cases, CaseDev* cases, const double* part_d, const double* part_r,
                        const int32_t* nranges, int32_t rstride, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_probe_final, dim3((n_cases + 127) / 128), dim3(128), 0, st, n_cases, cases, part_d, part_r,
                     nranges, rstride, out);
  check_launch("k_probe_final");
}

void launch_commit(const LaunchSet& ls, const CaseDev* cases, const float4* src, float4* dst, hipStream_t st) {
  if (ls.nblocks <= 0) return;
  hipLaunchKernelGGL(k_commit, dim3(ls.nblocks), dim3(ls.block), 0, st, ls.groups, ls.blk_group, cases, src, dst,
                     ls.xcd);
  check_launch("k_commit");
}

void launch_math_check(unsigned long long* counts, hipStream_t st) {
  // 2^15 blocks x 256 work-items x 256 patterns each = 2^31 patterns
  constexpr uint32_t blocks = 1u << 15, per_item = 1u << 8;
  static_assert((uint64_t)blocks * BLOCK * per_item == (1ull << 31), "covers every non-negative pattern");
  hipLaunchKernelGGL(k_math_check, dim3(blocks), dim3(BLOCK), 0, st, counts, per_item);
  check_launch("k_math_check");
}

void launch_copy(const float4* src, float4* dst, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, src, dst, n);
  check_launch("k_copy");
}

void launch_set_active(int32_t n_cases, CaseDev* cases, const CaseLoop* loops, int32_t all, hipStream_t st) {
  hipLaunchKernelGGL(k_set_active, dim3((n_cases + 127) / 128), dim3(128), 0, st, n_cases, cases, loops, all);
  check_launch("k_set_active");
}

void launch_zero_lf(int32_t n_cases, CaseDev* cases, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_lf, dim3((n_cases + 127) / 128), dim3(128), 0, st, n_cases, cases);
  check_launch("k_zero_lf");
}

void launch_count_active(int32_t n_cases, const CaseDev* cases, int32_t* count, hipStream_t st) {
  hipLaunchKernelGGL(k_count_active, dim3(1), dim3(BLOCK), 0, st, n_cases, cases, count);
  check_launch("k_count_active");
}

}  // namespace esr
