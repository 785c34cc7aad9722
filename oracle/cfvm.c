/* CPU oracle, C restatement (TEST INFRASTRUCTURE ONLY): the reference's
 * explicit finite-volume residual for HLLC or Roe + constant/MUSCL
 * reconstruction,
 * the sponge, get_dt and the SSP-RK2 step, in float32, parallel over cells
 * with OpenMP.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it, as the checker or as the timed CPU baseline; the product never
 * does. It follows the numpy oracle (oracle/fvm.py) operation for operation,
 * which in turn follows the reference:
 *   getPrimitive/getConserved   helper.py:25-46
 *   BC_state (walls, 3, 4, 5)   helper.py:168-256
 *   getgradientLSQ              helper.py:136-153 (2x2 solve: LU with partial
 *                               pivoting, as oracle/fvm.py solve2x2)
 *   venkatakrishnan, limiter    euler_solver.py:36-67
 *   MUSCL                       euler_solver.py:69-98
 *   build_reconstruction        euler_solver.py:106-122
 *   HLLC                        euler_solver.py:273-374
 *   Roe (R^T, S scaling)        euler_solver.py:145-199, Roe average
 *                               helper.py:101-127, entropy variables
 *                               helper.py:48-70
 *   getFlux, face lengths       euler_solver.py:141-143, 376-420
 *   sponge                      helper.py:262-295
 *   residual                    euler_solver.py:602-628
 *   time_step_RK2_SSP           euler_solver.py:635-643
 *   get_dt                      helper.py:8-18
 * Built with -ffp-contract=off and no fast-math, so every float operation
 * rounds once to float32 in the same order as numpy's elementwise kernels:
 * the results are bit-identical to oracle/fvm.py in float32 (checked by
 * tests/test_oracle_c.py) wherever only IEEE operations are involved (HLLC:
 * + - * / and sqrt). The Roe dissipation's entropy variables take logf and
 * powf from the C library, where numpy may use its own vectorised float32
 * log and power, so Roe agrees with oracle/fvm.py to within a few ulps of
 * those per evaluation (also checked there). Python-float constants of the
 * numpy code are formed in double and rounded to float32 where numpy's weak
 * promotion does.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <omp.h>

typedef struct {
  int64_t n;              /* cells */
  int64_t npts;           /* points */
  const float* points;    /* (npts, 2) */
  const float* bary;      /* (n, 2) */
  const float* area;      /* (n) */
  const int64_t* nb;      /* (n, 3) neighbours; -1 indexes the last cell (numpy) */
  const int64_t* fc;      /* (n, 3) face ids */
  const int64_t* fmark;   /* (nf) face markers */
  const int64_t* faces;   /* (nf, 2) vertex ids */
  const int64_t* fco;     /* (nf) face_connectivity_opposite */
  const float* surface;   /* (nf) face lengths */
  const float* normals;   /* (n, 3, 2) */
  const float* inlet_sub; /* (4) subsonic inlet primitives */
  const int64_t* walls;   /* wall markers */
  int32_t n_walls;
  int32_t bump;           /* metadata case == "bump": no sponge */
} cf_mesh;

typedef struct {
  double gamma, M;        /* Python floats of the keyword arguments */
  int32_t muscl;          /* 1: MUSCL, 0: constant */
  float value[4];         /* inlet state (float32) */
  int32_t has_width, has_strength;
  double sponge_width, sponge_strength;
  int32_t roe;            /* 1: Roe dissipation, 0: HLLC */
} cf_opts;

/* float32 forms of the Python-float expressions (numpy weak promotion) */
typedef struct {
  float g;        /* gamma */
  float M;        /* M */
  float M2;       /* M ** 2 (Python float) */
  float gm1;      /* gamma - 1 */
  float gm1M2;    /* (gamma - 1) * M ** 2 */
  float i2;       /* 1 / M ** 2 */
  float d_gm1;    /* default gamma 1.4 - 1 */
  float g2;       /* 2 * gamma */
  int roe;
} consts;

static consts make_consts(const cf_opts* o) {
  consts c;
  c.g = (float)o->gamma;
  c.M = (float)o->M;
  c.M2 = (float)(o->M * o->M);
  c.gm1 = (float)(o->gamma - 1.0);
  c.gm1M2 = (float)((o->gamma - 1.0) * (o->M * o->M));
  c.i2 = (float)(1.0 / (o->M * o->M));
  c.d_gm1 = (float)(1.4 - 1.0);
  c.g2 = (float)(2.0 * o->gamma);
  c.roe = o->roe;
  return c;
}

/* Divisions of the step: exact IEEE quotients by default (numpy's); built
 * with -DCF_RCP_DIV, a / b becomes a * (1 / b), the reciprocal rounded first
 * -- the rounding of the GPU kernels, which multiply by reciprocals (of the
 * hardware rcp, of host-precomputed constants) -- so the two builds measure
 * what that rounding alone moves after many steps. */
#ifdef CF_RCP_DIV
#define FDIV(a, b) ((a) * (1.0f / (b)))
#else
#define FDIV(a, b) ((a) / (b))
#endif

/* getPrimitive (helper.py:37-46): P = (E - 0.5 rho (u^2+v^2)) * M^2 * (gamma-1) */
static inline void prim_of(const float* W, float m2, float gm1, float* p) {
  const float rho = W[0];
  const float u = FDIV(W[1], rho);
  const float v = FDIV(W[2], rho);
  const float P = ((W[3] - 0.5f * rho * (u * u + v * v)) * m2) * gm1;
  p[0] = rho; p[1] = u; p[2] = v; p[3] = P;
}

/* getConserved (helper.py:25-35): E = P / ((gamma-1) M^2) + 0.5 rho (u^2+v^2) */
static inline void cons_of(const float* p, float gm1M2, float* W) {
  const float rho = p[0], u = p[1], v = p[2], P = p[3];
  W[0] = rho;
  W[1] = rho * u;
  W[2] = rho * v;
  W[3] = FDIV(P, gm1M2) + 0.5f * rho * (u * u + v * v);
}

static inline int64_t wrap(int64_t j, int64_t n) { return j < 0 ? j + n : j; }

/* numpy's minimum / maximum: a NaN operand gives NaN */
static inline float np_min(float a, float b) { return (a != a || b != b) ? NAN : (b < a ? b : a); }
static inline float np_max(float a, float b) { return (a != a || b != b) ? NAN : (b > a ? b : a); }

/* BC_state (helper.py:233-256): walls in order, then inflow (3), outflow (4),
 * subsonic inlet (5); WR is the neighbour's state, WL the cell's own. */
static inline void bc_state(const cf_mesh* m, const cf_opts* o, const consts* c, int64_t i, int k, int64_t mk,
                            const float* WL, float* WR) {
  for (int w = 0; w < m->n_walls; ++w) {
    if (mk != m->walls[w]) continue;
    /* BC_slipwall (helper.py:196-204), default gamma 1.4 and M = 1 */
    float p[4];
    prim_of(WL, 1.0f, c->d_gm1, p);
    const float nx = m->normals[(i * 3 + k) * 2 + 0], ny = m->normals[(i * 3 + k) * 2 + 1];
    const float vn = (p[1] - 0.0f) * nx + (p[2] - 0.0f) * ny;
    const float vbx = (p[1] - 0.0f) - 2.0f * vn * nx;
    const float vby = (p[2] - 0.0f) - 2.0f * vn * ny;
    float pb[4] = {p[0], vbx + 0.0f, vby + 0.0f, p[3]};
    cons_of(pb, c->d_gm1, WR);
  }
  if (mk == 3) memcpy(WR, o->value, 16);
  if (mk == 4) memcpy(WR, WL, 16);
  if (mk == 5) {
    float p[4];
    prim_of(WL, 1.0f, c->d_gm1, p);
    const float rho = m->inlet_sub[0], u = m->inlet_sub[1], v = m->inlet_sub[2], P = p[3];
    WR[0] = rho;
    WR[1] = rho * u;
    WR[2] = rho * v;
    WR[3] = P / c->d_gm1 + 0.5f * rho * (u * u + v * v);
  }
}

static inline void midpoint(const cf_mesh* m, int64_t f, float* x) {
  const int64_t a = m->faces[2 * f], b = m->faces[2 * f + 1];
  x[0] = (m->points[2 * a] + m->points[2 * b]) / 2.0f;
  x[1] = (m->points[2 * a + 1] + m->points[2 * b + 1]) / 2.0f;
}

/* venkatakrishnan with K = 0 (euler_solver.py:36-39) */
static inline float venkat(float a, float b) {
  return FDIV(a * a + 2.0f * a * b + 0.0f, a * a + 2.0f * (b * b) + a * b + 0.0f + 1e-7f);
}

/* The cell's neighbour state after the boundary conditions (conserved). */
static inline void right_state(const cf_mesh* m, const cf_opts* o, const consts* c, const float* W, int64_t i, int k,
                               float* WR) {
  const int64_t j = wrap(m->nb[i * 3 + k], m->n);
  memcpy(WR, W + 4 * j, 16);
  bc_state(m, o, c, i, k, m->fmark[m->fc[i * 3 + k]], W + 4 * i, WR);
}

/* Pass 1: LSQ gradient (getgradientLSQ) and limiter (getlimiting) of a cell. */
static void grad_phi(const cf_mesh* m, const cf_opts* o, const consts* c, const float* W, int64_t i, float* grad,
                     float* phi) {
  float PL[4], PR[3][4], dx[3][2];
  prim_of(W + 4 * i, c->M2, c->gm1, PL);
  const float bx = m->bary[2 * i], by = m->bary[2 * i + 1];
  for (int k = 0; k < 3; ++k) {
    float WR[4];
    right_state(m, o, c, W, i, k, WR);
    prim_of(WR, c->M2, c->gm1, PR[k]);
    const int64_t f = m->fc[i * 3 + k];
    if (m->fmark[f] > 0) {
      float mid[2];
      midpoint(m, f, mid);
      dx[k][0] = 2.0f * (mid[0] - bx);
      dx[k][1] = 2.0f * (mid[1] - by);
    } else {
      const int64_t j = wrap(m->nb[i * 3 + k], m->n);
      dx[k][0] = m->bary[2 * j] - bx;
      dx[k][1] = m->bary[2 * j + 1] - by;
    }
  }
  float wdx[3][2], w[3];
  for (int k = 0; k < 3; ++k) {
    const float nrm = sqrtf(dx[k][0] * dx[k][0] + dx[k][1] * dx[k][1]);
    w[k] = 1.0f / (nrm * nrm);
    wdx[k][0] = w[k] * dx[k][0];
    wdx[k][1] = w[k] * dx[k][1];
  }
  float A[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) A[a][b] = (wdx[0][a] * dx[0][b] + wdx[1][a] * dx[1][b]) + wdx[2][a] * dx[2][b];
  for (int v = 0; v < 4; ++v) {
    float wdW[3];
    for (int k = 0; k < 3; ++k) wdW[k] = w[k] * (PR[k][v] - PL[v]);
    const float b0 = (wdW[0] * dx[0][0] + wdW[1] * dx[1][0]) + wdW[2] * dx[2][0];
    const float b1 = (wdW[0] * dx[0][1] + wdW[1] * dx[1][1]) + wdW[2] * dx[2][1];
    /* solve2x2: LU with partial pivoting */
    const int swap = fabsf(A[1][0]) > fabsf(A[0][0]);
    const float p00 = swap ? A[1][0] : A[0][0], p01 = swap ? A[1][1] : A[0][1];
    const float p10 = swap ? A[0][0] : A[1][0], p11 = swap ? A[0][1] : A[1][1];
    const float c0 = swap ? b1 : b0, c1 = swap ? b0 : b1;
    const float l = p10 / p00;
    const float u11 = p11 - l * p01;
    const float y1 = c1 - l * c0;
    const float x1 = y1 / u11;
    const float x0 = (c0 - p01 * x1) / p00;
    grad[v * 2 + 0] = x0;
    grad[v * 2 + 1] = x1;
  }
  /* getlimiting: min/max over the cell (three times) and its three neighbours */
  float mid[3][2];
  for (int k = 0; k < 3; ++k) {
    midpoint(m, m->fc[i * 3 + k], mid[k]);
    mid[k][0] = mid[k][0] - bx;
    mid[k][1] = mid[k][1] - by;
  }
  for (int v = 0; v < 4; ++v) {
    float mn = PL[v], mx = PL[v];
    for (int k = 0; k < 3; ++k) {
      mn = np_min(mn, PR[k][v]);
      mx = np_max(mx, PR[k][v]);
    }
    float p = 1.0f;
    for (int k = 0; k < 3; ++k) {
      const float D = mid[k][0] * grad[v * 2 + 0] + mid[k][1] * grad[v * 2 + 1];
      float ph = 1.0f;
      if (D > 1e-7f) ph = venkat(mx - PL[v], D);
      if (D < -1e-7f) ph = venkat(mn - PL[v], D);
      p = (k == 0) ? ph : np_min(p, ph);
    }
    phi[v] = p;
  }
}

/* HLLC dissipation (euler_solver.py:273-374): returns central - F_HLLC. */
static inline void hllc(const float* WL, const float* WR, float nx, float ny, const consts* c, float* out) {
  float L[4], R[4];
  prim_of(WL, c->M2, c->gm1, L);
  prim_of(WR, c->M2, c->gm1, R);
  const float rL = L[0], uL = L[1], vL = L[2], PL = L[3];
  const float rR = R[0], uR = R[1], vR = R[2], PR = R[3];
  const float unL = uL * nx + vL * ny;
  const float unR = uR * nx + vR * ny;
  const float tx = -ny, ty = nx;
  const float utL = uL * tx + vL * ty;
  const float utR = uR * tx + vR * ty;
  const float aL = FDIV(sqrtf(fabsf(FDIV(c->g * PL, rL))), c->M);
  const float aR = FDIV(sqrtf(fabsf(FDIV(c->g * PR, rR))), c->M);
  float SL = fminf(unL - aL, unR - aR);
  float SR = fmaxf(unL + aL, unR + aR);
  if (SL == SR) SL = SL - 1e-6f;
  if (SL == SR) SR = SR + 1e-6f;
  const float EL = WL[3], ER = WR[3];
  const float FL[4] = {rL * unL, rL * uL * unL + FDIV(PL * nx, c->M2), rL * vL * unL + FDIV(PL * ny, c->M2),
                       (EL + FDIV(PL, c->M2)) * unL};
  const float FR[4] = {rR * unR, rR * uR * unR + FDIV(PR * nx, c->M2), rR * vR * unR + FDIV(PR * ny, c->M2),
                       (ER + FDIV(PR, c->M2)) * unR};
  float den = rL * (SL - unL) - rR * (SR - unR);
  if (den == 0.0f) den = 1e-12f;
  const float Ss = FDIV(PR - PL + rL * unL * (SL - unL) - rR * unR * (SR - unR), den);
  const float PsL = PL + rL * (SL - unL) * (Ss - unL);
  const float PsR = PR + rR * (SR - unR) * (Ss - unR);
  const float Ps = 0.5f * (PsL + PsR);
  const float rsL = FDIV(rL * (SL - unL), SL - Ss);
  const float usxL = Ss * nx + utL * tx;
  const float usyL = Ss * ny + utL * ty;
  const float EsL = FDIV((SL - unL) * EL - PL * unL + Ps * Ss, SL - Ss);
  const float UsL[4] = {rsL, rsL * usxL, rsL * usyL, EsL};
  const float rsR = FDIV(rR * (SR - unR), SR - Ss);
  const float usxR = Ss * nx + utR * tx;
  const float usyR = Ss * ny + utR * ty;
  const float EsR = FDIV((SR - unR) * ER - PR * unR + Ps * Ss, SR - Ss);
  const float UsR[4] = {rsR, rsR * usxR, rsR * usyR, EsR};
  for (int v = 0; v < 4; ++v) {
    float F;
    if (SL >= 0.0f) F = FL[v];
    else if (SL <= 0.0f && Ss >= 0.0f) F = FL[v] + SL * (UsL[v] - WL[v]);
    else if (Ss <= 0.0f && SR >= 0.0f) F = FR[v] + SR * (UsR[v] - WR[v]);
    else F = FR[v];
    out[v] = 0.5f * (FL[v] + FR[v]) - F;
  }
}

/* Entropy variables (helper.py:57-70, with get_specific_entropy 48-55):
 * s = log(P / rho^gamma), V = [(gamma - s) / (gamma - 1) - rho |u|^2 / (2P),
 * rho u / P, rho v / P, -rho / P]. */
static inline void entropy_vars(const float* W, const consts* c, float* V) {
  const float rho = W[0];
  const float u = FDIV(W[1], rho);
  const float v = FDIV(W[2], rho);
  const float P = (W[3] - 0.5f * rho * (u * u + v * v)) * c->gm1;
  const float s = logf(FDIV(P, powf(rho, c->g)));
  V[0] = FDIV(c->g - s, c->gm1) - FDIV(0.5f * rho * (u * u + v * v), P);
  V[1] = FDIV(W[1], P);
  V[2] = FDIV(W[2], P);
  V[3] = FDIV(-rho, P);
}

/* Roe dissipation 0.5 R |Lambda S| R^T dv (euler_solver.py:145-199) with the
 * Roe average of helper.py:101-127 (densities and pressures <= 0 clamped to
 * 1e-6; pressures without the M^2 factor). */
static inline void roe(const float* WL, const float* WR, float nx, float ny, const consts* c, float* out) {
  float rs[2], us[2], vs[2], Es[2], Ps[2];
  const float* Wk[2] = {WL, WR};
  for (int q = 0; q < 2; ++q) {
    float rho = Wk[q][0];
    if (rho <= 0.0f) rho = 1e-6f;
    const float u = FDIV(Wk[q][1], rho);
    const float v = FDIV(Wk[q][2], rho);
    const float E = Wk[q][3];
    float P = (E - 0.5f * rho * (u * u + v * v)) * c->gm1;
    if (P <= 0.0f) P = 1e-6f;
    rs[q] = rho; us[q] = u; vs[q] = v; Es[q] = E; Ps[q] = P;
  }
  const float sL = sqrtf(rs[0]), sR = sqrtf(rs[1]);
  const float r = sL * sR;
  const float u = FDIV(sL * us[0] + sR * us[1], sL + sR);
  const float v = FDIV(sL * vs[0] + sR * vs[1], sL + sR);
  const float HL = FDIV(Es[0] + Ps[0], rs[0]);
  const float HR = FDIV(Es[1] + Ps[1], rs[1]);
  const float H = FDIV(sL * HL + sR * HR, sL + sR);
  const float a = sqrtf(fabsf(c->gm1 * (H - 0.5f * (u * u + v * v))));
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
  float VL[4], VR[4], dv[4];
  entropy_vars(WR, c, VR);
  entropy_vars(WL, c, VL);
  for (int q = 0; q < 4; ++q) dv[q] = VR[q] - VL[q];
  const float S[4] = {FDIV(r, c->g2), FDIV(c->gm1 * r, c->g), FDIV(r, c->g), FDIV(r, c->g2)};
  float DRH[4];
  for (int i = 0; i < 4; ++i) {
    const float RH = ((Rm[0][i] * dv[0] + Rm[1][i] * dv[1]) + Rm[2][i] * dv[2]) + Rm[3][i] * dv[3];
    DRH[i] = fabsf(lam[i] * S[i]) * RH;
  }
  for (int i = 0; i < 4; ++i)
    out[i] = 0.5f * ((((Rm[i][0] * DRH[0] + Rm[i][1] * DRH[1]) + Rm[i][2] * DRH[2]) + Rm[i][3] * DRH[3]));
}

/* getFlux (euler_solver.py:376-420): central flux minus the dissipation. */
static inline void face_flux(const float* WL, const float* WR, float nx, float ny, const consts* c, float* F) {
  float L[4], R[4];
  prim_of(WL, c->M2, c->gm1, L);
  prim_of(WR, c->M2, c->gm1, R);
  const float i2 = c->i2;
  const float unL = L[1] * nx + L[2] * ny, unR = R[1] * nx + R[2] * ny;
  const float frL = L[0] * unL, frR = R[0] * unR;
  const float fuL = L[0] * L[1] * unL + i2 * L[3] * nx, fuR = R[0] * R[1] * unR + i2 * R[3] * nx;
  const float fvL = L[0] * L[2] * unL + i2 * L[3] * ny, fvR = R[0] * R[2] * unR + i2 * R[3] * ny;
  const float fEL = (WL[3] + i2 * L[3]) * unL, fER = (WR[3] + i2 * R[3]) * unR;
  float D[4];
  if (c->roe)
    roe(WL, WR, nx, ny, c, D);
  else
    hllc(WL, WR, nx, ny, c, D);
  F[0] = 0.5f * (frL + frR) - D[0];
  F[1] = 0.5f * (fuL + fuR) - D[1];
  F[2] = 0.5f * (fvL + fvR) - D[2];
  F[3] = 0.5f * (fEL + fER) - D[3];
}

/* Sponge coefficients (helper.py:262-295): per-call constants. */
typedef struct { float ymin, ymax, width, strength; } sponge_c;

static sponge_c sponge_consts(const cf_mesh* m, const cf_opts* o, const consts* c) {
  sponge_c s;
  float ymin = m->points[1], ymax = m->points[1];
  for (int64_t p = 1; p < m->npts; ++p) {
    ymin = fminf(ymin, m->points[2 * p + 1]);
    ymax = fmaxf(ymax, m->points[2 * p + 1]);
  }
  const float Ly = ymax - ymin;
  s.ymin = ymin;
  s.ymax = ymax;
  s.width = o->has_width ? (float)o->sponge_width : 0.2f * Ly;
  float pinf[4];
  prim_of(o->value, c->M2, c->gm1, pinf);
  const float U = sqrtf(pinf[1] * pinf[1] + pinf[2] * pinf[2]);
  const float a = sqrtf(c->g * pinf[3] / pinf[0]) / c->M;
  s.strength = o->has_strength ? (float)o->sponge_strength : 1.0f * (U + a) / s.width;
  return s;
}

/* Pass 2: face states, fluxes and the residual of one cell. */
static void cell_residual(const cf_mesh* m, const cf_opts* o, const consts* c, const sponge_c* sp, const float* W,
                          const float* grad, const float* phi, int64_t i, float* R) {
  const float* Wi = W + 4 * i;
  float PL[4];
  prim_of(Wi, c->M2, c->gm1, PL);
  const float bx = m->bary[2 * i], by = m->bary[2 * i + 1];
  float Fs[3][4];
  for (int k = 0; k < 3; ++k) {
    float WR[4], WLc[4], WRc[4];
    right_state(m, o, c, W, i, k, WR);
    const int64_t f = m->fc[i * 3 + k];
    if (o->muscl) {
      float PR[4];
      prim_of(WR, c->M2, c->gm1, PR);
      const int64_t j = wrap(m->nb[i * 3 + k], m->n);
      const int bnd = m->fmark[f] > 1;
      float mid[2], midn[2];
      midpoint(m, f, mid);
      const float dx0 = mid[0] - bx, dx1 = mid[1] - by;
      midpoint(m, m->fco[f], midn);
      float dn0 = midn[0] - m->bary[2 * j], dn1 = midn[1] - m->bary[2 * j + 1];
      if (bnd) { dn0 = -dx0; dn1 = -dx1; }
      float qL[4], qR[4];
      for (int v = 0; v < 4; ++v) {
        const float D = dx0 * grad[(4 * i + v) * 2] + dx1 * grad[(4 * i + v) * 2 + 1];
        const float Dn = dn0 * grad[(4 * j + v) * 2] + dn1 * grad[(4 * j + v) * 2 + 1];
        const float phL = bnd ? 0.0f : phi[4 * i + v];
        const float phR = bnd ? 0.0f : phi[4 * j + v];
        qL[v] = PL[v] + phL * D;
        qR[v] = PR[v] + phR * Dn;
      }
      cons_of(qL, c->gm1M2, WLc);
      cons_of(qR, c->gm1M2, WRc);
    } else {
      memcpy(WLc, Wi, 16);
      memcpy(WRc, WR, 16);
    }
    const float nx = m->normals[(i * 3 + k) * 2], ny = m->normals[(i * 3 + k) * 2 + 1];
    float F[4];
    face_flux(WLc, WRc, nx, ny, c, F);
    const float s = m->surface[f];
    for (int v = 0; v < 4; ++v) Fs[k][v] = s * F[v];
  }
  float sig = 0.0f;
  if (!m->bump) {
    const float y = m->bary[2 * i + 1];
    const float dist = fminf(y - sp->ymin, sp->ymax - y);
    float blend = (sp->width - dist) / sp->width;
    blend = fminf(fmaxf(blend, 0.0f), 1.0f);
    sig = sp->strength * (blend * blend * blend);
  }
  for (int v = 0; v < 4; ++v) {
    const float F = (Fs[0][v] + Fs[1][v]) + Fs[2][v];
    const float src = m->bump ? 0.0f : sig * (o->value[v] - Wi[v]);
    R[v] = FDIV(F, m->area[i]) - src;
  }
}

/* residual(W) into R; grad (n*8) and phi (n*4) are caller scratch. */
void cf_residual(const cf_mesh* m, const cf_opts* o, const float* W, float* R, float* grad, float* phi) {
  const consts c = make_consts(o);
  const sponge_c sp = sponge_consts(m, o, &c);
  const int64_t n = m->n;
  if (o->muscl) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) grad_phi(m, o, &c, W, i, grad + 8 * i, phi + 4 * i);
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) cell_residual(m, o, &c, &sp, W, grad, phi, i, R + 4 * i);
}

/* n_steps SSP-RK2 steps (euler_solver.py:635-643) of W in place;
 * scratch: W1, k (4n each), grad (8n), phi (4n). */
void cf_srk2(const cf_mesh* m, const cf_opts* o, float* W, float dt, int64_t n_steps, float* W1, float* k,
             float* grad, float* phi) {
  const int64_t n4 = 4 * m->n;
  for (int64_t s = 0; s < n_steps; ++s) {
    cf_residual(m, o, W, k, grad, phi);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n4; ++e) W1[e] = W[e] - dt * k[e];
    cf_residual(m, o, W1, k, grad, phi);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n4; ++e) W[e] = 0.5f * (W[e] + W1[e] - dt * k[e]);
  }
}

/* get_dt (helper.py:8-18): min over cells of area / sum(lambda len), x CFL. */
float cf_get_dt(const cf_mesh* m, const float* W, double CFL, double gamma, double M) {
  cf_opts o;
  memset(&o, 0, sizeof o);
  o.gamma = gamma;
  o.M = M;
  const consts c = make_consts(&o);
  float best = INFINITY;
#pragma omp parallel for reduction(min : best) schedule(static)
  for (int64_t i = 0; i < m->n; ++i) {
    float p[4];
    prim_of(W + 4 * i, c.M2, c.gm1, p);
    const float cs = sqrtf(fabsf(c.g * p[3] / p[0])) / c.M;
    float t[3];
    for (int k = 0; k < 3; ++k) {
      const float un = p[1] * m->normals[(i * 3 + k) * 2] + p[2] * m->normals[(i * 3 + k) * 2 + 1];
      t[k] = (cs + fabsf(un)) * m->surface[m->fc[i * 3 + k]];
    }
    const float d = m->area[i] / ((t[0] + t[1]) + t[2]);
    best = fminf(best, d);
  }
  return best * (float)CFL;
}

void cf_set_threads(int n) { omp_set_num_threads(n); }
int cf_max_threads(void) { return omp_get_max_threads(); }
