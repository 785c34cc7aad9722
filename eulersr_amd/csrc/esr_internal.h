// Internal definitions shared by the host mesh preparation (esr_mesh.cpp) and
// the device code (esr_kernels.hip). Not part of the public ABI.
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/eulersr.h"

namespace esr {

// Per (cell, local face) reference word.
//   interior / periodic : bits 30-31 = local slot of the shared face in the
//                         neighbour, bit 29 = periodic, bits 0-28 = neighbour
//   boundary            : bits 30-31 = 3, bits 27-29 = ghost kind,
//                         bits 0-26 = cell read by the WRAP kind
constexpr uint32_t FREF_BND = 3u << 30;
constexpr uint32_t FREF_PER = 1u << 29;
constexpr uint32_t FREF_IDX = (1u << 29) - 1;
constexpr uint32_t FREF_BIDX = (1u << 27) - 1;
constexpr int FREF_KIND_SHIFT = 27;

// Ghost states of helper.BC_state (helper.py:233-256), after the precedence
// walls < inflow(3) < outflow(4) < subsonic(5). WRAP: a boundary marker
// with no condition keeps W[neighbors] (the last cell for -1, euler_solver.py:108).
enum Ghost : uint32_t { G_WALL = 0, G_INLET = 1, G_OUTLET = 2, G_SUBSONIC = 3, G_WRAP = 4 };

__host__ __device__ inline bool fref_is_bnd(uint32_t f) { return (f & FREF_BND) == FREF_BND; }

// Device view of one renumbered mesh (structure of arrays).
//
// Cell arrays are indexed by the renumbered cell, [k][n] arrays by local
// face k then cell. Faces are "logical faces": one per interior face pair,
// one per periodic face pair, one per boundary face; each is owned by the
// first cell (in renumbered order) that touches it and carries the owner's
// outward normal. A cell reads the flux of its local face k from
// cface[k][i] (bit 31 set: the cell is the partner and takes -F).
struct MeshDev {
  int32_t n;             // cells
  int32_t nf;            // logical faces
  int32_t n_wall;        // force-marker faces
  // reconstruction (per cell)
  const uint32_t* fref;  // [3][n] neighbour word (layout above)
  const float2* dxm;     // [3][n] face midpoint - barycentre (float32, reference order)
  const float2* wdl;     // [3][n] LSQ offset scaled by its weight, w_k * dx_k
  const float4* ainv;    // [n]    inverse of the LSQ normal matrix (a00, a01, a10, a11)
  const float2* nrm;     // [3][n] outward normals (wall ghosts)
  // flux (per logical face)
  const uint32_t* f_own;  // [nf] (slot << 30) | owner cell
  const uint32_t* f_ref;  // [nf] partner word of the owner's side (fref layout)
  const float2* f_nrm;    // [nf] owner's outward normal
  const float* f_len;     // [nf] face length
  // update (per cell)
  const uint32_t* cface;  // [3][n] logical face id | sign bit
  const float* inv_area;  // [n]
  // time step (helper.get_dt)
  const float* len;       // [3][n] face lengths seen from the cell
  const float* area;      // [n]
  const float* sponge_g;  // [n] blend^3 for the default sponge width
  const int32_t* wall_cell;  // [n_wall]
  const float* wall_nx;      // [n_wall]
  const float* wall_ny;      // [n_wall]
  const float* wall_ds;      // [n_wall]
  float4 inlet_subsonic;     // primitives (helper.py:181)
};
constexpr uint32_t CFACE_NEG = 1u << 31;

// Physical constants of a group, derived in double precision on the host and
// rounded to float32 once (as Python floats meet float32 arrays in the
// reference: weak typing casts the double result).
struct Phys {
  float g;          // gamma
  float gm1;        // gamma - 1
  float M2;         // M**2
  float gm1M2;      // (gamma - 1) * M**2
  float invM;       // 1 / M
  float invM2;      // 1 / M**2
  float two_g;      // 2 * gamma
  float c_ir1;      // (gamma + 1) / (2 gamma)
  float c_ir2;      // (gamma - 1) / (2 gamma)
  float one_m_g;    // 1 - gamma
  float k_exp;      // 1 / (1 - gamma)
  float inv_gm1M2;  // 1 / ((gamma - 1) M**2)
  float inv_gm1;    // 1 / (gamma - 1)
  float inv_g;      // 1 / gamma
  float inv_2g;     // 1 / (2 gamma)
};

// A group: C cases sharing one mesh and one kernel configuration. The state
// of a group is stored case-innermost: element (cell, c) at st_off + cell*C + c.
struct GroupDev {
  MeshDev m;
  int32_t C;
  uint32_t cmagic;       // t / C = (h + ((t - h) >> 1)) >> (cshift - 1), h = umulhi(cmagic, t); C > 1
  int32_t cshift;
  int32_t case0;         // global index of the group's first case
  int64_t st_off;        // float4 offset of the group's state slab
  int64_t rec_off;       // float4 offset of the group's reconstruction slab ([3][n][C])
  int64_t fx_off;        // float4 offset of the group's face-flux slab ([nf][C])
  int32_t flux;          // esr_flux of the group
  Phys ph;
  int32_t sponge;
  const float* sponge_g; // mesh default or a custom width
};

// One entry of a launch's block table: the group a block works on and the
// block's index inside that group (group -1: padding). Any subset of groups
// can form a launch: the flux kernels launch per flux class, the
// reconstruction and the update once over all groups.
struct BlockRef {
  int32_t group;
  int32_t block;
};

// One block of a reduction launch over several groups: the group, its cell
// range (blockIdx.x of a one-group launch) and case chunk (blockIdx.y).
struct RedRef {
  int32_t group;
  int32_t range;
  int32_t chunk;
  int32_t pad;
};

// Per-case parameters read by the stage kernels.
struct CaseDev {
  float4 value;          // inlet conserved state
  float dt, dt_half, dt_sixth;
  float strength;        // sponge strength
  uint32_t lf_cmax;      // float bits of the Lax-Friedrichs global wave speed
  int32_t active;        // 0 once stopped or past N
  int32_t check;         // 1 when the current step is a stationarity check
  float cfl_local;       // > 0: local time stepping with this CFL (each cell its own dt); 0: global dt
};

// Per-case solve-loop state (device), main.py:141-199.
struct CaseLoop {
  int64_t N;
  int64_t ref_step;
  int64_t stopping_step;
  int64_t streak_start;
  int64_t eng_streak_start;
  int64_t eng_stopping_step;
  int64_t steps_run;
  double last_rel;
  double threshold;
  double q_ref;
  double target_cd, target_cl, cd_tol, cl_tol;
  double cd, cl;
  int32_t streak;
  int32_t patience;
  int32_t stop;
  int32_t done;
  int32_t track_eng;
  int32_t eng_streak;
  int32_t eng_patience;
  int32_t eng_stop;
  int32_t diverged;
  int32_t n_checks;
  double recent[16];  // check residuals, ring buffer indexed by n_checks % 16
  int64_t start;      // solver step at which the case entered its slot (replaced cases)
};
constexpr int32_t RECENT_CHECKS = 16;

enum StageMode : int {
  MODE_RESID = 0,   // out = R
  MODE_EE = 1,      // out = W0 - dt R
  MODE_S1 = 2,      // out = W0 - dt R                 (SRK2 / RK2 stage 1)
  MODE_SRK2_2 = 3,  // out = 0.5 ((W0 + W1) - dt R)
  MODE_RK2_2 = 4,   // out = 0.5 ((W0 - dt R) + W1)
  MODE_RK4_1 = 5,   // out = W0 - dt/2 R ; acc = R
  MODE_RK4_2 = 6,   // out = W0 - dt/2 R ; acc += 2R
  MODE_RK4_3 = 7,   // out = W0 - dt R   ; acc += 2R
  MODE_RK4_4 = 8    // out = W0 - dt/6 (acc + R)
};

struct Error : std::runtime_error {
  esr_status code;
  Error(esr_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(ESR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Device allocation owned by a handle.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(ESR_ERR_MEMORY, std::string("hipMalloc(") + std::to_string(b) + "): " + hipGetErrorString(e));
    }
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
void upload(DevBuf& b, const std::vector<T>& v) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) hip_check(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

}  // namespace esr

// Opaque mesh handle (host side).
struct esr_mesh {
  int32_t n = 0, n_faces = 0, n_points = 0;
  std::vector<int32_t> perm;    // perm[new] = old
  std::vector<int32_t> iperm;   // iperm[old] = new
  std::vector<float> bary_y;    // barycentre y, new order (custom sponge widths)
  float y_min = 0.f, y_max = 0.f;
  int32_t sponge = 1;
  esr::MeshDev dev{};
  esr::DevBuf b_fref, b_dxm, b_wdl, b_ainv, b_nrm, b_fown, b_fref2, b_fnrm, b_flen, b_cface, b_iarea, b_sponge,
      b_perm, b_len, b_area;
  esr::DevBuf b_wcell, b_wnx, b_wny, b_wds;
  int64_t device_bytes = 0;
  struct Workspace* ws = nullptr;  // single-state scratch (esr_kernels.hip)
  ~esr_mesh();
};

namespace esr {
// esr_mesh.cpp
esr_mesh* build_mesh(const esr_mesh_desc* d);
std::vector<int32_t> rcm_order(int32_t n, const int32_t* neighbors, const int32_t* seed_order);
std::vector<float> sponge_weights(const esr_mesh* m, float width);
void set_last_error(const std::string& s);
const char* last_error_cstr();
void free_workspace(Workspace* ws);
}  // namespace esr
