// Launch wrappers of the gfx950 kernels (esr_kernels.hip), called by the
// host orchestration (esr_api.cpp). Internal.
#pragma once

#include <vector>

#include "esr_internal.h"

namespace esr {

// Launch class of the flux kernels holding HLLC and Roe groups together: the
// flux is chosen per group at run time (uniform over a block), so both
// run in one launch.
constexpr int32_t FLUX_HLLC_ROE = 100;

// A launch covers a set of groups with the same kernel configuration.
struct LaunchSet {
  const GroupDev* groups = nullptr;   // device array
  const BlockRef* blk_group = nullptr;  // device array: cell block -> (group, block in group)
  int32_t nblocks = 0;
  const BlockRef* fblk_group = nullptr; // device array: face block -> (group, block in group)
  int32_t fnblocks = 0;
  int32_t flux = ESR_FLUX_HLLC;
  int32_t recon = ESR_RECON_MUSCL;
  int32_t entropy = 0;
  int32_t block = 256;  // threads per block of the cell and face kernels (64, 128 or 256)
  int32_t xcd = 1;      // XCD-contiguous block order (the flag after each block table)
};

constexpr int BLOCK = 256;  // maximum block of the stage kernels; block of the reductions

// Host helpers (esr_api.cpp): physical constants of a group, option checks,
// the sponge strength, the stage sequence of a time scheme, the per-case
// parameters of one case, the stage path of engines built from now on.
Phys make_phys(double gamma, double M);
void validate_opts(const esr_opts* o);
float sponge_strength(const esr_opts& o, float width);
struct Stage {
  int mode;
  int ws;   // input state: 0 W0, 1 W1, 2 W2
  int out;  // output state: 0 W0, 1 W1, 2 W2
};
std::vector<Stage> stages_of(int scheme);
CaseDev case_params(const esr_mesh* m, const esr_opts& o);
int32_t resolved_stage_path();

void launch_recon(const LaunchSet& ls, const CaseDev* cases, const float4* Ws, float4* rec, hipStream_t st);
void launch_lf_cmax(const LaunchSet& ls, const CaseDev* cases, const float4* Ws, const float4* rec,
                    hipStream_t st);
void launch_face(const LaunchSet& ls, const CaseDev* cases, const float4* Ws, const float4* rec, float4* fx,
                 hipStream_t st);
// local_dt: every cell with its own dt (CaseDev.cfl_local; the cell-centric
// kernel has no such variant, so local time stepping runs the face path)
void launch_update(const LaunchSet& ls, int mode, bool local_dt, const CaseDev* cases, const float4* Ws,
                   const float4* W0, const float4* fx, float4* out, float4* acc, hipStream_t st);
// cell-centric stage (face fluxes of each cell + residual + update in one
// launch); the alternative to launch_face + launch_update
void launch_cell(const LaunchSet& ls, int mode, const CaseDev* cases, const float4* Ws, const float4* W0,
                 const float4* rec, float4* out, float4* acc, hipStream_t st);

// Per-case reductions over one group (host-side group descriptor `g`,
// device copy at `gdev`). Partial results at part[case * rstride + r].
int32_t reduce_ranges(const GroupDev& g);
// Arguments of a stationarity check fused into the last norms launch.
struct CheckArgs {
  int32_t n_cases;
  CaseDev* cases;
  CaseLoop* loops;
  const int32_t* nranges;
  const double* forces;
  int64_t step;
  int64_t* dstep;
  uint32_t* counter;  // zero between launches; null: no fused check
};
// Block tables of the reductions over a set of groups (host side).
void red_table(const GroupDev& g, int32_t gi, std::vector<RedRef>& out);
// One launch over every block of `tbl` (n entries): stationarity norms of
// each case (and the fused check on the launch's last block), or the FP64
// wall force sums (one block per case: BlockRef {group, case in group}).
void launch_norms(const GroupDev* groups, const RedRef* tbl, int32_t n, const CaseDev* cases, const float4* Wn,
                  float4* Wref, double* part_d, double* part_r, int32_t rstride, int copy_ref, hipStream_t st,
                  const CheckArgs* fused = nullptr);
void launch_dt(const GroupDev& g, const GroupDev* gdev, const CaseDev* cases, const float4* W, float* part,
               int32_t rstride, int local, hipStream_t st);
void launch_forces(const GroupDev* groups, const BlockRef* tbl, int32_t n, const CaseDev* cases, const float4* W,
                   double* cd_cl, hipStream_t st);

// Finalisation of a stationarity check for every case (main.py:149-177).
void launch_check(int32_t n_cases, CaseDev* cases, CaseLoop* loops, const double* part_d,
                  const double* part_r, const int32_t* nranges, int32_t rstride, const double* forces,
                  int64_t step, int32_t K, hipStream_t st, int64_t* dstep = nullptr);
// dt finalisation: dt = min * cfl; writes dt, dt/2, dt/6 into the case params.
void launch_dt_final(int32_t n_cases, CaseDev* cases, const float* part, const int32_t* nranges,
                     int32_t rstride, const float* cfl, float* dt_out, hipStream_t st);

// Layout conversion between a user array (original order, (n,4)) and a
// group slab (renumbered, case-innermost).
void launch_gather_in_dev(int32_t n, const GroupDev* gdev, int32_t c, const int32_t* perm, const float4* user,
                          float4* slab, hipStream_t st);
void launch_scatter_out_dev(int32_t n, const GroupDev* gdev, int32_t c, const int32_t* perm, const float4* slab,
                            float4* user, hipStream_t st);
// stage-ordered parameter updates (kernel arguments are copied at launch)
void launch_set_case(CaseDev* dst, const CaseDev& v, hipStream_t st);
void launch_set_group(GroupDev* dst, const GroupDev& v, hipStream_t st);
// check flags: all == 1 marks every case, otherwise active cases with
// step % K == 0 or step == N.
// dstep != nullptr: the step is dstep[0] + K (graph replays) and is
// recorded in dstep[1]; k_check then reads dstep[1] and stores it in dstep[0].
void launch_mark_check(int32_t n_cases, CaseDev* cases, const CaseLoop* loops, int64_t step, int32_t K, int32_t all,
                       hipStream_t st, int64_t* dstep = nullptr);
void launch_set_step(int64_t* dstep, int64_t v, hipStream_t st);
// final one-step probe residual (main.py:228-230)
void launch_probe_final(int32_t n_cases, CaseDev* cases, const double* part_d, const double* part_r,
                        const int32_t* nranges, int32_t rstride, double* out, hipStream_t st);
// commit out -> W0 for the active work-items of a launch set (explicit Euler)
void launch_commit(const LaunchSet& ls, const CaseDev* cases, const float4* src, float4* dst, hipStream_t st);
void launch_copy(const float4* src, float4* dst, int64_t n, hipStream_t st);
// exhaustive check of the flux kernels' square root and logarithm against
// sqrtf / logf; counts: 4 zeroed device uint64 (see esr_check_device_math)
void launch_math_check(unsigned long long* counts, hipStream_t st);
void launch_set_active(int32_t n_cases, CaseDev* cases, const CaseLoop* loops, int32_t all, hipStream_t st);
void launch_zero_lf(int32_t n_cases, CaseDev* cases, hipStream_t st);
void launch_count_active(int32_t n_cases, const CaseDev* cases, int32_t* count, hipStream_t st);

}  // namespace esr
