// Host orchestration and the C ABI (include/eulersr.h).
//
// An Engine holds the device state of a set of cases: cases sharing a mesh
// and a kernel configuration form a group stored case-innermost; groups with
// the same kernel configuration are launched together. The single-state
// entry points (esr_residual, esr_step, ...) use a one-case Engine cached on
// the mesh; the batched solver (esr_solver_*) owns its own Engine.
//
// Solve loop semantics follow euler/main.py:116-230: global dt computed once
// from the initial state, N = int(tf/dt) + 1, a stationarity check every
// K steps and at step N with patience, states frozen once stopped, a final
// one-step probe residual.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <tuple>

#include "esr_launch.h"

using namespace esr;

// Helpers shared with the allocation-free entry point (esr_noalloc.hip);
// their float arithmetic is compiled here, without contraction.
namespace esr {

Phys make_phys(double gamma, double M) {
  Phys p;
  p.g = (float)gamma;
  p.gm1 = (float)(gamma - 1.0);
  p.M2 = (float)(M * M);
  p.gm1M2 = (float)((gamma - 1.0) * (M * M));
  p.invM = (float)(1.0 / M);
  p.invM2 = (float)(1.0 / (M * M));
  p.two_g = (float)(2.0 * gamma);
  p.c_ir1 = (float)((gamma + 1.0) / (2.0 * gamma));
  p.c_ir2 = (float)((gamma - 1.0) / (2.0 * gamma));
  p.one_m_g = (float)(1.0 - gamma);
  p.k_exp = (float)(1.0 / (1.0 - gamma));
  p.inv_gm1M2 = (float)(1.0 / ((gamma - 1.0) * (M * M)));
  p.inv_gm1 = (float)(1.0 / (gamma - 1.0));
  p.inv_g = (float)(1.0 / gamma);
  p.inv_2g = (float)(1.0 / (2.0 * gamma));
  return p;
}

void validate_opts(const esr_opts* o) {
  if (!o) throw Error(ESR_ERR_VALUE, "null options");
  if (o->flux < ESR_FLUX_LF || o->flux > ESR_FLUX_ROE)
    throw Error(ESR_ERR_VALUE, "Unknown flux type. Expected one of: 'lf', 'rusanov', 'tadmor', 'ausm', 'hll', 'hllc', 'roe'.");
  if (o->recon != ESR_RECON_CONSTANT && o->recon != ESR_RECON_MUSCL)
    throw Error(ESR_ERR_VALUE, "Unknown reconstruction scheme. Expected 'constant' or 'muscl'.");
  if (!(o->gamma > 1.0)) throw Error(ESR_ERR_VALUE, "gamma must be > 1");
  if (!(o->M > 0.0)) throw Error(ESR_ERR_VALUE, "M must be > 0");
}

// Sponge strength of helper.py:279-285 in float32: (|U_inf| + a_inf) / width
// from the primitives of the inlet state.
float sponge_strength(const esr_opts& o, float width) {
  const Phys ph = make_phys(o.gamma, o.M);
  const float r = o.value[0];
  const float u = o.value[1] / r;
  const float v = o.value[2] / r;
  const float P = (o.value[3] - 0.5f * r * (u * u + v * v)) * ph.M2 * ph.gm1;
  const float U = std::sqrt(u * u + v * v);
  const float a = std::sqrt(ph.g * P / r) * ph.invM;
  return 1.0f * (U + a) / width;
}

std::vector<Stage> stages_of(int scheme) {
  switch (scheme) {
    case ESR_SCHEME_EE: return {{MODE_EE, 0, 1}};
    case ESR_SCHEME_SRK2: return {{MODE_S1, 0, 1}, {MODE_SRK2_2, 1, 0}};
    case ESR_SCHEME_RK2: return {{MODE_S1, 0, 1}, {MODE_RK2_2, 1, 0}};
    case ESR_SCHEME_RK4:
      return {{MODE_RK4_1, 0, 1}, {MODE_RK4_2, 1, 2}, {MODE_RK4_3, 2, 1}, {MODE_RK4_4, 1, 0}};
    default: throw Error(ESR_ERR_KEY, "unknown time scheme (expected EE, RK2, RK4, SRK2, SSP_RK2)");
  }
}

// parameters of a case from its options (dt set later)
CaseDev case_params(const esr_mesh* m, const esr_opts& o) {
  CaseDev cd{};
  cd.value = make_float4(o.value[0], o.value[1], o.value[2], o.value[3]);
  const float width = o.sponge_width > 0 ? o.sponge_width : 0.2f * (m->y_max - m->y_min);
  cd.strength = o.sponge_strength > 0 ? o.sponge_strength : sponge_strength(o, width);
  cd.dt = cd.dt_half = cd.dt_sixth = 0.f;
  cd.active = 1;
  return cd;
}

}  // namespace esr

namespace {

struct CaseSpec {
  const esr_mesh* mesh;
  esr_opts opts;
  int32_t tag = 0;  // esr_case.group: different tags never share a group
};

}  // namespace

// Device state of a set of cases.
// autotuned launch choices by launch shape (esr_solver_autotune)
struct TuneChoice {
  int32_t path = ESR_STAGE_CELL, block = 256, xcd = 1;
  double ms = 0.0;
};
std::mutex g_tune_mu;
std::map<std::vector<int64_t>, TuneChoice> g_tune;

// stage path of engines built from now on (esr_set_stage_path)
std::atomic<int32_t> g_stage_path{ESR_STAGE_AUTO};
int32_t esr::resolved_stage_path() {
  const int32_t p = g_stage_path.load();
  return p == ESR_STAGE_FACE ? ESR_STAGE_FACE : ESR_STAGE_CELL;
}

struct Engine {
  int32_t n_cases = 0;
  int scheme = ESR_SCHEME_SRK2;
  int32_t stage_path = ESR_STAGE_CELL;
  int32_t cur_block = BLOCK, cur_xcd = 1;
  // local time stepping (esr_solve_opts.local_dt): the update kernel's local
  // variant, on the face path (the cell-centric kernel has none)
  bool local_dt = false;
  std::vector<uint8_t> group_on;      // empty: every group launched
  std::vector<GroupDev> groups;       // host copies
  std::vector<const esr_mesh*> gmesh; // mesh of each group
  std::vector<int32_t> case_group;    // global case -> group
  std::vector<int32_t> case_slot;     // global case -> c within group
  std::vector<CaseDev> cases_h;       // host copies of the case params
  DevBuf groups_d, cases_d, loops_d;
  struct Cls {
    LaunchSet ls;
    DevBuf blk, fblk;
    std::vector<int32_t> gids;  // groups of the class
  };
  std::vector<std::unique_ptr<Cls>> classes;  // per (flux, recon, entropy): the flux kernels
  Cls rec_all;   // every MUSCL group: the reconstruction (flux-independent), one launch
  Cls cell_all;  // every group: the update and the commit (flux-independent), one launch
  DevBuf red_tbl, frc_tbl;        // reduction block tables over the launched groups
  int32_t n_red = 0, n_frc = 0;

  // norms / forces block tables of the launched groups
  void reduction_tables() {
    std::vector<RedRef> red;
    std::vector<BlockRef> frc;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      if (!on(gi)) continue;
      red_table(groups[gi], (int32_t)gi, red);
      for (int32_t c = 0; c < groups[gi].C; ++c) frc.push_back({(int32_t)gi, c});
    }
    upload(red_tbl, red);
    upload(frc_tbl, frc);
    n_red = (int32_t)red.size();
    n_frc = (int32_t)frc.size();
  }
  bool any_muscl = false, any_lf = false;
  int64_t elems = 0, rec_elems = 0, fx_elems = 0;
  DevBuf W0, W1, W2, acc, Wref, rec, fx;
  int32_t rstride = 1;
  DevBuf part_d, part_r, part_f, nranges, cfl_d, dt_d, forces_d, probe_d, count_d;
  DevBuf counter_d;  // uint32 blocks-done counter of the fused check (zero between launches)
  std::vector<std::unique_ptr<DevBuf>> custom_sponge;
  int32_t* count_h = nullptr;  // pinned

  ~Engine() {
    if (count_h) (void)hipHostFree(count_h);
  }

  // cell / face block tables of a class for `block` threads per block
  // (host group copies updated; the caller uploads groups_d)
  void block_tables(Cls& c, int32_t block, int32_t xcd = 1) {
    std::vector<BlockRef> blk, fblk;
    for (int32_t gi : c.gids) {
      if (!group_on.empty() && !group_on[gi]) continue;  // every case of the group is done
      const GroupDev& g = groups[gi];
      const int64_t nblk = ((int64_t)g.m.n * g.C + block - 1) / block;
      for (int32_t b = 0; b < (int32_t)nblk; ++b) blk.push_back({gi, b});
      const int64_t fnblk = ((int64_t)g.m.nf * g.C + block - 1) / block;
      for (int32_t b = 0; b < (int32_t)fnblk; ++b) fblk.push_back({gi, b});
    }
    // launches cover a multiple of 8 blocks (one run per XCD); padding group -1
    while (blk.size() % 8) blk.push_back({-1, 0});
    while (fblk.size() % 8) fblk.push_back({-1, 0});
    const int32_t nb = (int32_t)blk.size(), nfb = (int32_t)fblk.size();
    upload(c.blk, blk);
    upload(c.fblk, fblk);
    c.ls.blk_group = c.blk.as<BlockRef>();
    c.ls.nblocks = nb;
    c.ls.fblk_group = c.fblk.as<BlockRef>();
    c.ls.fnblocks = nfb;
    c.ls.block = block;
    c.ls.xcd = xcd ? 1 : 0;
  }

  // one launch configuration for every class: stage path and block size
  void configure(int32_t path, int32_t block, int32_t xcd, hipStream_t st) {
    hip_check(hipStreamSynchronize(st), "configure sync");
    for (auto& c : classes) block_tables(*c, block, xcd);
    block_tables(rec_all, block, xcd);
    block_tables(cell_all, block, xcd);
    reduction_tables();
    stage_path = local_dt ? ESR_STAGE_FACE : path;
    cur_block = block;
    cur_xcd = xcd;
  }

  // launches cover only the groups with mask[g] = 1 (continuous batching:
  // a group whose cases are all done is skipped); true when it changed
  bool set_group_mask(const std::vector<uint8_t>& mask, hipStream_t st) {
    if (mask == group_on) return false;
    group_on = mask;
    configure(stage_path, cur_block, cur_xcd, st);
    return true;
  }
  bool on(size_t gi) const { return group_on.empty() || group_on[gi]; }

  float4* buf(int which) {
    return which == 0 ? W0.as<float4>() : which == 1 ? W1.as<float4>() : W2.as<float4>();
  }

  // Build groups from case specs. Cases with the same (mesh, flux, recon,
  // entropy, gamma, M, sponge width) share a group; groups with the same
  // (flux, recon, entropy) share a launch.
  void build(const std::vector<CaseSpec>& specs, int scheme_) {
    scheme = scheme_;
    stage_path = resolved_stage_path();
    (void)stages_of(scheme);
    n_cases = (int32_t)specs.size();
    if (n_cases <= 0) throw Error(ESR_ERR_VALUE, "no cases");
    using Key = std::tuple<const esr_mesh*, int, int, int, double, double, float, int32_t>;
    std::map<Key, std::vector<int32_t>> gm;
    std::vector<Key> order;
    for (int32_t i = 0; i < n_cases; ++i) {
      const auto& s = specs[i];
      if (!s.mesh) throw Error(ESR_ERR_VALUE, "case without a mesh");
      validate_opts(&s.opts);
      Key k{s.mesh, s.opts.flux, s.opts.recon, s.opts.entropy ? 1 : 0, s.opts.gamma, s.opts.M,
            s.opts.sponge_width > 0 ? s.opts.sponge_width : -1.0f, s.tag};
      if (!gm.count(k)) order.push_back(k);
      gm[k].push_back(i);
    }
    // global case order: group by group
    std::vector<int32_t> user_of_global;
    int32_t case0 = 0;
    int64_t st = 0, rc = 0, fo = 0;
    for (const Key& k : order) {
      const auto& members = gm[k];
      const esr_mesh* m = std::get<0>(k);
      GroupDev g{};
      g.m = m->dev;
      g.C = (int32_t)members.size();
      // magic number of the unsigned division by C (round-up method)
      if (g.C > 1) {
        int32_t l = 0;
        while ((1u << l) < (uint32_t)g.C) ++l;
        g.cshift = l;
        g.cmagic = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - (uint64_t)g.C)) / (uint64_t)g.C + 1);
      } else {
        g.cshift = 0;
        g.cmagic = 0;
      }
      // in-group offsets are 32-bit byte offsets: every slab of the group
      // (state n C, reconstruction 3 n C, face flux nf C float4 elements)
      // must stay below 4 GiB, i.e. (3 n + nf) C < 2^28
      if ((3 * (int64_t)m->n + (int64_t)m->dev.nf) * g.C >= ((int64_t)1 << 28))
        throw Error(ESR_ERR_VALUE, "too many cases on one mesh with one kernel configuration for one solver");
      g.case0 = case0;
      g.st_off = st;
      g.ph = make_phys(std::get<4>(k), std::get<5>(k));
      g.sponge = m->sponge;
      const float width = std::get<6>(k);
      if (width > 0) {
        auto b = std::make_unique<DevBuf>();
        upload(*b, sponge_weights(m, width));
        g.sponge_g = b->as<float>();
        custom_sponge.push_back(std::move(b));
      } else {
        g.sponge_g = m->dev.sponge_g;
      }
      st += (int64_t)m->n * g.C;
      g.flux = std::get<1>(k);
      g.fx_off = fo;
      fo += (int64_t)m->dev.nf * g.C;
      if (std::get<2>(k) == ESR_RECON_MUSCL) {
        g.rec_off = rc;
        rc += 3 * (int64_t)m->n * g.C;
        any_muscl = true;
      } else {
        g.rec_off = 0;
      }
      if (std::get<1>(k) == ESR_FLUX_LF) any_lf = true;
      for (int32_t c = 0; c < g.C; ++c) {
        user_of_global.push_back(members[c]);
        case_group.push_back((int32_t)groups.size());
        case_slot.push_back(c);
      }
      case0 += g.C;
      groups.push_back(g);
      gmesh.push_back(m);
    }
    elems = st;
    rec_elems = rc;
    fx_elems = fo;
    // launch classes of the flux kernels: HLLC and Roe groups share one
    // (the flux is picked per group inside the kernel)
    std::map<std::tuple<int, int, int>, std::vector<int32_t>> cls;
    for (size_t gi = 0; gi < order.size(); ++gi) {
      const Key& k = order[gi];
      const int fl = std::get<1>(k);
      const int lc = (fl == ESR_FLUX_HLLC || fl == ESR_FLUX_ROE) ? FLUX_HLLC_ROE : fl;
      cls[{lc, std::get<2>(k), std::get<3>(k)}].push_back((int32_t)gi);
    }
    for (auto& kv : cls) {
      auto c = std::make_unique<Cls>();
      c->gids = kv.second;
      c->ls.flux = std::get<0>(kv.first);
      c->ls.recon = std::get<1>(kv.first);
      c->ls.entropy = std::get<2>(kv.first);
      block_tables(*c, BLOCK);
      classes.push_back(std::move(c));
    }
    for (size_t gi = 0; gi < order.size(); ++gi) {
      if (std::get<2>(order[gi]) == ESR_RECON_MUSCL) rec_all.gids.push_back((int32_t)gi);
      cell_all.gids.push_back((int32_t)gi);
    }
    rec_all.ls.recon = ESR_RECON_MUSCL;
    block_tables(rec_all, BLOCK);
    block_tables(cell_all, BLOCK);
    upload(groups_d, groups);
    for (auto& c : classes) c->ls.groups = groups_d.as<GroupDev>();
    rec_all.ls.groups = groups_d.as<GroupDev>();
    cell_all.ls.groups = groups_d.as<GroupDev>();
    reduction_tables();

    // case params (user order -> global order)
    cases_h.assign(n_cases, CaseDev{});
    for (int32_t gc = 0; gc < n_cases; ++gc) {
      const CaseSpec& s = specs[user_of_global[gc]];
      cases_h[gc] = case_params(s.mesh, s.opts);
    }
    upload(cases_d, cases_h);
    global_of_user.assign(n_cases, 0);
    for (int32_t gc = 0; gc < n_cases; ++gc) global_of_user[user_of_global[gc]] = gc;

    // state
    W0.alloc(elems * sizeof(float4));
    W1.alloc(elems * sizeof(float4));
    Wref.alloc(elems * sizeof(float4));
    if (scheme == ESR_SCHEME_RK4) {
      W2.alloc(elems * sizeof(float4));
      acc.alloc(elems * sizeof(float4));
    }
    rec.alloc((rec_elems > 0 ? rec_elems : 1) * sizeof(float4));
    fx.alloc((fx_elems > 0 ? fx_elems : 1) * sizeof(float4));
    // reductions
    rstride = 1;
    std::vector<int32_t> nr(n_cases);
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const int32_t r = reduce_ranges(groups[gi]);
      rstride = std::max(rstride, r);
      for (int32_t c = 0; c < groups[gi].C; ++c) nr[groups[gi].case0 + c] = r;
    }
    upload(nranges, nr);
    part_d.alloc((size_t)n_cases * rstride * sizeof(double));
    part_r.alloc((size_t)n_cases * rstride * sizeof(double));
    part_f.alloc((size_t)n_cases * rstride * sizeof(float));
    cfl_d.alloc(n_cases * sizeof(float));
    dt_d.alloc(n_cases * sizeof(float));
    forces_d.alloc(2 * (size_t)n_cases * sizeof(double));
    probe_d.alloc(n_cases * sizeof(double));
    count_d.alloc(sizeof(int32_t));
    counter_d.alloc(sizeof(uint32_t));
    hip_check(hipMemset(counter_d.p, 0, sizeof(uint32_t)), "counter memset");
    hip_check(hipHostMalloc((void**)&count_h, sizeof(int32_t), hipHostMallocDefault), "hipHostMalloc");
    *count_h = 0;
  }

  std::vector<int32_t> global_of_user;

  CaseDev* cases() { return cases_d.as<CaseDev>(); }
  const GroupDev* gdev(int32_t gi) const { return groups_d.as<GroupDev>() + gi; }

  // one stage: reconstruction, face fluxes, gather + update into `out`
  void stage(int mode, const float4* Ws, float4* out, hipStream_t st) {
    launch_recon(rec_all.ls, cases(), Ws, rec.as<float4>(), st);
    for (auto& c : classes) {
      if (c->ls.flux == ESR_FLUX_LF) {
        launch_zero_lf(n_cases, cases(), st);
        launch_lf_cmax(c->ls, cases(), Ws, rec.as<float4>(), st);
      }
      if (stage_path == ESR_STAGE_FACE) launch_face(c->ls, cases(), Ws, rec.as<float4>(), fx.as<float4>(), st);
    }
    if (stage_path == ESR_STAGE_FACE)
      launch_update(cell_all.ls, mode, local_dt, cases(), Ws, W0.as<float4>(), fx.as<float4>(), out, acc.as<float4>(),
                    st);
    else
      for (auto& c : classes)
        launch_cell(c->ls, mode, cases(), Ws, W0.as<float4>(), rec.as<float4>(), out, acc.as<float4>(), st);
  }

  // one Runge-Kutta step of every active case
  void step(hipStream_t st) {
    for (const Stage& sg : stages_of(scheme)) stage(sg.mode, buf(sg.ws), buf(sg.out), st);
    if (scheme == ESR_SCHEME_EE)
      launch_commit(cell_all.ls, cases(), W1.as<float4>(), W0.as<float4>(), st);
  }

  // residual of W0 into W1 (one case engines)
  void residual(hipStream_t st) { stage(MODE_RESID, W0.as<float4>(), W1.as<float4>(), st); }

  void gather_in(int32_t user_case, const float* W, float4* slab, hipStream_t st) {
    const int32_t gc = global_of_user.at(user_case);
    const int32_t gi = case_group[gc];
    launch_gather_in_dev(groups[gi].m.n, gdev(gi), case_slot[gc], gmesh[gi]->b_perm.as<int32_t>(),
                         reinterpret_cast<const float4*>(W), slab, st);
  }
  void scatter_out(int32_t user_case, const float4* slab, float* W, hipStream_t st) {
    const int32_t gc = global_of_user.at(user_case);
    const int32_t gi = case_group[gc];
    launch_scatter_out_dev(groups[gi].m.n, gdev(gi), case_slot[gc], gmesh[gi]->b_perm.as<int32_t>(), slab,
                           reinterpret_cast<float4*>(W), st);
  }

  static CaseDev case_params(const esr_mesh* m, const esr_opts& o) { return esr::case_params(m, o); }

  // dt of global case gc alone from its current state (the other cases keep theirs)
  float compute_dt_one(int32_t gc, float cfl, hipStream_t st) {
    hip_check(hipMemcpyAsync(cfl_d.as<float>() + gc, &cfl, sizeof(float), hipMemcpyHostToDevice, st), "cfl H2D");
    const int32_t gi = case_group[gc];
    launch_dt(groups[gi], gdev(gi), cases(), W0.as<float4>(), part_f.as<float>(), rstride, 0, st);
    launch_dt_final(1, cases() + gc, part_f.as<float>() + (int64_t)gc * rstride, nranges.as<int32_t>() + gc, rstride,
                    cfl_d.as<float>() + gc, dt_d.as<float>() + gc, st);
    float dt = 0.f;
    hip_check(hipMemcpyAsync(&dt, dt_d.as<float>() + gc, sizeof(float), hipMemcpyDeviceToHost, st), "dt D2H");
    hip_check(hipStreamSynchronize(st), "compute_dt_one sync");
    return dt;
  }

  void compute_dt(const std::vector<float>& cfl_user, hipStream_t st) {
    std::vector<float> cfl(n_cases);
    for (int32_t u = 0; u < n_cases; ++u) cfl[global_of_user[u]] = cfl_user[u];
    hip_check(hipMemcpyAsync(cfl_d.p, cfl.data(), n_cases * sizeof(float), hipMemcpyHostToDevice, st), "cfl H2D");
    for (size_t gi = 0; gi < groups.size(); ++gi)
      launch_dt(groups[gi], gdev((int32_t)gi), cases(), W0.as<float4>(), part_f.as<float>(), rstride, 0, st);
    launch_dt_final(n_cases, cases(), part_f.as<float>(), nranges.as<int32_t>(), rstride, cfl_d.as<float>(),
                    dt_d.as<float>(), st);
    hip_check(hipStreamSynchronize(st), "compute_dt sync");
  }

  // norms of (Wn - Wref) and Wref for cases flagged 'check'
  // fused: the check arguments, run by the last block of the last launch
  // one launch over every launched group
  void norms(const float4* Wn, float4* Wref_, int copy_ref, hipStream_t st, const CheckArgs* fused = nullptr) {
    launch_norms(groups_d.as<GroupDev>(), red_tbl.as<RedRef>(), n_red, cases(), Wn, Wref_, part_d.as<double>(),
                 part_r.as<double>(), rstride, copy_ref, st, fused);
  }
  void forces(const float4* W, hipStream_t st) {
    launch_forces(groups_d.as<GroupDev>(), frc_tbl.as<BlockRef>(), n_frc, cases(), W, forces_d.as<double>(), st);
  }
};

// single-state workspace attached to a mesh
struct Workspace {
  std::unique_ptr<Engine> eng;
  esr_opts opts{};
  int scheme = -1;
  bool valid = false;
};

namespace esr {
void free_workspace(Workspace* ws) { delete ws; }
}  // namespace esr

namespace {

bool same_opts(const esr_opts& a, const esr_opts& b) {
  return a.flux == b.flux && a.recon == b.recon && a.entropy == b.entropy && a.gamma == b.gamma && a.M == b.M &&
         a.sponge_width == b.sponge_width && a.sponge_strength == b.sponge_strength &&
         std::memcmp(a.value, b.value, sizeof(a.value)) == 0;
}

// One-case engine cached on the mesh, rebuilt when the options change.
Engine& mesh_engine(esr_mesh* m, const esr_opts& o, int scheme, hipStream_t st) {
  if (!m->ws) m->ws = new Workspace();
  Workspace* ws = m->ws;
  const bool need_rk4 = (scheme == ESR_SCHEME_RK4);
  if (!ws->valid || !ws->eng || (need_rk4 && ws->scheme != ESR_SCHEME_RK4) ||
      ws->eng->groups[0].ph.g != make_phys(o.gamma, o.M).g || o.flux != ws->opts.flux ||
      o.recon != ws->opts.recon || o.entropy != ws->opts.entropy || o.gamma != ws->opts.gamma ||
      o.M != ws->opts.M || o.sponge_width != ws->opts.sponge_width) {
    hip_check(hipStreamSynchronize(st), "workspace sync");
    ws->eng.reset(new Engine());
    ws->eng->build({CaseSpec{m, o}}, need_rk4 ? ESR_SCHEME_RK4 : ESR_SCHEME_SRK2);
    ws->opts = o;
    ws->scheme = ws->eng->scheme;
    ws->valid = true;
  } else if (!same_opts(o, ws->opts)) {
    // only the inlet value / sponge strength changed
    CaseDev cd = ws->eng->cases_h[0];
    cd.value = make_float4(o.value[0], o.value[1], o.value[2], o.value[3]);
    const float width = o.sponge_width > 0 ? o.sponge_width : 0.2f * (m->y_max - m->y_min);
    cd.strength = o.sponge_strength > 0 ? o.sponge_strength : sponge_strength(o, width);
    ws->eng->cases_h[0] = cd;
    launch_set_case(ws->eng->cases(), cd, st);
    ws->opts = o;
  }
  Engine& e = *ws->eng;
  e.scheme = scheme;
  e.stage_path = resolved_stage_path();
  return e;
}

void set_dt(Engine& e, float dt, hipStream_t st) {
  CaseDev cd = e.cases_h[0];
  cd.dt = dt;
  cd.dt_half = dt / 2.0f;
  cd.dt_sixth = dt / 6.0f;
  cd.active = 1;
  cd.check = 0;
  e.cases_h[0] = cd;
  launch_set_case(e.cases(), cd, st);
}

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

}  // namespace

// ---------------------------------------------------------------------------
// batched solver
// ---------------------------------------------------------------------------

struct esr_solver {
  Engine eng;
  esr_solve_opts so{};
  std::vector<esr_case> cases;   // user order
  std::vector<float> dt;         // user order
  std::vector<int64_t> N;        // user order
  std::vector<double> final_residual, cd, cl;
  int64_t step = 0;
  bool prepared = false;
  bool finalized = false;
  // solve-loop stream (default priority) and the captured chunk of
  // check_every steps + one stationarity check
  hipStream_t gs = nullptr;
  hipGraphExec_t chunk = nullptr;
  DevBuf dstep;  // int64[2]: step before / after the chunk
  int64_t dstep_at = -1;  // value of dstep[0] on the device (-1: unknown)
  ~esr_solver() {
    if (chunk) (void)hipGraphExecDestroy(chunk);
    if (gs) (void)hipStreamDestroy(gs);
  }
};

namespace {

// loop state of user case u entering its slot at solver step `start`
// (s->N[u] is the absolute step N of the case)
CaseLoop initial_loop(const esr_solver* s, int32_t u, int64_t start) {
  const esr_case& c = s->cases[u];
  CaseLoop l;
  std::memset(&l, 0, sizeof(l));
  l.N = s->N[u];
  l.start = start;
  l.ref_step = start;
  l.stopping_step = s->N[u];
  l.last_rel = INFINITY;
  l.threshold = c.threshold > 0 ? c.threshold : s->so.threshold;
  l.patience = std::max(1, s->so.patience);
  l.q_ref = c.q_ref;
  l.track_eng = c.track_engineering ? 1 : 0;
  l.target_cd = c.target_cd;
  l.target_cl = c.target_cl;
  l.cd_tol = c.cd_tol_abs;
  l.cl_tol = c.cl_tol_abs;
  l.eng_patience = std::max(1, c.engineering_patience);
  l.eng_stopping_step = s->N[u];
  return l;
}

void solver_set_loops(esr_solver* s, hipStream_t st) {
  Engine& e = s->eng;
  std::vector<CaseLoop> L(e.n_cases);
  for (int32_t u = 0; u < e.n_cases; ++u) L[e.global_of_user[u]] = initial_loop(s, u, 0);
  e.loops_d.alloc(L.size() * sizeof(CaseLoop));
  hip_check(hipMemcpyAsync(e.loops_d.p, L.data(), L.size() * sizeof(CaseLoop), hipMemcpyHostToDevice, st),
            "loops H2D");
  hip_check(hipStreamSynchronize(st), "loops sync");
}

bool any_engineering(const esr_solver* s) {
  for (const auto& c : s->cases)
    if (c.track_engineering) return true;
  return false;
}

}  // namespace

extern "C" {

const char* esr_last_error(void) { return esr::last_error_cstr(); }

int32_t esr_version(void) { return 100; }

esr_status esr_mesh_create(const esr_mesh_desc* desc, esr_mesh** out) {
  return guard([&] {
    if (!out) throw Error(ESR_ERR_VALUE, "null output handle");
    *out = build_mesh(desc);
  });
}

esr_status esr_mesh_destroy(esr_mesh* mesh) {
  return guard([&] { delete mesh; });
}

esr_status esr_mesh_permutation(const esr_mesh* mesh, int32_t* perm) {
  return guard([&] {
    if (!mesh || !perm) throw Error(ESR_ERR_VALUE, "null argument");
    std::memcpy(perm, mesh->perm.data(), mesh->perm.size() * sizeof(int32_t));
  });
}

esr_status esr_mesh_info(const esr_mesh* mesh, int32_t* n_cells, int32_t* n_wall_faces, int64_t* device_bytes) {
  return guard([&] {
    if (!mesh) throw Error(ESR_ERR_VALUE, "null mesh");
    if (n_cells) *n_cells = mesh->n;
    if (n_wall_faces) *n_wall_faces = mesh->dev.n_wall;
    if (device_bytes) *device_bytes = mesh->device_bytes;
  });
}

esr_status esr_rcm(int32_t n_cells, const int32_t* neighbors, const int32_t* seed_order, int32_t* perm) {
  return guard([&] {
    if (n_cells <= 0 || !neighbors || !perm) throw Error(ESR_ERR_VALUE, "bad arguments");
    for (int64_t t = 0; t < 3 * (int64_t)n_cells; ++t)
      if (neighbors[t] < -1 || neighbors[t] >= n_cells) throw Error(ESR_ERR_VALUE, "neighbors out of range");
    const std::vector<int32_t> p = rcm_order(n_cells, neighbors, seed_order);
    std::memcpy(perm, p.data(), p.size() * sizeof(int32_t));
  });
}

esr_status esr_residual(esr_mesh* mesh, const esr_opts* opts, const float* W, float* R, void* stream) {
  return guard([&] {
    if (!mesh || !W || !R) throw Error(ESR_ERR_VALUE, "null argument");
    validate_opts(opts);
    hipStream_t st = S(stream);
    Engine& e = mesh_engine(mesh, *opts, ESR_SCHEME_SRK2, st);
    set_dt(e, 0.f, st);
    e.gather_in(0, W, e.W0.as<float4>(), st);
    e.residual(st);
    e.scatter_out(0, e.W1.as<float4>(), R, st);
  });
}

esr_status esr_step(esr_mesh* mesh, const esr_opts* opts, int32_t scheme, float dt, const float* W, float* W_out,
                    void* stream) {
  return guard([&] {
    if (!mesh || !W || !W_out) throw Error(ESR_ERR_VALUE, "null argument");
    validate_opts(opts);
    (void)stages_of(scheme);
    hipStream_t st = S(stream);
    Engine& e = mesh_engine(mesh, *opts, scheme, st);
    set_dt(e, dt, st);
    e.gather_in(0, W, e.W0.as<float4>(), st);
    e.step(st);
    e.scatter_out(0, e.W0.as<float4>(), W_out, st);
  });
}

esr_status esr_get_dt(esr_mesh* mesh, const float* W, float cfl, double gamma, double M, float* dt_out,
                      void* stream) {
  return guard([&] {
    if (!mesh || !W || !dt_out) throw Error(ESR_ERR_VALUE, "null argument");
    esr_opts o{};
    o.flux = ESR_FLUX_HLLC;
    o.recon = ESR_RECON_CONSTANT;
    o.gamma = gamma;
    o.M = M;
    o.value[0] = 1.f;
    o.value[1] = 1.f;
    o.value[3] = 1.f;
    o.sponge_width = -1.f;
    o.sponge_strength = -1.f;
    validate_opts(&o);
    hipStream_t st = S(stream);
    Engine& e = mesh_engine(mesh, o, ESR_SCHEME_SRK2, st);
    set_dt(e, 0.f, st);
    e.gather_in(0, W, e.W0.as<float4>(), st);
    e.compute_dt({cfl}, st);
    hip_check(hipMemcpy(dt_out, e.dt_d.p, sizeof(float), hipMemcpyDeviceToHost), "dt D2H");
  });
}

esr_status esr_forces(esr_mesh* mesh, const float* W, double q_ref, double* cd, double* cl, void* stream) {
  return guard([&] {
    if (!mesh || !W || !cd || !cl) throw Error(ESR_ERR_VALUE, "null argument");
    if (mesh->dev.n_wall == 0) throw Error(ESR_ERR_VALUE, "no wall face carries the force marker");
    hipStream_t st = S(stream);
    esr_opts o{};
    o.flux = ESR_FLUX_HLLC;
    o.recon = ESR_RECON_CONSTANT;
    o.gamma = 1.4;
    o.M = 1.0;
    o.value[0] = 1.f;
    o.value[1] = 1.f;
    o.value[3] = 1.f;
    o.sponge_width = -1.f;
    o.sponge_strength = -1.f;
    Engine& e = mesh_engine(mesh, mesh->ws && mesh->ws->valid ? mesh->ws->opts : o, ESR_SCHEME_SRK2, st);
    e.gather_in(0, W, e.W0.as<float4>(), st);
    launch_mark_check(1, e.cases(), nullptr, 0, 1, 1, st);
    e.forces(e.W0.as<float4>(), st);
    double f[2];
    hip_check(hipMemcpyAsync(f, e.forces_d.p, sizeof(f), hipMemcpyDeviceToHost, st), "forces D2H");
    hip_check(hipStreamSynchronize(st), "forces sync");
    launch_set_active(1, e.cases(), nullptr, 1, st);
    *cd = f[0] / q_ref;
    *cl = f[1] / q_ref;
  });
}

esr_status esr_norms(esr_mesh* mesh, const float* A, const float* B, double* diff_norm, double* ref_norm,
                     void* stream) {
  return guard([&] {
    if (!mesh || !A || !B) throw Error(ESR_ERR_VALUE, "null argument");
    hipStream_t st = S(stream);
    esr_opts o = mesh->ws && mesh->ws->valid ? mesh->ws->opts : esr_opts{};
    if (!(mesh->ws && mesh->ws->valid)) {
      o.flux = ESR_FLUX_HLLC;
      o.recon = ESR_RECON_CONSTANT;
      o.gamma = 1.4;
      o.M = 1.0;
      o.value[0] = 1.f;
      o.value[1] = 1.f;
      o.value[3] = 1.f;
      o.sponge_width = -1.f;
      o.sponge_strength = -1.f;
    }
    Engine& e = mesh_engine(mesh, o, ESR_SCHEME_SRK2, st);
    e.gather_in(0, A, e.W1.as<float4>(), st);
    e.gather_in(0, B, e.Wref.as<float4>(), st);
    launch_mark_check(1, e.cases(), nullptr, 0, 1, 1, st);
    e.norms(e.W1.as<float4>(), e.Wref.as<float4>(), 0, st);
    launch_probe_final(1, e.cases(), e.part_d.as<double>(), e.part_r.as<double>(), e.nranges.as<int32_t>(),
                       e.rstride, e.probe_d.as<double>(), st);
    std::vector<double> pd(e.rstride), pr(e.rstride);
    int32_t nr = 0;
    hip_check(hipMemcpyAsync(pd.data(), e.part_d.p, e.rstride * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipMemcpyAsync(pr.data(), e.part_r.p, e.rstride * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipMemcpyAsync(&nr, e.nranges.p, sizeof(int32_t), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "norms sync");
    double d = 0.0, r = 0.0;
    for (int32_t q = 0; q < nr; ++q) {
      d += pd[q];
      r += pr[q];
    }
    if (diff_norm) *diff_norm = std::sqrt(d);
    if (ref_norm) *ref_norm = std::sqrt(r);
  });
}

esr_status esr_solver_create(int32_t n_cases, const esr_case* cases, const esr_solve_opts* so, esr_solver** out) {
  return guard([&] {
    if (!out || !cases || !so || n_cases <= 0) throw Error(ESR_ERR_VALUE, "bad solver arguments");
    if (so->check_every < 1) throw Error(ESR_ERR_VALUE, "check_every must be >= 1");
    std::unique_ptr<esr_solver> s(new esr_solver());
    s->so = *so;
    s->cases.assign(cases, cases + n_cases);
    std::vector<CaseSpec> specs;
    for (int32_t i = 0; i < n_cases; ++i) specs.push_back(CaseSpec{cases[i].mesh, cases[i].opts, cases[i].group});
    s->eng.build(specs, so->scheme);
    if (so->local_dt) {
      // local time stepping: each case's cells advance with their own dt at
      // the case's CFL (the stage kernels read it from the case parameters)
      Engine& e = s->eng;
      e.local_dt = true;
      e.stage_path = ESR_STAGE_FACE;
      for (int32_t u = 0; u < n_cases; ++u) {
        if (!(cases[u].cfl > 0.f)) throw Error(ESR_ERR_VALUE, "cfl must be > 0");
        e.cases_h[e.global_of_user[u]].cfl_local = cases[u].cfl;
      }
      hip_check(hipMemcpy(e.cases_d.p, e.cases_h.data(), e.cases_h.size() * sizeof(CaseDev),
                          hipMemcpyHostToDevice), "case params H2D");
    }
    s->dt.assign(n_cases, 0.f);
    s->N.assign(n_cases, 0);
    s->final_residual.assign(n_cases, NAN);
    s->cd.assign(n_cases, NAN);
    s->cl.assign(n_cases, NAN);
    *out = s.release();
  });
}

esr_status esr_solver_destroy(esr_solver* s) {
  return guard([&] { delete s; });
}

esr_status esr_solver_set_state(esr_solver* s, int32_t c, const float* W, void* stream) {
  return guard([&] {
    if (!s || !W || c < 0 || c >= s->eng.n_cases) throw Error(ESR_ERR_VALUE, "bad case index");
    s->eng.gather_in(c, W, s->eng.W0.as<float4>(), S(stream));
    s->prepared = false;
  });
}

esr_status esr_solver_get_state(esr_solver* s, int32_t c, float* W, void* stream) {
  return guard([&] {
    if (!s || !W || c < 0 || c >= s->eng.n_cases) throw Error(ESR_ERR_VALUE, "bad case index");
    s->eng.scatter_out(c, s->eng.W0.as<float4>(), W, S(stream));
  });
}

esr_status esr_solver_prepare(esr_solver* s, void* stream) {
  return guard([&] {
    if (!s) throw Error(ESR_ERR_VALUE, "null solver");
    hipStream_t st = S(stream);
    Engine& e = s->eng;
    std::vector<float> cfl(e.n_cases);
    for (int32_t u = 0; u < e.n_cases; ++u) cfl[u] = s->cases[u].cfl;
    launch_set_active(e.n_cases, e.cases(), nullptr, 1, st);
    e.compute_dt(cfl, st);
    std::vector<float> dtg(e.n_cases);
    hip_check(hipMemcpy(dtg.data(), e.dt_d.p, e.n_cases * sizeof(float), hipMemcpyDeviceToHost), "dt D2H");
    for (int32_t u = 0; u < e.n_cases; ++u) {
      const float dt = dtg[e.global_of_user[u]];
      s->dt[u] = dt;
      const float q = s->cases[u].tf / dt;  // float32 division, as tf / dt in the reference
      if (!(q >= 0.f) || !std::isfinite(q)) throw Error(ESR_ERR_VALUE, "non-finite time step");
      s->N[u] = (int64_t)q + 1;
      if (s->cases[u].max_steps > 0) s->N[u] = std::min(s->N[u], s->cases[u].max_steps);
    }
    solver_set_loops(s, st);
    if (!e.group_on.empty() && e.set_group_mask(std::vector<uint8_t>(e.groups.size(), 1), st) && s->chunk) {
      (void)hipGraphExecDestroy(s->chunk);
      s->chunk = nullptr;
    }
    launch_copy(e.W0.as<float4>(), e.Wref.as<float4>(), e.elems, st);
    s->step = 0;
    s->prepared = true;
    s->dstep_at = -1;
    s->finalized = false;
  });
}

}  // extern "C"

namespace {

// one stationarity check after step k (host step, or the device counter)
// (forces first, then the norms, then the check)
constexpr int64_t kFusedCheckPartials = 4096;
void enqueue_check(esr_solver* s, int64_t k, bool eng_on, hipStream_t st, int64_t* dstep) {
  Engine& e = s->eng;
  const int32_t K = s->so.check_every;
  launch_mark_check(e.n_cases, e.cases(), e.loops_d.as<CaseLoop>(), k, K, 0, st, dstep);
  if (eng_on) e.forces(e.W0.as<float4>(), st);
  const double* forces = eng_on ? e.forces_d.as<double>() : nullptr;
  // small batches: the last block of the norms runs the checks (one launch
  // less); large ones: a wave per case in a launch of its own
  if ((int64_t)e.n_cases * e.rstride <= kFusedCheckPartials) {
    CheckArgs ca{};
    ca.n_cases = e.n_cases;
    ca.cases = e.cases();
    ca.loops = e.loops_d.as<CaseLoop>();
    ca.nranges = e.nranges.as<int32_t>();
    ca.forces = forces;
    ca.step = k;
    ca.dstep = dstep;
    ca.counter = e.counter_d.as<uint32_t>();
    e.norms(e.W0.as<float4>(), e.Wref.as<float4>(), 1, st, &ca);
  } else {
    e.norms(e.W0.as<float4>(), e.Wref.as<float4>(), 1, st);
    launch_check(e.n_cases, e.cases(), e.loops_d.as<CaseLoop>(), e.part_d.as<double>(), e.part_r.as<double>(),
                 e.nranges.as<int32_t>(), e.rstride, forces, k, K, st, dstep);
  }
}

void capture_chunk(esr_solver* s, bool eng_on) {
  Engine& e = s->eng;
  const int32_t K = s->so.check_every;
  hipGraph_t g = nullptr;
  hip_check(hipStreamBeginCapture(s->gs, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    for (int32_t q = 0; q < K; ++q) e.step(s->gs);
    enqueue_check(s, 0, eng_on, s->gs, s->dstep.as<int64_t>());
  } catch (...) {
    (void)hipStreamEndCapture(s->gs, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hip_check(hipStreamEndCapture(s->gs, &g), "end capture");
  hipError_t err = hipGraphInstantiate(&s->chunk, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  hip_check(err, "graph instantiate");
}

}  // namespace

extern "C" esr_status esr_solver_run(esr_solver* s, int64_t until_step, int64_t* steps, void* stream) {
  return guard([&] {
    if (!s) throw Error(ESR_ERR_VALUE, "null solver");
    if (!s->prepared) throw Error(ESR_ERR_STATE, "esr_solver_prepare must be called first");
    Engine& e = s->eng;
    // everything before this call on the caller's stream (states, prepare)
    hip_check(hipStreamSynchronize(S(stream)), "caller stream sync");
    if (!s->gs) hip_check(hipStreamCreateWithFlags(&s->gs, hipStreamNonBlocking), "stream create");
    if (!s->dstep.p) s->dstep.alloc(2 * sizeof(int64_t));
    hipStream_t st = s->gs;
    const int32_t K = s->so.check_every;
    int64_t maxN = 0;
    std::set<int64_t> Ns;
    for (int64_t n : s->N) {
      maxN = std::max(maxN, n);
      Ns.insert(n);
    }
    int64_t limit = maxN;
    if (s->so.max_steps > 0) limit = std::min(limit, s->so.max_steps);
    if (until_step > 0) limit = std::min(limit, until_step);
    const bool eng_on = any_engineering(s);
    const bool use_graph = !s->so.no_graph;
    const int poll_every = std::max(1, 64 / K);
    int checks = 0;
    bool all_done = false;
    while (s->step < limit && !all_done) {
      const int64_t k0 = s->step;
      auto nextN = Ns.upper_bound(k0);
      const bool n_inside = nextN != Ns.end() && *nextN < k0 + K;
      bool checked = false;
      if (use_graph && k0 % K == 0 && k0 + K <= limit && !n_inside) {
        if (!s->chunk) capture_chunk(s, eng_on);
        // the device step counter already holds k0 after a chunk
        if (s->dstep_at != k0) launch_set_step(s->dstep.as<int64_t>(), k0, st);
        hip_check(hipGraphLaunch(s->chunk, st), "graph launch");
        s->step = k0 + K;
        s->dstep_at = s->step;
        checked = true;
      } else {
        s->step = k0 + 1;
        e.step(st);
        const int64_t k = s->step;
        if (k % K == 0 || Ns.count(k)) {
          enqueue_check(s, k, eng_on, st, nullptr);
          checked = true;
        }
      }
      if (checked && ++checks % poll_every == 0) {
        launch_count_active(e.n_cases, e.cases(), e.count_d.as<int32_t>(), st);
        hip_check(hipMemcpyAsync(e.count_h, e.count_d.p, sizeof(int32_t), hipMemcpyDeviceToHost, st), "count");
        hip_check(hipStreamSynchronize(st), "poll sync");
        all_done = (*e.count_h == 0);
      }
    }
    hip_check(hipStreamSynchronize(st), "run sync");
    if (steps) *steps = s->step;
  });
}

extern "C" {

esr_status esr_solver_finalize(esr_solver* s, void* stream) {
  return guard([&] {
    if (!s) throw Error(ESR_ERR_VALUE, "null solver");
    if (!s->prepared) throw Error(ESR_ERR_STATE, "esr_solver_prepare must be called first");
    hipStream_t st = S(stream);
    Engine& e = s->eng;
    // the probe and the forces cover every case: groups that
    // esr_solver_compact dropped from the launches are launched again
    if (s->gs) hip_check(hipStreamSynchronize(s->gs), "solver stream sync");
    if (!e.group_on.empty() && e.set_group_mask(std::vector<uint8_t>(e.groups.size(), 1), st) && s->chunk) {
      (void)hipGraphExecDestroy(s->chunk);
      s->chunk = nullptr;
    }
    // probe: one more step from the final state (main.py:228-230)
    launch_copy(e.W0.as<float4>(), e.Wref.as<float4>(), e.elems, st);
    launch_set_active(e.n_cases, e.cases(), nullptr, 1, st);
    e.step(st);
    launch_mark_check(e.n_cases, e.cases(), nullptr, 0, 1, 1, st);
    e.norms(e.W0.as<float4>(), e.Wref.as<float4>(), 0, st);
    launch_probe_final(e.n_cases, e.cases(), e.part_d.as<double>(), e.part_r.as<double>(), e.nranges.as<int32_t>(),
                       e.rstride, e.probe_d.as<double>(), st);
    launch_copy(e.Wref.as<float4>(), e.W0.as<float4>(), e.elems, st);
    // forces of the final state
    launch_mark_check(e.n_cases, e.cases(), nullptr, 0, 1, 1, st);
    e.forces(e.W0.as<float4>(), st);
    launch_set_active(e.n_cases, e.cases(), e.loops_d.as<CaseLoop>(), 0, st);
    std::vector<double> pr(e.n_cases), f(2 * (size_t)e.n_cases);
    hip_check(hipMemcpyAsync(pr.data(), e.probe_d.p, pr.size() * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipMemcpyAsync(f.data(), e.forces_d.p, f.size() * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "finalize sync");
    for (int32_t u = 0; u < e.n_cases; ++u) {
      const int32_t gc = e.global_of_user[u];
      s->final_residual[u] = pr[gc];
      const double q = s->cases[u].q_ref;
      s->cd[u] = q > 0 ? f[2 * (size_t)gc] / q : NAN;
      s->cl[u] = q > 0 ? f[2 * (size_t)gc + 1] / q : NAN;
    }
    s->finalized = true;
  });
}

namespace {
void fill_report(const esr_solver* s, int32_t c, const CaseLoop& L, esr_case_report* out);
}

esr_status esr_solver_reports(const esr_solver* s, int32_t n, esr_case_report* out) {
  return guard([&] {
    if (!s || !out || n != s->eng.n_cases) throw Error(ESR_ERR_VALUE, "n must equal the number of cases");
    if (!s->prepared) throw Error(ESR_ERR_STATE, "esr_solver_prepare must be called first");
    const Engine& e = s->eng;
    std::vector<CaseLoop> L(e.n_cases);
    hip_check(hipMemcpy(L.data(), e.loops_d.p, L.size() * sizeof(CaseLoop), hipMemcpyDeviceToHost), "loops D2H");
    for (int32_t c = 0; c < n; ++c) fill_report(s, c, L[e.global_of_user[c]], out + c);
  });
}

esr_status esr_solver_report(const esr_solver* s, int32_t c, esr_case_report* out) {
  return guard([&] {
    if (!s || !out || c < 0 || c >= s->eng.n_cases) throw Error(ESR_ERR_VALUE, "bad case index");
    if (!s->prepared) throw Error(ESR_ERR_STATE, "esr_solver_prepare must be called first");
    const Engine& e = s->eng;
    CaseLoop L;
    const int32_t gc = e.global_of_user[c];
    hip_check(hipMemcpy(&L, e.loops_d.as<CaseLoop>() + gc, sizeof(L), hipMemcpyDeviceToHost), "loop D2H");
    fill_report(s, c, L, out);
  });
}

namespace {
void fill_report(const esr_solver* s, int32_t c, const CaseLoop& L, esr_case_report* out) {
    std::memset(out, 0, sizeof(*out));
    out->dt = s->dt[c];
    out->n_steps = s->N[c] - L.start;
    out->converged = L.stop;
    out->stopping_step = L.stopping_step - L.start;
    out->steps_run = (L.done ? L.steps_run : s->step) - L.start;
    out->last_rel = L.last_rel;
    out->final_residual = s->final_residual[c];
    // after finalize: forces of the final state; before: the last in-loop
    // check's (cases tracking the engineering criterion), else NaN
    const bool fin = !std::isnan(s->cd[c]);
    out->cd = fin ? s->cd[c] : (L.track_eng ? L.cd : s->cd[c]);
    out->cl = fin ? s->cl[c] : (L.track_eng ? L.cl : s->cl[c]);
    out->engineering_converged = L.eng_stop;
    out->engineering_stopping_step = L.eng_stopping_step - L.start;
    out->diverged = L.diverged;
    const int32_t nr = std::min(L.n_checks, RECENT_CHECKS);
    double rm = nr > 0 ? 0.0 : NAN;
    for (int32_t q = 0; q < nr; ++q) rm = std::max(rm, L.recent[q]);
    out->recent_rel_max = rm;
}
}  // namespace

esr_status esr_solver_info(const esr_solver* s, int64_t* cell_items, int64_t* group_cells, int64_t* face_items,
                           int64_t* group_faces) {
  return guard([&] {
    if (!s) throw Error(ESR_ERR_VALUE, "null solver");
    int64_t items = 0, cells = 0, fitems = 0, faces = 0;
    for (const GroupDev& g : s->eng.groups) {
      items += (int64_t)g.m.n * g.C;
      cells += g.m.n;
      fitems += (int64_t)g.m.nf * g.C;
      faces += g.m.nf;
    }
    if (cell_items) *cell_items = items;
    if (group_cells) *group_cells = cells;
    if (face_items) *face_items = fitems;
    if (group_faces) *group_faces = faces;
  });
}

esr_status esr_solver_replace_case(esr_solver* s, int32_t c, const esr_case* spec, const float* W, void* stream) {
  return guard([&] {
    if (!s || !spec || !W || c < 0 || c >= s->eng.n_cases) throw Error(ESR_ERR_VALUE, "bad case index");
    if (!s->prepared) throw Error(ESR_ERR_STATE, "esr_solver_prepare must be called first");
    const esr_case& old = s->cases[c];
    const esr_opts &a = old.opts, &b = spec->opts;
    validate_opts(&b);
    if (spec->mesh != old.mesh || a.flux != b.flux || a.recon != b.recon || (a.entropy != 0) != (b.entropy != 0) ||
        a.gamma != b.gamma || a.M != b.M || a.sponge_width != b.sponge_width)
      throw Error(ESR_ERR_VALUE, "a replacement case must use the slot's mesh and kernel configuration "
                                 "(flux, reconstruction, entropy, gamma, M, sponge width)");
    if (!(spec->cfl > 0.f) || !(spec->tf > 0.0)) throw Error(ESR_ERR_VALUE, "cfl and tf must be > 0");
    hipStream_t st = S(stream);
    if (s->gs) hip_check(hipStreamSynchronize(s->gs), "solver stream sync");
    hip_check(hipStreamSynchronize(st), "caller stream sync");
    Engine& e = s->eng;
    const int32_t gc = e.global_of_user[c];
    e.gather_in(c, W, e.W0.as<float4>(), st);
    e.gather_in(c, W, e.Wref.as<float4>(), st);
    CaseDev cd = Engine::case_params(spec->mesh, b);
    cd.cfl_local = s->so.local_dt ? spec->cfl : 0.f;
    e.cases_h[gc] = cd;
    launch_set_case(e.cases() + gc, cd, st);
    const float dt = e.compute_dt_one(gc, spec->cfl, st);
    const float q = (float)spec->tf / dt;  // float32 division, as tf / dt in the reference
    if (!(q >= 0.f) || !std::isfinite(q)) throw Error(ESR_ERR_VALUE, "non-finite time step");
    s->cases[c] = *spec;
    s->dt[c] = dt;
    int64_t n_rel = (int64_t)q + 1;
    if (spec->max_steps > 0) n_rel = std::min(n_rel, spec->max_steps);
    s->N[c] = s->step + n_rel;
    s->final_residual[c] = NAN;
    s->cd[c] = NAN;
    s->cl[c] = NAN;
    const CaseLoop l = initial_loop(s, c, s->step);
    hip_check(hipMemcpyAsync(e.loops_d.as<CaseLoop>() + gc, &l, sizeof(l), hipMemcpyHostToDevice, st), "loop H2D");
    hip_check(hipStreamSynchronize(st), "replace sync");
    // the slot's group takes part in the launches again
    if (!e.group_on.empty() && !e.group_on[e.case_group[gc]]) {
      std::vector<uint8_t> mask = e.group_on;
      mask[e.case_group[gc]] = 1;
      e.set_group_mask(mask, st);
      if (s->chunk) {
        (void)hipGraphExecDestroy(s->chunk);
        s->chunk = nullptr;
      }
    }
  });
}

esr_status esr_solver_compact(esr_solver* s, int32_t* groups_on, void* stream) {
  return guard([&] {
    if (!s) throw Error(ESR_ERR_VALUE, "null solver");
    if (!s->prepared) throw Error(ESR_ERR_STATE, "esr_solver_prepare must be called first");
    Engine& e = s->eng;
    if (s->gs) hip_check(hipStreamSynchronize(s->gs), "solver stream sync");
    std::vector<CaseLoop> L(e.n_cases);
    hip_check(hipMemcpy(L.data(), e.loops_d.p, L.size() * sizeof(CaseLoop), hipMemcpyDeviceToHost), "loops D2H");
    std::vector<uint8_t> mask(e.groups.size(), 0);
    for (int32_t gc = 0; gc < e.n_cases; ++gc)
      if (!L[gc].done) mask[e.case_group[gc]] = 1;
    if (e.set_group_mask(mask, S(stream)) && s->chunk) {
      (void)hipGraphExecDestroy(s->chunk);
      s->chunk = nullptr;
    }
    if (groups_on) {
      int32_t n = 0;
      for (uint8_t m : mask) n += m;
      *groups_on = n;
    }
  });
}

esr_status esr_solver_autotune(esr_solver* s, int32_t reps, int32_t* path, int32_t* block, int32_t* xcd_order,
                               double* us_per_stage, void* stream) {
  return guard([&] {
    if (!s || reps < 1) throw Error(ESR_ERR_VALUE, "bad arguments");
    if (!s->prepared) throw Error(ESR_ERR_STATE, "esr_solver_prepare must be called first");
    hipStream_t st = S(stream);
    Engine& e = s->eng;
    if (s->gs) hip_check(hipStreamSynchronize(s->gs), "solver stream sync");
    // launch shape of the solver: per class its kernel configuration and the
    // sizes of its launched groups
    std::vector<int64_t> key{(int64_t)e.scheme};
    for (auto& c : e.classes) {
      int64_t items = 0, fitems = 0, ng = 0;
      for (int32_t gi : c->gids) {
        if (!e.on(gi)) continue;
        items += (int64_t)e.groups[gi].m.n * e.groups[gi].C;
        fitems += (int64_t)e.groups[gi].m.nf * e.groups[gi].C;
        ++ng;
      }
      key.insert(key.end(), {c->ls.flux, c->ls.recon, c->ls.entropy, ng, items, fitems});
    }
    TuneChoice best{};
    bool cached = false;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune.find(key);
      if (it != g_tune.end()) {
        best = it->second;
        cached = true;
      }
    }
    if (!cached) {
      launch_set_active(e.n_cases, e.cases(), nullptr, 1, st);
      hipEvent_t ev[2];
      for (auto& x : ev) hip_check(hipEventCreate(&x), "event");
      // every configuration timed in two sweeps (the best of the two counts),
      // so that clock ramps and neighbours' noise do not pick a slower one
      std::vector<TuneChoice> cand;
      for (int32_t p : {ESR_STAGE_CELL, ESR_STAGE_FACE}) {
        if (e.local_dt && p == ESR_STAGE_CELL) continue;  // local time stepping runs the face path
        for (int32_t b : {64, 128, 256})
          for (int32_t x : {1, 0}) cand.push_back(TuneChoice{p, b, x, INFINITY});
      }
      for (int sweep = 0; sweep < 2; ++sweep) {
        for (TuneChoice& c : cand) {
          e.configure(c.path, c.block, c.xcd, st);
          e.stage(MODE_S1, e.W0.as<float4>(), e.W1.as<float4>(), st);  // code objects loaded
          hip_check(hipEventRecord(ev[0], st), "record");
          for (int32_t r = 0; r < reps; ++r) e.stage(MODE_S1, e.W0.as<float4>(), e.W1.as<float4>(), st);
          hip_check(hipEventRecord(ev[1], st), "record");
          hip_check(hipEventSynchronize(ev[1]), "event sync");
          float t = 0.f;
          hip_check(hipEventElapsedTime(&t, ev[0], ev[1]), "elapsed");
          c.ms = std::min(c.ms, (double)t / reps);
        }
      }
      best.ms = INFINITY;
      for (const TuneChoice& c : cand)
        if (c.ms < best.ms) best = c;
      for (auto& x : ev) (void)hipEventDestroy(x);
      launch_set_active(e.n_cases, e.cases(), e.loops_d.as<CaseLoop>(), 0, st);
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune[key] = best;
    }
    e.configure(best.path, best.block, best.xcd, st);
    // the captured chunk holds the old launch shapes
    if (s->chunk) {
      (void)hipGraphExecDestroy(s->chunk);
      s->chunk = nullptr;
    }
    if (path) *path = best.path;
    if (block) *block = best.block;
    if (xcd_order) *xcd_order = best.xcd;
    if (us_per_stage) *us_per_stage = best.ms * 1e3;
  });
}

esr_status esr_solver_set_launch(esr_solver* s, int32_t path, int32_t block, int32_t xcd_order, void* stream) {
  return guard([&] {
    if (!s) throw Error(ESR_ERR_VALUE, "null solver");
    if (path != ESR_STAGE_CELL && path != ESR_STAGE_FACE)
      throw Error(ESR_ERR_VALUE, "stage path must be ESR_STAGE_CELL or ESR_STAGE_FACE");
    if (path == ESR_STAGE_CELL && s->eng.local_dt)
      throw Error(ESR_ERR_VALUE, "local time stepping runs the face path (ESR_STAGE_FACE)");
    if (block != 64 && block != 128 && block != 256) throw Error(ESR_ERR_VALUE, "block must be 64, 128 or 256");
    if (s->gs) hip_check(hipStreamSynchronize(s->gs), "solver stream sync");
    s->eng.configure(path, block, xcd_order ? 1 : 0, S(stream));
    if (s->chunk) {
      (void)hipGraphExecDestroy(s->chunk);
      s->chunk = nullptr;
    }
  });
}

esr_status esr_set_stage_path(int32_t path) {
  return guard([&] {
    if (path != ESR_STAGE_AUTO && path != ESR_STAGE_CELL && path != ESR_STAGE_FACE)
      throw Error(ESR_ERR_VALUE, "stage path must be ESR_STAGE_AUTO, ESR_STAGE_CELL or ESR_STAGE_FACE");
    g_stage_path.store(path);
  });
}

int32_t esr_get_stage_path(void) { return resolved_stage_path(); }

esr_status esr_check_device_math(uint64_t* counts, void* stream) {
  return guard([&] {
    if (!counts) throw Error(ESR_ERR_VALUE, "counts is null");
    hipStream_t st = S(stream);
    DevBuf d;
    d.alloc(4 * sizeof(unsigned long long));
    hip_check(hipMemsetAsync(d.p, 0, 4 * sizeof(unsigned long long), st), "memset");
    launch_math_check(d.as<unsigned long long>(), st);
    unsigned long long h[4];
    hip_check(hipMemcpyAsync(h, d.p, sizeof(h), hipMemcpyDeviceToHost, st), "copy");
    hip_check(hipStreamSynchronize(st), "sync");
    for (int k = 0; k < 4; ++k) counts[k] = (uint64_t)h[k];
  });
}

esr_status esr_solver_time_kernels(esr_solver* s, int32_t reps, double* ms, void* stream) {
  return guard([&] {
    if (!s || reps <= 0 || !ms) throw Error(ESR_ERR_VALUE, "bad arguments");
    hipStream_t st = S(stream);
    Engine& e = s->eng;
    launch_set_active(e.n_cases, e.cases(), nullptr, 1, st);
    auto recon = [&] { launch_recon(e.rec_all.ls, e.cases(), e.W0.as<float4>(), e.rec.as<float4>(), st); };
    auto face = [&] {
      for (auto& cl : e.classes) launch_face(cl->ls, e.cases(), e.W0.as<float4>(), e.rec.as<float4>(),
                                             e.fx.as<float4>(), st);
    };
    auto update = [&] {
      launch_update(e.cell_all.ls, MODE_S1, e.local_dt, e.cases(), e.W0.as<float4>(), e.W0.as<float4>(),
                    e.fx.as<float4>(),
                    e.W1.as<float4>(), e.acc.as<float4>(), st);
    };
    auto cell = [&] {
      for (auto& cl : e.classes)
        launch_cell(cl->ls, MODE_S1, e.cases(), e.W0.as<float4>(), e.W0.as<float4>(), e.rec.as<float4>(),
                    e.W1.as<float4>(), e.acc.as<float4>(), st);
    };
    recon();
    face();
    update();
    cell();  // code objects loaded, buffers valid
    // Each kernel timed inside the stage sequence it belongs to (face path:
    // recon, face, update; cell path: recon, cell), with an event between
    // kernels, so every launch meets the cache state of a real stage.
    // events without the system-scope fence of a default event; the time
    // between two events is the kernel's duration plus the dispatch gap
    // before it (a profiler's kernel trace counts the duration alone)
    std::vector<hipEvent_t> ev(6 * (size_t)reps);
    for (auto& x : ev) hip_check(hipEventCreateWithFlags(&x, hipEventDisableSystemFence), "event");
    for (int32_t r = 0; r < reps; ++r) {
      hipEvent_t* q = ev.data() + 6 * (size_t)r;
      hip_check(hipEventRecord(q[0], st), "record");
      recon();
      hip_check(hipEventRecord(q[1], st), "record");
      face();
      hip_check(hipEventRecord(q[2], st), "record");
      update();
      hip_check(hipEventRecord(q[3], st), "record");
      recon();
      hip_check(hipEventRecord(q[4], st), "record");
      cell();
      hip_check(hipEventRecord(q[5], st), "record");
    }
    hip_check(hipEventSynchronize(ev.back()), "event sync");
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int32_t r = 0; r < reps; ++r) {
      hipEvent_t* q = ev.data() + 6 * (size_t)r;
      float t = 0.f;
      const int pairs[4][2] = {{0, 1}, {1, 2}, {2, 3}, {4, 5}};
      for (int k = 0; k < 4; ++k) {
        hip_check(hipEventElapsedTime(&t, q[pairs[k][0]], q[pairs[k][1]]), "elapsed");
        acc[k] += (double)t;
      }
    }
    for (int k = 0; k < 4; ++k) ms[k] = acc[k] / reps;
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (s->prepared) launch_set_active(e.n_cases, e.cases(), e.loops_d.as<CaseLoop>(), 0, st);
  });
}

esr_status esr_solver_time_kernel_runs(esr_solver* s, int32_t reps, double* ms, void* stream) {
  return guard([&] {
    if (!s || reps <= 0 || !ms) throw Error(ESR_ERR_VALUE, "bad arguments");
    hipStream_t st = S(stream);
    Engine& e = s->eng;
    launch_set_active(e.n_cases, e.cases(), nullptr, 1, st);
    // each launch set reads what the one before it in a stage wrote (rec,
    // fx) and rewrites its own outputs, so repeating it is idempotent
    auto recon = [&] { launch_recon(e.rec_all.ls, e.cases(), e.W0.as<float4>(), e.rec.as<float4>(), st); };
    auto face = [&] {
      for (auto& cl : e.classes) launch_face(cl->ls, e.cases(), e.W0.as<float4>(), e.rec.as<float4>(),
                                             e.fx.as<float4>(), st);
    };
    auto update = [&] {
      launch_update(e.cell_all.ls, MODE_S1, e.local_dt, e.cases(), e.W0.as<float4>(), e.W0.as<float4>(),
                    e.fx.as<float4>(), e.W1.as<float4>(), e.acc.as<float4>(), st);
    };
    auto cell = [&] {
      for (auto& cl : e.classes)
        launch_cell(cl->ls, MODE_S1, e.cases(), e.W0.as<float4>(), e.W0.as<float4>(), e.rec.as<float4>(),
                    e.W1.as<float4>(), e.acc.as<float4>(), st);
    };
    const std::function<void()> kern[4] = {recon, face, update, cell};
    recon();
    face();
    update();
    cell();  // code objects loaded, buffers valid
    hipEvent_t ev[2];
    for (auto& x : ev) hip_check(hipEventCreateWithFlags(&x, hipEventDisableSystemFence), "event");
    for (int k = 0; k < 4; ++k) {
      // the kernels before it in its stage, so its inputs are the stage's
      if (k >= 1) recon();
      if (k == 2) face();
      hip_check(hipEventRecord(ev[0], st), "record");
      for (int32_t r = 0; r < reps; ++r) kern[k]();
      hip_check(hipEventRecord(ev[1], st), "record");
      hip_check(hipEventSynchronize(ev[1]), "event sync");
      float t = 0.f;
      hip_check(hipEventElapsedTime(&t, ev[0], ev[1]), "elapsed");
      ms[k] = (double)t / reps;
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (s->prepared) launch_set_active(e.n_cases, e.cases(), e.loops_d.as<CaseLoop>(), 0, st);
  });
}

}  // extern "C"
