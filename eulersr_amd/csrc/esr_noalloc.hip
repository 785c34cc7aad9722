// Allocation-free single-state step (esr_step_noalloc).
//
// The same stage kernels as esr_step on the same one-case group layout, but
// every device buffer of the call -- the group and case descriptors, the
// launch block tables, the stage states W0/W1/W2, the RK4 accumulator, the
// reconstructed face states and the face fluxes -- lives in one file-scope
// device array, `scratch`, sized for meshes of up to NA_MAX_CELLS cells and
// NA_MAX_FACES logical faces. Descriptors are written by one-thread kernels
// that take them as arguments and block tables by a fill kernel, so the call
// makes no device allocation and no host-to-device copy.
//
// The host code reaches the array only through the address
// hipGetSymbolAddress returns for the current device (looked up on the
// entry point's first call on that device and kept per device): the array's
// name in host code is a host-side stand-in, not a device address.
//
// Every call on a device shares that device's array: calls that overlap on
// one device (two streams, two host threads) race. Nothing in the array has
// to start at zero: each call writes every descriptor and table it reads,
// and every stage writes its outputs before a later stage reads them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "esr_launch.h"

using namespace esr;

namespace {

constexpr int32_t NA_MAX_CELLS = 1 << 17;   // 131,072 cells
constexpr int32_t NA_MAX_FACES = 1 << 18;   // 262,144 logical faces
constexpr int32_t NA_BLOCK = BLOCK;         // the block of esr_step's one-case engine
constexpr int32_t pad8(int32_t b) { return (b + 7) / 8 * 8; }
constexpr int32_t NA_CELL_BLOCKS = pad8((NA_MAX_CELLS + NA_BLOCK - 1) / NA_BLOCK);
constexpr int32_t NA_FACE_BLOCKS = pad8((NA_MAX_FACES + NA_BLOCK - 1) / NA_BLOCK);
constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t STATE_BYTES = (size_t)NA_MAX_CELLS * sizeof(float4);
// layout of the array (byte offsets, 256-aligned)
constexpr size_t OFF_GROUP = 0;
constexpr size_t OFF_CASE = al256(OFF_GROUP + sizeof(GroupDev));
constexpr size_t OFF_CBLK = al256(OFF_CASE + sizeof(CaseDev));
constexpr size_t OFF_FBLK = al256(OFF_CBLK + NA_CELL_BLOCKS * sizeof(BlockRef));
constexpr size_t OFF_W0 = al256(OFF_FBLK + NA_FACE_BLOCKS * sizeof(BlockRef));
constexpr size_t OFF_W1 = OFF_W0 + STATE_BYTES;
constexpr size_t OFF_W2 = OFF_W1 + STATE_BYTES;
constexpr size_t OFF_ACC = OFF_W2 + STATE_BYTES;
constexpr size_t OFF_REC = OFF_ACC + STATE_BYTES;              // [3][n] face states
constexpr size_t OFF_FX = OFF_REC + 3 * STATE_BYTES;           // [nf] face fluxes
constexpr size_t SCRATCH_BYTES = OFF_FX + (size_t)NA_MAX_FACES * sizeof(float4);
static_assert(SCRATCH_BYTES == 18887424, "scratch size is the one include/eulersr.h states");

}  // namespace

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

__global__ void k_na_blocks(BlockRef* tbl, int32_t used, int32_t total) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) {
    BlockRef r;
    r.group = i < used ? 0 : -1;
    r.block = i < used ? i : 0;
    tbl[i] = r;
  }
}

void fill_blocks(BlockRef* tbl, int32_t used, int32_t total, hipStream_t st) {
  hipLaunchKernelGGL(k_na_blocks, dim3((total + 255) / 256), dim3(256), 0, st, tbl, used, total);
  hip_check(hipGetLastError(), "k_na_blocks");
}

// device address of `scratch` per device, looked up once per device
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_ptr[MAX_DEVICES] = {};

unsigned char* scratch_ptr() {
  int dev = -1;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= MAX_DEVICES) throw Error(ESR_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(ESR_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_ptr[dev];
}

template <class F>
esr_status guard(F&& f) {
  try {
    f();
    return ESR_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return ESR_ERR_DEVICE;
  }
}

}  // namespace

extern "C" esr_status esr_step_noalloc(esr_mesh* mesh, const esr_opts* opts, int32_t scheme, float dt,
                                       const float* W, float* W_out, void* stream) {
  return guard([&] {
    if (!mesh || !W || !W_out) throw Error(ESR_ERR_VALUE, "null argument");
    validate_opts(opts);
    const std::vector<Stage> stages = stages_of(scheme);
    if (mesh->n > NA_MAX_CELLS || mesh->dev.nf > NA_MAX_FACES)
      throw Error(ESR_ERR_VALUE, "esr_step_noalloc: the mesh has " + std::to_string(mesh->n) + " cells and " +
                                     std::to_string(mesh->dev.nf) + " faces; at most " +
                                     std::to_string(NA_MAX_CELLS) + " cells and " + std::to_string(NA_MAX_FACES) +
                                     " faces fit its fixed scratch array (use esr_step)");
    if (opts->sponge_width > 0)
      throw Error(ESR_ERR_VALUE, "esr_step_noalloc: a custom sponge width needs its own weights (use esr_step)");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    unsigned char* base = scratch_ptr();
    GroupDev* gdev = reinterpret_cast<GroupDev*>(base + OFF_GROUP);
    CaseDev* cdev = reinterpret_cast<CaseDev*>(base + OFF_CASE);
    BlockRef* cblk = reinterpret_cast<BlockRef*>(base + OFF_CBLK);
    BlockRef* fblk = reinterpret_cast<BlockRef*>(base + OFF_FBLK);
    float4* Wb[3] = {reinterpret_cast<float4*>(base + OFF_W0), reinterpret_cast<float4*>(base + OFF_W1),
                     reinterpret_cast<float4*>(base + OFF_W2)};
    float4* acc = reinterpret_cast<float4*>(base + OFF_ACC);
    float4* rec = reinterpret_cast<float4*>(base + OFF_REC);
    float4* fx = reinterpret_cast<float4*>(base + OFF_FX);

    // the one-case group of esr_step's engine (C = 1, every slab at offset 0)
    GroupDev g{};
    g.m = mesh->dev;
    g.C = 1;
    g.cmagic = 0;
    g.cshift = 0;
    g.case0 = 0;
    g.st_off = 0;
    g.rec_off = 0;
    g.fx_off = 0;
    g.flux = opts->flux;
    g.ph = make_phys(opts->gamma, opts->M);
    g.sponge = mesh->sponge;
    g.sponge_g = mesh->dev.sponge_g;
    CaseDev cd = case_params(mesh, *opts);
    cd.dt = dt;
    cd.dt_half = dt / 2.0f;
    cd.dt_sixth = dt / 6.0f;
    cd.active = 1;
    cd.check = 0;
    launch_set_group(gdev, g, st);
    launch_set_case(cdev, cd, st);
    const int32_t nb = pad8((mesh->n + NA_BLOCK - 1) / NA_BLOCK);
    const int32_t nfb = pad8((mesh->dev.nf + NA_BLOCK - 1) / NA_BLOCK);
    fill_blocks(cblk, (mesh->n + NA_BLOCK - 1) / NA_BLOCK, nb, st);
    fill_blocks(fblk, (mesh->dev.nf + NA_BLOCK - 1) / NA_BLOCK, nfb, st);

    LaunchSet all;   // the update and the commit (flux-independent)
    all.groups = gdev;
    all.blk_group = cblk;
    all.nblocks = nb;
    all.fblk_group = fblk;
    all.fnblocks = nfb;
    all.block = NA_BLOCK;
    all.xcd = 1;
    LaunchSet cls = all;   // the flux kernels of the case's class
    cls.flux = (opts->flux == ESR_FLUX_HLLC || opts->flux == ESR_FLUX_ROE) ? FLUX_HLLC_ROE : opts->flux;
    cls.recon = opts->recon;
    cls.entropy = opts->entropy ? 1 : 0;
    const bool muscl = opts->recon == ESR_RECON_MUSCL;
    const bool face_path = resolved_stage_path() == ESR_STAGE_FACE;

    launch_gather_in_dev(mesh->n, gdev, 0, mesh->b_perm.as<int32_t>(), reinterpret_cast<const float4*>(W), Wb[0],
                         st);
    for (const Stage& sg : stages) {
      const float4* Ws = Wb[sg.ws];
      float4* out = Wb[sg.out];
      if (muscl) launch_recon(all, cdev, Ws, rec, st);
      if (cls.flux == ESR_FLUX_LF) {
        launch_zero_lf(1, cdev, st);
        launch_lf_cmax(cls, cdev, Ws, rec, st);
      }
      if (face_path) {
        launch_face(cls, cdev, Ws, rec, fx, st);
        launch_update(all, sg.mode, false, cdev, Ws, Wb[0], fx, out, acc, st);
      } else {
        launch_cell(cls, sg.mode, cdev, Ws, Wb[0], rec, out, acc, st);
      }
    }
    if (scheme == ESR_SCHEME_EE) launch_commit(all, cdev, Wb[1], Wb[0], st);
    launch_scatter_out_dev(mesh->n, gdev, 0, mesh->b_perm.as<int32_t>(), Wb[0], reinterpret_cast<float4*>(W_out),
                           st);
  });
}
