/*
 * eulersr.h -- C ABI of the MI355X explicit 2D unstructured finite-volume
 * Euler solver (libeulersr.so, gfx950).
 *
 * The reference exposes no FFI: its boundary is a set of Python functions
 * (euler/jax_fvm/src/euler_solver.py, helper.py, euler/main.py). Each entry
 * point below names the reference function it replaces; the Python package
 * eulersr_amd binds them with ctypes and keeps the reference signatures.
 *
 * Conventions
 *   - State arrays W are float32, shape (n_cells, 4) row-major, conserved
 *     variables (rho, rho u, rho v, E), cells in the mesh's ORIGINAL order.
 *     Pointers passed to compute entry points are DEVICE pointers; the
 *     library renumbers internally (reverse Cuthill-McKee) and undoes it.
 *   - Mesh descriptions are HOST pointers; the library copies them.
 *   - `stream` is a hipStream_t (NULL = default stream). Entry points that
 *     return host values synchronise that stream.
 *   - Every entry point returns an esr_status; esr_last_error() gives the
 *     message of the calling thread's last failure.
 *   - Handles are not thread-safe; use one handle per host thread.
 */
#ifndef EULERSR_H
#define EULERSR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  ESR_OK = 0,
  ESR_ERR_VALUE = 1,    /* bad argument (Python ValueError)            */
  ESR_ERR_KEY = 2,      /* unknown option key (Python KeyError)        */
  ESR_ERR_DEVICE = 3,   /* HIP runtime failure (Python RuntimeError)   */
  ESR_ERR_MEMORY = 4,   /* allocation failure (Python MemoryError)     */
  ESR_ERR_STATE = 5     /* call not valid in the handle's state        */
} esr_status;

/* numerical flux: euler_solver.py:124-139 ('hll' maps to HLLC) */
typedef enum {
  ESR_FLUX_LF = 0, ESR_FLUX_RUSANOV = 1, ESR_FLUX_TADMOR = 2,
  ESR_FLUX_AUSM = 3, ESR_FLUX_HLLC = 4, ESR_FLUX_ROE = 5
} esr_flux;

/* reconstruction: euler_solver.py:106-122 */
typedef enum { ESR_RECON_CONSTANT = 0, ESR_RECON_MUSCL = 1 } esr_recon;

/* time scheme: euler_solver.py:630-667 (dispatch table main.py:50-51) */
typedef enum {
  ESR_SCHEME_EE = 0, ESR_SCHEME_RK2 = 1, ESR_SCHEME_SRK2 = 2, ESR_SCHEME_RK4 = 3
} esr_scheme;

typedef struct esr_mesh esr_mesh;
typedef struct esr_solver esr_solver;

/* Host description of a mesh, fields of the reference Mesh pytree
 * (jax_fvm/src/mesh.py:28-83); geometry already cast to float32. */
typedef struct {
  int32_t n_points, n_cells, n_faces;
  const float* points;                      /* (n_points,2)              */
  const int32_t* tris;                      /* (n_cells,3)               */
  const int32_t* neighbors;                 /* (n_cells,3), -1 boundary  */
  const int32_t* faces;                     /* (n_faces,2)               */
  const int32_t* face_markers;              /* (n_faces,)                */
  const int32_t* face_connectivity;         /* (n_cells,3)               */
  const int32_t* face_connectivity_opposite;/* (n_faces,)                */
  const float* barycenter;                  /* (n_cells,2)               */
  const float* area;                        /* (n_cells,)                */
  const float* surface;                     /* (n_faces,)                */
  const float* normals;                     /* (n_cells,3,2)             */
  const float* inlet_subsonic;              /* (4,) primitives           */
  int32_t n_wall_markers;                   /* helper.py:237             */
  const int32_t* wall_markers;
  int32_t force_marker;                     /* helper.py:370 (default 2) */
  int32_t sponge;                           /* 0 when metadata case=='bump' (helper.py:266) */
  int32_t renumber;                         /* 1: reverse Cuthill-McKee  */
  const int32_t* rcm_seed_order;            /* (n_cells,) order in which RCM
                                               seeds components, as
                                               numpy.argsort(degree) in scipy's
                                               reverse_cuthill_mckee; NULL: by
                                               degree, ties by lowest index  */
} esr_mesh_desc;

/* Solver keyword arguments of residual()/time_step_*() (euler_solver.py:602-667,
 * helper.py:262-295). sponge_width/strength < 0 select the defaults. */
typedef struct {
  int32_t flux;          /* esr_flux  */
  int32_t recon;         /* esr_recon */
  int32_t entropy;       /* 1: Ismail-Roe central flux (getFlux_Tadmor) */
  double gamma;          /* double: derived constants (gamma-1, ...) are   */
  double M;              /* rounded to float32 once, as Python does        */
  float value[4];        /* inlet conserved state ('value' kwarg) */
  float sponge_width;
  float sponge_strength;
} esr_opts;

/* -- mesh ------------------------------------------------------------------ */

/* Replaces Mesh.mesh_generator_from_points + load_mesh (mesh.py:131-150,
 * 337-359) on the device side: validates, renumbers (RCM) and uploads. */
esr_status esr_mesh_create(const esr_mesh_desc* desc, esr_mesh** out);
esr_status esr_mesh_destroy(esr_mesh* mesh);
/* perm[new] = old cell index (identity when renumber == 0). */
esr_status esr_mesh_permutation(const esr_mesh* mesh, int32_t* perm);
esr_status esr_mesh_info(const esr_mesh* mesh, int32_t* n_cells, int32_t* n_wall_faces,
                         int64_t* device_bytes);
/* Host-only: the reverse Cuthill-McKee order esr_mesh_create uses, from a
 * (n_cells,3) neighbour table (-1 = boundary); the algorithm of
 * scipy.sparse.csgraph.reverse_cuthill_mckee (scipy 1.15.3, symmetric mode),
 * seeding components in `seed_order` (NULL: by degree, ties by index).
 * perm[new] = old. No GPU needed. No reference counterpart (the reference
 * keeps the mesher's order). */
esr_status esr_rcm(int32_t n_cells, const int32_t* neighbors, const int32_t* seed_order, int32_t* perm);

/* -- single-state operations (device W, original cell order) -------------- */

/* residual(W, mesh, **kw) -- euler_solver.py:602-628. R = sum(F*len)/area - sponge. */
esr_status esr_residual(esr_mesh* mesh, const esr_opts* opts, const float* W, float* R,
                        void* stream);

/* time_step_{Euler,RK2,RK2_SSP,RK4}(W, mesh, dt, **kw) -- euler_solver.py:630-667.
 * W_out may equal W. */
esr_status esr_step(esr_mesh* mesh, const esr_opts* opts, int32_t scheme, float dt,
                    const float* W, float* W_out, void* stream);

/* esr_step without any device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * kernels, same results bit for bit; every buffer of the call lives in one
 * fixed 18,887,424-byte device array per GPU (held until the process ends),
 * which limits meshes to 131,072 cells and 262,144 logical faces and the
 * sponge to its default width (ESR_ERR_VALUE beyond). Calls that overlap on
 * one GPU share that array and race. The mesh's device must be current. */
esr_status esr_step_noalloc(esr_mesh* mesh, const esr_opts* opts, int32_t scheme, float dt,
                            const float* W, float* W_out, void* stream);

/* helper.get_dt(W, mesh, CFL, gamma, M) -- helper.py:8-18 (synchronises). */
esr_status esr_get_dt(esr_mesh* mesh, const float* W, float cfl, double gamma, double M,
                      float* dt_out, void* stream);

/* get_drag_coefficient / get_lift_coefficient -- helper.py:366-415, wall sums in
 * float64; q_ref = 0.5 rho_inf U_inf^2 L_ref (synchronises). */
esr_status esr_forces(esr_mesh* mesh, const float* W, double q_ref, double* cd, double* cl,
                      void* stream);

/* ||A - B||_F and ||B||_F in float64 over n_cells x 4 (the norms of main.py:152, 230). */
esr_status esr_norms(esr_mesh* mesh, const float* A, const float* B, double* diff_norm,
                     double* ref_norm, void* stream);

/* -- batched solve loop (main.py:116-230 for many cases) ------------------ */

typedef struct {
  const esr_mesh* mesh;
  esr_opts opts;
  float cfl;                 /* after the transonic clamp of main.py:117-120 */
  float tf;                  /* N = int(tf/dt) + 1 (main.py:122)             */
  double q_ref;              /* 0.5 rho U^2 L for CD/CL (<= 0: no forces)    */
  int32_t track_engineering; /* engineering CD/CL criterion (main.py:66-114) */
  double target_cd, target_cl, cd_tol_abs, cl_tol_abs;
  int32_t engineering_patience;
  double threshold;          /* > 0: per-case stationarity threshold (the
                                HR-probe x 1.001 rule of eval/convergence_case.py:61-66);
                                <= 0: esr_solve_opts.threshold */
  int64_t max_steps;         /* > 0: N = min(int(tf/dt) + 1, max_steps), a per-case
                                step cap of the campaign harness (no reference
                                counterpart); 0: none */
  int32_t group;             /* cases with different values never share a launch
                                group (a scheduler's sub-groups of one mesh and
                                kernel configuration; no reference counterpart); 0 */
} esr_case;

typedef struct {
  int32_t scheme;            /* esr_scheme, shared by all cases */
  double threshold;          /* stationarity_threshold          */
  int32_t check_every;       /* stationarity_check_every        */
  int32_t patience;          /* stationarity_patience (>= 1)    */
  int64_t max_steps;         /* extra cap on steps (<= 0: none) */
  int32_t local_dt;          /* 0: one global dt per case from its initial state
                                (main.py:116-122, helper.py:8-18). 1: local time
                                stepping, an acceleration mode with no reference
                                counterpart: every cell advances with its own
                                dt = cfl * area / sum_k (c + |u . n_k|) len_k of
                                the step's start state (the same dt in every stage;
                                the global dt is its minimum over cells). A steady
                                state is the same; the path to it, and therefore
                                the stationarity residual, is not -- measure such
                                solves by CD/CL (engineering targets). */
  int32_t no_graph;          /* 1: launch every kernel eagerly instead of replaying
                                a captured HIP graph of check_every steps */
} esr_solve_opts;

typedef struct {
  float dt;
  int64_t n_steps;           /* N */
  int32_t converged;
  int64_t stopping_step;
  int64_t steps_run;
  double last_rel;
  double final_residual;     /* one-step probe, main.py:228-230 */
  double cd, cl;
  int32_t engineering_converged;
  int64_t engineering_stopping_step;
  int32_t diverged;          /* non-finite residual at a check: the case stops there, unconverged */
  double recent_rel_max;     /* largest stationarity residual of the last min(16, checks) checks */
} esr_case_report;

esr_status esr_solver_create(int32_t n_cases, const esr_case* cases, const esr_solve_opts* so,
                             esr_solver** out);
esr_status esr_solver_destroy(esr_solver* s);
/* Initial state of case c (device pointer, original cell order, (n,4) float32). */
esr_status esr_solver_set_state(esr_solver* s, int32_t c, const float* W, void* stream);
esr_status esr_solver_get_state(esr_solver* s, int32_t c, float* W, void* stream);
/* Computes dt per case from the current states (helper.py:8-18 once per run,
 * main.py:121) and N = int(tf/dt)+1. */
esr_status esr_solver_prepare(esr_solver* s, void* stream);
/* Advances until every case has stopped or reached N (or max_steps, or the
 * absolute step `until_step` when > 0); *steps = steps taken since prepare.
 * Successive calls continue the same loop. */
esr_status esr_solver_run(esr_solver* s, int64_t until_step, int64_t* steps, void* stream);
/* Continuous batching: puts a new case into slot c of a prepared solver,
 * between two esr_solver_run calls, typically once the slot's case is done.
 * The new case must use the slot's mesh and kernel configuration (flux,
 * reconstruction, entropy, gamma, M, sponge width); its parameters, dt
 * (from W, device pointer, original cell order), N, stationarity state and
 * reference state are set afresh, and its steps count from the current
 * solver step (reports give steps relative to that start). The other cases
 * are untouched. No counterpart in the reference (one case per process). */
esr_status esr_solver_replace_case(esr_solver* s, int32_t c, const esr_case* spec, const float* W, void* stream);
/* Continuous batching: drops from the launches every group (cases sharing a
 * mesh and kernel configuration) whose cases are all done, until a case
 * enters one of its slots again (esr_solver_replace_case). *groups_on (may be
 * null) receives the number of groups still launched. No counterpart in the
 * reference. */
esr_status esr_solver_compact(esr_solver* s, int32_t* groups_on, void* stream);
/* One-step probe residual + CD/CL of the current states, for every case
 * (groups dropped by esr_solver_compact are launched again). */
esr_status esr_solver_finalize(esr_solver* s, void* stream);
esr_status esr_solver_report(const esr_solver* s, int32_t c, esr_case_report* out);
/* Reports of all n cases (n = the solver's case count) with one device copy. */
esr_status esr_solver_reports(const esr_solver* s, int32_t n, esr_case_report* out);

/* -- learned warm start ----------------------------------------------------- */
/* The reference's warm start is a Flax network over kNN/IDW features
 * (models/dam.py:302-304: IDW baseline + learned delta; kNN k=6 LR->HR,
 * preprocessing/knn.py:20-40; IDW p=2, utils/metrics.py:9-14; wall features,
 * utils/aero.py:154-181). Here: a multilayer perceptron (ReLU hidden layers,
 * linear output) trained and run by hand-written BF16 MFMA kernels with FP32
 * accumulation. Parameters are float32, per layer W (out x in, row-major)
 * then b (out). X is bfloat16 (rows x widths[0]) on the device. */
typedef struct esr_mlp esr_mlp;

esr_status esr_mlp_create(int32_t n_layers, const int32_t* widths, int64_t max_rows, esr_mlp** out);
esr_status esr_mlp_destroy(esr_mlp* m);
esr_status esr_mlp_param_count(const esr_mlp* m, int64_t* n);
esr_status esr_mlp_set_params(esr_mlp* m, const float* host);   /* also resets the AdamW moments */
esr_status esr_mlp_get_params(const esr_mlp* m, float* host);
/* Y (rows x widths[n_layers], float32, device) = network(X); any row count. */
esr_status esr_mlp_forward(esr_mlp* m, int64_t rows, const uint16_t* X, float* Y, void* stream);
/* Forward, mean-squared error against T (rows x out, float32, device) and
 * backward; gradients kept on the device (and copied to grads_host if not
 * null); rows <= max_rows. Synchronises (returns the loss). */
esr_status esr_mlp_loss_grad(esr_mlp* m, int64_t rows, const uint16_t* X, const float* T, double* loss,
                             float* grads_host, void* stream);
/* One AdamW update (decoupled weight decay on W only) with step count t >= 1. */
esr_status esr_mlp_adamw(esr_mlp* m, float lr, float beta1, float beta2, float eps, float weight_decay,
                         int32_t t, void* stream);

/* LR -> HR neighbour map, device arrays in original cell order. */
typedef struct {
  int32_t n_hr, n_lr, k;
  const int32_t* knn_idx;   /* (n_hr, k) LR cell of each neighbour          */
  const float* knn_rel;     /* (n_hr, k, 2) (LR barycentre - HR) / h_LR      */
  const float* knn_w;       /* (n_hr, k) normalised IDW weights              */
  const float* wall;        /* (n_hr, 4) [wd/L, exp(-wd/(2 h_LR)), n_x, n_y] */
} esr_knn_map;

/* Network inputs for n_cases LR states (n_cases x n_lr x 4 conserved,
 * device): X (n_cases*n_hr rows x in_pad, bf16) and the IDW prediction
 * (n_cases*n_hr x 4 primitives). case_scal: (n_cases, 2) normalised
 * (Mach, alpha) on the device; mu/sig: host primitive normalisation. */
esr_status esr_warm_features(const esr_knn_map* map, int32_t n_cases, const float* lr_states,
                             const float* case_scal, const float* mu, const float* sig, double gamma,
                             int32_t in_pad, uint16_t* X, float* idw, void* stream);
/* Warm-start state: prim = idw + sig * delta (delta null: IDW only) ->
 * conserved W_out (rows x 4); bad[case] |= 1 when a cell of the case has
 * rho <= 0, p <= 0 or a non-finite value (the freestream fallback test). */
esr_status esr_warm_apply(int64_t rows, int32_t n_hr, const float* idw, const float* delta, const float* sig,
                          double gamma, float* W_out, int32_t* bad, void* stream);

/* -- diagnostics ----------------------------------------------------------- */

/* Sizes of one stage launch: cell work-items (cell x case), distinct group
 * cells, face work-items (logical face x case) and distinct group faces. */
esr_status esr_solver_info(const esr_solver* s, int64_t* cell_items, int64_t* group_cells, int64_t* face_items,
                           int64_t* group_faces);
/* Average device time (ms) of one launch set of each stage kernel over the
 * solver's current states: ms[0] reconstruction, ms[1] face flux, ms[2]
 * gather + stage-1 update (the per-face path), ms[3] cell-centric flux +
 * stage-1 update (the cell path). Each is timed between HIP events on
 * `stream` inside the stage sequence it belongs to (recon, face, update;
 * recon, cell), `reps` sequences each, so a time includes the dispatch gap
 * before the kernel; all are idempotent (they rewrite scratch buffers and
 * the stage buffer only). ms holds 4 entries. */
esr_status esr_solver_time_kernels(esr_solver* s, int32_t reps, double* ms, void* stream);
/* The same four kernels each launched `reps` times back to back between two
 * HIP events on `stream` (after the kernels before it in its stage, so its
 * inputs are a stage's): ms[k] = the average per launch, the kernel's
 * duration plus only the gap between two queued launches, where
 * esr_solver_time_kernels' intervals also hold each launch's dispatch after
 * an event. Idempotent as esr_solver_time_kernels. No counterpart in the
 * reference. */
esr_status esr_solver_time_kernel_runs(esr_solver* s, int32_t reps, double* ms, void* stream);

/* Exhaustive self-check of the device arithmetic: the square root and the
 * natural logarithm that the flux kernels evaluate (sound speeds, Roe
 * averages, entropy jumps) against sqrtf and logf, bit for bit, over all 2^31
 * non-negative float bit patterns. counts[0] = square-root mismatches at
 * inputs of at least 2^-96, counts[1] = below 2^-96 (tiny and denormal
 * inputs, which those quantities never take); counts[2], counts[3] the same
 * for the logarithm. Synchronous on `stream`. No counterpart in the reference
 * (a check of this implementation's arithmetic). */
esr_status esr_check_device_math(uint64_t* counts, void* stream);

/* Stage path of the engines built after the call (solvers created, and the
 * per-mesh engine of esr_residual / esr_step at its next call):
 * ESR_STAGE_CELL -- per stage a reconstruction launch and one cell-centric
 * launch that evaluates the three face fluxes of each cell (each interior
 * face from both of its cells, like the reference's per-cell flux arrays),
 * the residual and the update; ESR_STAGE_FACE -- reconstruction, one flux per
 * logical face, then a gather + update launch (each face evaluated once, the
 * neighbour takes its negation). ESR_STAGE_AUTO is the cell path. No
 * counterpart in the reference (a tuning choice of this implementation). */
enum esr_stage_path { ESR_STAGE_AUTO = 0, ESR_STAGE_CELL = 1, ESR_STAGE_FACE = 2 };
esr_status esr_set_stage_path(int32_t path);
int32_t esr_get_stage_path(void);

/* Autotuner: times one SSP-RK2 stage of the prepared solver on its current
 * states for each stage path (cell, face) x block size (64, 128, 256 threads)
 * x block order (XCD-contiguous runs or plain) with HIP events, `reps`
 * launches each, and keeps the fastest; the choice is cached per launch
 * shape (kernel configurations and work-item counts) for the process, so
 * later solvers of the same shape reuse it without timing. Outputs (any may
 * be null): the chosen path, block and order, and the device time of one
 * stage in microseconds. No counterpart in the reference. */
esr_status esr_solver_autotune(esr_solver* s, int32_t reps, int32_t* path, int32_t* block, int32_t* xcd_order,
                               double* us_per_stage, void* stream);
/* Pins the launch configuration of a solver: stage path (ESR_STAGE_CELL or
 * ESR_STAGE_FACE), block size (64, 128 or 256) and block order (1: each XCD
 * takes a contiguous run of blocks, 0: the hardware's round-robin order).
 * Results do not depend on the block size or order. No counterpart in the
 * reference. */
esr_status esr_solver_set_launch(esr_solver* s, int32_t path, int32_t block, int32_t xcd_order, void* stream);
const char* esr_last_error(void);
int32_t esr_version(void);

#ifdef __cplusplus
}
#endif
#endif /* EULERSR_H */
