/*
 * swe_hip.h -- C ABI of the MI355X (gfx950) shallow-water backend, libswe_hip.so.
 *
 * The reference has no foreign-function interface: its backends are C++ classes selected by
 * SolverAssembly (include/solver_assembly/solver_assembly.hpp:34-120).  The replaced
 * single-GPU backend, FastHLLMUSCLBathyCUDASolver (include/backend/cuda/cuda_solver.hpp),
 * keeps its device arrays and kernels private.  This header exposes the same operations as
 * a C ABI (plain pointers and sizes, no C++ or torch types) so the host-only C++ backend
 * class, Python (ctypes) and any other FFI can drive them.  Each entry point names the
 * reference code it replaces.
 *
 * Layout.  Every array is float64 of n_total = (nx + 2*ng) * (ny + 2*ng) elements in the
 * reference layout: row-major, k = i*(ny + 2*ng) + j, i = x index (slow), j = y index
 * (contiguous), interior i, j in [ng, ng + n).  (omp_kernels.hpp:33-35, grid.hpp:11-12)
 *
 * Conventions.  Every int-returning function returns SWE_OK (0) or an swe_status code;
 * swe_hip_last_error() gives the message for the most recent failure on that context.
 * Host pointers are caller-owned and not retained.  Calls on one context are not
 * thread-safe; distinct contexts are independent.  All device work of a context runs on
 * its stream; functions that return host data synchronize that stream.
 */
#ifndef SWE_HIP_H
#define SWE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWE_HIP_ABI_VERSION 1

typedef enum {
    SWE_OK = 0,
    SWE_ERR_INVALID = 1,    /* bad argument or descriptor */
    SWE_ERR_HIP = 2,        /* HIP runtime error (message in swe_hip_last_error) */
    SWE_ERR_NONFINITE = 3,  /* non-finite h/hu/hv after a stage (omp_kernels.hpp:322-330) */
    SWE_ERR_CFL = 4,        /* dt exceeds the CFL limit (omp_solver.hpp:159-162) */
    SWE_ERR_OOM = 5,        /* device allocation failed */
    SWE_ERR_NO_DEVICE = 6   /* no usable HIP device */
} swe_status;

typedef enum {
    SWE_PRECISION_EXACT = 0, /* the reference's arithmetic, bitwise equal to omp_solver */
    SWE_PRECISION_FAST = 1   /* reciprocal-based velocities; differs from the reference by a
                                few ulps per step (bounded by the tests after N steps) */
} swe_precision;

typedef enum {
    SWE_BC_REFLECTING = 0,  /* reference walls (omp_solver.hpp:340-432) */
    SWE_BC_TRANSMISSIVE = 1 /* zero-gradient copy; extension, no reference counterpart */
} swe_bc_kind;

/* Scheme family of the stage kernels (swe_hip_set_scheme). */
typedef enum {
    SWE_SCHEME_OPTIMIZED = 0,    /* MUSCL + HLL + hydrostatic reconstruction + SSP-RK3 of the
                                    optimized backends (omp_kernels.hpp); the default */
    SWE_SCHEME_SERIAL = 1,       /* the modular SerialSolver with bathymetry
                                    (flux_assembly_bathy.hpp), HLL only, MUSCL or PC */
    SWE_SCHEME_SERIAL_PLAIN = 2  /* the modular SerialSolver without bathymetry
                                    (flux_assembly.hpp), any reconstruction x Riemann solver */
} swe_scheme;

typedef enum {
    SWE_RECON_MUSCL = 0,  /* reconstruction/muscl.hpp */
    SWE_RECON_PC = 1      /* reconstruction/piecewise_const.hpp */
} swe_recon;

typedef enum {
    SWE_RIEMANN_HLL = 0,      /* riemann/HLL.hpp */
    SWE_RIEMANN_RUSANOV = 1,  /* riemann/rusanov.hpp */
    SWE_RIEMANN_ROE = 2       /* riemann/ROE.hpp */
} swe_riemann;

/* Grid descriptor.  Replaces Grid (include/core/grid.hpp:3-31) + GridView
 * (omp_kernels.hpp:22-31) + constants::g (include/constants.hpp:4).  dx = Lx/Nx. */
typedef struct {
    int32_t nx, ny, ng; /* interior cells per axis; ghost width (>= 2) */
    double dx, dy;
    double g;           /* gravity; the reference uses 9.80665 */
    int32_t bc_kind;    /* swe_bc_kind */
} swe_grid_desc;

typedef struct swe_hip_ctx swe_hip_ctx;

int swe_hip_abi_version(void);

/* Status code -> static string. */
const char *swe_hip_error_string(int code);

/* Detailed message of the last failure on ctx ("" if none).  With ctx == NULL: the message
 * of the calling thread's last failed swe_hip_create (which leaves no context behind). */
const char *swe_hip_last_error(const swe_hip_ctx *ctx);

/* Outcome of the last stepping call on ctx (swe_hip_step, swe_hip_step_adaptive[_noalloc],
 * swe_hip_time_steps): *step = steps it started, *stage = the stage (1..3) whose output was
 * non-finite, or 0.  When *stage != 0 the call returned SWE_ERR_NONFINITE at step *step
 * (1-based within the call) and did nothing after that stage, as the reference's
 * rk3_step throws there (omp_kernels.hpp:322-330, omp_solver.hpp:250-338): the state is
 * U^n of that step when stage 1 or 2 failed, the positivity-floored stage-3 output when
 * stage 3 did.  Replaces what that exception ends (the run loop, omp_solver.hpp:185-214).
 * Either output may be NULL. */
int swe_hip_last_status(const swe_hip_ctx *ctx, int64_t *step, int *stage);

/* Number of visible HIP devices (0 when there is no GPU). */
int swe_hip_device_count(void);

/* Allocate a context on `device`: U, U1, U2 (3 arrays each), B, scalars.
 * Replaces FastHLLMUSCLBathyCUDASolver::allocate_device (cuda_solver.hpp:233-252). */
int swe_hip_create(const swe_grid_desc *desc, int device, swe_hip_ctx **out);

/* Free all device memory of ctx (cuda_solver.hpp:254-273).  NULL is a no-op. */
void swe_hip_destroy(swe_hip_ctx *ctx);

/* Upload state and bathymetry from host (n_total each; ghosts are ignored and rebuilt:
 * B gets the scalar ghost fill of omp_solver.hpp:395-432 as at solver construction).
 * Replaces upload_initial_state_to_device (cuda_solver.hpp:277-286). */
int swe_hip_upload(swe_hip_ctx *ctx, const double *h, const double *hu, const double *hv,
                   const double *b);

/* Download the current state with its ghost cells filled as after apply_reflecting_bc.
 * Replaces download_state_from_device (cuda_solver.hpp:288-296). */
int swe_hip_download(swe_hip_ctx *ctx, double *h, double *hu, double *hv);

/* Same as upload/download with device pointers (e.g. torch tensors on ctx's device);
 * the copies are ordered on ctx's stream. */
int swe_hip_upload_device(swe_hip_ctx *ctx, const double *h, const double *hu, const double *hv,
                          const double *b);
int swe_hip_download_device(swe_hip_ctx *ctx, double *h, double *hu, double *hv);

/* compute_stable_dt of the current state (omp_kernels.hpp:288-307).  Either output may be
 * NULL. */
int swe_hip_stable_dt(swe_hip_ctx *ctx, double cfl, double *dt_out, double *max_speed_out);

/* Advance nsteps SSP-RK3 steps with a fixed dt: the reference's run loop
 * (omp_solver.hpp:185-214 / rk3_step :250-338; cuda_solver.hpp:195-227 / :400-417).
 * A non-finite stage output is detected on the device and stops the call there, as the
 * reference's throw does: the later launches of the call do nothing, the call returns
 * SWE_ERR_NONFINITE and swe_hip_last_status names the step and stage. */
int swe_hip_step(swe_hip_ctx *ctx, double dt, int64_t nsteps);

/* Advance nsteps steps, each with dt = compute_stable_dt(U^n, cfl) evaluated on the device
 * (extension: the reference fixes dt = end_time/steps, omp_solver.hpp:99).  dt_hist (may be
 * NULL) receives the nsteps dt values used (quiet NaN after a non-finite stop, whose step
 * and stage swe_hip_last_status gives). */
int swe_hip_step_adaptive(swe_hip_ctx *ctx, double cfl, int64_t nsteps, double *dt_hist);

/* swe_hip_step_adaptive that allocates nothing: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and results.  The dt history goes to
 * dt_hist_device, a caller-owned device array of nsteps doubles (NULL: not recorded); the
 * CFL reduction's per-block maxima and the history counter live in a fixed scratch array
 * of the library (one per GPU, held until the process ends), so the stage launch may have
 * at most 65536 blocks (SWE_ERR_INVALID with a message beyond that).  Steps are launched
 * directly (no graph).  Calls that overlap on one GPU share that scratch and race. */
int swe_hip_step_adaptive_noalloc(swe_hip_ctx *ctx, double cfl, int64_t nsteps,
                                  double *dt_hist_device);

/* compute_rhs of the current state (omp_solver.hpp:220-248): rhs arrays of n_total on the
 * host, interior written, ghosts 0 (as the reference's zero-initialized vectors). */
int swe_hip_rhs(swe_hip_ctx *ctx, double *rhs_h, double *rhs_hu, double *rhs_hv);

/* Launch shape of the fused stage kernel: interior rows marched per wave (>= 1; values
 * above nx are stored as nx, one row segment).  Default picked by swe_hip_create from the
 * grid size. */
int swe_hip_set_tile_rows(swe_hip_ctx *ctx, int tile_rows);
int swe_hip_get_tile_rows(const swe_hip_ctx *ctx);
/* Tile height of one stage's launches (stage 1..3; swe_hip_set_tile_rows sets all three,
 * swe_hip_get_tile_rows returns stage 1's; the right-hand side alone runs with stage 1's).
 * The default gives stages 2 and 3 shorter tiles than stage 1 on large grids.  Results do
 * not depend on it.  No counterpart in the reference (a launch-shape setting). */
int swe_hip_set_stage_tile_rows(swe_hip_ctx *ctx, int stage, int tile_rows);
int swe_hip_get_stage_tile_rows(const swe_hip_ctx *ctx, int stage);

/* Bottom that is +0.0 in every cell (ghosts included, checked at every upload): with mode 1
 * (the default) the exact-mode stage launches then take kernels that do not read B, with
 * the same results; mode 0 always reads it.  get returns 1 when those kernels are in use.
 * No counterpart in the reference (whose flat bathymetry is an array of zeros read like
 * any other). */
int swe_hip_set_flat_bottom(swe_hip_ctx *ctx, int mode);
int swe_hip_get_flat_bottom(const swe_hip_ctx *ctx);

/* Arithmetic of the stage kernels (swe_precision); default SWE_PRECISION_EXACT.  No
 * counterpart in the reference: FAST trades bitwise parity for fewer FP64 divisions. */
int swe_hip_set_precision(swe_hip_ctx *ctx, int mode);
int swe_hip_get_precision(const swe_hip_ctx *ctx);

/* Scheme family of swe_hip_step / swe_hip_rhs / swe_hip_time_steps (swe_scheme).
 * SWE_SCHEME_SERIAL* replace the modular SerialSolver<ReflectingWalls, Recon, Riemann,
 * SSPRK3, Bathymetry> (include/backend/serial/serial_solver.hpp) for the combinations its
 * SolverAssembly accepts (solver_assembly.hpp:315-500) and are bitwise equal to it: reflecting
 * walls, fixed dt, no FAST precision.  recon / riemann are ignored for SWE_SCHEME_OPTIMIZED.
 * SWE_ERR_INVALID for other combinations, a transmissive context, or while a batch/adjoint
 * needs the optimized scheme.  With the plain scheme B is not read. */
int swe_hip_set_scheme(swe_hip_ctx *ctx, int scheme, int recon, int riemann);
int swe_hip_get_scheme(const swe_hip_ctx *ctx, int *scheme, int *recon, int *riemann);

/* One launch per step instead of three: the three SSP-RK3 stages fused (temporal
 * blocking across the stages), `rows` interior rows per tile (>= 4, clamped to [4, nx]),
 * 0 = off (the default).  No counterpart in the reference, whose rk3_step
 * (omp_solver.hpp:250-338) runs the stages one after the other; results are bitwise the
 * same.  Each wave recomputes stages 1 and 2 on a 4- and 2-cell halo of its 52-column,
 * `rows`-row tile, so the HBM traffic falls from 216 to about 100 bytes per cell-update
 * while the face arithmetic grows by about a fifth.  Used by swe_hip_step and
 * swe_hip_step_adaptive with the optimized scheme (launched directly, no graph); needs
 * ng = 2 and a second state buffer (3 arrays, allocated on the first call). */
int swe_hip_set_step_fusion(swe_hip_ctx *ctx, int rows);
int swe_hip_get_step_fusion(const swe_hip_ctx *ctx);

/* Use a caller stream (hipStream_t passed as void*; NULL = the context's own stream). */
int swe_hip_set_stream(swe_hip_ctx *ctx, void *stream);

/* Replay-mode switch: 1 = capture 32 steps in a hipGraph and replay it for every 32 steps of
 * a call, launching the rest directly (default), 0 = launch kernels directly. */
int swe_hip_set_graph(swe_hip_ctx *ctx, int enabled);

/* Timing helper for benchmarks: runs nsteps fixed-dt steps with direct launches and HIP
 * events recorded on ctx's stream around every stage launch; stage_ms[0..2] receive the
 * mean duration of stages 1..3 and *total_ms the event time of the whole loop. */
int swe_hip_time_steps(swe_hip_ctx *ctx, double dt, int64_t nsteps, double *stage_ms,
                       double *total_ms);

/* Wait for all work on ctx's stream. */
int swe_hip_sync(swe_hip_ctx *ctx);

/* ---- Batched ensemble --------------------------------------------------------------------
 * No counterpart in the reference (which runs one grid per process): `count` independent
 * grids of uneven size advanced together -- every step is three launches over a tile
 * work-list covering all instances, each instance with its own dt.  All descriptors must
 * have ng = 2 and share g and bc_kind.  Arrays passed per instance use that instance's
 * reference layout ((nx+4)*(ny+4) doubles). */
typedef struct swe_hip_batch swe_hip_batch;

int swe_hip_batch_create(const swe_grid_desc *descs, int32_t count, int device,
                         swe_hip_batch **out);
void swe_hip_batch_destroy(swe_hip_batch *batch);
const char *swe_hip_batch_last_error(const swe_hip_batch *batch);
int swe_hip_batch_count(const swe_hip_batch *batch);

/* Per-instance host transfers (ghosts rebuilt as in swe_hip_upload / swe_hip_download). */
int swe_hip_batch_upload(swe_hip_batch *batch, int32_t i, const double *h, const double *hu,
                         const double *hv, const double *b);
int swe_hip_batch_download(swe_hip_batch *batch, int32_t i, double *h, double *hu, double *hv);

/* Fixed per-instance dt (count values) for swe_hip_batch_step. */
int swe_hip_batch_set_dt(swe_hip_batch *batch, const double *dt);
/* compute_stable_dt of every instance's current state into dt_out[count] (and into the
 * per-instance dt used by swe_hip_batch_step). */
int swe_hip_batch_stable_dt(swe_hip_batch *batch, double cfl, double *dt_out);
/* nsteps fixed-dt steps of every instance. */
int swe_hip_batch_step(swe_hip_batch *batch, int64_t nsteps);
/* nsteps steps with per-instance dt = compute_stable_dt(U_i^n, cfl) before each step (a
 * segmented max reduction on the device); dt_hist (nsteps x count, row-major) may be NULL. */
int swe_hip_batch_step_adaptive(swe_hip_batch *batch, double cfl, int64_t nsteps,
                                double *dt_hist);
/* Per-instance outcome of the last step call (count ints): 0, or the stage (1..3) whose
 * output was non-finite -- that instance stopped there as a single grid does (its later
 * launches did nothing; its dt_hist entries after that are quiet NaN) while the others
 * went on. */
int swe_hip_batch_nonfinite(swe_hip_batch *batch, int32_t *flags);

/* Launch shape: interior rows per tile and persistent waves per CU (0 = one tile per wave).
 * swe_hip_batch_autotune times candidates on the current state without modifying it and
 * keeps the fastest. */
int swe_hip_batch_set_launch(swe_hip_batch *batch, int tile_rows, int waves_per_cu);
int swe_hip_batch_get_launch(const swe_hip_batch *batch, int *tile_rows, int *waves_per_cu);
int swe_hip_batch_autotune(swe_hip_batch *batch, int *tile_rows, int *waves_per_cu);
int swe_hip_batch_set_precision(swe_hip_batch *batch, int mode);
int swe_hip_batch_set_graph(swe_hip_batch *batch, int enabled);
int swe_hip_batch_sync(swe_hip_batch *batch);

/* ---- discrete adjoint -------------------------------------------------------------------
 * The reference has no differentiation (its solvers only step forward); these entry points
 * are this backend's own.  They reverse the reference's SSP-RK3 step
 * (omp_solver.hpp:250-338) stage by stage with dt frozen from the forward pass, under the
 * one-sided derivative conventions documented in csrc/swe_adjoint.hip, and recompute the
 * forward from binomial checkpoints (optimal schedule for the given number of stored
 * states) held in device memory.
 *
 * Usage: create on a context whose state holds U^0 (after swe_hip_upload), forward (runs
 * nsteps, leaves U^N in the context), backward once with dL/dU^N.  A backward consumes the
 * checkpoints; call forward again before another backward.  The adjoint object uses the
 * context's stream, B and stage buffers: destroy it before the context.  ng must be 2. */
typedef struct swe_hip_adj swe_hip_adj;

typedef enum {
    SWE_DT_FIXED = 0,    /* one dt for every step */
    SWE_DT_ADAPTIVE = 1, /* dt from the CFL number at each step's start state (recorded) */
    SWE_DT_GIVEN = 2     /* a caller-given dt per step */
} swe_dt_mode;

/* nsnap = number of stored states, U^0 included (>= 1): binomial checkpoints, each reversed
 * step recomputing its two stage states (for nsteps >= 2 on a second stream, one reversal
 * ahead of the adjoint kernels, in 9 more arrays of working memory when they can be
 * allocated).  -1 stores every step together with its two stage
 * states (the backward then recomputes nothing; 9 arrays per step).  0 picks the storage
 * plan of swe_hip_adj_create_budget for 60% of the free device memory. */
int swe_hip_adj_create(swe_hip_ctx *ctx, int64_t nsteps, int32_t nsnap, swe_hip_adj **out);

/* Adjoint whose stored states take at most max_bytes of device memory.  Plan: the last K
 * steps keep their stage records (U^p, U1, U2: 9 arrays) from the forward pass, and the
 * steps before them a checkpoint every m steps (3 arrays); the backward recomputes each of
 * those steps once, segment by segment into the record buffer.  K (the most that fits) and
 * m are chosen for the budget; K = nsteps when every record fits.  Below the smallest such
 * plan: binomial checkpoints in as many states as fit.  SWE_ERR_OOM when not one state fits.
 * No counterpart in the reference. */
int swe_hip_adj_create_budget(swe_hip_ctx *ctx, int64_t nsteps, double max_bytes,
                              swe_hip_adj **out);

/* The plan swe_hip_adj_create_budget would pick for nsteps steps of a grid with n_total
 * cells per array, computed on the host (no device needed): recorded steps K (0: binomial
 * fallback), checkpoint interval m (0 for binomial) and stored states.  SWE_ERR_OOM when not
 * one state fits. */
int swe_hip_adj_plan_budget(int64_t nsteps, int64_t n_total, double max_bytes,
                            int64_t *record_steps, int64_t *ckpt_every, int64_t *stored_states);

/* The storage plan in use: recorded steps K (0 for binomial checkpoints), checkpoint
 * interval m, stored states (records + checkpoints, or binomial slots), their bytes, and
 * the forward work the backward pass recomputes, in steps.  Any pointer may be NULL. */
int swe_hip_adj_plan(const swe_hip_adj *adj, int64_t *record_steps, int64_t *ckpt_every,
                     int64_t *stored_states, double *store_bytes, double *recomputed_steps);
void swe_hip_adj_destroy(swe_hip_adj *adj);
const char *swe_hip_adj_last_error(const swe_hip_adj *adj);

/* Forward pass from the context's current state.  dt_mode FIXED: value = dt; ADAPTIVE:
 * value = CFL number; GIVEN: dt_hist holds nsteps dts (input).  For FIXED and ADAPTIVE,
 * dt_hist (may be NULL) receives the nsteps dts used.  Leaves U^N in the context. */
int swe_hip_adj_forward(swe_hip_adj *adj, int32_t dt_mode, double value, double *dt_hist);

/* Backward pass: lam_* = dL/dU^N (n_total host arrays; ghost entries ignored) ->
 * grad_* = dL/dU^0 and grad_b = dL/dB (interior entries; ghost entries 0).  grad_b may be
 * NULL. */
int swe_hip_adj_backward(swe_hip_adj *adj, const double *lam_h, const double *lam_hu,
                         const double *lam_hv, double *grad_h, double *grad_hu,
                         double *grad_hv, double *grad_b);
/* Same with device pointers (on the context's device). */
int swe_hip_adj_backward_device(swe_hip_adj *adj, const double *lam_h, const double *lam_hu,
                                const double *lam_hv, double *grad_h, double *grad_hu,
                                double *grad_hv, double *grad_b);

/* Launch shape of the adjoint stages.  block_waves 2, 4 or 8: one fused kernel per stage,
 * blocks of that many waves with at most `rows` rows each (default 4 x 4; the block's LDS,
 * (block_waves*rows + 3) * 4 KiB, must stay within 160 KiB).  The launch splits the rows
 * into blocks of equal height, as many as fill the device's resident blocks in the fewest
 * rounds.  block_waves 0: the two-kernel path (x-face kernel with `rows` faces per wave
 * writing to device memory, then the y-face kernel with `y_rows` rows per wave), whose
 * buffers (9 doubles per cell) are allocated on first selection.  0 for rows / y_rows keeps
 * the current value.  Every shape gives the same bits. */
int swe_hip_adj_set_launch(swe_hip_adj *adj, int block_waves, int rows, int y_rows);

/* The checkpoint schedule for (nsteps, nsnap), computed on the host (no device needed):
 * total forward steps advanced (the first sweep included), the most times any one step is
 * advanced, and the number of schedule actions. */
int swe_hip_adj_schedule(int64_t nsteps, int32_t nsnap, int64_t *advances, int32_t *max_reps,
                         int64_t *actions);

/* The number of stored states in use, whether stage records are stored (1/0: the records
 * plan / binomial checkpoints), and the wall-clock milliseconds of the last forward and
 * backward calls.  Any pointer may be NULL. */
int swe_hip_adj_info(const swe_hip_adj *adj, int32_t *nsnap, int32_t *records,
                     double *forward_ms, double *backward_ms);

#ifdef __cplusplus
}
#endif

#endif /* SWE_HIP_H */
