/*
 * mosscap_hip.h - C ABI of the MI355X (gfx950) finite-volume Euler hot path.
 *
 * Drop-in for the operator slots Mosscap's driver calls (Simulation.hpp:38-62):
 * compute_dt (Hydro.hpp:16, Hydro.cpp:208-242), time_step
 * (TimeStepping.cpp:39-308) and fill_bcs (Boundaries.hpp:132-149). Arrays
 * use the reference's layout: Q[var][k][j][i], x fastest, ghost cells
 * included, var = rho, m_x[, m_y[, m_z]], E (State.hpp:10-26).
 *
 * Conventions: every function returns MH_OK (0) or an MH_ERR_* code and never
 * throws; mh_last_error() gives the message. Calls are synchronous on return,
 * like the reference's fenced kernels (LoopUtils.hpp:253-255), except
 * mh_run, which returns once its steps have completed. A context is bound to
 * the HIP device current at mh_create and owns its device buffers and stream;
 * it must not be used from two threads at once.
 */
#ifndef MOSSCAP_HIP_H
#define MOSSCAP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MH_OK 0
#define MH_ERR_INVALID 1     /* bad argument; the reference throws std::runtime_error (SetupModules.hpp:146) */
#define MH_ERR_UNSUPPORTED 2 /* valid for the reference but outside this library (e.g. 3-D userfn faces, 1-D tracers) */
#define MH_ERR_HIP 3         /* HIP runtime or launch failure */

#define MH_FORM_FP64 0  /* drop-in: FP64 state and arithmetic, bitwise the reference CPU build */
#define MH_FORM_FP32C 1 /* precision form: FP32 reconstruction/Riemann, compensated hi+lo FP32 state,
                           pressure recovery and CFL speed in FP64 where kinetic energy exceeds half
                           of E (FP32 elsewhere), FP64 dt reduction */

typedef struct mh_ctx mh_ctx;

/* GridSize, State.hpp:101-106: counts include 2*ng ghosts; unused axes are 1. */
typedef struct {
    int32_t xc, yc, zc, ng;
} mh_grid_size;

/* GridLoc, State.hpp:108-112. */
typedef struct {
    double x, y, z;
} mh_grid_loc;

/* Boundaries, State.hpp:82-98; values are BoundaryType (State.hpp:62-70):
 * 0 wall, 1 periodic, 2 symmetric, 3 symmetricoutflowdiode, 4 zerograd,
 * 5 constant, 6 userfn (ghosts left untouched for the caller's hook). */
typedef struct {
    int32_t xs, xe, ys, ye, zs, ze;
    double xs_const[5], xe_const[5], ys_const[5], ye_const[5], zs_const[5], ze_const[5];
} mh_boundaries;

/* NumericalSchemes, Simulation.hpp:17-22: reconstruction 0 fog, 1 muscl,
 * 2 ppm, 3 weno5z; slope_limit 0 vanleer, 1 monotonizedcentral, 2 minmod;
 * riemann_solver 0 rusanov, 1 hll, 2 hllc; time_stepper 0 rk2, 1 ssprk3,
 * 2 ssprk4 (the 4-stage third-order SSPRK(4,3)). */
typedef struct {
    int32_t reconstruction, slope_limit, riemann_solver, time_stepper;
} mh_schemes;

/* Library options (no counterpart in the reference). Zero-initialised
 * options select FP64 and the tuned defaults. */
typedef struct {
    int32_t precision_form; /* MH_FORM_* */
    int32_t span_rows;      /* rows marched per wavefront in 2-D (the rows are split into
                             * ceil(ny / span_rows) spans whose heights differ by at most
                             * one); <= 0: the default layout, the most spans whose
                             * wavefronts fit the GPU in one round at 4 per SIMD */
    float rho_floor;        /* FP32C positivity floors on the reconstruction input (cells and
                             * faces) and on each stage's new state (a cell below rho_floor is
                             * put at rest at (rho_floor, p_floor); one below p_floor gets its
                             * energy raised to p_floor); <= 0 disables. No counterpart in the
                             * reference (whose plain method has no floors). */
    float p_floor;
} mh_options;

typedef struct {
    uint64_t floor_hits; /* FP32C: (interior cell, stage) pairs whose input or new state took a
                          * floor (each pair counted once) */
    int64_t steps;       /* steps taken by this context */
    double last_dt;      /* dt of the last step */
    double time;         /* device clock after mh_run */
    double prev_output_time;
} mh_stats;

/* Build a context (replaces setup_grid's allocations, SetupModules.hpp:9-64,
 * plus select_hydro_fns / select_timestepper, Hydro.cpp:244-323 and
 * TimeStepping.cpp:331-344). num_dim 1, 2 or 3; 3-D contexts take every boundary
 * type but userfn (the z faces use zs/ze and their constants). num_tracers
 * (simulation.n_extra_fields, 0..16, 2-D only) passive fields follow the hydro
 * variables in Q: conserved tracer densities, advected with the upwinded density
 * flux (Hydro.cpp:27-33, 155-162); constant boundaries are refused with tracers. */
int mh_create(const mh_grid_size* sz, double dx, const mh_grid_loc* loc, const mh_boundaries* bdry,
              int32_t num_dim, int32_t num_tracers, const mh_schemes* schemes, double gamma,
              const mh_options* opts, mh_ctx** out);
void mh_destroy(mh_ctx* ctx);
/* Message of the last failure on ctx, or of the last failed mh_create on this thread when ctx is NULL. */
const char* mh_last_error(const mh_ctx* ctx);

/* Number of doubles in Q: nvar * zc * yc * xc. */
int64_t mh_num_values(const mh_ctx* ctx);

/* Move the conserved state Q (ghosts included) between host and device. */
int mh_upload(mh_ctx* ctx, const double* q_host);
int mh_download(mh_ctx* ctx, double* q_host);
/* Same, from/to device memory on the context's device (no PCIe transfer). */
int mh_upload_device(mh_ctx* ctx, const double* q_dev);
int mh_download_device(mh_ctx* ctx, double* q_dev);

/* fill_bcs(sim), Boundaries.hpp:132-149; runs the user_bc hook (below) when set. */
int mh_fill_bcs(mh_ctx* ctx);

/* compute_dt(sim), Hydro.cpp:208-242: CFL reduction over all cells incl.
 * ghosts, times max_cfl, clamped to min(max_time, prev_output_time + delta). */
int mh_compute_dt(mh_ctx* ctx, double max_cfl, double time, double max_time, double prev_output_time,
                  double delta, double* dt_out);

/* sim.time_step(sim, dt), TimeStepping.cpp:39-308: every stage, ghost fill
 * after each, ending with ghosts filled. On a 2-D or 3-D grid whose faces are
 * all periodic or zerograd (no user_bc hook, no tracers, no tabulated EOS),
 * the stages read ghost cells from the cells they copy instead, and one fill
 * at the end of the step leaves the same values (the same bits in FP64). */
int mh_time_step(mh_ctx* ctx, double dt);

/* One stage of the scheme (stage 0 snapshots Q^n; later stages use it);
 * ghosts of the result are filled. mh_time_step == stages 0..n-1. */
int mh_stage(mh_ctx* ctx, int32_t stage, double dt);

/* main.cpp:54-71 step loop on the device: each step computes dt (fused into
 * the previous step's last stage), advances and does the output-time
 * bookkeeping, for nsteps steps (steps past max_time have dt = 0). *time and
 * *prev_output_time are read and updated. With use_graph, steps replay a
 * captured HIP graph. The first dt of a call comes from a CFL pass over the
 * state, except right after another mh_run, whose last stage left the
 * current state's maximum signal speed behind (every other call that changes
 * the state clears it). A caller that writes the state through
 * mh_device_state outside a user_bc hook calls mh_fill_bcs afterwards, as the
 * reference's setup does, which clears it too. */
int mh_run(mh_ctx* ctx, int64_t nsteps, double max_cfl, double* time, double max_time, double* prev_output_time,
           double delta, int32_t use_graph);

/* Per-kernel timing of a device step loop (no counterpart in the reference,
 * which only has YAKL auto-profile timers, LoopUtils.hpp:200-202): as mh_run
 * without graphs, with HIP events recorded around every launch on the
 * context's stream. Totals in milliseconds. */
typedef struct {
    double stage_ms;        /* fused stage kernels */
    int64_t stage_launches;
    double ghost_ms;        /* ghost-fill kernels */
    int64_t ghost_launches;
    double clock_ms;        /* dt finalize kernels */
    int64_t clock_launches;
    double step_ms;         /* whole loop, first event to last */
} mh_timing;

int mh_run_timed(mh_ctx* ctx, int64_t nsteps, double max_cfl, double* time, double max_time,
                 double* prev_output_time, double delta, mh_timing* out);

/* Autotuner (no counterpart in the reference): times `steps` device steps of
 * the current state for each candidate span height (rows marched per
 * wavefront in 2-D, as mh_options.span_rows; -1 the default layout;
 * 0-terminated list, NULL = built-in candidates, the default layout
 * included) with HIP events, keeps the fastest for this context and returns
 * it in *best_span (and the per-candidate milliseconds per step in ms_out
 * when non-NULL). The
 * state, the tabulated EOS's fields and the floor-hit count are restored
 * afterwards; the record of the last stage's input is cleared (the trials
 * overwrote that buffer), so mh_eos_fields(which = 1) then reads the current
 * state. A user_bc hook runs during the trials like in any step. */
int mh_autotune(mh_ctx* ctx, int32_t steps, const int32_t* candidates, int32_t* best_span, double* ms_out);

int mh_get_stats(mh_ctx* ctx, mh_stats* out);

/* HIP stream the context launches on (a hipStream_t), for event timing. */
void* mh_stream(mh_ctx* ctx);

/* ---- user boundary hook: Simulation::user_bc (Simulation.hpp:55), called at
 * the end of fill_bcs (Boundaries.hpp:146-148) ----------------------------------
 * The hook runs on the calling host thread after every ghost fill of the
 * state: mh_fill_bcs, each stage of mh_time_step / mh_stage / mh_run. Inside
 * it, mh_device_state gives the buffer just filled; the hook may enqueue
 * kernels on mh_stream(ctx) or copy synchronously. With the compensated form
 * it must write both planes (hi = (float)v, lo = (float)(v - hi)). A non-zero
 * return fails the calling function with MH_ERR_INVALID. While a hook is set,
 * the CFL reduction runs as a separate pass after it (so the hook's ghost
 * values enter dt, as in Hydro.cpp:214-228), mh_run launches step by step
 * instead of replaying a graph, and the adjoint is refused. fn = NULL removes
 * the hook. */
/* ---- gravity source term: setup_gravity (Gravity.cpp:43-70), registered in
 * sim.compute_source_terms ----------------------------------------------------
 * With enable != 0, every stage adds S*dt with S = (0, rho*g, m.g) of the
 * stage's own input state (gravity_kernel, Gravity.cpp:14-40; the stage
 * lambdas' `source`, TimeStepping.cpp). Components beyond num_dim must be 0.
 * enable = 0 removes it. The adjoint differentiates it. */
int mh_set_gravity(mh_ctx* ctx, int32_t enable, double gx, double gy, double gz);

/* ---- analytic LTE hydrogen EOS fields: eos.type = analyticlteh, the update_eos
 * pass of AnalyticLteH.hpp:51-125 (Eos.cpp:23-27, 42-63) ---------------------
 * Writes, for every cell (ghosts included, xc*yc*zc values each), the ionisation
 * fraction y and temperature T (bisection on T, Saha equation for y) of the
 * state the reference's update_eos last saw: which = 1 takes the input of the
 * most recent stage (update_eos runs at the start of each stage's fluxes,
 * Hydro.cpp:325-328), which = 0 the current state. The ghosts of that input
 * hold the values the stage read (a step that read ghosts from their source
 * cells fills them first). The hydro step does not use them. An inverted
 * temperature bracket (the reference aborts) returns MH_ERR_INVALID. */
int mh_eos_fields(mh_ctx* ctx, int32_t which, double avg_mass, int32_t ionisation_e, double* y, double* T);

typedef int (*mh_user_bc_fn)(mh_ctx* ctx, void* user);
int mh_set_user_bc(mh_ctx* ctx, mh_user_bc_fn fn, void* user);

/* ---- tabulated LTE hydrogen EOS: eos.type = tabulatedlteh (TabulatedLteH.hpp,
 * Eos.cpp:66-88; the table file's variables, TabulatedLteH.hpp:30-55) ------- */
typedef struct {
    double log_eint_rho_start, log_eint_rho_step; /* log10(eint / rho) axis, eint in J m^-3 */
    int32_t n_eint_rho;
    double log_rho_start, log_rho_step;           /* log10(rho) axis */
    int32_t n_rho;
    const double* y_table;                        /* [n_rho][n_eint_rho] ionisation fraction */
    const double* log_T_table;                    /* [n_rho][n_eint_rho] log10 temperature */
} mh_lte_table;
/* Installs the table (copied to the device) and sets the y and T fields to -1.
 * From then on every stage starts with update_eos on its input
 * (TabulatedLteH.hpp:120-161, called from compute_hydro_fluxes, Hydro.cpp:325-328):
 * per cell, ghosts included, clamped bilinear interpolation of y and log10 T at
 * (log10 eint/rho, log10 rho), and E += rho chi_H / m_H (y - y_prev) once y_prev
 * is set. Stage 0's Q^n is the state before that update, as the reference copies
 * Q_old first. A negative internal energy (the reference aborts) makes the step
 * return MH_ERR_INVALID "eint negative". table = NULL removes the EOS. */
int mh_set_tabulated_eos(mh_ctx* ctx, const mh_lte_table* table);
/* update_eos on the current state, in place: what setup_sim does once before
 * its first ghost fill (SetupModules.hpp:250-255). */
int mh_tabulated_eos_update(mh_ctx* ctx);
/* The y and T fields of the last update_eos, xc*yc*zc values each. */
int mh_tabulated_eos_fields(mh_ctx* ctx, double* y, double* T);

typedef struct {
    int32_t precision_form; /* MH_FORM_* */
    double* q;              /* FP64 state, or NULL */
    float* q_hi;            /* compensated state planes, or NULL */
    float* q_lo;
    int64_t num_values;     /* values per plane: nvar * zc * yc * xc */
} mh_device_view;
/* Device pointers of the state buffer the hook must act on (inside a hook) or
 * of the current state (elsewhere); laid out like Q. */
int mh_device_state(mh_ctx* ctx, mh_device_view* out);

/* ---- discrete adjoint (no counterpart in the reference) ----------------------
 * FP64 2-D contexts; every boundary type but symmetricoutflowdiode. Time step
 * sizes are frozen: the backward pass differentiates the map
 * Q^0 -> Q^N of the recorded steps with their dt held fixed.
 *
 * mh_adjoint_forward runs nsteps steps from the current state, keeping a
 * checkpoint every `checkpoint_every` steps. With dts_in, those step sizes are
 * used; otherwise each dt is the CFL dt of its state (Hydro.cpp:208-229,
 * no output-time clamp). The sizes used are written to dts_out (nsteps
 * values, may be NULL).
 *
 * mh_adjoint_backward takes dJ/dQ^N (lambda_final, host array laid out like
 * Q, ghost entries ignored) and returns dJ/dQ^0 (lambda_initial, host, ghost
 * entries zero), recomputing each checkpoint segment's stage states and
 * running the reverse stages (hand-written VJPs of cons->prim, WENO5-Z and
 * Rusanov; forward-mode dual numbers for the other functors). Its scratch
 * (4 + checkpoint_every x stages states and a pinned host copy of one) is
 * allocated by the first call and kept by the context until mh_destroy. */
int mh_adjoint_forward(mh_ctx* ctx, int64_t nsteps, int32_t checkpoint_every, double max_cfl, const double* dts_in,
                       double* dts_out);
int mh_adjoint_backward(mh_ctx* ctx, const double* lambda_final, double* lambda_initial);
/* Tangent-linear run of the same recorded steps (forward mode, no
 * counterpart in the reference): from the first checkpoint and a direction
 * dq0 (host array laid out like Q, ghost entries ignored) it evaluates every
 * stage on dual numbers, with the drop-in's expressions and the recorded dt,
 * and returns dQ^N/dQ^0 . dq0 in dq_final (ghost-filled); q_final (may be
 * NULL) receives the value part, Q^N, bit for bit the forward result. The
 * exact directional derivative that mh_adjoint_backward's gradient must
 * match: <lambda_initial, dq0> = <lambda_final, dq_final>. Constant boundaries
 * are refused. Unoptimised (one thread per cell reloads its stencil). */
int mh_adjoint_tangent(mh_ctx* ctx, const double* dq0, double* dq_final, double* q_final);
/* Device time (HIP events on the context's stream) of the last
 * mh_adjoint_backward: segment recomputation plus reverse stages, the two
 * host transfers excluded. */
int mh_adjoint_backward_ms(const mh_ctx* ctx, double* ms);
/* ---- one step without device allocations (no counterpart in the reference) --
 * For callers that cannot allocate device memory: tight serving loops,
 * harnesses that forbid hipMalloc, code that runs beside a graph-captured
 * pipeline (the call itself synchronizes its stream before returning, so it
 * cannot be captured into a graph). The step of
 * mh_compute_dt + mh_time_step (FP64 drop-in arithmetic, bit for bit the same
 * result) on a caller-owned state, with no context: q_in is Q^n on the device
 * ([var][k][j][i] FP64, ghosts filled), q_out receives Q^{n+1} (a distinct
 * device buffer of the same size; ghosts filled). dt > 0 takes that step;
 * dt <= 0 takes the CFL step max_cfl * dx / max(c_s + |v|) of q_in, ghosts
 * included (Hydro.cpp:208-229, no output-time clamp). *dt_used (may be NULL)
 * receives it. 1-D and 2-D grids, no tracers, gravity or userfn faces.
 * The step clock and the intermediate stage states live in a fixed device
 * array of MH_NOALLOC_SCRATCH_BYTES per device, held until the process ends:
 * grids needing more than 2 states of (bytes - 16.6 KB) / 16 values each
 * (about 1 M cells of 4 variables) are refused with MH_ERR_INVALID. Calls
 * that overlap on one device (other threads, other streams) race on that
 * array. Runs on `stream` (a hipStream_t; NULL = the null stream) and is
 * synchronous on return. Errors: mh_last_error(NULL). */
#define MH_NOALLOC_SCRATCH_BYTES (64u << 20)
int mh_step_noalloc(const mh_grid_size* sz, double dx, const mh_boundaries* bdry, int32_t num_dim,
                    const mh_schemes* schemes, double gamma, double max_cfl, double dt, const double* q_in,
                    double* q_out, double* dt_used, void* stream);

/* probe of the hand-written VJPs against dual-number Jacobians (what: 0/1 WENO5-Z
 * hand/dual, in[n][5], cot[n][2] -> out[n][5]; 2/3 Rusanov x hand/dual, 4/5 Rusanov y,
 * 6/7 HLLC x, 8/9 HLLC y: in[n][8] (wL, wR), cot[n][4] -> out[n][8]). */
int mh_probe_vjp(int32_t what, int64_t n, const double* in, const double* cot, double* out, double gamma);

/* ---- batched ensemble of independent 1-D problems -------------------------
 * No counterpart in the reference (one Simulation per process there). Each
 * problem is a 1-D grid of lengths[p] interior cells on [0, x_dim] with ng
 * ghosts, sharing schemes, gamma and the x boundaries of bdry; problem p's
 * state is the block Q_p[3][lengths[p] + 2 ng] at offset offsets[p] of one
 * packed array. Each problem keeps its own dt (CFL over its own cells,
 * Hydro.cpp:208-242) and time. */
typedef struct mh_batch mh_batch;

int mh_batch_create(int32_t nprob, const int32_t* lengths, int32_t ng, double x_dim, const mh_boundaries* bdry,
                    const mh_schemes* schemes, double gamma, const mh_options* opts, mh_batch** out);
void mh_batch_destroy(mh_batch* b);
const char* mh_batch_last_error(const mh_batch* b);
int64_t mh_batch_num_values(const mh_batch* b);
/* Element offset of every problem's block (nprob values). */
int mh_batch_offsets(const mh_batch* b, int64_t* offsets);
/* Move the packed state (ghosts included); upload resets every problem's time to 0. */
int mh_batch_upload(mh_batch* b, const double* q_host);
int mh_batch_download(mh_batch* b, double* q_host);
int mh_batch_fill_bcs(mh_batch* b);
/* nsteps steps of every problem. mode 0: one persistent launch per stage
 * covering all problems (work list of 62-cell segments); mode 1: one launch
 * per problem and stage (the comparison baseline). With use_graph, each step
 * replays a captured HIP graph. Steps past max_time have dt = 0. */
int mh_batch_run(mh_batch* b, int64_t nsteps, double max_cfl, double max_time, int32_t mode, int32_t use_graph);
/* Per-problem clock and last dt (nprob values each; either may be NULL). */
int mh_batch_times(mh_batch* b, double* times, double* last_dt);
/* Persistent grid size in wavefronts (0 = built-in default), and an
 * autotuner that times candidate sizes (0-terminated list of waves per CU,
 * NULL = built-in) over `steps` steps, restores the state and keeps the best. */
int mh_batch_set_waves(mh_batch* b, int32_t waves);
int mh_batch_autotune(mh_batch* b, int32_t steps, const int32_t* waves_per_cu, int32_t* best_waves, double* ms_out);
void* mh_batch_stream(mh_batch* b);
int mh_batch_stats(mh_batch* b, mh_stats* out);

/* ---- functor probes (run the device functors on n inputs; host arrays) ---- */
/* reconstruct<recon, limiter>: stencils[n][5] (s[-2..2]) -> wl[n], wr[n]; prec 0 FP64, 1 FP32. */
int mh_probe_reconstruct(int32_t recon, int32_t limiter, int32_t prec, int64_t n, const double* stencils,
                         double* wl, double* wr);
/* riemann_flux<solver, axis, nd>: wl, wr, flux are [n][2+nd] primitive / flux vectors. */
int mh_probe_riemann(int32_t solver, int32_t axis, int32_t nd, int32_t prec, double gamma, int64_t n,
                     const double* wl, const double* wr, double* flux);
/* op 0: a/b, 1: sqrt(a) in FP64 as the drop-in computes them; 2, 3: the FP32 forms. */
int mh_probe_math(int32_t op, int64_t n, const double* a, const double* b, double* out);

#ifdef __cplusplus
}
#endif

#endif /* MOSSCAP_HIP_H */
