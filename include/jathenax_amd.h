/*
 * jathenax_amd: MI355X (gfx950) per-RK-stage update of JathenAX's periodic isothermal
 * hydro solver, behind a C ABI.
 *
 * Conventions
 *  - Arrays are C-contiguous device pointers in the reference layout: the state is
 *    (ny, nx, nvar) with components (rho, rho*vx, rho*vy, rho*s...), forces (ny, nx, 2).
 *  - dtype 0 = float64, 1 = float32 (FP32 storage and arithmetic; time, box sums and the
 *    kick normalisation stay float64).
 *  - Every call is ordered on the context's HIP stream and never frees caller memory.
 *    Calls that return host scalars synchronise the stream.  A context is not
 *    thread-safe.
 *  - Return value: JX_OK, or an error code whose message is jx_last_error(ctx).
 *    JX_EINVAL maps to the reference's ValueError, JX_EKEY to its KeyError (unknown
 *    integrator / reconstruction / Riemann solver), JX_EHIP to a runtime failure.
 *  - Reference citations are to the JathenAX sources (the jathenax Python package).
 */
#ifndef JATHENAX_AMD_H
#define JATHENAX_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define JX_OK 0
#define JX_EINVAL 1
#define JX_EKEY 2
#define JX_EHIP 3

/* Reconstruction ids, in the order of RECONSTRUCTIONS (jathenax/fluxes.py:120-123). */
#define JX_RECON_DC 0
#define JX_RECON_PLM 1
#define JX_RECON_PPM4 2
#define JX_RECON_PPMX 3
#define JX_RECON_WENOZ 4
/* Riemann solver ids, in the order of RIEMANN_SOLVERS (jathenax/fluxes.py:204). */
#define JX_RS_LLF 0
#define JX_RS_HLLE 1
#define JX_RS_ROE 2

typedef struct jx_ctx jx_ctx;

/* Run parameters: the static and scalar arguments of isothermal_step
 * (jathenax/hydro.py:92-95) and evolve_isothermal (jathenax/run.py:24-27). */
typedef struct jx_run_params {
  int ny, nx, nvar;          /* state shape; nvar in {3, 4, 5} */
  int dtype;                 /* 0 float64, 1 float32 */
  int reconstruct, rsolver;  /* JX_RECON_*, JX_RS_* */
  int integrator;            /* number of RK stages: with rk_rows NULL, 2 = rk2, 3 = rk3,
                                4 = rk4 of the built-in table; else 1..JX_MAX_RK_STAGES */
  int nlim;                  /* cycle limit (evolve only) */
  double dx, dy, cs, dfloor, cfl, tlim;
  double viscosity, smagorinsky_cs, drag_rate;
  double dt_update, tdriv_start, dedt;   /* forcing schedule (evolve only) */
  /* The integrator's rows of RK_STAGES (jathenax/hydro.py:20-36), read on every call as
   * hydro.py:100 reads the table: `integrator` rows of (gam0, gam1, beta, delta), delta NaN
   * where the reference row holds None.  NULL selects the built-in rk2/rk3/rk4 rows.  The
   * batched engine accepts only tables equal to the built-in rows. */
  const double* rk_rows;
} jx_run_params;

#define JX_MAX_RK_STAGES 8

/* Forcing tables (jathenax/run.py:148-155): the OU coefficient table and the phase
 * factors of forcing_basis, all device arrays in the state dtype. */
typedef struct jx_forcing {
  const void* coefficients;  /* (n_updates, n_modes, 2, 2) */
  const void* cos_x;         /* (n_modes, nx) */
  const void* sin_x;
  const void* cos_y;         /* (n_modes, ny) */
  const void* sin_y;
  int n_updates, n_modes;
} jx_forcing;

/* Context: one device, one stream (0 = the device's null stream), owned scratch. */
int jx_ctx_create(jx_ctx** out, int device, void* stream);
void jx_ctx_destroy(jx_ctx* ctx);
const char* jx_last_error(const jx_ctx* ctx);
/* Tile shape of the fused stage kernel: block_x threads (64..256, multiple of 64) per
 * strip of block_x-6 columns, `rows` rows per block; 0 selects the default. */
int jx_set_tiles(jx_ctx* ctx, int block_x, int rows);
/* The tile shape the next step or run with these parameters will use (the context's
 * setting, or the occupancy-based default).  No reference counterpart: tile
 * introspection for the autotuner. */
int jx_get_tiles(jx_ctx* ctx, const jx_run_params* p, int* block_x, int* rows);
/* Measurement hooks (no reference counterpart): when on, every fused stage launch is
 * bracketed by HIP events on the context stream; jx_kernel_timing synchronises and
 * returns the number of timed launches and their summed duration, then resets. */
int jx_set_kernel_timing(jx_ctx* ctx, int on);
int jx_kernel_timing(jx_ctx* ctx, int* launches, double* total_ms);
/* FP32 forced runs with Smagorinsky evaluate the raw force per cell where the kick reads
 * it (from per-row sums of the basis's mode groups, at most 4 groups) instead of rendering
 * the (ny, nx, 2) array each step; the values are the rendered ones bit for bit.  On by
 * default; off renders the array.  No reference counterpart: a performance switch, whose
 * two settings give the same results. */
int jx_set_fused_forcing(jx_ctx* ctx, int on);

/* One step, all RK stages (replaces isothermal_step, jathenax/hydro.py:92-146).
 * U_in is not modified; the floored result goes to U_out (distinct from U_in).
 * raw_force may be NULL; otherwise the final stage is kicked with power `dedt`. */
int jx_isothermal_step(jx_ctx* ctx, const jx_run_params* p, const void* U_in, void* U_out,
                       double dt, const void* raw_force, double dedt);

/* jx_isothermal_step without allocation.  No reference counterpart: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The stage registers and partial sums live in a static
 * scratch array of 96 MiB per GPU instead of context buffers: states up to 47.5 MiB
 * (rk2/rk3) or 31.6 MiB (tables with a delta, e.g. rk4), kicked steps up to 18724 stage
 * blocks; larger inputs fail with JX_EINVAL before any work.  The caller passes U_in and
 * U_out as for jx_isothermal_step, and the result equals its result bit for bit.  All
 * calls on one GPU share that scratch: calls that overlap on one GPU (several streams or
 * threads) race and must be serialised by the caller. */
int jx_isothermal_step_noalloc(jx_ctx* ctx, const jx_run_params* p, const void* U_in,
                               void* U_out, double dt, const void* raw_force, double dedt);

/* Timestep limit from primitives W (replaces isothermal_timestep, hydro.py:63-89). */
int jx_isothermal_timestep(jx_ctx* ctx, const jx_run_params* p, const void* W,
                           double previous_dt, double time, double tlim, double* dt_out);

/* Device-resident time loop (replaces evolve_isothermal, jathenax/run.py:24-69).
 * forcing may be NULL.  Returns (time, last dt, cycle) like the reference. */
int jx_evolve_isothermal(jx_ctx* ctx, const jx_run_params* p, const void* U_in, void* U_out,
                         const jx_forcing* forcing, double* time_out, double* dt_out,
                         int* cycle_out);

/* recover_isothermal (hydro.py:54-60): floored U and primitives W; either may be NULL. */
int jx_recover_isothermal(jx_ctx* ctx, int dtype, long long ncells, int nvar, const void* U,
                          double dfloor, void* U_out, void* W_out);

/* rk_stage (hydro.py:39-51) on n values. U_saved_out may be NULL when has_delta == 0. */
int jx_rk_stage(jx_ctx* ctx, int dtype, long long n, const void* U, const void* U_saved,
                const void* div_flux, double dt, double gam0, double gam1, double beta,
                double delta, int has_delta, void* U_next, void* U_saved_out);

/* flux_divergence (fluxes.py:283-287) of (ny, nx, nvar) face fluxes. */
int jx_flux_divergence(jx_ctx* ctx, int dtype, int ny, int nx, int nvar, const void* flux_x,
                       const void* flux_y, double dx, double dy, void* out);

/* stress_face_fluxes (fluxes.py:269-280): (ny, nx, 3) stress to two (ny, nx, 2) arrays. */
int jx_stress_face_fluxes(jx_ctx* ctx, int dtype, int ny, int nx, const void* tau,
                          void* flux_x, void* flux_y);

/* render_force (forcing.py:111-122) of one OU row `amplitudes` (n_modes, 2, 2). */
int jx_render_force(jx_ctx* ctx, int dtype, int ny, int nx, int n_modes,
                    const void* amplitudes, const void* cos_x, const void* sin_x,
                    const void* cos_y, const void* sin_y, void* raw_out);

/* normalize_force (forcing.py:137-149): acceleration (ny, nx, 2). */
int jx_normalize_force(jx_ctx* ctx, int dtype, int ny, int nx, int nvar, const void* U,
                       const void* raw_force, double dt, double dedt, void* acc_out);

/* ---- Unfused stage operators --------------------------------------------------------
 * The traced stage path (isothermal_step(trace=True), hydro.py:134-142) and the
 * learned-stress path (hydro.py:111-119) run the stage as separate launches that keep
 * every intermediate array; each entry point restates one reference array function with
 * the fused kernel's scalar operators.  Arrays are (ny, nx, nvar) unless stated. */
/* reconstruct_faces (fluxes.py:124-131): left/right states at each cell's low face.
 * dfloor NaN skips the PPMX/WENO-Z face-density floor, which gives the registry functions
 * dc/plm/ppm4/ppmx/wenoz_face_states (fluxes.py:6-117). */
int jx_reconstruct_faces(jx_ctx* ctx, int dtype, int ny, int nx, int nvar, const void* W,
                         int axis, int reconstruct, double dfloor, void* left, void* right);
/* isothermal_flux (fluxes.py:207-219). */
int jx_isothermal_flux(jx_ctx* ctx, int dtype, int ny, int nx, int nvar, const void* left,
                       const void* right, double cs, int axis, int rsolver, void* flux);
/* viscous_momentum_flux (fluxes.py:248-266): out is (ny, nx, 2). */
int jx_viscous_momentum_flux(jx_ctx* ctx, int dtype, int ny, int nx, int nvar, const void* W,
                             double dx, double dy, double viscosity, double smagorinsky_cs,
                             int axis, void* out);
/* face_velocity_gradients (fluxes.py:227-238): dv_dx and dv_dy, each (ny, nx, 2), at each
 * cell's low face along axis (1 = x faces, 0 = y faces). */
int jx_face_velocity_gradients(jx_ctx* ctx, int dtype, int ny, int nx, int nvar, const void* W,
                               double dx, double dy, int axis, void* dv_dx, void* dv_dy);
/* smagorinsky_viscosity (fluxes.py:241-245) of n faces: nu (n) from dv_dx, dv_dy (n, 2). */
int jx_smagorinsky_viscosity(jx_ctx* ctx, int dtype, long long n, const void* dv_dx,
                             const void* dv_dy, double dx, double dy, double smagorinsky_cs,
                             void* nu);
/* flux[..., 1:3] += add, add is (ncells, 2) (hydro.py:106-110, 117-119). */
int jx_add_momentum_flux(jx_ctx* ctx, int dtype, long long ncells, int nvar, void* flux,
                         const void* add);
/* Drag in place: U[..., 1:3] -= (beta*dt*drag_rate*rho) * v with the stage's incoming
 * primitives W (hydro.py:122-125). */
int jx_apply_drag(jx_ctx* ctx, int dtype, long long ncells, int nvar, void* U, const void* W,
                  double beta, double dt, double drag_rate);
/* Kick in place: U[..., 1:3] += rho * (acceleration * dt) (hydro.py:131-132). */
int jx_apply_kick(jx_ctx* ctx, int dtype, long long ncells, int nvar, void* U,
                  const void* acceleration, double dt);

/* ---- Forced-run resume (no reference counterpart) -------------------------------------
 * The loop state of a run between steps.  Resuming evolve_isothermal from the state (and
 * the state array) another run ended with continues bit for bit as if uninterrupted: the
 * raw force is re-rendered from OU row `update` and the first dt is derived from the state
 * and prev_dt exactly as the uninterrupted loop derives it.  For serving pre-emption and
 * checkpoint/restart of long forced runs. */
typedef struct jx_loop_state {
  double time;     /* simulation time (float64) */
  double prev_dt;  /* dt of the last completed step; +inf before the first */
  int cycle;       /* completed steps */
  int update;      /* OU row of the current raw force; -1 = none */
} jx_loop_state;
/* evolve_isothermal from `start` (NULL = a fresh run: time 0, no step, no OU row); the
 * loop state at the end is written to `end` (may be NULL).  tlim and nlim are the run's
 * totals, as in the uninterrupted run. */
int jx_evolve_isothermal_resume(jx_ctx* ctx, const jx_run_params* p, const void* U_in,
                                void* U_out, const jx_forcing* f, const jx_loop_state* start,
                                jx_loop_state* end);

/* ---- Learned subgrid stress in the fused stage (hydro.py:111-119) -------------------
 * Before each RK stage the library writes the stage input's primitives (recover_isothermal,
 * hydro.py:54-60) to W (ny, nx, nvar) and calls fn(user, stage, W, tau); the callback must
 * fill tau (ny, nx, 3) = (tau_xx, tau_xy, tau_yy) with work ordered on the context stream
 * and return 0 (nonzero aborts with JX_EINVAL).  The fused stage averages tau onto the
 * faces and adds it to the momentum fluxes after the viscous flux.  W and tau are caller-
 * owned device buffers of the state dtype. */
typedef int (*jx_stress_fn)(void* user, int stage, const void* W, void* tau);
/* isothermal_step(predict_stress=...) (hydro.py:92-146). */
int jx_isothermal_step_sgs(jx_ctx* ctx, const jx_run_params* p, const void* U_in, void* U_out,
                           double dt, const void* raw_force, double dedt, void* W, void* tau,
                           jx_stress_fn fn, void* user);
/* evolve_isothermal(predict_stress=...) (run.py:20-69); the device loop, one step per
 * host round trip so that the callback runs between stages. */
int jx_evolve_isothermal_sgs(jx_ctx* ctx, const jx_run_params* p, const void* U_in,
                             void* U_out, const jx_forcing* f, void* W, void* tau,
                             jx_stress_fn fn, void* user, double* time_out, double* dt_out,
                             int* cycle_out);

/* Float64 box sums of each component (run.py:186-187 before the dx*dy factor). */
int jx_conserved_sums(jx_ctx* ctx, int dtype, int ny, int nx, int nvar, const void* U,
                      double* sums_out);

/* Seed path on the host, float64 (forcing.py:11-100). */
int jx_ran_st(long long seed, long long count, double* out);            /* forcing.py:11 */
int jx_ran_gaussian_st(long long seed, long long count, double* out);   /* forcing.py:35 */
/* cartesian_modes (forcing.py:48-60): writes up to `cap` (kx, ky) pairs, returns count
 * (or -JX_EINVAL). */
int jx_cartesian_modes(int nlow, int nhigh, int kx_min, int kx_max, int ky_min, int ky_max,
                       int* modes_out, int cap);
/* modal_coefficients (forcing.py:63-100): out is (updates, n_modes, 2, 2) float64. */
int jx_modal_coefficients(const int* modes, int n_modes, double lx, double ly, int nlow,
                          int nhigh, double npeak, double sol_fraction, long long rseed,
                          double tcorr, double dt_update, int updates, double* out);

/* ---- Batched ensemble engine -------------------------------------------------------
 * No counterpart in the reference (which runs one grid per call): many independent
 * instances of evolve_isothermal (jathenax/run.py:24-69), each with its own shape,
 * operators, integrator, physics and forcing, advanced together by persistent kernels.
 * One round = one RK stage of every active instance; rounds are replayed from a
 * hipGraph.  Instances live in one HBM pool of pool_bytes; nvar is 3 or 4.  The batch
 * owns its own HIP stream: jx_batch_add copies U0 (a device array ready on the caller's
 * stream) into the pool, so the caller may free U0 afterwards.  Results are bitwise
 * independent of which other instances share the batch. */
typedef struct jx_batch jx_batch;
int jx_batch_create(int device, int dtype, int max_instances, long long pool_bytes,
                    jx_batch** out);
void jx_batch_destroy(jx_batch* b);
const char* jx_batch_last_error(const jx_batch* b);
/* HBM bytes an instance of these parameters takes from the pool (admission control). */
long long jx_batch_footprint(const jx_batch* b, const jx_run_params* p, int forced);
long long jx_batch_pool_free(const jx_batch* b);
/* Register an instance (tlim, nlim, cfl and the forcing schedule come from p); `rows`
 * caps its tile height (rounded down to a multiple of 8, at least 8; 0 = 32): each round
 * the engine lowers the height of every tile when the active work is too small to fill
 * the GPU.  forcing may be NULL. */
int jx_batch_add(jx_batch* b, const jx_run_params* p, const void* U0, const jx_forcing* f,
                 int rows, int* id_out);
/* Register an instance that resumes from a loop state (see jx_loop_state): it continues bit
 * for bit as the run it was checkpointed from. */
int jx_batch_add_resume(jx_batch* b, const jx_run_params* p, const void* U0,
                        const jx_forcing* f, int rows, const jx_loop_state* start, int* id_out);
/* Copy an instance's state to U_out and return its loop state, between rounds: a mid-step
 * instance is reported at the start of its step, a finished step as committed.  Pre-empting
 * = checkpoint, remove, and later jx_batch_add_resume. */
int jx_batch_checkpoint(jx_batch* b, int id, void* U_out, jx_loop_state* state_out);
int jx_batch_remove(jx_batch* b, int id);
/* Run at least `rounds` rounds; returns the number of still-active instances. */
int jx_batch_run(jx_batch* b, int rounds, int* active_out);
/* jx_batch_run in two halves: enqueue the rounds on the batch's stream and return at once,
 * then wait for them.  Other streams (for example a single-grid run) can work on the GPU in
 * between.  No counterpart in the reference (batching is new). */
int jx_batch_launch(jx_batch* b, int rounds);
int jx_batch_wait(jx_batch* b, int* active_out);
/* Per slot: done (1), running (0) or empty (-1), time and cycle. */
int jx_batch_poll(jx_batch* b, int* done, double* time, int* cycle, int cap);
/* The same, and the active count, as of the end of the oldest jx_batch_launch not yet
 * polled (every launch snapshots the loop state; two are kept, a third drops the oldest;
 * jx_batch_wait, jx_batch_run and jx_batch_poll clear them).  Waits for that launch only,
 * so a caller that keeps the next launch queued never leaves the GPU idle while it reads.
 * A slot whose instance was added after that launch is reported running (0), at time NaN
 * and cycle -1.  No counterpart in the reference (batching is new). */
int jx_batch_poll_launched(jx_batch* b, int* done, double* time, int* cycle, int cap,
                           int* active_out);
/* Copy an instance's state to U_out (may be NULL) and return (time, last dt, cycle). */
int jx_batch_result(jx_batch* b, int id, void* U_out, double* time, double* dt, int* cycle);
int jx_batch_state(jx_batch* b, int id, void** state_out);
int jx_batch_set_graph(jx_batch* b, int use_graph, int rounds_per_graph);
/* Tiles per resident wave that each round's tile heights aim for (default 2); 0 keeps
 * every instance at its own cap.  Results do not depend on it. */
int jx_batch_set_fill(jx_batch* b, int fill);

#ifdef __cplusplus
}
#endif
#endif /* JATHENAX_AMD_H */
