/*
 * swm.h -- C ABI of the MI355X (gfx950) time-stepping library for ShallowWaters.jl.
 *
 * The library replaces the hot loop of ShallowWaters.jl,
 *     time_integration(Prog, Diag, S)            src/time_integration.jl:2-250
 * with hand-written HIP kernels.  Everything before it (Parameter, Grid, Constants,
 * Forcing, initial_conditions: src/run_model.jl:25-39) stays on the host and hands the
 * library resolved enums, pre-rounded constants and the static fields.
 *
 * Conventions
 *  - Plain C types only.  Every entry point returns an int status: SWM_OK (0) or a
 *    negative SWM_E_* code; swm_last_error(ctx) gives the message of the last failure.
 *  - Host arrays are Julia Array{T,2} layout: column-major, no pitch, halo included,
 *    sizes exactly as src/preallocate.jl allocates them (listed per entry point).  Host
 *    buffers are borrowed for the duration of the call only.
 *  - Floating-point host arrays are stored in the format named by the config: SWM_F64
 *    (double), SWM_F32 (float), SWM_F16 (IEEE binary16) or SWM_BF16 (bfloat16 bits).
 *  - A context owns its device buffers and one HIP stream; it is not re-entrant.
 *    Contexts are independent of each other.
 *  - No exceptions or callbacks cross the ABI.
 */
#ifndef SWM_H
#define SWM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWM_ABI_VERSION 2

/* status codes */
#define SWM_OK 0
#define SWM_E_ARG (-1)       /* invalid argument / shape (Julia: ArgumentError, BoundsError) */
#define SWM_E_UNSUPPORTED (-2) /* option not implemented by the GPU path */
#define SWM_E_HIP (-3)       /* HIP runtime failure */
#define SWM_E_STATE (-4)     /* call out of order (e.g. step before set_static) */

/* number formats: T (right-hand side), Tprog (prognostic), Tcomm (ghost copies)
 * src/default_parameters.jl:3-6 */
enum { SWM_F64 = 0, SWM_F32 = 1, SWM_F16 = 2, SWM_BF16 = 3 };
/* P.bc, src/default_parameters.jl:60 */
enum { SWM_PERIODIC = 0, SWM_NONPERIODIC = 1 };
/* P.time_scheme, src/default_parameters.jl:47-48 */
enum { SWM_RK = 0, SWM_SSPRK2 = 1, SWM_SSPRK3 = 2, SWM_4SSPRK3 = 3 };
/* P.bottom_drag, src/default_parameters.jl:71 */
enum { SWM_DRAG_QUADRATIC = 0, SWM_DRAG_LINEAR = 1, SWM_DRAG_NONE = 2 };
/* P.diffusion, src/default_parameters.jl:76 */
enum { SWM_DIFF_CONSTANT = 0, SWM_DIFF_SMAGORINSKY = 1 };
/* P.adv_scheme / P.dynamics, src/default_parameters.jl:66-67 */
enum { SWM_ADV_ARAKAWAHSU = 0, SWM_ADV_SADOURNY = 1 };
enum { SWM_NONLINEAR = 0, SWM_LINEAR = 1 };
/* continuity variant, src/continuity.jl:91-97 */
enum { SWM_CONT_ITSELF = 0, SWM_CONT_RELAX = 1, SWM_CONT_FORCING = 2 };

/* Resolved run configuration.  Values marked (Tprog) / (T) must already be rounded to
 * that format (the host computes them in Float64 exactly as src/grid.jl and
 * src/constants.jl do, then rounds); (F64) values stay Float64. */
typedef struct swm_config {
    int32_t abi_version;          /* SWM_ABI_VERSION */
    int32_t nx, ny;               /* grid.jl:4,22 */
    int32_t bc;                   /* SWM_PERIODIC (ep=1) or SWM_NONPERIODIC (ep=0) */
    int32_t T, Tprog, Tcomm;      /* SWM_F64 ... */
    int32_t time_scheme, RKo, RKs, RKn;
    int32_t nstep_diff, nstep_advcor;
    int32_t adv_scheme, dynamics, bottom_drag, diffusion, continuity;
    int32_t tracer_advection, tracer_consumption;
    int32_t seasonal_wind_x, seasonal_wind_y;
    /* u, v are carried inside the library as 2^scale_log2 · u, v (0: unscaled); the arrays
     * at this ABI stay unscaled.  A low-precision option that is not in the reference. */
    int32_t scale_log2;
    int64_t dtint, nadvstep, nadvstep_half, nout;   /* grid.jl:70,77-78,86 */
    /* (Tprog) constants.jl:69-75,81,86 and grid.jl:74 */
    double RKaDt[4], RKbDt[3];
    double Dt_Ds, Dt_D, Dt_D_half;
    int32_t ssp_n, ssp_s, ssp_kn, ssp_mn;            /* constants.jl:20-27 */
    double ssp_Dt_Dn, ssp_kna, ssp_knb, ssp_Dt_Dnc;
    double one_minus_alpha, Dt_diff;
    /* (T) constants.jl:87-105, grid.jl:80-83 */
    double g, cD, rD, gamma, cSmag, nuB;
    double dtadvu, dtadvv, half_dtadvu, half_dtadvv;
    double jSST, SSTmin;
    /* (F64) constants.jl:108-110 */
    double omegaFx, omegaFy, omegaFeta;
    /* classic RK: sum the stage increments apart and add them to the state with Kahan's
     * compensated summation (a persistent compensation per variable; not in the reference) */
    int32_t compensated;
    /* T = Float16 or BFloat16 with Tprog = Float32 only: T's values (the stage inputs, the
     * static fields, the (T) constants) in T, the arithmetic on them in Float32, results not
     * rounded back to T (a form of this library, not of the reference; nstep_advcor = 0) */
    int32_t t_arith32;
} swm_config;

typedef struct swm_ctx swm_ctx;

/* Library identification and device check.  *ndev receives hipGetDeviceCount. */
int swm_version(int32_t *abi_version);
int swm_device_count(int32_t *ndev);

/* Create a context on the current HIP device; validates the config.
 * Replaces the allocation half of preallocate (src/preallocate.jl:454-471). */
int swm_create(const swm_config *cfg, swm_ctx **out);
int swm_destroy(swm_ctx *ctx);
const char *swm_last_error(const swm_ctx *ctx);

/* Static fields of Forcing{T} and Grid (format T, Julia layout):
 *   H     (nx+2, ny+2)        forcing.jl:39-47
 *   Fx    (nux, nuy)          forcing.jl:15-25
 *   Fy    (nvx, nvy)          forcing.jl:27-37
 *   f_q   (nx+1, ny+1)        grid.jl:95
 *   eta_ref, Feta (nx, ny)    forcing.jl:49-50; may be NULL unless continuity needs them */
int swm_set_static(swm_ctx *ctx, const void *H, const void *Fx, const void *Fy,
                   const void *f_q, const void *eta_ref, const void *Feta);

/* The same static fields (and f_u, f_v with the linear dynamics) computed on the device
 * instead of uploaded: Forcing{T}(P, G) and Grid.f_q / f_u / f_v (forcing.jl:11-217,
 * grid.jl:91-108), each field evaluated in Float64 from the grid coordinates with the
 * reference's formulas and operation order and rounded once to T.  The setup below carries
 * the scalars the host derives from Parameter and Grid. */
enum { SWM_TOPO_FLAT = 0, SWM_TOPO_RIDGE = 1, SWM_TOPO_RIDGES = 2, SWM_TOPO_SEAMOUNT = 3 };
enum { SWM_WIND_CHANNEL = 0, SWM_WIND_SHEAR = 1, SWM_WIND_DOUBLE_GYRE = 2, SWM_WIND_CONSTANT = 3, SWM_WIND_NONE = 4 };
typedef struct swm_setup {
    int32_t topography;           /* SWM_TOPO_*, P.topography */
    int32_t wind_x, wind_y;       /* SWM_WIND_*, P.wind_forcing_x / _y (y: CONSTANT or NONE) */
    int32_t pad0;
    double Lx, Ly, Delta;         /* grid.jl:21-23 (F64) */
    double H, topo_height, topo_width;
    double Fx0, Fy0, rho;
    double f0, beta;              /* grid.jl:91-92 (F64) */
    double eta_refh, eta_refw;    /* forcing.jl:127-136 */
    double A, phik, wk, R, phi;   /* forcing.jl:198-212 */
} swm_setup;
int swm_set_static_device(swm_ctx *ctx, const swm_setup *setup);
/* The static fields as the context holds them, in the shapes swm_set_static takes (NULL
 * skips one; eta_ref / Feta only where continuity reads them; Fx, Fy, f_q multiplied by
 * 2^scale_log2 where the velocities are carried scaled). */
int swm_download_static(swm_ctx *ctx, void *H, void *Fx, void *Fy, void *f_q, void *eta_ref, void *Feta);

/* The Coriolis parameter on the u- and v-grid, read by the linear dynamics only
 * (rhs_linear!, src/rhs.jl:55-86; Grid.f_u, f_v, src/grid.jl:93-94), format T, Julia layout:
 *   f_u   (nux, nuy)
 *   f_v   (nvx, nvy)
 * Required once, after swm_create, when dynamics is SWM_LINEAR; an error otherwise. */
int swm_set_static_linear(swm_ctx *ctx, const void *f_u, const void *f_v);

/* Prognostic variables in format Tprog, halo included (PrognosticVars,
 * src/initial_conditions.jl:6-11):
 *   u (nux+4, nuy+4), v (nvx+4, nvy+4), eta (nx+2, ny+2), sst (nx+2, ny).
 * Upload applies ghost_points! and ghost_points_sst! as the first step of the loop would
 * (time_integration.jl:48; ghost_points.jl:182-242).  Halo cells that the reference never
 * writes must be zero (as add_halo leaves them); upload rejects other values. */
int swm_upload_prog(swm_ctx *ctx, const void *u, const void *v, const void *eta, const void *sst);
int swm_download_prog(swm_ctx *ctx, void *u, void *v, void *eta, void *sst);

/* Run time steps i = i0 .. i0+nsteps-1 of the loop of time_integration.jl:45-243 with
 * model time t0 at step i0 (the reference starts at i0=1, t0=0; t advances by dtint per
 * step).  NaN detection runs every nout steps as feedback.jl:136-148 does; on detection
 * the loop stops before the copy-back (time_integration.jl:235-237) and *nan_step
 * receives the step, else 0.  *steps_done receives the number of steps completed. */
int swm_step(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0,
             int64_t *steps_done, int64_t *nan_step);

/* Per-operator entry points (parity tests).
 * swm_stages: at step i0 with time t0, run the first n_stages stages of the time scheme
 * (time_integration.jl:53-196) from the current Prog, without ghost points, dissipation,
 * tracer or copy-back.  swm_download_rk then returns u0,v0,eta0 and u1,v1,eta1
 * (RungeKuttaVars, preallocate.jl:21-26, format Tprog; NULL skips an array). */
int swm_stages(swm_ctx *ctx, int32_t n_stages, int64_t i0, int64_t t0);
int swm_download_rk(swm_ctx *ctx, void *u0, void *v0, void *eta0, void *u1, void *v1, void *eta1);

/* swm_ghost: ghost_points! of (u, v, eta) and ghost_points_sst! of sst, on the current Prog
 * in place (ghost_points.jl:136-199; what time_integration.jl:48 applies each step). */
int swm_ghost(swm_ctx *ctx);

/* swm_dissipation: bottom_drag!, diffusion!, add_drag_diff_tendencies! and ghost_points_uv!
 * (time_integration.jl:216-221; bottom_drag.jl:2-119, diffusion.jl:2-251) with the current
 * Prog as u0, v0, eta0; the result becomes Prog (eta unchanged). */
int swm_dissipation(swm_ctx *ctx);

/* swm_tracer: tracer! at step i (tracer_advection.jl:1-44) with the current Prog velocities as
 * the post-dissipation velocities (time_integration.jl:226-228): the mid-point snapshot
 * um, vm <- u, v when (i + nadvstep_half) % nadvstep == 0, the semi-Lagrangian advection of
 * sst when i % nadvstep == 0.  A departure point that is NaN or outside the grid returns
 * SWM_E_ARG (the reference throws InexactError / BoundsError). */
int swm_tracer(swm_ctx *ctx, int64_t i);

/* FP64 reductions over the current Prog (no counterpart in the reference, which only sums
 * isnan, feedback.jl:68): mass = sum of eta over the nx*ny T-cells; energy = sum over
 * T-cells (x,y) of 0.5*g*eta^2 + 0.5*h*u^2 + 0.5*h*v^2 with h = eta + H of that cell, u
 * at its right face (x = 1..nx) and v at its top face (y = 1..ny-1). */
int swm_reduce(swm_ctx *ctx, double *mass, double *energy);

/* swm_reduce without allocating: no counterpart in the reference.  For callers that cannot
 * allocate device memory (serving loops, code under HIP graph capture, harnesses that forbid
 * hipMalloc).  The same mass and energy, bitwise, as swm_reduce; the row partials and the
 * compensated sums live in a static device array the library holds for the process (one copy
 * per GPU, 512 KiB).  At most 32768 rows (ny); larger grids get SWM_E_ARG.  Calls on one GPU
 * share that array: calls that overlap in time on one GPU (several threads, several
 * contexts' streams) race and must be serialised by the caller. */
int swm_reduce_noalloc(swm_ctx *ctx, double *mass, double *energy);

/* Measurement hook (no counterpart in the reference): launch kernel `which` `reps` times
 * back to back on the context's stream between two HIP events and return the average
 * duration in milliseconds.  which: 0 middle RK stage (stage 2), 1 first stage, 2 last
 * stage, 3 dissipation, 4 tracer advection, 5 fused stages 1+2, 6 fused stages 3+4
 * (nonperiodic).  Operates on the current buffers; the prognostic state is left unchanged
 * except the RK scratch slots. */
int swm_time_kernel(swm_ctx *ctx, int32_t which, int32_t reps, double *ms_avg);
/* Measurement hook (no counterpart in the reference): run steps i0 .. i0+nsteps-1 as
 * swm_step does, with HIP events recorded around each kernel launch, and return per kind of
 * kernel the average duration (ms) and the number of launches timed: [0] fused stages 1+2,
 * [1] fused stages 3+4, [2] dissipation, [3] tracer advection, [4] the whole step, [5] first
 * stage, [6] middle stage, [7] last stage (one kernel per stage).  ms_avg and counts hold 8
 * entries; kinds that did not run give 0.  The state advances as by swm_step. */
int swm_time_steps(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, double *ms_avg, int32_t *counts);

/* ---- Ensembles (no counterpart in the reference, whose "ensemble mode" runs separate
 * processes with run ids, README.md:88, src/output.jl:168-207).  An ensemble advances
 * several contexts, of possibly different grid sizes, together: each kernel of a step
 * (stage or fused stage pair, dissipation) is one batched launch over the tiles of every
 * member, and pairs of steps without host work (tracer, NaN check) replay one captured
 * hipGraph.  Members must share the number formats, boundary condition, time scheme
 * (classic RK), drag, diffusion, continuity variant and device, use nstep_diff = 1, and
 * have their static fields and a state uploaded; each member remains a context of its own
 * (download, step) between ensemble calls.  Results equal each member's own swm_step bit
 * for bit. */
typedef struct swm_ens swm_ens;
int swm_ens_create(int32_t n, swm_ctx *const *members, swm_ens **out);
int swm_ens_destroy(swm_ens *ens);
const char *swm_ens_last_error(const swm_ens *ens);
/* Autotuner: time the batched step for candidate rows per wave segment (4..48) of the stage,
 * fused-pair and dissipation kernels in turn and keep the fastest; the choices are returned
 * (pointers may be NULL). The timed steps leave every member as it was (state, and the kept
 * q and KEu+KEv of nstep_advcor > 0), so it may run before the first step or between steps.
 * Tuning changes only which rows per segment the kernels use. A step's results are the same
 * for any segment length. */
int swm_ens_tune(swm_ens *ens, int32_t reps, int32_t *rows_stage, int32_t *rows_fused, int32_t *rows_diss);
/* Steps i0 .. i0+nsteps-1 of every member, member m at model time (i-1)*dtint of m.  NaNs
 * are checked every nout steps of each member (feedback.jl:136-148); after a step in which
 * any member has them, all members stop: members with NaNs keep their state from before that
 * step and nan_steps[m] (n entries) receives the step, else 0.  *steps_done receives the
 * steps completed, *graph_steps how many of them ran as graph replays. */
int swm_ens_step(swm_ens *ens, int64_t nsteps, int64_t i0, int64_t *steps_done, int64_t *nan_steps,
                 int64_t *graph_steps);
/* Persistent launch (off by default; swm_ens_set_persistent(ens, 1) turns it on):
 * swm_ens_step runs all its steps in one cooperative launch that holds as many blocks as
 * stay resident on the device.  For every step it runs
 * the step's batched kernels (fused RK4 stage pairs, dissipation) over the tiles of every
 * member, then the members' tracer snapshot, advection and NaN check where each member's own
 * schedule has them, with grid-wide barriers between phases; a NaN found stops every member
 * after that step, as above.  Used when every member runs the fused RK4 pairs with
 * T = Tprog (Float64 or Float32), the continuity itself and Arakawa-Hsu advection (any
 * drag and diffusion), with a plan that repeats every two steps (no seasonal wind; a state
 * just uploaded takes one batched step first to settle its slot roles); other steps run as
 * batched launches and graph replays.  Results are the same bit for bit either way.  The batched launches are the
 * default because they run faster on an MI355X: the persistent kernel's phases share one
 * register allocation (the largest of the stage pair, seam and dissipation bodies, with
 * spills) and each grid barrier costs tens of microseconds of cross-XCD cache write-back and
 * atomics, where a graph replays its kernels back to back.  swm_ens_persistent_steps gives the steps of the last swm_ens_step that ran
 * in it (0 when that call used the other path) and, where it did not (why_not nullable), a
 * static string saying why.  A barrier that waits longer than 5 s (a
 * launch whose blocks are not all resident) makes every block leave and the call fails with
 * SWM_E_HIP; the members' states are then undefined. */
int swm_ens_set_persistent(swm_ens *ens, int32_t on);
/* Measurement hook: with swm_ens_set_persistent(ens, 2) the persistent launch records the
 * time block 0 leaves each of its first grid barriers; ms[0..3] then receive the average
 * duration in ms of the first fused stage pair, the second and the dissipation (each up to
 * the barrier after it), and of whole steps, over the steps of the last call's first launch
 * that follow a step without tracer or NaN-check work (-1 where none). */
int swm_ens_phase_ms(const swm_ens *ens, double *ms);
int swm_ens_persistent_steps(const swm_ens *ens, int64_t *steps, const char **why_not);

/* ---- Reverse mode (no counterpart in the reference, which has no derivatives; it is what a
 * ChainRulesCore.rrule of time_integration, src/time_integration.jl:2-250, would call).
 * Every time scheme of the reference: classic RK, SSPRK2 with RKs = 2 or 3 stages, SSPRK3
 * with RKn = 2 or 3, 4SSPRK3 (time_integration.jl:53-195; the taped step keeps every stage
 * input, so more stages are refused), with T = Tprog = Tcomm in {Float64, Float32}; the
 * nonlinear dynamics
 * (Arakawa-Hsu or Sadourny PV fluxes, any nstep_advcor: the kept q and KEu+KEv are
 * differentiated through to the states they are formed from) or the linear dynamics
 * (rhs.jl:55-86).  Kept fields formed before swm_adj_forward's first step (not by it) are
 * constants.
 * Derivatives are those of
 * the exact-arithmetic linearisation of each operation; at the kinks: sqrt'(0) = 0
 * (Smagorinsky ν, drag √KE), floor / clip / wrap of the tracer's departure points held
 * constant (tracer_advection.jl:211-239).
 *
 * swm_adj_forward: run steps i0 .. i0+nsteps-1 from the current state as swm_step does
 * (the NaN check stops it with SWM_E_STATE) and keep device snapshots of full states
 * (u, v, η, sst and the tracer's mid-point velocities) where a binomial checkpointing
 * schedule with `snapshots` snapshots places them (revolve: the steps are recomputed at most
 * r times for the least r with nsteps <= C(snapshots + r, snapshots); 0 snapshots recompute
 * from the start for every step; nsteps - 1 or more keep every state, so the reverse sweep
 * recomputes nothing beyond its taped steps).  snapshots < 0: as many as half of the device's
 * free memory holds, at most nsteps - 1 (on an MI355X every state of most runs).  The
 * snapshots live in one device block, reused by later calls of the same layout.  Leaves the
 * final state, as swm_step would. */
int swm_adj_forward(swm_ctx *ctx, int64_t nsteps, int64_t i0, int64_t t0, int32_t snapshots);
/* swm_adj_pullback: the vector-Jacobian product of the recorded steps.  ybar_* are the
 * cotangents of the final Prog (Julia layout and format Tprog, halo included, as
 * swm_download_prog returns it); xbar_* (nullable) receive the gradient of <ybar, y> with
 * respect to the Prog given to the forward run, halo cells zero (upload applies the ghost
 * copies, so the gradient lives on the interior).  The tracer's mid-point velocities at i0
 * are constants (never read when i0 = 1).  *forward_steps (nullable) receives the forward
 * steps recomputed, taped steps included; *sweep_ms (nullable) the device time of the
 * reverse sweep (host transfers excluded).  The context is left at the final state. */
int swm_adj_pullback(swm_ctx *ctx, const void *ybar_u, const void *ybar_v, const void *ybar_eta, const void *ybar_sst,
                     void *xbar_u, void *xbar_v, void *xbar_eta, void *xbar_sst, int64_t *forward_steps,
                     double *sweep_ms);
/* swm_adj_track_Fx: with on != 0, the following pullbacks also accumulate the gradient of
 * <ybar, y> with respect to the zonal wind stress Fx (Forcing.Fx, forcing.jl:15-25), which
 * enters every stage's du as Tprog(Fxt*Fx) (rhs.jl:242): the sum over the reversed steps and
 * stages of Fxt times the cotangent of du.  Off by default (one more field read and written
 * per stage adjoint).  swm_adj_download_Fx: that gradient from the last pullback, format T
 * (= Tprog), shape (nux, nuy) of Fx. */
int swm_adj_track_Fx(swm_ctx *ctx, int32_t on);
int swm_adj_download_Fx(swm_ctx *ctx, void *xbar_Fx);
/* Measurement hook: average duration in ms of `reps` launches of: 0 a forward step, 1 a taped
 * forward step (unfused stages, intermediates kept), 2 the reverse of one step (all its
 * adjoint kernels), 3 one stage adjoint kernel, 4 the dissipation adjoint kernel.  The state
 * is restored afterwards. */
int swm_adj_time(swm_ctx *ctx, int32_t which, int32_t reps, double *ms_avg);

#ifdef __cplusplus
}
#endif

#endif /* SWM_H */
