/* hypar_mi355x.h -- C-ABI of the MI355X-native NS3D WENO5 + RK4 path.
 *
 * libhypar_mi355x.so runs HyPar's explicit right-hand side and RK4 time step
 * for the 3D compressible Navier-Stokes model (NavierStokes3D) on one gfx950
 * GPU with hand-written HIP kernels.  Every entry point takes plain pointers
 * and sizes and returns an hm_status (0 = success); it never exits the
 * process.  hm_last_error() describes the most recent failure of the calling
 * thread.
 *
 * Device layout ("hm layout"): structure of arrays without ghost points,
 *   u[v*ncell + (k*ny + j)*nx + i],  v = 0..4 (rho, rho*u, rho*v, rho*w, e),
 *   ncell = nx*ny*nz, x fastest.
 * Ghost points of the reference layout are not stored: every boundary is
 * periodic (BCPeriodicU) and the kernels wrap indices, which gives the values
 * the reference's ghost fill would.  The reference CPU layout (AoS, variables
 * fastest, 3 ghost layers) is accepted at the host transfer entry points.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   hm_rk4, hm_ssprk3 .. HyPar::TimeIntegrate = TimeRK (include/hypar.h:220,
 *                        src/TimeIntegration/TimeRK.c:35-202)
 *   hm_rhs ............. TimeIntegration::RHSFunction = TimeRHSFunctionExplicit
 *                        (src/TimeIntegration/TimeRHSFunctionExplicit.c:30-98)
 *   hm_hyperbolic ...... HyPar::HyperbolicFunction (include/hypar.h:250-253,
 *                        src/HyParFunctions/HyperbolicFunction.c:31-126)
 *   hm_parabolic ....... HyPar::ParabolicFunction = NavierStokes3DParabolicFunction
 *                        (include/hypar.h:256, src/PhysicalModels/NavierStokes3D/
 *                        NavierStokes3DParabolicFunction.c:50-325)
 *   hm_cfl ............. HyPar::ComputeCFL = NavierStokes3DComputeCFL
 *                        (include/hypar.h:269, NavierStokes3DComputeCFL.c:16-55)
 *   hm_step_norm ....... TimePostStep norm (src/TimeIntegration/TimePostStep.c:43-63)
 *   hm_volume_integral . HyPar::VolumeIntegralFunction = VolumeIntegral
 *                        (include/hypar.h:388, src/HyParFunctions/VolumeIntegral.c:14-50)
 *   hm_flux ... hm_fill  the per-slot operators (see below)
 */
#ifndef HYPAR_MI355X_H
#define HYPAR_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HM_OK = 0,
  HM_ERR_ARG = 1,         /* invalid argument or size                    */
  HM_ERR_HIP = 2,         /* a HIP runtime call failed                   */
  HM_ERR_NONFINITE = 3,   /* a reduction found NaN/Inf in the state      */
  HM_ERR_UNSUPPORTED = 4, /* scheme or option not implemented            */
  HM_ERR_NOMEM = 5,       /* device allocation failed                    */
  HM_ERR_TIMEOUT = 6      /* a bounded wait inside a kernel gave up; the
                             results of that call are not valid          */
} hm_status;

/* hyp_interp_type of solver.inp (src/Simulation/ReadInputs.c) */
typedef enum { HM_INTERP_COMPONENTS = 0, HM_INTERP_CHARACTERISTIC = 1 } hm_interp_kind;
/* upwinding of physics.inp (NavierStokes3DInitialize.c:162-333) */
typedef enum { HM_UPWIND_ROE = 0, HM_UPWIND_RUSANOV = 1 } hm_upwind_kind;
/* WENO weight formulation of weno.inp (WENOInitialize.c:62-96; dispatch
 * WENOFifthOrderCalculateWeights.c:55-57): Jiang-Shu, mapped (Henrick,
 * Aslam & Powers), WENO-Z (Borges et al.), Yamaleev-Carpenter, and the
 * optimal weights 0.1, 0.6, 0.3 that no_limiting 1 freezes
 * (WENOFifthOrderInitializeWeights.c:33-228, InitializeSolvers.c:219) */
typedef enum {
  HM_WENO_JS = 0, HM_WENO_MAPPED = 1, HM_WENO_Z = 2, HM_WENO_YC = 3, HM_WENO_OPTIMAL = 4
} hm_weno_kind;
/* host array layouts accepted by the transfer calls */
typedef enum {
  HM_LAYOUT_REFERENCE = 0, /* AoS, variables fastest, 3 ghost layers (HyPar CPU) */
  HM_LAYOUT_AOS_INTERIOR = 1,/* AoS, variables fastest, interior only          */
  HM_LAYOUT_SOA_INTERIOR = 2 /* hm layout                                        */
} hm_layout_kind;

typedef struct {
  int dim[3];          /* interior grid size (each >= 5)                          */
  double gamma;        /* physics.inp gamma                                       */
  double Re;           /* physics.inp Re; <= 0 is inviscid.  Divided by Minf as
                          NavierStokes3DInitialize.c:368 does                     */
  double Minf;         /* physics.inp Minf                                        */
  double Pr;           /* physics.inp Pr                                          */
  int interp;          /* hm_interp_kind                                             */
  int upwind;          /* hm_upwind_kind                                             */
  double weno_eps;     /* weno.inp epsilon (p = 2)                                */
  const double *dxinv; /* optional host array, HyPar::dxinv layout: x, y, z
                          concatenated, each with 3 ghost entries per side.
                          NULL: uniform periodic grid x_i = 2*pi*i/n, metrics
                          computed exactly as InitialSolution.c:74-113        */
  int weno_type;       /* hm_weno_kind (0, the zero-initialised value, is JS)     */
} hm_config;

typedef struct hm_solver hm_solver;

/* Create a solver bound to the current HIP device; allocates its scratch. */
hm_status hm_create(const hm_config *cfg, hm_solver **out);
/* Free a solver (NULL is ignored). */
void hm_destroy(hm_solver *s);
/* Message of the last failure on the calling thread ("" if none). */
const char *hm_last_error(void);
/* Library version string. */
const char *hm_version(void);

/* Use `stream` (a hipStream_t; NULL = default stream) for all launches. */
hm_status hm_set_stream(hm_solver *s, void *stream);
/* Number of interior cells and doubles in one hm-layout state. */
size_t hm_ncell(const hm_solver *s);

/* Host <-> device state transfer (synchronous).  `dev` is an hm-layout
 * device array of 5*ncell doubles; `host` uses `layout`. */
hm_status hm_upload(hm_solver *s, double *dev, const double *host, int layout);
hm_status hm_download(hm_solver *s, double *host, const double *dev, int layout);

/* nsteps classical RK4 steps of size dt on the hm-layout device state u,
 * in place (TimeRK with the "rk 44" tableau). */
hm_status hm_rk4(hm_solver *s, double *u, double dt, int nsteps);
/* nsteps three-stage SSP Runge-Kutta steps, in place (TimeRK with the
 * "rk ssprk3" / "tvdrk3" tableau, TimeExplicitRKInitialize.c:69-79). */
hm_status hm_ssprk3(hm_solver *s, double *u, double dt, int nsteps);

/* rhs = -hyperbolic + parabolic (+ zero source) of the device state u.
 * rhs and u must be distinct arrays (HM_ERR_ARG otherwise). */
hm_status hm_rhs(hm_solver *s, double *rhs, const double *u);
/* hyperbolic term alone: sum over x,y,z of dxinv*(F_{i+1/2}-F_{i-1/2}).
 * hyp and u must be distinct arrays (HM_ERR_ARG otherwise). */
hm_status hm_hyperbolic(hm_solver *s, double *hyp, const double *u);
/* viscous term alone (zero when Re <= 0). */
hm_status hm_parabolic(hm_solver *s, double *par, const double *u);

/* Max over cells and directions of (|v_d|+c)*dt*dxinv_d. Sets *cfl. Returns
 * HM_ERR_NONFINITE when the state holds NaN/Inf. */
hm_status hm_cfl(hm_solver *s, const double *u, double dt, double *cfl);
/* sqrt(sum((a-b)^2) / ncell) over all 5 variables (TimePostStep norm). */
hm_status hm_step_norm(hm_solver *s, const double *a, const double *b, double *norm);
/* Volume integral of each conserved variable, out[v] = sum over interior
 * cells of u_v dV, dV = 1/dxinv_x 1/dxinv_y 1/dxinv_z (the conservation
 * check of TimePostStep.c:81-93).  Synchronises; HM_ERR_NONFINITE on NaN/Inf. */
hm_status hm_volume_integral(hm_solver *s, const double *u, double *out);

/* ---- an entry point that allocates nothing ----------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hm_cfl without a solver object.  The grid metrics and
 * the reduction partials live in a static device array of
 * HM_NOALLOC_SCRATCH_DOUBLES doubles (66 KiB) that each GPU the library
 * runs on holds until the process ends; grid sides above
 * HM_NOALLOC_MAX_SIDE are rejected with HM_ERR_ARG.  Enqueued on `stream`
 * (a hipStream_t; NULL = default stream), then synchronised.  Calls that
 * overlap on one GPU share that array and race.  The result equals
 * hm_cfl's for a solver created from the same cfg (its dxinv, or the
 * uniform periodic grid's metrics when cfg->dxinv is NULL). */
#define HM_NOALLOC_MAX_SIDE 2048
#define HM_NOALLOC_SCRATCH_DOUBLES 8448
hm_status hm_cfl_noalloc(const hm_config *cfg, const double *u, double dt, double *cfl, void *stream);

/* ---- device memory (so that host code needs no HIP headers) ---------- */
/* Allocate / free n doubles of device memory on the current device. */
hm_status hm_malloc(double **dev, size_t n);
void hm_free(double *dev);
/* Synchronous copies of n doubles: host->device, device->host, device->device. */
hm_status hm_copy_to_device(double *dev, const double *host, size_t n);
hm_status hm_copy_to_host(double *host, const double *dev, size_t n);
hm_status hm_copy_device(hm_solver *s, double *dst, const double *src, size_t n);
/* Wait for all work enqueued on the solver's stream. */
hm_status hm_synchronize(hm_solver *s);

/* ---- per-slot operators -------------------------------------------------
 * HyPar composes its hyperbolic term from function-pointer slots
 * (ReconstructHyperbolic, src/HyParFunctions/HyperbolicFunction.c:167-222).
 * Each entry below is one slot on hm-layout device arrays; the scheme
 * (component-wise/characteristic, Roe/Rusanov) is the solver's.  Interface
 * arrays store interface i-1/2 at the index of cell i (periodic lines hold
 * n interfaces). */
/* FFunction = NavierStokes3DFlux (NavierStokes3DFlux.c:23-57; include/hypar.h:276) */
hm_status hm_flux(hm_solver *s, double *f, const double *u, int dir);
/* UFunction = NavierStokes3DModifiedSolution (NavierStokes3DModifiedSolution.c:31-70;
 * include/hypar.h:321) */
hm_status hm_modified_solution(hm_solver *s, double *uC, const double *u);
/* InterpolateInterfacesHyp = Interp1PrimFifthOrderWENO[Char] with the weights
 * of SetInterpLimiterVar = WENOFifthOrderCalculateWeights[Char]
 * (include/hypar.h:224,234; Interp1PrimFifthOrderWENO.c:74-168,
 * Interp1PrimFifthOrderWENOChar.c:86-198).  upw > 0 left-biased, < 0
 * right-biased; uflag = 1 when fC is the (modified) solution, whose weights
 * then come from the raw solution u. */
hm_status hm_interpolate(hm_solver *s, double *fI, const double *fC, const double *u, int upw,
                         int dir, int uflag);
/* Upwind = NavierStokes3DUpwindRoe / NavierStokes3DUpwindRusanov
 * (NavierStokes3DUpwind.c:40-124,349-420; include/hypar.h:295) */
hm_status hm_upwind(hm_solver *s, double *fI, const double *fL, const double *fR,
                    const double *uL, const double *uR, const double *u, int dir);
/* hyp (+)= dxinv*(F_{i+1/2}-F_{i-1/2}) along dir (HyperbolicFunction.c:94-101) */
hm_status hm_flux_difference(hm_solver *s, double *hyp, const double *fI, int dir, int accumulate);
/* y += a*x and x = value on n doubles (_ArrayAXPY_, _ArraySetValue_,
 * include/arrayfunctions.h:271-275) */
hm_status hm_axpy(hm_solver *s, const double *x, double a, double *y, size_t n);
hm_status hm_fill(hm_solver *s, double *x, double value, size_t n);

/* ---- discrete adjoint ------------------------------------------------
 * No counterpart in the reference (it has no adjoint).  Reverse-mode
 * derivatives of this library's own discretisation, inviscid or viscous;
 * every WENO5 weight formulation (JS, mapped, Z, YC) with Roe or Rusanov
 * upwinding, component-wise or characteristic-wise, grid sides <= 512
 * (HM_ERR_UNSUPPORTED otherwise).  lam, out: hm-layout
 * device arrays. */
/* out = (d rhs / d u)^T lam at state u */
hm_status hm_rhs_vjp(hm_solver *s, double *out, const double *u, const double *lam);
/* lam <- (d u_{n+1} / d u_n)^T lam for one RK4 step of size dt from u_n */
hm_status hm_rk4_vjp(hm_solver *s, const double *u_n, double dt, double *lam);
/* Forward sweep of nsteps RK4 steps in place on u, keeping a checkpoint of
 * the state every `checkpoint_every` (K) steps in HBM (the solver owns
 * them) and, for the last `sets` segments of K steps, either the four RK
 * stage states of every step (stage level) or the state every step starts
 * from (step level).  The schedule is the cheapest that fits the tape
 * budget, in this order: stage level with 3, 2, 1 sets (12K, 8K, 4K
 * states), then step level with 3, 2, 1 sets (3K, 2K, K states, plus 3).
 * HM_ERR_NOMEM names the bytes the smallest schedule needs when none fits. */
hm_status hm_adjoint_forward(hm_solver *s, double *u, double dt, int nsteps, int checkpoint_every);
/* Reverse sweep of the last hm_adjoint_forward: lam holds dJ/du_N on entry
 * and dJ/du_0 on return.  Every segment the sets do not hold is recomputed
 * once from its checkpoint; at step level every step also runs forward
 * once more for its stage states before it is taken back.  The gradient is
 * the same, bit for bit, whatever the schedule. */
hm_status hm_adjoint_reverse(hm_solver *s, double *lam);
/* Bytes the tape of the next forward sweep may hold: checkpoints, sets,
 * scratch states and viscous work space.  0 (the default) takes the
 * device's free memory plus what the tape already holds, less a reserve of
 * 2 GiB or 2% of the device, whichever is larger. */
hm_status hm_set_tape_budget(hm_solver *s, size_t bytes);
/* The schedule of the recorded trajectory: sets (1..3), step_level (0/1)
 * and the bytes it holds (any pointer may be NULL). */
hm_status hm_tape_info(const hm_solver *s, int *sets, int *step_level, size_t *bytes);

/* Kernel timing for roofline reports (no counterpart in the reference; the
 * reference's GPU_STAT build prints per-kernel cudaEvent times,
 * HyperbolicFunction_GPU.cu:517-576).  Launches each kernel of one RK stage
 * `reps` times on the solver's stream with state u and stores the mean
 * duration in milliseconds, measured with HIP events on that stream:
 * ms[0..2] the fused x, y, z passes of an RK stage (the z pass in its
 * stage-1 form, which reads u and the accumulator and writes the next stage
 * state); ms[3] the viscous-term kernel and ms[4] the primitive-variable
 * kernel that hm_rk4 runs once per call (0 if inviscid).  u is not
 * modified. */
hm_status hm_profile_kernels(hm_solver *s, const double *u, int reps, double *ms);

/* ------------------------------------------------------------------------
 * Batched ensembles.
 *
 * Replaces the member loop of ensemble_simulations
 * (src/TimeIntegration/TimeRK.c:130-193, include/ensemble_simulations.h,
 * src/Simulation/ReadInputs.c:189-207), which advances the nsims members
 * one after another with a shared dt.  Here all members, of uneven grid
 * sizes and each with its own dt, advance together: every direction pass
 * of an RK stage is one persistent launch over the (member, tile) work
 * items of all members (or, with tile shapes 2 and 3, the three passes of
 * a stage are one launch), and one RK4 step is a captured HIP graph
 * replayed per step; viscous members add one launch per stage over the
 * (member, tile) items of the viscous term.  Members share the scheme
 * (interpolation, upwinding, WENO weights, gamma, eps, Re, Minf, Pr).
 * The batched state holds member m's hm-layout state at
 * u[5*offset_m .. 5*(offset_m + ncell_m)).
 * ------------------------------------------------------------------------ */
typedef struct hm_ensemble hm_ensemble;

hm_status hm_ensemble_create(const hm_config *members, int nmembers, hm_ensemble **out);
void hm_ensemble_destroy(hm_ensemble *e);
/* total cells of all members */
size_t hm_ensemble_ncell(const hm_ensemble *e);
/* first cell and cell count of member m in the batched state */
hm_status hm_ensemble_member(const hm_ensemble *e, int m, size_t *offset, size_t *ncell);
hm_status hm_ensemble_set_stream(hm_ensemble *e, void *stream);
/* per-member time steps dt[0..nmembers-1] (synchronous) */
hm_status hm_ensemble_set_dt(hm_ensemble *e, const double *dt);
/* nsteps RK4 steps of every member, in place (graph replays on the
 * stream).  With the one-launch-per-stage shapes (2, 3) it synchronises the
 * stream at the end and returns HM_ERR_TIMEOUT if a dependency wait inside
 * a stage kernel gave up. */
hm_status hm_ensemble_rk4(hm_ensemble *e, double *u, int nsteps);
/* Tile shape and persistent grid sizes.  Shapes 0 and 1: one launch per
 * direction pass, blocks of 4 (0) or 2 (1) wavefronts = x-lines / rows per
 * march step, grid[d] blocks for pass d (0: one block per work item).
 * Shapes 2 and 3: one launch per RK stage for the x, y and z work items of
 * all members (blocks of 4 or 2 wavefronts), grid[0] blocks (0: as many
 * as the GPU holds at once); viscous members add one launch per stage for
 * their viscous terms. */
hm_status hm_ensemble_set_tiles(hm_ensemble *e, int shape, const unsigned *grid);
hm_status hm_ensemble_get_tiles(const hm_ensemble *e, int *shape, unsigned *grid);
/* Autotuner: times `reps` graph-replayed steps of every candidate (shapes
 * 0 and 1 with four grid sizes each, shapes 2 and 3 with the co-resident
 * grid) on a scratch copy of u (u is not modified), keeps the fastest and
 * returns it. */
hm_status hm_ensemble_tune(hm_ensemble *e, const double *u, int reps, int *shape, unsigned *grid);
/* Polls after which a dependency wait of the one-launch-per-stage kernel
 * gives up and flags HM_ERR_TIMEOUT (default 2^21, well above a stage's
 * length).  For tests of that error path. */
hm_status hm_ensemble_set_wait_limit(hm_ensemble *e, unsigned polls);
/* Batched discrete adjoint (no counterpart in the reference): the adjoint
 * of hm_adjoint_forward / hm_adjoint_reverse for all members at once, each
 * with its own dt.  The forward sweep advances the batched state nsteps RK4
 * steps in place with a checkpoint every `checkpoint_every` steps and the
 * schedule of hm_adjoint_forward; the reverse sweep recomputes what the
 * schedule did not keep, takes lam = dJ/du_N of all members and returns
 * dJ/du_0.
 * Every VJP launch of a stage covers the work items of all members (groups
 * of whole lines of one member, one tile or one cell).  Built for the
 * schemes hm_rhs_vjp supports. */
hm_status hm_ensemble_adjoint_forward(hm_ensemble *e, double *u, int nsteps, int checkpoint_every);
hm_status hm_ensemble_adjoint_reverse(hm_ensemble *e, double *lam);
/* The tape budget and schedule of the batched adjoint, as hm_set_tape_budget
 * and hm_tape_info for one grid. */
hm_status hm_ensemble_set_tape_budget(hm_ensemble *e, size_t bytes);
hm_status hm_ensemble_tape_info(const hm_ensemble *e, int *sets, int *step_level, size_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
