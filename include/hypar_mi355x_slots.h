/* hypar_mi355x_slots.h -- HyPar's own function-pointer slots, exported with
 * their exact signatures by libhypar_mi355x.so.
 *
 * HyPar calls its operators through function pointers on the solver object
 * (include/hypar.h:210-359 of the reference) and on the time-integration
 * object (include/timeintegration_struct.h:96-98).  Each slot receives the
 * object as an opaque `void *`.  The entry points below have those
 * signatures, so a maintainer installs them with one assignment per slot:
 *
 *   hm_bind(solver, hm);                        once per solver object
 *   hm_bind_state(solver, solver->u, HM_SLOT_ARRAYS_REFERENCE_HOST);
 *   hm_bind_time(&TS, solver, TS.dt, HM_TIME_RK4);
 *   solver->HyperbolicFunction = hm_slot_HyperbolicFunction;
 *   solver->ParabolicFunction  = hm_slot_ParabolicFunction;
 *   solver->ComputeCFL         = hm_slot_ComputeCFL;
 *   TS.RHSFunction             = hm_slot_RHSFunction;
 *   TS.TimeIntegrate           = hm_slot_TimeIntegrate;
 *
 * The library keeps a table from each bound object pointer to its solver,
 * its state array and the convention of the arrays its slots receive.
 * Slots return what the reference's return: 0 on success and nonzero on
 * failure for the int slots, the CFL number (negative on failure) for
 * ComputeCFL.  hm_last_error() tells why a slot failed.  The table is
 * guarded by a lock; calls on the same solver from several threads race as
 * the solver's own entry points do.
 */
#ifndef HYPAR_MI355X_SLOTS_H
#define HYPAR_MI355X_SLOTS_H

#include "hypar_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The arrays the slots of a bound object receive. */
typedef enum {
  /* device arrays in hm layout (SoA without ghosts): the state stays in HBM
   * and the slots only enqueue work on the solver's stream */
  HM_SLOT_ARRAYS_DEVICE = 0,
  /* HyPar's own host arrays in the reference layout (AoS, variables
   * fastest, 3 ghost layers; the CPU build's solver->u, solver->hyp ...).
   * Every slot call uploads its input, computes on the device and
   * downloads its output (ghost entries come back as zeros), through two
   * device arrays of 5*ncell doubles that the binding allocates on first
   * use.  A drop-in for unmodified CPU HyPar; its time includes PCIe. */
  HM_SLOT_ARRAYS_REFERENCE_HOST = 1
} hm_slot_arrays;

/* The tableau hm_slot_TimeIntegrate advances a bound member with. */
typedef enum {
  HM_TIME_RK4 = 0,   /* "rk 44" (hm_rk4)                          */
  HM_TIME_SSPRK3 = 1 /* "rk ssprk3" / "rk tvdrk3" (hm_ssprk3)     */
} hm_time_scheme;

/* Bind HyPar's solver object `hypar` (the `void *s` its slots receive) to
 * the device solver `s`.  Binding an object again replaces its solver and
 * keeps its state binding.  The arrays default to HM_SLOT_ARRAYS_DEVICE.
 * hm_destroy(s) drops every binding of s.  No counterpart in the
 * reference: it replaces the cast `(HyPar*) s` inside each slot. */
hm_status hm_bind(void *hypar, hm_solver *s);
/* The array HyPar's solver->u points to (read by ComputeCFL and advanced by
 * TimeIntegrate) and the convention of every array the object's slots
 * receive (hm_slot_arrays).  `u` may be NULL when only the term slots
 * are used. */
hm_status hm_bind_state(void *hypar, double *u, int arrays);
/* Add the bound solver object `hypar` as a member of the time-integration
 * object `ts` (the `void *ts` of TimeIntegrate), stepped with `dt` and
 * `scheme` (hm_time_scheme).  Members advance in the order they were
 * added (the reference's member loop, TimeRK.c:130-193); adding a member
 * again updates its dt and scheme. */
hm_status hm_bind_time(void *ts, void *hypar, double dt, int scheme);
/* Remove the bindings of `obj` (a solver or a time-integration object)
 * and free its staging arrays.  Unknown objects are ignored. */
void hm_unbind(void *obj);

/* HyPar::TimeIntegrate (include/hypar.h:220; TimeIntegration::TimeIntegrate,
 * include/timeintegration_struct.h:96) = TimeRK (src/TimeIntegration/
 * TimeRK.c:35-202): one time step of every member bound to `ts`, in place
 * on each member's bound state. */
int hm_slot_TimeIntegrate(void *ts);

/* TimeIntegration::RHSFunction (include/timeintegration_struct.h:98) =
 * TimeRHSFunctionExplicit (src/TimeIntegration/TimeRHSFunctionExplicit.c:
 * 30-98): rhs = -hyperbolic + parabolic + source (zero) of u.  `s` is the
 * bound solver object; `m` (MPIVariables) and `t` are not used: one rank,
 * periodic boundaries, no time-dependent terms. */
int hm_slot_RHSFunction(double *rhs, double *u, void *s, void *m, double t);

/* HyPar::HyperbolicFunction (include/hypar.h:250-253) = HyperbolicFunction
 * (src/HyParFunctions/HyperbolicFunction.c:31-126): the hyperbolic term of
 * u with the bound solver's flux (NavierStokes3DFlux), WENO5 interpolation
 * and upwinding, the pointers HyPar passes for this path
 * (TimeRHSFunctionExplicit.c:63-70).  LimFlag must be 1: the weights are
 * evaluated from u at every call (HM_ERR_UNSUPPORTED otherwise).  As in
 * the reference, a NULL FFunction gives hyp = 0; any other FFunction and
 * UpwindFunction are not called. */
int hm_slot_HyperbolicFunction(double *hyp, double *u, void *s, void *m, double t, int LimFlag,
                               int (*FFunction)(double *, double *, int, void *, double),
                               int (*UpwindFunction)(double *, double *, double *, double *, double *,
                                                     double *, int, void *, double));

/* HyPar::ParabolicFunction (include/hypar.h:256) =
 * NavierStokes3DParabolicFunction (src/PhysicalModels/NavierStokes3D/
 * NavierStokes3DParabolicFunction.c:50-325): the viscous term of u. */
int hm_slot_ParabolicFunction(double *par, double *u, void *s, void *m, double t);

/* HyPar::ComputeCFL (include/hypar.h:269) = NavierStokes3DComputeCFL
 * (NavierStokes3DComputeCFL.c:16-55): the CFL number of the bound state
 * for time step dt; negative on failure. */
double hm_slot_ComputeCFL(void *s, void *m, double dt, double t);

#ifdef __cplusplus
}
#endif

#endif
