/* hypar_host.h -- C host layer of the MI355X NS3D WENO5-RK4 path.
 *
 * Keeps the shape of HyPar's host side: a solver object whose operators are
 * function-pointer slots with the reference's signatures
 * (include/hypar.h:210-359 of the reference), a physics plugin that installs
 * its slots (NavierStokes3DInitialize.c:388-506), the reference's input files
 * and a TimeRK that composes the slots (TimeRK.c:35-202).  All arrays the
 * slots see are device arrays in hm layout (include/hypar_mi355x.h); the
 * compute is done by libhypar_mi355x.so.  Where a slot is the library's
 * own (TimeIntegrate, RHSFunction, HyperbolicFunction, ParabolicFunction,
 * ComputeCFL), the exported hm_slot_* entry point is installed or called
 * (include/hypar_mi355x_slots.h), with each solver object bound to its
 * device solver.
 */
#ifndef HYPAR_HOST_H
#define HYPAR_HOST_H

#include "../../include/hypar_mi355x.h"
#include "../../include/hypar_mi355x_slots.h"

#define HY_STR 500

/* single-rank communicator (the GPU path runs one rank per GPU) */
typedef struct {
  int rank, nproc;
  int iproc[3], ip[3];
} MPIVariables;

typedef int (*hy_flux_fn)(double *, double *, int, void *, double);
typedef int (*hy_upwind_fn)(double *, double *, double *, double *, double *, double *, int,
                            void *, double);

typedef struct {
  double gamma, Re, Pr, Minf;       /* Re as read; the device solver divides by Minf */
  double grav[3];
  char upw_choice[HY_STR];
} NavierStokes3D;

typedef struct {
  int mapped, borges, yc, no_limiting;
  double eps, p;
} WENOParameters;

typedef struct HyPar {
  /* grid */
  int ndims, nvars, ghosts;
  int dim_global[3], dim_local[3];
  long npoints;             /* interior points                              */
  double *x;                /* host, x|y|z with ghosts (reference layout)  */
  double *dxinv;            /* host, x|y|z with ghosts                     */
  /* state and scratch on the device (hm layout) */
  double *u, *hyp, *par, *source, *uC, *fluxC, *fluxI, *fL, *fR, *uL, *uR;
  hm_solver *dev;
  hm_config cfg;            /* the configuration dev was created with      */
  int u_shared;             /* u lies inside the batched ensemble state    */
  /* solver.inp */
  int n_iter, restart_iter, screen_op_iter, file_op_iter;
  double dt;
  char time_scheme[HY_STR], time_scheme_type[HY_STR], spatial_scheme_hyp[HY_STR];
  char interp_type[HY_STR], spatial_type_par[HY_STR], spatial_scheme_par[HY_STR];
  char SplitHyperbolicFlux[HY_STR], ConservationCheck[HY_STR];
  char ip_file_type[HY_STR], input_mode[HY_STR], output_mode[HY_STR];
  char op_file_format[HY_STR], op_overwrite[HY_STR], model[HY_STR];
  char op_fname_root[HY_STR];
  int file_index;
  /* plugin objects */
  NavierStokes3D *physics;
  WENOParameters *interp;
  int nBoundaryZones, all_periodic;
  /* conservation check (include/hypar.h:376-392) and call counts */
  double VolumeIntegral[5], VolumeIntegralInitial[5], TotalBoundaryIntegral[5];
  double StepBoundaryIntegral[30], ConservationError[5];
  int count_hyp, count_par, count_sou;
  /* slots (reference signatures) */
  int (*HyperbolicFunction)(double *, double *, void *, void *, double, int, hy_flux_fn, hy_upwind_fn);
  int (*ParabolicFunction)(double *, double *, void *, void *, double);
  int (*SourceFunction)(double *, double *, void *, void *, double);
  int (*FFunction)(double *, double *, int, void *, double);
  int (*Upwind)(double *, double *, double *, double *, double *, double *, int, void *, double);
  int (*UFunction)(double *, double *, int, void *, void *, double);
  int (*SFunction)(double *, double *, void *, void *, double);
  int (*InterpolateInterfacesHyp)(double *, double *, double *, double *, int, int, void *, void *, int);
  int (*SetInterpLimiterVar)(double *, double *, double *, int, void *, void *);
  int (*ApplyBoundaryConditions)(void *, void *, double *, double *, double);
  double (*ComputeCFL)(void *, void *, double, double);
  int (*PreStep)(double *, void *, void *, double);
} HyPar;

typedef struct {
  int nstages;
  double A[16], b[4], c[4];
} ExplicitRKParameters;

typedef struct TimeIntegration {
  HyPar *sims;
  MPIVariables *mpi;
  int nsims, iter, n_iter, fused, ssp; /* ssp: the fused step is SSPRK3 */
  double waqt, dt, max_cfl, norm;
  ExplicitRKParameters rk;
  double *U[4], *Udot[4];   /* per member, generic path: stage arrays */
  double **Ustore, **Kstore;/* nsims*4 device arrays                   */
  double **uprev;           /* per member copy for the step norm       */
  hm_ensemble *ens;         /* fused ensembles: all members per launch  */
  double *ens_u;            /* batched state; member u point into it    */
  int (*RHSFunction)(double *, double *, void *, void *, double);
  int (*TimeIntegrate)(void *);
} TimeIntegration;

/* ---- inputs (ReadInputs.c, WENOInitialize.c, InitializeBoundaries.c) ---- */
int hy_read_nsims(const char *dir, int *nsims);
int hy_read_solver(const char *dir, HyPar *sims, int nsims);
int hy_read_weno(const char *dir, WENOParameters *w);
int hy_read_boundaries(const char *dir, HyPar *sim, int member, int nsims);
int hy_read_initial(const char *dir, HyPar *sim, int member, int nsims, double **u_host_ghosted);
int hy_write_binary(const char *fname, HyPar *sim, const double *u_host_ghosted);
void hy_member_index(int i, int nsims, char *out);

/* ---- initialisation ---- */
int hy_initialize(HyPar *sim);                     /* sizes and host arrays      */
void hy_compute_metrics(HyPar *sim);               /* dxinv from the coordinates */
int hy_initialize_solvers(HyPar *sim, const char *dir);
int NavierStokes3DInitialize(void *s, void *m, const char *dir);
void hy_cleanup(HyPar *sim);

/* ---- slots ---- */
int NavierStokes3DFlux(double *f, double *u, int dir, void *s, double t);
int NavierStokes3DUpwindRoe(double *fI, double *fL, double *fR, double *uL, double *uR, double *u,
                            int dir, void *s, double t);
int NavierStokes3DUpwindRusanov(double *fI, double *fL, double *fR, double *uL, double *uR, double *u,
                                int dir, void *s, double t);
int NavierStokes3DModifiedSolution(double *uC, double *u, int d, void *s, void *m, double t);
int NavierStokes3DSource(double *source, double *u, void *s, void *m, double t);
int NavierStokes3DParabolicFunction(double *par, double *u, void *s, void *m, double t);
int NavierStokes3DPreStep(double *u, void *s, void *m, double t);
int Interp1PrimFifthOrderWENO(double *fI, double *fC, double *u, double *x, int upw, int dir, void *s,
                              void *m, int uflag);
int WENOFifthOrderCalculateWeights(double *fC, double *uC, double *x, int dir, void *s, void *m);
int HyperbolicFunction(double *hyp, double *u, void *s, void *m, double t, int LimFlag,
                       hy_flux_fn FFunction, hy_upwind_fn UpwindFunction);
int SourceFunction(double *source, double *u, void *s, void *m, double t);

/* ---- conservation check and error files ---- */
int VolumeIntegral(double *VolumeIntegral, double *u, void *s, void *m);
int BoundaryIntegral(void *s, void *m);
int CalculateConservationError(void *s, void *m);
int SimulationWriteErrors(const char *dir, HyPar *sims, int nsims, double solver_runtime,
                          double main_runtime);
int ApplyBoundaryConditions(void *s, void *m, double *x, double *xref, double waqt);

/* ---- time integration (TimeInitialize.c, TimeRK.c, TimeRHSFunctionExplicit.c) ---- */
int TimeExplicitRKInitialize(const char *cls, const char *type, ExplicitRKParameters *p);
int TimeInitialize(TimeIntegration *TS, HyPar *sims, MPIVariables *mpi, int nsims, int fused);
int EnsembleInitialize(TimeIntegration *TS);
int TimeRHSFunctionExplicit(double *rhs, double *u, void *s, void *m, double t);
int TimeRK(void *ts);
int TimeRKFused(void *ts);
int TimePreStep(TimeIntegration *TS);
int TimePostStep(TimeIntegration *TS);
void TimeCleanup(TimeIntegration *TS);

/* error reporting */
int hy_error(const char *fmt, ...);

#endif
