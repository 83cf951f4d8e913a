/* solver.c -- HyPar-style solver object on top of libhypar_mi355x.so.
 *
 * Initialisation follows src/Simulation/Initialize.c, InitialSolution.c:74-113
 * (dxinv), InitializeSolvers.c:52-492 (slot wiring) and
 * src/PhysicalModels/NavierStokes3D/NavierStokes3DInitialize.c:116-527
 * (physics.inp, physics slots).  The hyperbolic operator composes its slots
 * exactly as ReconstructHyperbolic (HyperbolicFunction.c:167-222) does, or,
 * when every slot is the stock NavierStokes3D/WENO5 one, calls the fused
 * device operator that computes the same thing in one kernel per direction.
 * Time integration follows TimeRK.c:126-195 (any explicit tableau of
 * TimeExplicitRKInitialize.c) or, for "rk 44" and "rk ssprk3", can use the
 * fused steps of the device library.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "hypar_host.h"

#define NV 5

static int hm_check(hm_status st, const char *what) {
  if (st == HM_OK) return 0;
  return hy_error("%s: %s (status %d)", what, hm_last_error(), (int)st);
}

static size_t n5(const HyPar *s) { return (size_t)NV * (size_t)s->npoints; }

/* ------------------------------------------------------------------ */
/* initialisation                                                     */
/* ------------------------------------------------------------------ */
int hy_initialize(HyPar *s) {
  if (s->ndims != 3) return hy_error("ndims must be 3 for navierstokes3d");
  if (s->nvars != NV) return hy_error("nvars must be 5 for navierstokes3d");
  if (s->ghosts < 3) return hy_error("WENO5 needs at least 3 ghost points");
  for (int d = 0; d < 3; d++)
    if (s->dim_global[d] < 5) return hy_error("grid size %d along dimension %d is below 5", s->dim_global[d], d);
  s->npoints = (long)s->dim_global[0] * s->dim_global[1] * s->dim_global[2];
  long nx = 0;
  for (int d = 0; d < 3; d++) nx += s->dim_global[d] + 2 * s->ghosts;
  s->x = (double *)calloc(nx, sizeof(double));
  s->dxinv = (double *)calloc(nx, sizeof(double));
  if (!s->x || !s->dxinv) return hy_error("out of host memory");
  return 0;
}

/* dxinv = 2/(x[i+1]-x[i-1]) inside, edge copies in the ghosts
 * (InitialSolution.c:74-113); needs the ghost coordinates read first */
void hy_compute_metrics(HyPar *s) {
  const int g = s->ghosts;
  long off = 0;
  for (int d = 0; d < 3; d++) {
    const int n = s->dim_global[d];
    double *x = s->x + off, *dx = s->dxinv + off;
    for (int i = 0; i < n; i++) dx[i + g] = 2.0 / (x[i + 1 + g] - x[i - 1 + g]);
    for (int i = 0; i < g; i++) dx[i] = dx[g];
    for (int i = n + g; i < n + 2 * g; i++) dx[i] = dx[n + g - 1];
    off += n + 2 * g;
  }
}

static int alloc_device(HyPar *s) {
  double **bufs[] = {&s->u, &s->hyp, &s->par, &s->source, &s->uC, &s->fluxC,
                     &s->fluxI, &s->fL, &s->fR, &s->uL, &s->uR};
  for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++)
    if (hm_check(hm_malloc(bufs[i], n5(s)), "hm_malloc")) return 1;
  return 0;
}

void hy_cleanup(HyPar *s) {
  double *bufs[] = {s->u_shared ? NULL : s->u, s->hyp, s->par, s->source, s->uC, s->fluxC, s->fluxI,
                    s->fL, s->fR, s->uL, s->uR};
  for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++) hm_free(bufs[i]);
  if (s->dev) hm_destroy(s->dev);
  free(s->x);
  free(s->dxinv);
  free(s->physics);
  free(s->interp);
  memset(s, 0, sizeof(*s));
}

/* InitializeSolvers.c:52-492 for the WENO5 hyperbolic scheme */
int hy_initialize_solvers(HyPar *s, const char *dir) {
  if (strcmp(s->spatial_scheme_hyp, "weno5"))
    return hy_error("hyp_space_scheme %s is not supported (weno5 only)", s->spatial_scheme_hyp);
  if (strcmp(s->interp_type, "components") && strcmp(s->interp_type, "characteristic"))
    return hy_error("%s is not a supported interpolation type", s->interp_type);
  if (strcmp(s->SplitHyperbolicFlux, "no"))
    return hy_error("hyp_flux_split yes is not supported by the GPU path");
  s->interp = (WENOParameters *)calloc(1, sizeof(WENOParameters));
  if (hy_read_weno(dir, s->interp)) return 1;
  s->HyperbolicFunction = HyperbolicFunction;
  s->SourceFunction = SourceFunction;
  s->InterpolateInterfacesHyp = Interp1PrimFifthOrderWENO;
  s->SetInterpLimiterVar = WENOFifthOrderCalculateWeights;
  s->ApplyBoundaryConditions = ApplyBoundaryConditions;
  return 0;
}

/* ------------------------------------------------------------------ */
/* NavierStokes3D plugin                                              */
/* ------------------------------------------------------------------ */
static int read_physics(const char *dir, NavierStokes3D *p) {
  p->gamma = 1.4;
  p->Pr = 0.72;
  p->Re = -1;
  p->Minf = 1.0;
  p->grav[0] = p->grav[1] = p->grav[2] = 0.0;
  strcpy(p->upw_choice, "roe");
  char path[2 * HY_STR];
  snprintf(path, sizeof(path), "%s/physics.inp", dir);
  FILE *in = fopen(path, "r");
  if (!in) {
    printf("Warning: File \"physics.inp\" not found. Using default values.\n");
    return 0;
  }
  char word[HY_STR];
  if (fscanf(in, "%s", word) != 1 || strcmp(word, "begin")) {
    fclose(in);
    return hy_error("illegal format in physics.inp");
  }
  while (fscanf(in, "%s", word) == 1 && strcmp(word, "end")) {
    int ferr = 1;
    if (!strcmp(word, "gamma")) ferr = fscanf(in, "%lf", &p->gamma);
    else if (!strcmp(word, "upwinding")) ferr = fscanf(in, "%s", p->upw_choice);
    else if (!strcmp(word, "Pr")) ferr = fscanf(in, "%lf", &p->Pr);
    else if (!strcmp(word, "Re")) ferr = fscanf(in, "%lf", &p->Re);
    else if (!strcmp(word, "Minf")) ferr = fscanf(in, "%lf", &p->Minf);
    else if (!strcmp(word, "gravity"))
      ferr = fscanf(in, "%lf %lf %lf", &p->grav[0], &p->grav[1], &p->grav[2]) == 3;
    else {
      char useless[HY_STR];
      ferr = fscanf(in, "%s", useless);
    }
    if (ferr != 1) {
      fclose(in);
      return hy_error("read error in physics.inp near keyword %s", word);
    }
  }
  fclose(in);
  return 0;
}

int NavierStokes3DInitialize(void *sv, void *m, const char *dir) {
  (void)m;
  HyPar *s = (HyPar *)sv;
  if (strcmp(s->model, "navierstokes3d")) return hy_error("model %s is not supported", s->model);
  s->physics = (NavierStokes3D *)calloc(1, sizeof(NavierStokes3D));
  NavierStokes3D *p = s->physics;
  if (read_physics(dir, p)) return 1;
  if (p->grav[0] != 0 || p->grav[1] != 0 || p->grav[2] != 0)
    return hy_error("gravity is not supported by the GPU path");
  if (strcmp(p->upw_choice, "roe") && strcmp(p->upw_choice, "rusanov"))
    return hy_error("upwinding %s is not supported (roe, rusanov)", p->upw_choice);
  if (p->Re > 0 && strcmp(s->spatial_type_par, "nonconservative-2stage"))
    return hy_error("Parabolic term spatial discretization must be \"nonconservative-2stage\"");
  if (p->Re > 0 && strcmp(s->spatial_scheme_par, "4"))
    return hy_error("par_space_scheme %s is not supported (4)", s->spatial_scheme_par);
  hm_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  for (int d = 0; d < 3; d++) cfg.dim[d] = s->dim_global[d];
  cfg.gamma = p->gamma;
  cfg.Re = p->Re;
  cfg.Minf = p->Minf;
  cfg.Pr = p->Pr;
  cfg.interp = strcmp(s->interp_type, "characteristic") ? HM_INTERP_COMPONENTS : HM_INTERP_CHARACTERISTIC;
  cfg.upwind = strcmp(p->upw_choice, "roe") ? HM_UPWIND_RUSANOV : HM_UPWIND_ROE;
  cfg.weno_eps = s->interp->eps;
  /* the reference's precedence (WENOFifthOrderCalculateWeights.c:55-57);
   * no_limiting keeps the initial optimal weights whatever the formulation
   * (InitializeSolvers.c:219, HyperbolicFunction.c:62) */
  cfg.weno_type = s->interp->no_limiting ? HM_WENO_OPTIMAL
                  : (s->interp->yc ? HM_WENO_YC
                                   : (s->interp->borges ? HM_WENO_Z
                                                        : (s->interp->mapped ? HM_WENO_MAPPED : HM_WENO_JS)));
  cfg.dxinv = s->dxinv;
  if (hm_check(hm_create(&cfg, &s->dev), "hm_create")) return 1;
  s->cfg = cfg;
  if (alloc_device(s)) return 1;
  /* the library's slots find the device solver and the state of this
   * solver object through the binding */
  if (hm_check(hm_bind(s, s->dev), "hm_bind") ||
      hm_check(hm_bind_state(s, s->u, HM_SLOT_ARRAYS_DEVICE), "hm_bind_state"))
    return 1;
  /* physics slots (NavierStokes3DInitialize.c:388-506) */
  s->PreStep = NavierStokes3DPreStep;
  s->ComputeCFL = hm_slot_ComputeCFL;
  s->FFunction = NavierStokes3DFlux;
  s->SFunction = NavierStokes3DSource;
  s->UFunction = NavierStokes3DModifiedSolution;
  s->Upwind = strcmp(p->upw_choice, "roe") ? NavierStokes3DUpwindRusanov : NavierStokes3DUpwindRoe;
  s->ParabolicFunction = NavierStokes3DParabolicFunction;
  return 0;
}

int NavierStokes3DFlux(double *f, double *u, int dir, void *s, double t) {
  (void)t;
  return hm_check(hm_flux(((HyPar *)s)->dev, f, u, dir), "NavierStokes3DFlux");
}

static int upwind_slot(const char *want, double *fI, double *fL, double *fR, double *uL, double *uR,
                       double *u, int dir, HyPar *s) {
  if (strcmp(s->physics->upw_choice, want))
    return hy_error("upwinding slot %s called on a solver set up for %s", want, s->physics->upw_choice);
  return hm_check(hm_upwind(s->dev, fI, fL, fR, uL, uR, u, dir), "Upwind");
}

int NavierStokes3DUpwindRoe(double *fI, double *fL, double *fR, double *uL, double *uR, double *u,
                            int dir, void *s, double t) {
  (void)t;
  return upwind_slot("roe", fI, fL, fR, uL, uR, u, dir, (HyPar *)s);
}

int NavierStokes3DUpwindRusanov(double *fI, double *fL, double *fR, double *uL, double *uR, double *u,
                                int dir, void *s, double t) {
  (void)t;
  return upwind_slot("rusanov", fI, fL, fR, uL, uR, u, dir, (HyPar *)s);
}

int NavierStokes3DModifiedSolution(double *uC, double *u, int d, void *s, void *m, double t) {
  (void)d; (void)m; (void)t;
  return hm_check(hm_modified_solution(((HyPar *)s)->dev, uC, u), "NavierStokes3DModifiedSolution");
}

/* no gravity: the source is zero (NavierStokes3DSource.c:50-51) */
int NavierStokes3DSource(double *source, double *u, void *s, void *m, double t) {
  (void)u; (void)m; (void)t;
  HyPar *h = (HyPar *)s;
  return hm_check(hm_fill(h->dev, source, 0.0, n5(h)), "NavierStokes3DSource");
}

/* the library's slot, plus the call count the reference keeps
 * (NavierStokes3DParabolicFunction.c:71-73) */
int NavierStokes3DParabolicFunction(double *par, double *u, void *s, void *m, double t) {
  HyPar *h = (HyPar *)s;
  if (h->physics->Re > 0) h->count_par++;
  return hm_check(hm_slot_ParabolicFunction(par, u, s, m, t), "ParabolicFunction");
}

/* The reference's PreStep builds a split-flux Jacobian that only
 * hyp_flux_split uses (NavierStokes3DPreStep.c:32-102); with the flux not
 * split it leaves the solution untouched, so there is nothing to do. */
int NavierStokes3DPreStep(double *u, void *s, void *m, double t) {
  (void)u; (void)s; (void)m; (void)t;
  return 0;
}

/* ------------------------------------------------------------------ */
/* numerics slots                                                     */
/* ------------------------------------------------------------------ */
int Interp1PrimFifthOrderWENO(double *fI, double *fC, double *u, double *x, int upw, int dir, void *s,
                              void *m, int uflag) {
  (void)x; (void)m;
  return hm_check(hm_interpolate(((HyPar *)s)->dev, fI, fC, u, upw, dir, uflag), "InterpolateInterfacesHyp");
}

/* The device interpolation evaluates the WENO weights of its stencil from
 * the same inputs the reference's weight pass would use, so there is
 * nothing to precompute. */
int WENOFifthOrderCalculateWeights(double *fC, double *uC, double *x, int dir, void *s, void *m) {
  (void)fC; (void)uC; (void)x; (void)dir; (void)s; (void)m;
  return 0;
}

/* periodic faces need no work: hm-layout arrays hold no ghost points and
 * the kernels wrap their stencils (BCPeriodic.c:36-57) */
int ApplyBoundaryConditions(void *s, void *m, double *x, double *xref, double waqt) {
  (void)m; (void)x; (void)xref; (void)waqt;
  HyPar *h = (HyPar *)s;
  return h->all_periodic ? 0 : hy_error("non-periodic boundaries are not supported");
}

int SourceFunction(double *source, double *u, void *s, void *m, double t) {
  HyPar *h = (HyPar *)s;
  if (hm_check(hm_fill(h->dev, source, 0.0, n5(h)), "SourceFunction")) return 1;
  if (h->SFunction) {
    h->count_sou++;
    /* the physics source accumulates into a scratch array */
    if (h->SFunction(h->uC, u, s, m, t)) return 1;
    if (hm_check(hm_axpy(h->dev, h->uC, 1.0, source, n5(h)), "SourceFunction")) return 1;
  }
  return 0;
}

static int is_stock(const HyPar *s, hy_flux_fn F, hy_upwind_fn U) {
  return F == NavierStokes3DFlux &&
         (U == NavierStokes3DUpwindRoe || U == NavierStokes3DUpwindRusanov) && U == s->Upwind &&
         s->UFunction == NavierStokes3DModifiedSolution &&
         s->InterpolateInterfacesHyp == Interp1PrimFifthOrderWENO;
}

/* HyperbolicFunction.c:31-126 with ReconstructHyperbolic :167-222 */
int HyperbolicFunction(double *hyp, double *u, void *sv, void *m, double t, int LimFlag,
                       hy_flux_fn FFunction, hy_upwind_fn UpwindFunction) {
  HyPar *s = (HyPar *)sv;
  if (!FFunction) return hm_check(hm_fill(s->dev, hyp, 0.0, n5(s)), "HyperbolicFunction");
  s->count_hyp++;  /* HyperbolicFunction.c:67 */
  if (is_stock(s, FFunction, UpwindFunction))
    return hm_check(hm_slot_HyperbolicFunction(hyp, u, sv, m, t, LimFlag, FFunction, UpwindFunction),
                    "HyperbolicFunction");
  for (int d = 0; d < 3; d++) {
    if (FFunction(s->fluxC, u, d, s, t)) return 1;
    if (LimFlag && s->SetInterpLimiterVar && s->SetInterpLimiterVar(s->fluxC, u, s->x, d, s, m)) return 1;
    double *uC = u;
    if (s->UFunction) {
      if (s->UFunction(s->uC, u, d, s, m, t)) return 1;
      uC = s->uC;
    }
    if (s->InterpolateInterfacesHyp(s->uL, uC, u, s->x, 1, d, s, m, 1)) return 1;
    if (s->InterpolateInterfacesHyp(s->uR, uC, u, s->x, -1, d, s, m, 1)) return 1;
    if (s->InterpolateInterfacesHyp(s->fL, s->fluxC, u, s->x, 1, d, s, m, 0)) return 1;
    if (s->InterpolateInterfacesHyp(s->fR, s->fluxC, u, s->x, -1, d, s, m, 0)) return 1;
    if (UpwindFunction) {
      if (UpwindFunction(s->fluxI, s->fL, s->fR, s->uL, s->uR, u, d, s, t)) return 1;
    } else {
      /* DefaultUpwinding: arithmetic mean */
      if (hm_check(hm_fill(s->dev, s->fluxI, 0.0, n5(s)), "DefaultUpwinding") ||
          hm_check(hm_axpy(s->dev, s->fL, 0.5, s->fluxI, n5(s)), "DefaultUpwinding") ||
          hm_check(hm_axpy(s->dev, s->fR, 0.5, s->fluxI, n5(s)), "DefaultUpwinding"))
        return 1;
    }
    if (hm_check(hm_flux_difference(s->dev, hyp, s->fluxI, d, d > 0), "flux difference")) return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* time integration                                                   */
/* ------------------------------------------------------------------ */
int TimeExplicitRKInitialize(const char *cls, const char *type, ExplicitRKParameters *p) {
  memset(p, 0, sizeof(*p));
  if (strcmp(cls, "rk")) return hy_error("time_scheme %s is not supported (rk)", cls);
  if (!strcmp(type, "1fe")) {
    p->nstages = 1; p->b[0] = 1.0;
  } else if (!strcmp(type, "22")) {
    p->nstages = 2; p->A[2] = 1.0; p->c[1] = 1.0; p->b[0] = p->b[1] = 0.5;
  } else if (!strcmp(type, "33")) {
    p->nstages = 3; p->A[3] = 2.0 / 3.0; p->A[6] = 2.0 / 3.0 - 1.0 / 4.0; p->A[7] = 1.0 / 4.0;
    p->c[1] = 2.0 / 3.0; p->c[2] = 2.0 / 3.0; p->b[0] = 1.0 / 4.0; p->b[1] = -1.0 / 4.0; p->b[2] = 1.0;
  } else if (!strcmp(type, "44")) {
    p->nstages = 4; p->A[4] = 0.5; p->A[9] = 0.5; p->A[14] = 1.0;
    p->c[1] = p->c[2] = 0.5; p->c[3] = 1.0;
    p->b[0] = 1.0 / 6.0; p->b[1] = 1.0 / 3.0; p->b[2] = 1.0 / 3.0; p->b[3] = 1.0 / 6.0;
  } else if (!strcmp(type, "ssprk3") || !strcmp(type, "tvdrk3")) {
    p->nstages = 3; p->A[3] = 1.0; p->A[6] = 0.25; p->A[7] = 0.25; p->c[1] = 1.0; p->c[2] = 0.5;
    p->b[0] = p->b[1] = 1.0 / 6.0; p->b[2] = 2.0 / 3.0;
  } else {
    return hy_error("%s is not a supported multi-stage time integration scheme of class %s", type, cls);
  }
  return 0;
}

/* TimeRHSFunctionExplicit.c:30-98 */
int TimeRHSFunctionExplicit(double *rhs, double *u, void *sv, void *m, double t) {
  HyPar *s = (HyPar *)sv;
  if (s->ApplyBoundaryConditions(s, m, u, NULL, t)) return 1;
  if (s->HyperbolicFunction(s->hyp, u, s, m, t, 1, s->FFunction, s->Upwind)) return 1;
  if (s->ParabolicFunction) {
    if (s->ParabolicFunction(s->par, u, s, m, t)) return 1;
  } else if (hm_check(hm_fill(s->dev, s->par, 0.0, n5(s)), "par")) {
    return 1;
  }
  if (s->SourceFunction(s->source, u, s, m, t)) return 1;
  if (hm_check(hm_fill(s->dev, rhs, 0.0, n5(s)), "rhs") ||
      hm_check(hm_axpy(s->dev, s->hyp, -1.0, rhs, n5(s)), "rhs") ||
      hm_check(hm_axpy(s->dev, s->par, 1.0, rhs, n5(s)), "rhs") ||
      hm_check(hm_axpy(s->dev, s->source, 1.0, rhs, n5(s)), "rhs"))
    return 1;
  return 0;
}

int TimeInitialize(TimeIntegration *TS, HyPar *sims, MPIVariables *mpi, int nsims, int fused) {
  memset(TS, 0, sizeof(*TS));
  TS->sims = sims;
  TS->mpi = mpi;
  TS->nsims = nsims;
  TS->dt = sims[0].dt;            /* the ensemble shares sim[0]'s dt (TimeInitialize.c:49-52) */
  TS->n_iter = sims[0].n_iter;
  TS->waqt = sims[0].restart_iter * sims[0].dt;
  if (TimeExplicitRKInitialize(sims[0].time_scheme, sims[0].time_scheme_type, &TS->rk)) return 1;
  const char *type = sims[0].time_scheme_type;
  TS->ssp = !strcmp(type, "ssprk3") || !strcmp(type, "tvdrk3");
  TS->fused = fused && ((TS->rk.nstages == 4 && !strcmp(type, "44")) || TS->ssp);
  int stock = 1;
  for (int n = 0; n < nsims; n++)
    if (!is_stock(&sims[n], sims[n].FFunction, sims[n].Upwind)) stock = 0;
  TS->fused = TS->fused && stock;
  /* Fused steps go through the library's TimeIntegrate slot.  Without
   * --slots, the other tableaux take TimeRK with the library's fused
   * right-hand side; with --slots (fused == 0), TimeRHSFunctionExplicit
   * composes the slots. */
  TS->RHSFunction = (fused && stock) ? hm_slot_RHSFunction : TimeRHSFunctionExplicit;
  TS->TimeIntegrate = TS->fused ? TimeRKFused : TimeRK;
  if (TS->fused)
    for (int n = 0; n < nsims; n++)
      if (hm_check(hm_bind_time(TS, &sims[n], TS->dt, TS->ssp ? HM_TIME_SSPRK3 : HM_TIME_RK4), "hm_bind_time"))
        return 1;
  TS->Ustore = (double **)calloc((size_t)4 * nsims, sizeof(double *));
  TS->Kstore = (double **)calloc((size_t)4 * nsims, sizeof(double *));
  TS->uprev = (double **)calloc((size_t)nsims, sizeof(double *));
  for (int n = 0; n < nsims; n++) {
    if (hm_check(hm_malloc(&TS->uprev[n], n5(&sims[n])), "hm_malloc")) return 1;
    if (TS->fused) continue;
    for (int st = 0; st < TS->rk.nstages; st++) {
      if (hm_check(hm_malloc(&TS->Ustore[4 * n + st], n5(&sims[n])), "hm_malloc")) return 1;
      if (hm_check(hm_malloc(&TS->Kstore[4 * n + st], n5(&sims[n])), "hm_malloc")) return 1;
    }
  }
  /* the batched ensemble advances RK4 members; SSPRK3 members step one
   * after another */
  if (TS->fused && nsims > 1 && !TS->ssp) return EnsembleInitialize(TS);
  return 0;
}

/* Fused ensembles: the members (which share the scheme, TimeInitialize.c:49-52
 * and ReadInputs.c:189-207) advance together, every direction pass of an RK
 * stage one launch over all members.  Each member's state moves into the
 * batched state and its u points there, so the per-member work (CFL, norm,
 * output, conservation) reads it as before. */
int EnsembleInitialize(TimeIntegration *TS) {
  const int nsims = TS->nsims;
  hm_config *cfgs = (hm_config *)calloc((size_t)nsims, sizeof(hm_config));
  double *dts = (double *)calloc((size_t)nsims, sizeof(double));
  for (int n = 0; n < nsims; n++) {
    cfgs[n] = TS->sims[n].cfg;
    dts[n] = TS->dt;
  }
  const hm_status st = hm_ensemble_create(cfgs, nsims, &TS->ens);
  free(cfgs);
  if (st != HM_OK) {
    printf("Note: members advance one after another (%s)\n", hm_last_error());
    TS->ens = NULL;
    free(dts);
    return 0;
  }
  if (hm_check(hm_malloc(&TS->ens_u, (size_t)NV * hm_ensemble_ncell(TS->ens)), "hm_malloc") ||
      hm_check(hm_ensemble_set_dt(TS->ens, dts), "hm_ensemble_set_dt")) {
    free(dts);
    return 1;
  }
  free(dts);
  for (int n = 0; n < nsims; n++) {
    HyPar *s = &TS->sims[n];
    size_t off, nc;
    if (hm_check(hm_ensemble_member(TS->ens, n, &off, &nc), "hm_ensemble_member")) return 1;
    double *dst = TS->ens_u + (size_t)NV * off;
    if (hm_check(hm_copy_device(s->dev, dst, s->u, n5(s)), "ensemble state") ||
        hm_check(hm_synchronize(s->dev), "ensemble state"))
      return 1;
    hm_free(s->u);
    s->u = dst;
    s->u_shared = 1;
    if (hm_check(hm_bind_state(s, s->u, HM_SLOT_ARRAYS_DEVICE), "hm_bind_state")) return 1;
  }
  return 0;
}

void TimeCleanup(TimeIntegration *TS) {
  hm_unbind(TS);
  if (TS->ens) hm_ensemble_destroy(TS->ens);
  hm_free(TS->ens_u);
  for (int i = 0; i < 4 * TS->nsims; i++) {
    if (TS->Ustore) hm_free(TS->Ustore[i]);
    if (TS->Kstore) hm_free(TS->Kstore[i]);
  }
  for (int n = 0; n < TS->nsims; n++)
    if (TS->uprev) hm_free(TS->uprev[n]);
  free(TS->Ustore);
  free(TS->Kstore);
  free(TS->uprev);
}

/* The call counts of `nrhs` right-hand sides evaluated inside the library:
 * one hyperbolic, one parabolic (viscous only) and one source term each
 * (HyperbolicFunction.c:67, NavierStokes3DParabolicFunction.c:71-73,
 * SourceFunction.c:46-49). */
static void count_rhs(HyPar *s, int nrhs) {
  s->count_hyp += nrhs;
  if (s->physics->Re > 0) s->count_par += nrhs;
  s->count_sou += nrhs;
}

/* TimeRK.c:126-195: stage values, right-hand sides, step completion */
int TimeRK(void *ts) {
  TimeIntegration *TS = (TimeIntegration *)ts;
  const ExplicitRKParameters *p = &TS->rk;
  for (int n = 0; n < TS->nsims; n++) {
    HyPar *s = &TS->sims[n];
    double **U = &TS->Ustore[4 * n], **K = &TS->Kstore[4 * n];
    for (int stage = 0; stage < p->nstages; stage++) {
      double stagetime = TS->waqt + p->c[stage] * TS->dt;
      if (hm_check(hm_copy_device(s->dev, U[stage], s->u, n5(s)), "TimeRK copy")) return 1;
      for (int i = 0; i < stage; i++)
        if (hm_check(hm_axpy(s->dev, K[i], TS->dt * p->A[stage * p->nstages + i], U[stage], n5(s)),
                     "TimeRK axpy"))
          return 1;
      if (TS->RHSFunction(K[stage], U[stage], s, &TS->mpi[n], stagetime)) return 1;
      if (TS->RHSFunction == hm_slot_RHSFunction) count_rhs(s, 1);
    }
    for (int stage = 0; stage < p->nstages; stage++)
      if (hm_check(hm_axpy(s->dev, K[stage], TS->dt * p->b[stage], s->u, n5(s)), "TimeRK update"))
        return 1;
  }
  return 0;
}

/* one fused RK4 or SSPRK3 step per member (same operation order as
 * TimeRK); for RK4 ensembles one batched step of all members */
int TimeRKFused(void *ts) {
  TimeIntegration *TS = (TimeIntegration *)ts;
  if (TS->ens) {
    if (hm_check(hm_ensemble_rk4(TS->ens, TS->ens_u, 1), "hm_ensemble_rk4")) return 1;
  } else if (hm_check(hm_slot_TimeIntegrate(ts), "TimeIntegrate")) {
    return 1;
  }
  for (int n = 0; n < TS->nsims; n++) count_rhs(&TS->sims[n], TS->rk.nstages);
  return 0;
}

/* TimePreStep.c:22-129 */
int TimePreStep(TimeIntegration *TS) {
  for (int n = 0; n < TS->nsims; n++) {
    HyPar *s = &TS->sims[n];
    if (s->ApplyBoundaryConditions(s, &TS->mpi[n], s->u, NULL, TS->waqt)) return 1;
    if ((TS->iter + 1) % s->screen_op_iter == 0) {
      if (hm_check(hm_copy_device(s->dev, TS->uprev[n], s->u, n5(s)), "TimePreStep copy")) return 1;
      if (s->ComputeCFL) {
        double c = s->ComputeCFL(s, &TS->mpi[n], TS->dt, TS->waqt);
        if (c < 0) return 1;
        if (n == 0 || c > TS->max_cfl) TS->max_cfl = c;
      } else {
        TS->max_cfl = -1;
      }
    }
    if (s->PreStep && s->PreStep(s->u, s, &TS->mpi[n], TS->waqt)) return 1;
  }
  return 0;
}

/* TimePostStep.c:28-63: t += dt and the RMS norm of u_{n+1} - u_n */
int TimePostStep(TimeIntegration *TS) {
  TS->waqt += TS->dt;
  if ((TS->iter + 1) % TS->sims[0].screen_op_iter == 0) {
    double sum = 0.0, npts = 0.0;
    for (int n = 0; n < TS->nsims; n++) {
      HyPar *s = &TS->sims[n];
      double norm;
      if (hm_check(hm_step_norm(s->dev, s->u, TS->uprev[n], &norm), "TimePostStep")) return 1;
      sum += norm * norm * (double)s->npoints;
      npts += (double)s->npoints;
    }
    TS->norm = sqrt(sum / npts);
  }
  /* TimePostStep.c:78-93 */
  for (int n = 0; n < TS->nsims; n++) {
    HyPar *s = &TS->sims[n];
    if (!strcmp(s->ConservationCheck, "yes")) {
      if (VolumeIntegral(s->VolumeIntegral, s->u, s, &TS->mpi[n])) return 1;
      if (BoundaryIntegral(s, &TS->mpi[n])) return 1;
      if (CalculateConservationError(s, &TS->mpi[n])) return 1;
    }
  }
  return 0;
}

/* ---- conservation check -------------------------------------------------- */

/* VolumeIntegral.c:14-50 (the sum over cells runs on the device) */
int VolumeIntegral(double *VolumeIntegral, double *u, void *s, void *m) {
  (void)m;
  HyPar *h = (HyPar *)s;
  return hm_check(hm_volume_integral(h->dev, u, VolumeIntegral), "VolumeIntegral");
}

/* BoundaryIntegral.c:17-58: adds the step's boundary flux integral, times
 * the face areas, to the running total.  The step integral is the dt b_s
 * weighted sum of the stage integrals (TimeRK.c:182-195), each of which
 * HyperbolicFunction.c:103-106 forms as -F at the low face plus F at the
 * high face of every line.  Every boundary of this path is periodic: both
 * faces of a line are the same interface and carry the same flux value, so
 * each stage integral is exactly zero and StepBoundaryIntegral stays zero. */
int BoundaryIntegral(void *s, void *m) {
  (void)m;
  HyPar *h = (HyPar *)s;
  const int nv = h->nvars;
  double local[5] = {0, 0, 0, 0, 0};
  for (int d = 0; d < h->ndims; d++) {
    for (int v = 0; v < nv; v++) {
      double dS = 1.0;
      for (int k = 0; k < h->ndims; k++) {
        if (k == d) continue;
        /* uniform grid assumed: dxinv at the middle of the line */
        int off = 0;
        for (int q = 0; q < k; q++) off += h->dim_local[q] + 2 * h->ghosts;
        dS *= 1.0 / h->dxinv[off + h->ghosts + h->dim_local[k] / 2];
      }
      local[v] += h->StepBoundaryIntegral[(2 * d + 0) * nv + v] * dS;
      local[v] += h->StepBoundaryIntegral[(2 * d + 1) * nv + v] * dS;
    }
  }
  for (int v = 0; v < nv; v++) h->TotalBoundaryIntegral[v] += local[v];
  return 0;
}

/* CalculateConservationError.c:16-38 */
int CalculateConservationError(void *s, void *m) {
  (void)m;
  HyPar *h = (HyPar *)s;
  for (int v = 0; v < h->nvars; v++) {
    const double a = fabs(h->VolumeIntegralInitial[v]);
    const double base = a > 1.0 ? a : 1.0;
    const double e = h->VolumeIntegral[v] + h->TotalBoundaryIntegral[v] - h->VolumeIntegralInitial[v];
    h->ConservationError[v] = sqrt(e * e) / base;
  }
  return 0;
}

/* SimulationWriteErrors.c:17-134: errors, conservation errors and call
 * counts per member (no exact solution is read, so the solution errors are
 * -1 as CalculateError.c:61 leaves them) */
int SimulationWriteErrors(const char *dir, HyPar *sims, int nsims, double solver_runtime,
                          double main_runtime) {
  if (nsims > 1) printf("\n");
  for (int n = 0; n < nsims; n++) {
    HyPar *h = &sims[n];
    char idx[HY_STR] = "", name[3 * HY_STR];
    if (nsims > 1) {
      char tmp[HY_STR];
      hy_member_index(n, nsims, tmp);
      snprintf(idx, sizeof idx, "_%s", tmp);
    }
    snprintf(name, sizeof name, "%s/errors%s.dat", dir, idx);
    FILE *out = fopen(name, "w");
    if (!out) return hy_error("cannot write %s", name);
    for (int d = 0; d < h->ndims; d++) fprintf(out, "%4d ", h->dim_global[d]);
    for (int d = 0; d < h->ndims; d++) fprintf(out, "%4d ", 1);
    fprintf(out, "%1.16E  ", h->dt);
    fprintf(out, "%1.16E %1.16E %1.16E   ", -1.0, -1.0, -1.0);
    fprintf(out, "%1.16E %1.16E\n", solver_runtime, main_runtime);
    fclose(out);
    snprintf(name, sizeof name, "%s/conservation%s.dat", dir, idx);
    if (!(out = fopen(name, "w"))) return hy_error("cannot write %s", name);
    for (int d = 0; d < h->ndims; d++) fprintf(out, "%4d ", h->dim_global[d]);
    for (int d = 0; d < h->ndims; d++) fprintf(out, "%4d ", 1);
    fprintf(out, "%1.16E  ", h->dt);
    for (int v = 0; v < h->nvars; v++) fprintf(out, "%1.16E ", h->ConservationError[v]);
    fprintf(out, "\n");
    fclose(out);
    snprintf(name, sizeof name, "%s/function_counts%s.dat", dir, idx);
    if (!(out = fopen(name, "w"))) return hy_error("cannot write %s", name);
    fprintf(out, "%d\n%d\n%d\n%d\n", h->n_iter, h->count_hyp, h->count_par, h->count_sou);
    fclose(out);
    if (!strcmp(h->ConservationCheck, "yes")) {
      printf("Conservation Errors:\n");
      for (int v = 0; v < h->nvars; v++) printf("\t%1.16E\n", h->ConservationError[v]);
    }
  }
  return 0;
}
