/* ns3d_oracle.h -- CPU restatement of HyPar's explicit NS3D WENO5 + RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the CPU reference that the HIP path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (libhypar_mi355x.so) never links or calls it.
 *
 * Arrays use the reference CPU layout: AoS with the 5 conserved variables
 * fastest, then x, y, z, with `ghosts` ghost layers on every side
 * (include/arrayfunctions.h:40-47 of the reference).
 */
#ifndef NS3D_ORACLE_H
#define NS3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_INTERP_COMPONENTS = 0, ORC_INTERP_CHARACTERISTIC = 1 };
enum { ORC_UPWIND_ROE = 0, ORC_UPWIND_RUSANOV = 1 };
/* WENO weight formulations (weno.inp: mapped / borges / yc; default JS) */
enum { ORC_WENO_JS = 0, ORC_WENO_MAPPED = 1, ORC_WENO_Z = 2, ORC_WENO_YC = 3, ORC_WENO_OPTIMAL = 4 };

typedef struct {
  int dim[3];          /* interior grid size                                  */
  int ghosts;          /* ghost layers (3 for WENO5)                          */
  double gamma;        /* ratio of heat capacities                            */
  double Re;           /* Reynolds number as the solver uses it (Re/Minf);
                          <= 0 means inviscid                                 */
  double Pr;           /* Prandtl number                                      */
  int interp;          /* ORC_INTERP_*                                        */
  int upwind;          /* ORC_UPWIND_*                                        */
  double eps;          /* WENO epsilon                                        */
  const double *dxinv; /* (dim0+2g)+(dim1+2g)+(dim2+2g) inverse spacings      */
  int weno_type;       /* ORC_WENO_*                                          */
} orc_params;

/* Number of doubles in one ghosted state array. */
long orc_size(const orc_params *p);

/* Periodic ghost fill of the six face slabs (BCPeriodicU). */
void orc_periodic_bc(const orc_params *p, double *u);

/* rhs = -hyp + par + source for state u (its ghosts are refilled). */
int orc_rhs(const orc_params *p, double *rhs, double *u);

/* Hyperbolic term alone (HyperbolicFunction), u's ghosts must be filled. */
int orc_hyperbolic(const orc_params *p, double *hyp, const double *u);

/* Viscous term alone (NavierStokes3DParabolicFunction), u's ghosts filled. */
int orc_parabolic(const orc_params *p, double *par, const double *u);

/* nsteps classical RK4 steps of size dt in place (TimeRK). */
int orc_rk4(const orc_params *p, double *u, double dt, int nsteps);
/* nsteps SSPRK3 steps ("rk ssprk3", TimeExplicitRKInitialize.c:69-79) */
int orc_ssprk3(const orc_params *p, double *u, double dt, int nsteps);
/* nsteps steps of the explicit tableau A (ns x ns, row-major) and b, the
 * stage loop of TimeRK.c:126-195; ns <= 4 (returns 1 otherwise). */
int orc_rk_tableau(const orc_params *p, double *u, double dt, int nsteps, int ns, const double *A,
                   const double *b);

/* Max CFL over the interior (NavierStokes3DComputeCFL). */
double orc_cfl(const orc_params *p, const double *u, double dt);

/* Volume integral of each variable over the interior (VolumeIntegral). */
void orc_volume_integral(const orc_params *p, const double *u, double *out);

/* Component-wise WENO5-JS weights of one 5-point stencil (macro
 * _WENOWeights_v_JS_), exposed for unit tests. */
void orc_weno_js_weights(double m3, double m2, double m1, double p1, double p2,
                         double eps, double w[3]);

/* Weights of one 5-point stencil for a weight formulation ORC_WENO_*
 * (macros _WENOWeights_v_JS_/_M_/_Z_/_YC_), exposed for unit tests. */
void orc_weno_weights(int type, double m3, double m2, double m1, double p1, double p2, double eps,
                      double w[3]);

/* Component-wise WENO5-JS interpolant at the interface right of f[2] from
 * the stencil f[0..4] (left-biased). */
double orc_weno5_interp(const double *f, double eps);

/* 4th-order first derivative of a 1-D array of n+2g points (ghosts
 * included), not divided by the grid spacing. */
void orc_first_derivative4_1d(int n, int g, const double *f, double *Df);

/* Roe average, left and right eigenvectors of one pair of states. */
void orc_roe_average(double *uavg, const double *uL, const double *uR, double gamma);
void orc_left_eigenvectors(const double *u, double *L, double gamma, int dir);
void orc_right_eigenvectors(const double *u, double *R, double gamma, int dir);

#ifdef __cplusplus
}
#endif
#endif
