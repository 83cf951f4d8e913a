/* ns3d_oracle.c -- CPU restatement of HyPar's explicit NS3D WENO5 + RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY (see ns3d_oracle.h).  Every expression keeps the
 * reference's operation order so that, built with -O3 -ffp-contract=off on
 * x86-64, results are bit-identical to the reference binary; the parity
 * tests pin that against golden vectors written by the reference itself.
 *
 * Reference files followed (paths relative to the reference root):
 *   TimeRK ........................ src/TimeIntegration/TimeRK.c:126-195,
 *                                   tableau TimeExplicitRKInitialize.c:58-68
 *   RHS ........................... src/TimeIntegration/TimeRHSFunctionExplicit.c:46-93
 *   periodic ghost fill ........... src/BoundaryConditions/BCPeriodic.c:36-57
 *   hyperbolic term ............... src/HyParFunctions/HyperbolicFunction.c:64-123,167-222
 *   flux, Roe average, eigenvectors include/physicalmodels/navierstokes3d.h:98-470
 *   modified solution ............. src/PhysicalModels/NavierStokes3D/NavierStokes3DModifiedSolution.c:31-70
 *   WENO5-JS weights .............. src/InterpolationFunctions/WENOFifthOrderCalculateWeights.c:111-236,760-896
 *                                   include/interpolation.h:265-285
 *   WENO5 interpolation ........... src/InterpolationFunctions/Interp1PrimFifthOrderWENO.c:74-168,
 *                                   Interp1PrimFifthOrderWENOChar.c:86-198
 *   Roe / Rusanov upwinding ....... src/PhysicalModels/NavierStokes3D/NavierStokes3DUpwind.c:40-124,349-420,
 *                                   include/matmult_native.h:82-165
 *   viscous term .................. src/PhysicalModels/NavierStokes3D/NavierStokes3DParabolicFunction.c:50-325
 *   4th-order first derivative .... src/FirstDerivative/FirstDerivativeFourthOrder.c:36-129
 *   CFL ........................... src/PhysicalModels/NavierStokes3D/NavierStokes3DComputeCFL.c:16-55
 */
#include "ns3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NV 5

/* ------------------------------------------------------------------ */
/* indexing                                                           */
/* ------------------------------------------------------------------ */
typedef struct {
  int n[3], g, ng[3];
  long stride[3];
} grid_t;

static grid_t mkgrid(const orc_params *p) {
  grid_t G;
  G.g = p->ghosts;
  for (int d = 0; d < 3; d++) { G.n[d] = p->dim[d]; G.ng[d] = p->dim[d] + 2 * p->ghosts; }
  G.stride[0] = 1;
  G.stride[1] = G.ng[0];
  G.stride[2] = (long)G.ng[0] * G.ng[1];
  return G;
}

/* point index of (i,j,k), each in [-g, n+g) */
static inline long pidx(const grid_t *G, int i, int j, int k) {
  return ((long)(k + G->g) * G->ng[1] + (j + G->g)) * G->ng[0] + (i + G->g);
}

long orc_size(const orc_params *p) {
  grid_t G = mkgrid(p);
  return (long)G.ng[0] * G.ng[1] * G.ng[2] * NV;
}

static const double *dxinv_dir(const orc_params *p, int d) {
  int off = 0;
  for (int e = 0; e < d; e++) off += p->dim[e] + 2 * p->ghosts;
  return p->dxinv + off + p->ghosts; /* indexable by interior index, ghosts at <0 */
}

/* ------------------------------------------------------------------ */
/* physics macros (navierstokes3d.h:98-470)                           */
/* ------------------------------------------------------------------ */
static inline void flowvar(const double *u, double gamma, double *rho, double *vx,
                           double *vy, double *vz, double *e, double *P) {
  double vsq;
  *rho = u[0];
  *vx = (*rho == 0) ? 0 : u[1] / *rho;
  *vy = (*rho == 0) ? 0 : u[2] / *rho;
  *vz = (*rho == 0) ? 0 : u[3] / *rho;
  *e = u[4];
  vsq = (*vx) * (*vx) + (*vy) * (*vy) + (*vz) * (*vz);
  *P = (*e - 0.5 * (*rho) * vsq) * (gamma - 1.0);
}

static inline void setflux(double *f, double rho, double vx, double vy, double vz,
                           double e, double P, int dir) {
  if (dir == 0) {
    f[0] = rho * vx;
    f[1] = rho * vx * vx + P;
    f[2] = rho * vx * vy;
    f[3] = rho * vx * vz;
    f[4] = (e + P) * vx;
  } else if (dir == 1) {
    f[0] = rho * vy;
    f[1] = rho * vy * vx;
    f[2] = rho * vy * vy + P;
    f[3] = rho * vy * vz;
    f[4] = (e + P) * vy;
  } else {
    f[0] = rho * vz;
    f[1] = rho * vz * vx;
    f[2] = rho * vz * vy;
    f[3] = rho * vz * vz + P;
    f[4] = (e + P) * vz;
  }
}

void orc_roe_average(double *uavg, const double *uL, const double *uR, double gamma) {
  double rho, vx, vy, vz, e, P, H;
  double rhoL, vxL, vyL, vzL, eL, PL, HL, cLsq;
  double rhoR, vxR, vyR, vzR, eR, PR, HR, cRsq;
  flowvar(uL, gamma, &rhoL, &vxL, &vyL, &vzL, &eL, &PL);
  cLsq = gamma * PL / rhoL;
  HL = 0.5 * (vxL * vxL + vyL * vyL + vzL * vzL) + cLsq / (gamma - 1.0);
  flowvar(uR, gamma, &rhoR, &vxR, &vyR, &vzR, &eR, &PR);
  cRsq = gamma * PR / rhoR;
  HR = 0.5 * (vxR * vxR + vyR * vyR + vzR * vzR) + cRsq / (gamma - 1.0);
  double tL = sqrt(rhoL);
  double tR = sqrt(rhoR);
  rho = tL * tR;
  vx = (tL * vxL + tR * vxR) / (tL + tR);
  vy = (tL * vyL + tR * vyR) / (tL + tR);
  vz = (tL * vzL + tR * vzR) / (tL + tR);
  H = (tL * HL + tR * HR) / (tL + tR);
  P = (H - 0.5 * (vx * vx + vy * vy + vz * vz)) * (rho * (gamma - 1.0)) / gamma;
  e = P / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz);
  uavg[0] = rho;
  uavg[1] = rho * vx;
  uavg[2] = rho * vy;
  uavg[3] = rho * vz;
  uavg[4] = e;
}

static void eigenvalues(const double *u, double *D, double gamma, int dir) {
  double rho, vx, vy, vz, e, P, c;
  flowvar(u, gamma, &rho, &vx, &vy, &vz, &e, &P);
  c = sqrt(gamma * P / rho);
  for (int k = 0; k < 25; k++) D[k] = 0;
  if (dir == 0) {
    D[0] = vx; D[6] = vx - c; D[12] = vx; D[18] = vx; D[24] = vx + c;
  } else if (dir == 1) {
    D[0] = vy; D[6] = vy; D[12] = vy - c; D[18] = vy; D[24] = vy + c;
  } else {
    D[0] = vz; D[6] = vz; D[12] = vz; D[18] = vz - c; D[24] = vz + c;
  }
}

void orc_left_eigenvectors(const double *u, double *L, double ga, int dir) {
  double gm1 = ga - 1.0;
  double rho, vx, vy, vz, e, P, a, ek;
  flowvar(u, ga, &rho, &vx, &vy, &vz, &e, &P);
  ek = 0.5 * (vx * vx + vy * vy + vz * vz);
  a = sqrt(ga * P / rho);
  if (dir == 0) {
    L[1 * NV + 0] = (gm1 * ek + a * vx) / (2 * a * a);
    L[1 * NV + 1] = ((-gm1) * vx - a) / (2 * a * a);
    L[1 * NV + 2] = ((-gm1) * vy) / (2 * a * a);
    L[1 * NV + 3] = ((-gm1) * vz) / (2 * a * a);
    L[1 * NV + 4] = gm1 / (2 * a * a);
    L[0 * NV + 0] = (a * a - gm1 * ek) / (a * a);
    L[0 * NV + 1] = (gm1 * vx) / (a * a);
    L[0 * NV + 2] = (gm1 * vy) / (a * a);
    L[0 * NV + 3] = (gm1 * vz) / (a * a);
    L[0 * NV + 4] = (-gm1) / (a * a);
    L[4 * NV + 0] = (gm1 * ek - a * vx) / (2 * a * a);
    L[4 * NV + 1] = ((-gm1) * vx + a) / (2 * a * a);
    L[4 * NV + 2] = ((-gm1) * vy) / (2 * a * a);
    L[4 * NV + 3] = ((-gm1) * vz) / (2 * a * a);
    L[4 * NV + 4] = gm1 / (2 * a * a);
    L[2 * NV + 0] = vy;  L[2 * NV + 1] = 0.0; L[2 * NV + 2] = -1.0; L[2 * NV + 3] = 0.0; L[2 * NV + 4] = 0.0;
    L[3 * NV + 0] = -vz; L[3 * NV + 1] = 0.0; L[3 * NV + 2] = 0.0;  L[3 * NV + 3] = 1.0; L[3 * NV + 4] = 0.0;
  } else if (dir == 1) {
    L[2 * NV + 0] = (gm1 * ek + a * vy) / (2 * a * a);
    L[2 * NV + 1] = ((1.0 - ga) * vx) / (2 * a * a);
    L[2 * NV + 2] = ((1.0 - ga) * vy - a) / (2 * a * a);
    L[2 * NV + 3] = ((1.0 - ga) * vz) / (2 * a * a);
    L[2 * NV + 4] = gm1 / (2 * a * a);
    L[0 * NV + 0] = (a * a - gm1 * ek) / (a * a);
    L[0 * NV + 1] = gm1 * vx / (a * a);
    L[0 * NV + 2] = gm1 * vy / (a * a);
    L[0 * NV + 3] = gm1 * vz / (a * a);
    L[0 * NV + 4] = (1.0 - ga) / (a * a);
    L[4 * NV + 0] = (gm1 * ek - a * vy) / (2 * a * a);
    L[4 * NV + 1] = ((1.0 - ga) * vx) / (2 * a * a);
    L[4 * NV + 2] = ((1.0 - ga) * vy + a) / (2 * a * a);
    L[4 * NV + 3] = ((1.0 - ga) * vz) / (2 * a * a);
    L[4 * NV + 4] = gm1 / (2 * a * a);
    L[1 * NV + 0] = -vx; L[1 * NV + 1] = 1.0; L[1 * NV + 2] = 0.0; L[1 * NV + 3] = 0.0;  L[1 * NV + 4] = 0;
    L[3 * NV + 0] = vz;  L[3 * NV + 1] = 0.0; L[3 * NV + 2] = 0.0; L[3 * NV + 3] = -1.0; L[3 * NV + 4] = 0;
  } else {
    L[3 * NV + 0] = (gm1 * ek + a * vz) / (2 * a * a);
    L[3 * NV + 1] = ((1.0 - ga) * vx) / (2 * a * a);
    L[3 * NV + 2] = ((1.0 - ga) * vy) / (2 * a * a);
    L[3 * NV + 3] = ((1.0 - ga) * vz - a) / (2 * a * a);
    L[3 * NV + 4] = gm1 / (2 * a * a);
    L[0 * NV + 0] = (a * a - gm1 * ek) / (a * a);
    L[0 * NV + 1] = gm1 * vx / (a * a);
    L[0 * NV + 2] = gm1 * vy / (a * a);
    L[0 * NV + 3] = gm1 * vz / (a * a);
    L[0 * NV + 4] = (1.0 - ga) / (a * a);
    L[4 * NV + 0] = (gm1 * ek - a * vz) / (2 * a * a);
    L[4 * NV + 1] = ((1.0 - ga) * vx) / (2 * a * a);
    L[4 * NV + 2] = ((1.0 - ga) * vy) / (2 * a * a);
    L[4 * NV + 3] = ((1.0 - ga) * vz + a) / (2 * a * a);
    L[4 * NV + 4] = gm1 / (2 * a * a);
    L[1 * NV + 0] = vx;  L[1 * NV + 1] = -1.0; L[1 * NV + 2] = 0.0; L[1 * NV + 3] = 0.0; L[1 * NV + 4] = 0;
    L[2 * NV + 0] = -vy; L[2 * NV + 1] = 0.0;  L[2 * NV + 2] = 1.0; L[2 * NV + 3] = 0.0; L[2 * NV + 4] = 0;
  }
}

void orc_right_eigenvectors(const double *u, double *R, double ga, int dir) {
  double gm1 = ga - 1.0;
  double rho, vx, vy, vz, e, P, ek, a, h0;
  flowvar(u, ga, &rho, &vx, &vy, &vz, &e, &P);
  ek = 0.5 * (vx * vx + vy * vy + vz * vz);
  a = sqrt(ga * P / rho);
  h0 = a * a / gm1 + ek;
  if (dir == 0) {
    R[0 * NV + 1] = 1.0; R[1 * NV + 1] = vx - a; R[2 * NV + 1] = vy; R[3 * NV + 1] = vz; R[4 * NV + 1] = h0 - a * vx;
    R[0 * NV + 0] = 1.0; R[1 * NV + 0] = vx;     R[2 * NV + 0] = vy; R[3 * NV + 0] = vz; R[4 * NV + 0] = ek;
    R[0 * NV + 4] = 1.0; R[1 * NV + 4] = vx + a; R[2 * NV + 4] = vy; R[3 * NV + 4] = vz; R[4 * NV + 4] = h0 + a * vx;
    R[0 * NV + 2] = 0.0; R[1 * NV + 2] = 0.0; R[2 * NV + 2] = -1.0; R[3 * NV + 2] = 0.0; R[4 * NV + 2] = -vy;
    R[0 * NV + 3] = 0.0; R[1 * NV + 3] = 0.0; R[2 * NV + 3] = 0.0;  R[3 * NV + 3] = 1.0; R[4 * NV + 3] = vz;
  } else if (dir == 1) {
    R[0 * NV + 2] = 1.0; R[1 * NV + 2] = vx; R[2 * NV + 2] = vy - a; R[3 * NV + 2] = vz; R[4 * NV + 2] = h0 - a * vy;
    R[0 * NV + 0] = 1.0; R[1 * NV + 0] = vx; R[2 * NV + 0] = vy;     R[3 * NV + 0] = vz; R[4 * NV + 0] = ek;
    R[0 * NV + 4] = 1.0; R[1 * NV + 4] = vx; R[2 * NV + 4] = vy + a; R[3 * NV + 4] = vz; R[4 * NV + 4] = h0 + a * vy;
    R[0 * NV + 1] = 0.0; R[1 * NV + 1] = 1.0; R[2 * NV + 1] = 0.0; R[3 * NV + 1] = 0.0;  R[4 * NV + 1] = vx;
    R[0 * NV + 3] = 0.0; R[1 * NV + 3] = 0.0; R[2 * NV + 3] = 0.0; R[3 * NV + 3] = -1.0; R[4 * NV + 3] = -vz;
  } else {
    R[0 * NV + 3] = 1.0; R[1 * NV + 3] = vx; R[2 * NV + 3] = vy; R[3 * NV + 3] = vz - a; R[4 * NV + 3] = h0 - a * vz;
    R[0 * NV + 0] = 1.0; R[1 * NV + 0] = vx; R[2 * NV + 0] = vy; R[3 * NV + 0] = vz;     R[4 * NV + 0] = ek;
    R[0 * NV + 4] = 1.0; R[1 * NV + 4] = vx; R[2 * NV + 4] = vy; R[3 * NV + 4] = vz + a; R[4 * NV + 4] = h0 + a * vz;
    R[0 * NV + 1] = 0.0; R[1 * NV + 1] = -1.0; R[2 * NV + 1] = 0.0; R[3 * NV + 1] = 0.0; R[4 * NV + 1] = -vx;
    R[0 * NV + 2] = 0.0; R[1 * NV + 2] = 0.0;  R[2 * NV + 2] = 1.0; R[3 * NV + 2] = 0.0; R[4 * NV + 2] = vy;
  }
}

/* MatMult5 / MatVecMult5 / MatVecMult (matmult_native.h) */
static inline void matmult5(double *A, const double *X, const double *Y) {
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++)
      A[i * 5 + j] = X[i * 5 + 0] * Y[0 * 5 + j] + X[i * 5 + 1] * Y[1 * 5 + j] +
                     X[i * 5 + 2] * Y[2 * 5 + j] + X[i * 5 + 3] * Y[3 * 5 + j] +
                     X[i * 5 + 4] * Y[4 * 5 + j];
}
static inline void matvecmult5(double *y, const double *A, const double *x) {
  for (int i = 0; i < 5; i++)
    y[i] = A[i * 5 + 0] * x[0] + A[i * 5 + 1] * x[1] + A[i * 5 + 2] * x[2] +
           A[i * 5 + 3] * x[3] + A[i * 5 + 4] * x[4];
}
static inline void matvecmult(double *y, const double *A, const double *x) {
  for (int i = 0; i < 5; i++) {
    y[i] = 0;
    for (int j = 0; j < 5; j++) y[i] += A[i * 5 + j] * x[j];
  }
}

/* ------------------------------------------------------------------ */
/* WENO5-JS                                                           */
/* ------------------------------------------------------------------ */
static const double thirteen_by_twelve = 13.0 / 12.0;
static const double one_fourth = 1.0 / 4.0;
static const double one_sixth = 1.0 / 6.0;

static inline void weno_js(double *w1, double *w2, double *w3, double c1, double c2, double c3,
                           double m3, double m2, double m1, double p1, double p2, double eps) {
  double b1, b2, b3, a1, a2, a3, a_sum_inv;
  b1 = thirteen_by_twelve * (m3 - 2 * m2 + m1) * (m3 - 2 * m2 + m1) +
       one_fourth * (m3 - 4 * m2 + 3 * m1) * (m3 - 4 * m2 + 3 * m1);
  a1 = c1 / ((b1 + eps) * (b1 + eps));
  b2 = thirteen_by_twelve * (m2 - 2 * m1 + p1) * (m2 - 2 * m1 + p1) +
       one_fourth * (m2 - p1) * (m2 - p1);
  a2 = c2 / ((b2 + eps) * (b2 + eps));
  b3 = thirteen_by_twelve * (m1 - 2 * p1 + p2) * (m1 - 2 * p1 + p2) +
       one_fourth * (3 * m1 - 4 * p1 + p2) * (3 * m1 - 4 * p1 + p2);
  a3 = c3 / ((b3 + eps) * (b3 + eps));
  a_sum_inv = 1.0 / (a1 + a2 + a3);
  *w1 = a1 * a_sum_inv;
  *w2 = a2 * a_sum_inv;
  *w3 = a3 * a_sum_inv;
}

void orc_weno_js_weights(double m3, double m2, double m1, double p1, double p2, double eps,
                         double w[3]) {
  weno_js(&w[0], &w[1], &w[2], 0.1, 0.6, 0.3, m3, m2, m1, p1, p2, eps);
}

/* Mapped weights of Henrick, Aslam & Powers (macro _WENOWeights_v_M_,
 * include/interpolation.h:313-344): the JS weights mapped by
 * g_k(w) = w (c_k + c_k^2 - 3 c_k w + w^2) / (c_k^2 + w (1 - 2 c_k)) and
 * normalised again. */
static inline void weno_mapped(double *w1, double *w2, double *w3, double c1, double c2, double c3,
                               double m3, double m2, double m1, double p1, double p2, double eps) {
  double a1, a2, a3, a_sum_inv;
  weno_js(w1, w2, w3, c1, c2, c3, m3, m2, m1, p1, p2, eps);
  a1 = *w1 * (c1 + c1 * c1 - 3 * c1 * *w1 + *w1 * *w1) / (c1 * c1 + *w1 * (1.0 - 2.0 * c1));
  a2 = *w2 * (c2 + c2 * c2 - 3 * c2 * *w2 + *w2 * *w2) / (c2 * c2 + *w2 * (1.0 - 2.0 * c2));
  a3 = *w3 * (c3 + c3 * c3 - 3 * c3 * *w3 + *w3 * *w3) / (c3 * c3 + *w3 * (1.0 - 2.0 * c3));
  a_sum_inv = 1.0 / (a1 + a2 + a3);
  *w1 = a1 * a_sum_inv;
  *w2 = a2 * a_sum_inv;
  *w3 = a3 * a_sum_inv;
}

/* WENO-Z of Borges et al. (tau = |b1 - b3|, macro _WENOWeights_v_Z_,
 * include/interpolation.h:426-447) and of Yamaleev & Carpenter (tau = the
 * squared 4th undivided difference, macro _WENOWeights_v_YC_,
 * include/interpolation.h:479-500): a_k = c_k (1 + (tau/(b_k+eps))^2). */
static inline void weno_tau(int yc, double *w1, double *w2, double *w3, double c1, double c2, double c3,
                            double m3, double m2, double m1, double p1, double p2, double eps) {
  double b1, b2, b3, a1, a2, a3, a_sum_inv, tau;
  b1 = thirteen_by_twelve * (m3 - 2 * m2 + m1) * (m3 - 2 * m2 + m1) +
       one_fourth * (m3 - 4 * m2 + 3 * m1) * (m3 - 4 * m2 + 3 * m1);
  b2 = thirteen_by_twelve * (m2 - 2 * m1 + p1) * (m2 - 2 * m1 + p1) +
       one_fourth * (m2 - p1) * (m2 - p1);
  b3 = thirteen_by_twelve * (m1 - 2 * p1 + p2) * (m1 - 2 * p1 + p2) +
       one_fourth * (3 * m1 - 4 * p1 + p2) * (3 * m1 - 4 * p1 + p2);
  if (yc)
    tau = (m3 - 4 * m2 + 6 * m1 - 4 * p1 + p2) * (m3 - 4 * m2 + 6 * m1 - 4 * p1 + p2);
  else
    tau = fabs(b3 - b1);
  a1 = c1 * (1.0 + (tau / (b1 + eps)) * (tau / (b1 + eps)));
  a2 = c2 * (1.0 + (tau / (b2 + eps)) * (tau / (b2 + eps)));
  a3 = c3 * (1.0 + (tau / (b3 + eps)) * (tau / (b3 + eps)));
  a_sum_inv = 1.0 / (a1 + a2 + a3);
  *w1 = a1 * a_sum_inv;
  *w2 = a2 * a_sum_inv;
  *w3 = a3 * a_sum_inv;
}

/* dispatch on the weight formulation (WENOFifthOrderCalculateWeights.c:55-57, 79-81);
 * with no_limiting the weights keep the optimal values they are initialised
 * to (WENOFifthOrderInitializeWeights.c:33-228, InitializeSolvers.c:219,
 * HyperbolicFunction.c:62) */
static inline void weno_weights(int type, double *w1, double *w2, double *w3, double c1, double c2, double c3,
                                double m3, double m2, double m1, double p1, double p2, double eps) {
  if (type == ORC_WENO_OPTIMAL) {
    *w1 = c1;
    *w2 = c2;
    *w3 = c3;
  } else if (type == ORC_WENO_YC) weno_tau(1, w1, w2, w3, c1, c2, c3, m3, m2, m1, p1, p2, eps);
  else if (type == ORC_WENO_Z) weno_tau(0, w1, w2, w3, c1, c2, c3, m3, m2, m1, p1, p2, eps);
  else if (type == ORC_WENO_MAPPED) weno_mapped(w1, w2, w3, c1, c2, c3, m3, m2, m1, p1, p2, eps);
  else weno_js(w1, w2, w3, c1, c2, c3, m3, m2, m1, p1, p2, eps);
}

void orc_weno_weights(int type, double m3, double m2, double m1, double p1, double p2, double eps,
                      double w[3]) {
  weno_weights(type, &w[0], &w[1], &w[2], 0.1, 0.6, 0.3, m3, m2, m1, p1, p2, eps);
}

/* fifth-order candidate interpolants and their convex combination,
 * component-wise form (Interp1PrimFifthOrderWENO.c:146-158) */
static inline double weno_interp_comp(double w1, double w2, double w3, double fm3, double fm2,
                                      double fm1, double fp1, double fp2) {
  double f1 = (2 * one_sixth) * fm3 + (-7 * one_sixth) * fm2 + (11 * one_sixth) * fm1;
  double f2 = (-one_sixth) * fm2 + (5 * one_sixth) * fm1 + (2 * one_sixth) * fp1;
  double f3 = (2 * one_sixth) * fm1 + (5 * one_sixth) * fp1 + (-one_sixth) * fp2;
  return w1 * f1 + w2 * f2 + w3 * f3;
}

/* characteristic form (Interp1PrimFifthOrderWENOChar.c:167-182) */
static inline double weno_interp_char(double w1, double w2, double w3, double fm3, double fm2,
                                      double fm1, double fp1, double fp2) {
  double f1 = (2 * one_sixth) * fm3 - (7.0 * one_sixth) * fm2 + (11.0 * one_sixth) * fm1;
  double f2 = (-one_sixth) * fm2 + (5.0 * one_sixth) * fm1 + (2 * one_sixth) * fp1;
  double f3 = (2 * one_sixth) * fm1 + (5 * one_sixth) * fp1 - (one_sixth) * fp2;
  return w1 * f1 + w2 * f2 + w3 * f3;
}

double orc_weno5_interp(const double *f, double eps) {
  double w[3];
  orc_weno_js_weights(f[0], f[1], f[2], f[3], f[4], eps, w);
  return weno_interp_comp(w[0], w[1], w[2], f[0], f[1], f[2], f[3], f[4]);
}

/* ------------------------------------------------------------------ */
/* periodic boundary                                                  */
/* ------------------------------------------------------------------ */
void orc_periodic_bc(const orc_params *p, double *u) {
  grid_t G = mkgrid(p);
  int g = G.g;
  /* boundary.inp order: dim 0 face +1, face -1, dim 1 ..., dim 2 ... ;
   * only face slabs (other indices interior) are filled */
  for (int d = 0; d < 3; d++) {
    int n = G.n[d];
    for (int face = 0; face < 2; face++) {
      int lo[3] = {0, 0, 0}, hi[3] = {G.n[0], G.n[1], G.n[2]};
      if (face == 0) { lo[d] = -g; hi[d] = 0; } else { lo[d] = n; hi[d] = n + g; }
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++) {
            int s[3] = {i, j, k};
            s[d] += (face == 0) ? n : -n;
            long p1 = pidx(&G, i, j, k), p2 = pidx(&G, s[0], s[1], s[2]);
            for (int v = 0; v < NV; v++) u[NV * p1 + v] = u[NV * p2 + v];
          }
    }
  }
}

/* ------------------------------------------------------------------ */
/* hyperbolic term                                                    */
/* ------------------------------------------------------------------ */
static void cell_flux_all(const grid_t *G, const orc_params *p, double *fC, const double *u,
                          int dir) {
  long npts = (long)G->ng[0] * G->ng[1] * G->ng[2];
#pragma omp parallel for schedule(static)
  for (long q = 0; q < npts; q++) {
    double rho, vx, vy, vz, e, P;
    flowvar(u + NV * q, p->gamma, &rho, &vx, &vy, &vz, &e, &P);
    setflux(fC + NV * q, rho, vx, vy, vz, e, P, dir);
  }
}

static void modified_solution_all(const grid_t *G, const orc_params *p, double *uC,
                                  const double *u) {
  long npts = (long)G->ng[0] * G->ng[1] * G->ng[2];
  double inv_gamma_m1 = 1.0 / (p->gamma - 1.0);
  const double grav_f = 1.0, grav_g = 1.0; /* no gravity: f = g = exp(0) */
#pragma omp parallel for schedule(static)
  for (long q = 0; q < npts; q++) {
    double rho, uvel, vvel, wvel, E, P;
    flowvar(u + NV * q, p->gamma, &rho, &uvel, &vvel, &wvel, &E, &P);
    uC[NV * q + 0] = u[NV * q + 0] * grav_f;
    uC[NV * q + 1] = u[NV * q + 1] * grav_f;
    uC[NV * q + 2] = u[NV * q + 2] * grav_f;
    uC[NV * q + 3] = u[NV * q + 3] * grav_f;
    uC[NV * q + 4] = (P * inv_gamma_m1) * (1.0 / grav_g) +
                     (0.5 * rho * (uvel * uvel + vvel * vvel + wvel * wvel)) * grav_f;
  }
}

/* interface flux at interface I (between cells I-1 and I) of one line.
 * line(c) gives the point index of cell c along the line. */
static void interface_flux(const orc_params *p, const double *u, const double *uC,
                           const double *fC, const long *q /* cells I-3..I+2 -> q[0..5] */,
                           int dir, double *fI) {
  const double eps = p->eps;
  const double c1 = 0.1, c2 = 0.6, c3 = 0.3;
  double fL[NV], fR[NV], uL[NV], uR[NV];
  /* stencil point indices: left-biased m3..p2 = I-3..I+1 -> q[0..4];
   * right-biased m3..p2 = I+2..I-2 -> q[5],q[4],q[3],q[2],q[1] */
  const long qm3L = q[0], qm2L = q[1], qm1L = q[2], qp1L = q[3], qp2L = q[4];
  const long qm3R = q[5], qm2R = q[4], qm1R = q[3], qp1R = q[2], qp2R = q[1];

  if (p->interp == ORC_INTERP_COMPONENTS) {
    for (int v = 0; v < NV; v++) {
      double w1LF, w2LF, w3LF, w1RF, w2RF, w3RF, w1LU, w2LU, w3LU, w1RU, w2RU, w3RU;
      weno_weights(p->weno_type, &w1LF, &w2LF, &w3LF, c1, c2, c3, fC[qm3L * NV + v], fC[qm2L * NV + v],
              fC[qm1L * NV + v], fC[qp1L * NV + v], fC[qp2L * NV + v], eps);
      weno_weights(p->weno_type, &w1RF, &w2RF, &w3RF, c1, c2, c3, fC[qm3R * NV + v], fC[qm2R * NV + v],
              fC[qm1R * NV + v], fC[qp1R * NV + v], fC[qp2R * NV + v], eps);
      weno_weights(p->weno_type, &w1LU, &w2LU, &w3LU, c1, c2, c3, u[qm3L * NV + v], u[qm2L * NV + v],
              u[qm1L * NV + v], u[qp1L * NV + v], u[qp2L * NV + v], eps);
      weno_weights(p->weno_type, &w1RU, &w2RU, &w3RU, c1, c2, c3, u[qm3R * NV + v], u[qm2R * NV + v],
              u[qm1R * NV + v], u[qp1R * NV + v], u[qp2R * NV + v], eps);
      uL[v] = weno_interp_comp(w1LU, w2LU, w3LU, uC[qm3L * NV + v], uC[qm2L * NV + v],
                               uC[qm1L * NV + v], uC[qp1L * NV + v], uC[qp2L * NV + v]);
      uR[v] = weno_interp_comp(w1RU, w2RU, w3RU, uC[qm3R * NV + v], uC[qm2R * NV + v],
                               uC[qm1R * NV + v], uC[qp1R * NV + v], uC[qp2R * NV + v]);
      fL[v] = weno_interp_comp(w1LF, w2LF, w3LF, fC[qm3L * NV + v], fC[qm2L * NV + v],
                               fC[qm1L * NV + v], fC[qp1L * NV + v], fC[qp2L * NV + v]);
      fR[v] = weno_interp_comp(w1RF, w2RF, w3RF, fC[qm3R * NV + v], fC[qm2R * NV + v],
                               fC[qm1R * NV + v], fC[qp1R * NV + v], fC[qp2R * NV + v]);
    }
  } else {
    /* weights: L from the Roe average of the raw solution at I-1, I */
    double uavg[NV], L[25], R[25];
    double wLF[3][NV], wRF[3][NV], wLU[3][NV], wRU[3][NV];
    orc_roe_average(uavg, u + NV * qm1L, u + NV * qp1L, p->gamma);
    orc_left_eigenvectors(uavg, L, p->gamma, dir);
    {
      double m3LF[NV], m2LF[NV], m1LF[NV], p1LF[NV], p2LF[NV];
      double m3RF[NV], m2RF[NV], m1RF[NV], p1RF[NV], p2RF[NV];
      double m3LU[NV], m2LU[NV], m1LU[NV], p1LU[NV], p2LU[NV];
      double m3RU[NV], m2RU[NV], m1RU[NV], p1RU[NV], p2RU[NV];
      matvecmult(m3LF, L, fC + NV * qm3L); matvecmult(m2LF, L, fC + NV * qm2L);
      matvecmult(m1LF, L, fC + NV * qm1L); matvecmult(p1LF, L, fC + NV * qp1L);
      matvecmult(p2LF, L, fC + NV * qp2L);
      matvecmult(m3RF, L, fC + NV * qm3R); matvecmult(m2RF, L, fC + NV * qm2R);
      matvecmult(m1RF, L, fC + NV * qm1R); matvecmult(p1RF, L, fC + NV * qp1R);
      matvecmult(p2RF, L, fC + NV * qp2R);
      matvecmult(m3LU, L, u + NV * qm3L); matvecmult(m2LU, L, u + NV * qm2L);
      matvecmult(m1LU, L, u + NV * qm1L); matvecmult(p1LU, L, u + NV * qp1L);
      matvecmult(p2LU, L, u + NV * qp2L);
      matvecmult(m3RU, L, u + NV * qm3R); matvecmult(m2RU, L, u + NV * qm2R);
      matvecmult(m1RU, L, u + NV * qm1R); matvecmult(p1RU, L, u + NV * qp1R);
      matvecmult(p2RU, L, u + NV * qp2R);
      for (int v = 0; v < NV; v++) {
        weno_weights(p->weno_type, &wLF[0][v], &wLF[1][v], &wLF[2][v], c1, c2, c3, m3LF[v], m2LF[v], m1LF[v], p1LF[v], p2LF[v], eps);
        weno_weights(p->weno_type, &wRF[0][v], &wRF[1][v], &wRF[2][v], c1, c2, c3, m3RF[v], m2RF[v], m1RF[v], p1RF[v], p2RF[v], eps);
        weno_weights(p->weno_type, &wLU[0][v], &wLU[1][v], &wLU[2][v], c1, c2, c3, m3LU[v], m2LU[v], m1LU[v], p1LU[v], p2LU[v], eps);
        weno_weights(p->weno_type, &wRU[0][v], &wRU[1][v], &wRU[2][v], c1, c2, c3, m3RU[v], m2RU[v], m1RU[v], p1RU[v], p2RU[v], eps);
      }
    }
    /* four characteristic interpolations; the averaged state takes
     * (u[qm1], u[qp1]) in stencil order, i.e. swapped for right bias */
    for (int pass = 0; pass < 4; pass++) {
      int upw = (pass == 0 || pass == 2) ? 1 : -1;     /* uL, uR, fL, fR */
      int uflag = (pass < 2);
      const double *F = uflag ? uC : fC;
      double (*W)[NV] = uflag ? (upw > 0 ? wLU : wRU) : (upw > 0 ? wLF : wRF);
      long qm3, qm2, qm1, qp1, qp2;
      if (upw > 0) { qm3 = q[0]; qm2 = q[1]; qm1 = q[2]; qp1 = q[3]; qp2 = q[4]; }
      else         { qm3 = q[5]; qm2 = q[4]; qm1 = q[3]; qp1 = q[2]; qp2 = q[1]; }
      double uav[NV], Lp[25], fchar[NV];
      orc_roe_average(uav, u + NV * qm1, u + NV * qp1, p->gamma);
      orc_left_eigenvectors(uav, Lp, p->gamma, dir);
      orc_right_eigenvectors(uav, R, p->gamma, dir);
      for (int v = 0; v < NV; v++) {
        double fm3 = 0, fm2 = 0, fm1 = 0, fp1 = 0, fp2 = 0;
        for (int k = 0; k < NV; k++) {
          fm3 += Lp[v * NV + k] * F[qm3 * NV + k];
          fm2 += Lp[v * NV + k] * F[qm2 * NV + k];
          fm1 += Lp[v * NV + k] * F[qm1 * NV + k];
          fp1 += Lp[v * NV + k] * F[qp1 * NV + k];
          fp2 += Lp[v * NV + k] * F[qp2 * NV + k];
        }
        fchar[v] = weno_interp_char(W[0][v], W[1][v], W[2][v], fm3, fm2, fm1, fp1, fp2);
      }
      double *out = (pass == 0) ? uL : (pass == 1) ? uR : (pass == 2) ? fL : fR;
      matvecmult(out, R, fchar);
    }
  }

  /* upwinding with the cell-centred states at I-1 (pL) and I (pR) */
  const double *uPL = u + NV * q[2], *uPR = u + NV * q[3];
  double udiff[NV], uavg[NV];
  for (int v = 0; v < NV; v++) udiff[v] = 0.5 * (uR[v] - uL[v]);
  orc_roe_average(uavg, uPL, uPR, p->gamma);
  if (p->upwind == ORC_UPWIND_ROE) {
    double D[25], L[25], R[25], DL[25], modA[25], udiss[NV];
    eigenvalues(uavg, D, p->gamma, dir);
    orc_left_eigenvectors(uavg, L, p->gamma, dir);
    orc_right_eigenvectors(uavg, R, p->gamma, dir);
    double delta = 0.000001, delta2 = delta * delta;
    for (int k = 0; k < 25; k += 6) {
      double ad = D[k] < 0 ? -D[k] : D[k];
      D[k] = (ad < delta ? (D[k] * D[k] + delta2) / (2 * delta) : ad);
    }
    matmult5(DL, D, L);
    matmult5(modA, R, DL);
    matvecmult5(udiss, modA, udiff);
    for (int v = 0; v < NV; v++) fI[v] = 0.5 * (fL[v] + fR[v]) - udiss[v];
  } else {
    double c, vel[3], rho, E, P;
    flowvar(uPL, p->gamma, &rho, &vel[0], &vel[1], &vel[2], &E, &P);
    c = sqrt(p->gamma * P / rho);
    double alphaL = c + (vel[dir] < 0 ? -vel[dir] : vel[dir]);
    flowvar(uPR, p->gamma, &rho, &vel[0], &vel[1], &vel[2], &E, &P);
    c = sqrt(p->gamma * P / rho);
    double alphaR = c + (vel[dir] < 0 ? -vel[dir] : vel[dir]);
    flowvar(uavg, p->gamma, &rho, &vel[0], &vel[1], &vel[2], &E, &P);
    c = sqrt(p->gamma * P / rho);
    double alphaavg = c + (vel[dir] < 0 ? -vel[dir] : vel[dir]);
    double kappa = 1.0; /* max(g_L, g_R) with g = 1 without gravity */
    double m01 = alphaL > alphaR ? alphaL : alphaR;
    double m12 = alphaR > alphaavg ? alphaR : alphaavg;
    double alpha = kappa * (m01 > m12 ? m01 : m12);
    for (int v = 0; v < NV; v++) fI[v] = 0.5 * (fL[v] + fR[v]) - alpha * udiff[v];
  }
}

int orc_hyperbolic(const orc_params *p, double *hyp, const double *u) {
  grid_t G = mkgrid(p);
  long npts = (long)G.ng[0] * G.ng[1] * G.ng[2];
  double *fC = (double *)malloc(sizeof(double) * NV * npts);
  double *uC = (double *)malloc(sizeof(double) * NV * npts);
  if (!fC || !uC) { free(fC); free(uC); return 1; }
  memset(hyp, 0, sizeof(double) * NV * npts);
  for (int d = 0; d < 3; d++) {
    cell_flux_all(&G, p, fC, u, d);
    modified_solution_all(&G, p, uC, u);
    const double *dxi = dxinv_dir(p, d);
    int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
    int n = G.n[d];
    long nlines = (long)G.n[e1] * G.n[e2];
#pragma omp parallel for schedule(dynamic, 16)
    for (long l = 0; l < nlines; l++) {
      int a = (int)(l % G.n[e1]), b = (int)(l / G.n[e1]);
      double *FI = (double *)malloc(sizeof(double) * NV * (n + 1));
      int idx[3];
      idx[e1] = a; idx[e2] = b;
      for (int I = 0; I <= n; I++) {
        long q[6];
        for (int s = 0; s < 6; s++) {
          idx[d] = I - 3 + s;
          q[s] = pidx(&G, idx[0], idx[1], idx[2]);
        }
        interface_flux(p, u, uC, fC, q, d, FI + NV * I);
      }
      for (int i = 0; i < n; i++) {
        idx[d] = i;
        long pp = pidx(&G, idx[0], idx[1], idx[2]);
        for (int v = 0; v < NV; v++)
          hyp[NV * pp + v] += dxi[i] * (FI[NV * (i + 1) + v] - FI[NV * i + v]);
      }
      free(FI);
    }
  }
  free(fC);
  free(uC);
  return 0;
}

/* ------------------------------------------------------------------ */
/* viscous term                                                       */
/* ------------------------------------------------------------------ */
static const double one_twelve = 1.0 / 12.0;

/* one point of FirstDerivativeFourthOrderCentral (not divided by dx):
 * one-sided and biased stencils in the two outermost layers, central inside */
static inline double d4_point(const double *F, long st, int i, int n, int g) {
  if (i == -g)
    return (-25 * F[0] + 48 * F[st] - 36 * F[2 * st] + 16 * F[3 * st] - 3 * F[4 * st]) * one_twelve;
  if (i == -g + 1)
    return (-3 * F[-st] - 10 * F[0] + 18 * F[st] - 6 * F[2 * st] + F[3 * st]) * one_twelve;
  if (i < n + g - 2)
    return (F[-2 * st] - 8 * F[-st] + 8 * F[st] - F[2 * st]) * one_twelve;
  if (i == n + g - 2)
    return (-F[-3 * st] + 6 * F[-2 * st] - 18 * F[-st] + 10 * F[0] + 3 * F[st]) * one_twelve;
  return (3 * F[-4 * st] - 16 * F[-3 * st] + 36 * F[-2 * st] - 48 * F[-st] + 25 * F[0]) * one_twelve;
}

/* FirstDerivativeFourthOrderCentral over lines whose other indices are
 * interior; the derivative array is written at every point of those lines */
static void first_derivative4(const grid_t *G, double *Df, const double *f, int dir) {
  int g = G->g, n = G->n[dir];
  int e1 = (dir + 1) % 3, e2 = (dir + 2) % 3;
  long nlines = (long)G->n[e1] * G->n[e2];
  long st = G->stride[dir] * NV;
#pragma omp parallel for schedule(static)
  for (long l = 0; l < nlines; l++) {
    int idx[3];
    idx[e1] = (int)(l % G->n[e1]); idx[e2] = (int)(l / G->n[e1]);
    for (int i = -g; i < n + g; i++) {
      idx[dir] = i;
      long qC = pidx(G, idx[0], idx[1], idx[2]) * NV;
      for (int v = 0; v < NV; v++) Df[qC + v] = d4_point(f + qC + v, st, i, n, g);
    }
  }
}

void orc_first_derivative4_1d(int n, int g, const double *f, double *Df) {
  for (int i = -g; i < n + g; i++) Df[i + g] = d4_point(f + i + g, 1, i, n, g);
}

int orc_parabolic(const orc_params *p, double *par, const double *u) {
  grid_t G = mkgrid(p);
  int g = G.g, imax = G.n[0], jmax = G.n[1], kmax = G.n[2];
  long npts = (long)G.ng[0] * G.ng[1] * G.ng[2];
  memset(par, 0, sizeof(double) * NV * npts);
  if (p->Re <= 0) return 0;
  static const double two_third = 2.0 / 3.0;
  double inv_gamma_m1 = 1.0 / (p->gamma - 1.0);
  double inv_Re = 1.0 / p->Re;
  double inv_Pr = 1.0 / p->Pr;
  const double *dxi[3] = {dxinv_dir(p, 0), dxinv_dir(p, 1), dxinv_dir(p, 2)};

  double *Q = (double *)calloc(NV * npts, sizeof(double));
  double *QD[3];
  for (int d = 0; d < 3; d++) QD[d] = (double *)calloc(NV * npts, sizeof(double));
  double *FV = (double *)calloc(NV * npts, sizeof(double));
  double *FD = (double *)calloc(NV * npts, sizeof(double));

#pragma omp parallel for schedule(static)
  for (long q = 0; q < npts; q++) {
    double energy, pressure;
    flowvar(u + NV * q, p->gamma, &Q[NV * q + 0], &Q[NV * q + 1], &Q[NV * q + 2], &Q[NV * q + 3],
            &energy, &pressure);
    Q[NV * q + 4] = p->gamma * pressure / Q[NV * q + 0];
  }
  for (int d = 0; d < 3; d++) first_derivative4(&G, QD[d], Q, d);
#pragma omp parallel for schedule(static)
  for (int k = -g; k < kmax + g; k++)
    for (int j = -g; j < jmax + g; j++)
      for (int i = -g; i < imax + g; i++) {
        long q = pidx(&G, i, j, k) * NV;
        double dx = dxi[0][i], dy = dxi[1][j], dz = dxi[2][k];
        for (int v = 0; v < NV; v++) {
          QD[0][q + v] *= dx;
          QD[1][q + v] *= dy;
          QD[2][q + v] *= dz;
        }
      }

  for (int dir = 0; dir < 3; dir++) {
#pragma omp parallel for schedule(static)
    for (int k = -g; k < kmax + g; k++)
      for (int j = -g; j < jmax + g; j++)
        for (int i = -g; i < imax + g; i++) {
          long q = pidx(&G, i, j, k) * NV;
          double uvel = Q[q + 1], vvel = Q[q + 2], wvel = Q[q + 3], T = Q[q + 4];
          double mu = exp(0.76 * log(T));
          double *F = FV + q;
          if (dir == 0) {
            double Tx = QD[0][q + 4], ux = QD[0][q + 1], vx = QD[0][q + 2], wx = QD[0][q + 3];
            double uy = QD[1][q + 1], vy = QD[1][q + 2];
            double uz = QD[2][q + 1], wz = QD[2][q + 3];
            double tau_xx = two_third * (mu * inv_Re) * (2 * ux - vy - wz);
            double tau_xy = (mu * inv_Re) * (uy + vx);
            double tau_xz = (mu * inv_Re) * (uz + wx);
            double qx = (mu * inv_Re * inv_gamma_m1 * inv_Pr) * Tx;
            F[0] = 0.0; F[1] = tau_xx; F[2] = tau_xy; F[3] = tau_xz;
            F[4] = uvel * tau_xx + vvel * tau_xy + wvel * tau_xz + qx;
          } else if (dir == 1) {
            double Ty = QD[1][q + 4], ux = QD[0][q + 1], vx = QD[0][q + 2];
            double uy = QD[1][q + 1], vy = QD[1][q + 2], wy = QD[1][q + 3];
            double vz = QD[2][q + 2], wz = QD[2][q + 3];
            double tau_yx = (mu * inv_Re) * (uy + vx);
            double tau_yy = two_third * (mu * inv_Re) * (-ux + 2 * vy - wz);
            double tau_yz = (mu * inv_Re) * (vz + wy);
            double qy = (mu * inv_Re * inv_gamma_m1 * inv_Pr) * Ty;
            F[0] = 0.0; F[1] = tau_yx; F[2] = tau_yy; F[3] = tau_yz;
            F[4] = uvel * tau_yx + vvel * tau_yy + wvel * tau_yz + qy;
          } else {
            double Tz = QD[2][q + 4], ux = QD[0][q + 1], wx = QD[0][q + 3];
            double vy = QD[1][q + 2], wy = QD[1][q + 3];
            double uz = QD[2][q + 1], vz = QD[2][q + 2], wz = QD[2][q + 3];
            double tau_zx = (mu * inv_Re) * (uz + wx);
            double tau_zy = (mu * inv_Re) * (vz + wy);
            double tau_zz = two_third * (mu * inv_Re) * (-ux - vy + 2 * wz);
            double qz = (mu * inv_Re * inv_gamma_m1 * inv_Pr) * Tz;
            F[0] = 0.0; F[1] = tau_zx; F[2] = tau_zy; F[3] = tau_zz;
            F[4] = uvel * tau_zx + vvel * tau_zy + wvel * tau_zz + qz;
          }
        }
    first_derivative4(&G, FD, FV, dir);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < kmax; k++)
      for (int j = 0; j < jmax; j++)
        for (int i = 0; i < imax; i++) {
          long q = pidx(&G, i, j, k) * NV;
          int ii = dir == 0 ? i : dir == 1 ? j : k;
          double dxinv = dxi[dir][ii];
          for (int v = 0; v < NV; v++) par[q + v] += (dxinv * FD[q + v]);
        }
  }
  free(Q); free(QD[0]); free(QD[1]); free(QD[2]); free(FV); free(FD);
  return 0;
}

/* ------------------------------------------------------------------ */
/* RHS, RK4, CFL                                                      */
/* ------------------------------------------------------------------ */
int orc_rhs(const orc_params *p, double *rhs, double *u) {
  long n = orc_size(p);
  double *hyp = (double *)malloc(sizeof(double) * n);
  double *par = (double *)malloc(sizeof(double) * n);
  if (!hyp || !par) { free(hyp); free(par); return 1; }
  orc_periodic_bc(p, u);
  int ierr = orc_hyperbolic(p, hyp, u);
  ierr |= orc_parabolic(p, par, u);
  const double source = 0.0;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    double r = 0.0;
    r += -1.0 * hyp[i];
    r += 1.0 * par[i];
    r += 1.0 * source;
    rhs[i] = r;
  }
  free(hyp);
  free(par);
  return ierr;
}

/* TimeRK.c:126-195 for the explicit tableau A (ns x ns, row-major, every
 * entry below the diagonal applied, zeros included) and b */
static int orc_rk(const orc_params *p, double *u, double dt, int nsteps, int ns, const double *A,
                  const double *b) {
  long n = orc_size(p);
  double *U = (double *)malloc(sizeof(double) * n);
  double *K[4];
  for (int s = 0; s < ns; s++) K[s] = (double *)malloc(sizeof(double) * n);
  int ierr = 0;
  for (int it = 0; it < nsteps && !ierr; it++) {
    orc_periodic_bc(p, u); /* TimePreStep */
    for (int stage = 0; stage < ns; stage++) {
      memcpy(U, u, sizeof(double) * n);
      for (int i = 0; i < stage; i++) {
        double a = dt * A[stage * ns + i];
#pragma omp parallel for schedule(static)
        for (long q = 0; q < n; q++) U[q] += a * K[i][q];
      }
      ierr |= orc_rhs(p, K[stage], U);
    }
    for (int stage = 0; stage < ns; stage++) {
      double a = dt * b[stage];
#pragma omp parallel for schedule(static)
      for (long q = 0; q < n; q++) u[q] += a * K[stage][q];
    }
  }
  free(U);
  for (int s = 0; s < ns; s++) free(K[s]);
  return ierr;
}

/* "rk 44" (TimeExplicitRKInitialize.c:58-68) */
int orc_rk4(const orc_params *p, double *u, double dt, int nsteps) {
  const double A[16] = {0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1.0, 0};
  const double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  return orc_rk(p, u, dt, nsteps, 4, A, b);
}

/* any explicit tableau of TimeExplicitRKInitialize.c:26-79 (ns <= 4) */
int orc_rk_tableau(const orc_params *p, double *u, double dt, int nsteps, int ns, const double *A,
                   const double *b) {
  if (ns < 1 || ns > 4) return 1;
  return orc_rk(p, u, dt, nsteps, ns, A, b);
}

/* "rk ssprk3" / "tvdrk3" (TimeExplicitRKInitialize.c:69-79) */
int orc_ssprk3(const orc_params *p, double *u, double dt, int nsteps) {
  const double A[9] = {0, 0, 0, 1.0, 0, 0, 0.25, 0.25, 0};
  const double b[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
  return orc_rk(p, u, dt, nsteps, 3, A, b);
}

double orc_cfl(const orc_params *p, const double *u, double dt) {
  grid_t G = mkgrid(p);
  const double *dxi[3] = {dxinv_dir(p, 0), dxinv_dir(p, 1), dxinv_dir(p, 2)};
  double max_cfl = 0;
  for (int k = 0; k < G.n[2]; k++)
    for (int j = 0; j < G.n[1]; j++)
      for (int i = 0; i < G.n[0]; i++) {
        double rho, vx, vy, vz, e, P, c;
        flowvar(u + NV * pidx(&G, i, j, k), p->gamma, &rho, &vx, &vy, &vz, &e, &P);
        c = sqrt(p->gamma * P / rho);
        double lx = ((vx < 0 ? -vx : vx) + c) * dt * dxi[0][i];
        double ly = ((vy < 0 ? -vy : vy) + c) * dt * dxi[1][j];
        double lz = ((vz < 0 ? -vz : vz) + c) * dt * dxi[2][k];
        if (lx > max_cfl) max_cfl = lx;
        if (ly > max_cfl) max_cfl = ly;
        if (lz > max_cfl) max_cfl = lz;
      }
  return max_cfl;
}

/* VolumeIntegral.c:14-50: sum of u dV over the interior, x fastest, with
 * dV = 1/dxinv_x 1/dxinv_y 1/dxinv_z multiplied in that order */
void orc_volume_integral(const orc_params *p, const double *u, double *out) {
  grid_t G = mkgrid(p);
  const double *dxi[3] = {dxinv_dir(p, 0), dxinv_dir(p, 1), dxinv_dir(p, 2)};
  for (int v = 0; v < NV; v++) out[v] = 0.0;
  for (int k = 0; k < G.n[2]; k++)
    for (int j = 0; j < G.n[1]; j++)
      for (int i = 0; i < G.n[0]; i++) {
        const double *q = u + NV * pidx(&G, i, j, k);
        double dV = 1.0;
        dV *= (1.0 / dxi[0][i]);
        dV *= (1.0 / dxi[1][j]);
        dV *= (1.0 / dxi[2][k]);
        for (int v = 0; v < NV; v++) out[v] += (q[v] * dV);
      }
}
