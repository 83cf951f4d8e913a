// hm_device.h -- device math of the NS3D WENO5 path for gfx950 (FP64).
//
// Physics follows include/physicalmodels/navierstokes3d.h:98-470 of the
// reference (flow variables, Euler flux, Roe average, eigenvalues and
// eigenvectors of the flux Jacobian).  WENO5 follows the Jiang-Shu
// formulation of include/interpolation.h:265-285 and
// src/InterpolationFunctions/Interp1PrimFifthOrderWENO.c:146-158.
//
// Restatements (equal in exact arithmetic, rounding-level differences only):
//  * The Jiang-Shu weights are applied as a single ratio
//      sum_k a_k f_k / sum_k a_k,  a_k = c_k * prod_{j!=k} (b_j+eps)^2,
//    which is w_k = (c_k/(b_k+eps)^2) / sum_j (c_j/(b_j+eps)^2) multiplied
//    through by prod_j (b_j+eps)^2: one division per interpolant instead of
//    four; the WENO-Z / Yamaleev-Carpenter weights likewise.
//  * Eigenvector entries divided by a^2 or 2a^2 use one reciprocal.
//  * Characteristic upwinding uses L*R = I: with uL = R cL and uR = R cR,
//    R|Lambda|L (uR-uL)/2 = R |Lambda| (cR-cL)/2, so the interface flux is
//    R (0.5 (fcL + fcR) - 0.5 |lambda| (ucR - ucL)).
#pragma once

#include <hip/hip_runtime.h>

namespace hm {

constexpr int NV = 5;

struct Prim {
  double rho, vx, vy, vz, e, P;
};

// _NavierStokes3DGetFlowVar_ (navierstokes3d.h:98-108)
__device__ __forceinline__ Prim flowvar(const double *u, double gamma) {
  Prim q;
  q.rho = u[0];
  q.vx = (q.rho == 0) ? 0.0 : u[1] / q.rho;
  q.vy = (q.rho == 0) ? 0.0 : u[2] / q.rho;
  q.vz = (q.rho == 0) ? 0.0 : u[3] / q.rho;
  q.e = u[4];
  double vsq = q.vx * q.vx + q.vy * q.vy + q.vz * q.vz;
  q.P = (q.e - 0.5 * q.rho * vsq) * (gamma - 1.0);
  return q;
}

template <int DIR> __device__ __forceinline__ double vel(const Prim &q) {
  return DIR == 0 ? q.vx : (DIR == 1 ? q.vy : q.vz);
}

// _NavierStokes3DSetFlux_ (navierstokes3d.h:118-139)
template <int DIR> __device__ __forceinline__ void euler_flux(double f[NV], const Prim &q) {
  const double vn = vel<DIR>(q);
  const double m = q.rho * vn;
  f[0] = m;
  f[1] = m * q.vx;
  f[2] = m * q.vy;
  f[3] = m * q.vz;
  f[DIR + 1] += q.P;
  f[4] = (q.e + q.P) * vn;
}

// energy of the modified solution uC (NavierStokes3DModifiedSolution.c:65
// with the gravity fields f = g = 1): P/(gamma-1) + rho|v|^2/2
__device__ __forceinline__ double modified_energy(const Prim &q, double inv_gm1) {
  return (q.P * inv_gm1) + (0.5 * q.rho * (q.vx * q.vx + q.vy * q.vy + q.vz * q.vz));
}

// _NavierStokes3DRoeAverage_ (navierstokes3d.h:144-169), returned as
// primitive variables of the averaged state.
__device__ __forceinline__ Prim roe_average(const double *uL, const double *uR, double gamma) {
  const double gm1 = gamma - 1.0;
  Prim l = flowvar(uL, gamma), r = flowvar(uR, gamma);
  double HL = 0.5 * (l.vx * l.vx + l.vy * l.vy + l.vz * l.vz) + (gamma * l.P / l.rho) / gm1;
  double HR = 0.5 * (r.vx * r.vx + r.vy * r.vy + r.vz * r.vz) + (gamma * r.P / r.rho) / gm1;
  double tL = sqrt(l.rho), tR = sqrt(r.rho);
  double inv = 1.0 / (tL + tR);
  Prim a;
  a.rho = tL * tR;
  a.vx = (tL * l.vx + tR * r.vx) * inv;
  a.vy = (tL * l.vy + tR * r.vy) * inv;
  a.vz = (tL * l.vz + tR * r.vz) * inv;
  double H = (tL * HL + tR * HR) * inv;
  double vsq = a.vx * a.vx + a.vy * a.vy + a.vz * a.vz;
  a.P = (H - 0.5 * vsq) * (a.rho * gm1) / gamma;
  a.e = a.P / gm1 + 0.5 * a.rho * vsq;
  return a;
}

// Eigen-decomposition of the x/y/z flux Jacobian at a state given by its
// primitive variables (navierstokes3d.h:254-470).  Row r of L and column r
// of R belong to eigenvalue lam[r]; the row order is the reference's.
struct Eigen {
  double L[NV][NV];
  double R[NV][NV];  // R[row][col]
  double lam[NV];
};

template <int DIR>
__device__ __forceinline__ void eigen(Eigen &E, const Prim &q, double gamma) {
  const double gm1 = gamma - 1.0;
  const double vx = q.vx, vy = q.vy, vz = q.vz;
  const double ek = 0.5 * (vx * vx + vy * vy + vz * vz);
  const double a = sqrt(gamma * q.P / q.rho);
  const double ia2 = 1.0 / (a * a);
  const double ih2 = 0.5 * ia2;
  const double h0 = a * a / gm1 + ek;
  const double vn = DIR == 0 ? vx : (DIR == 1 ? vy : vz);
  const int m = DIR + 1;  // row/column of the u-a wave
  // entropy wave: row 0
  E.L[0][0] = (a * a - gm1 * ek) * ia2;
  E.L[0][1] = (gm1 * vx) * ia2;
  E.L[0][2] = (gm1 * vy) * ia2;
  E.L[0][3] = (gm1 * vz) * ia2;
  E.L[0][4] = (-gm1) * ia2;
  // u-a (row m) and u+a (row 4)
  E.L[m][0] = (gm1 * ek + a * vn) * ih2;
  E.L[4][0] = (gm1 * ek - a * vn) * ih2;
  E.L[m][1] = ((-gm1) * vx) * ih2;
  E.L[m][2] = ((-gm1) * vy) * ih2;
  E.L[m][3] = ((-gm1) * vz) * ih2;
  E.L[4][1] = E.L[m][1];
  E.L[4][2] = E.L[m][2];
  E.L[4][3] = E.L[m][3];
  E.L[m][DIR + 1] = ((-gm1) * vn - a) * ih2;
  E.L[4][DIR + 1] = ((-gm1) * vn + a) * ih2;
  E.L[m][4] = gm1 * ih2;
  E.L[4][4] = gm1 * ih2;
  // R columns 0, m, 4
  E.R[0][0] = 1.0; E.R[1][0] = vx; E.R[2][0] = vy; E.R[3][0] = vz; E.R[4][0] = ek;
  E.R[0][m] = 1.0; E.R[1][m] = vx; E.R[2][m] = vy; E.R[3][m] = vz; E.R[4][m] = h0 - a * vn;
  E.R[0][4] = 1.0; E.R[1][4] = vx; E.R[2][4] = vy; E.R[3][4] = vz; E.R[4][4] = h0 + a * vn;
  E.R[DIR + 1][m] = vn - a;
  E.R[DIR + 1][4] = vn + a;
  // shear waves
  if (DIR == 0) {
    E.L[2][0] = vy;  E.L[2][1] = 0; E.L[2][2] = -1; E.L[2][3] = 0; E.L[2][4] = 0;
    E.L[3][0] = -vz; E.L[3][1] = 0; E.L[3][2] = 0;  E.L[3][3] = 1; E.L[3][4] = 0;
    E.R[0][2] = 0; E.R[1][2] = 0; E.R[2][2] = -1; E.R[3][2] = 0; E.R[4][2] = -vy;
    E.R[0][3] = 0; E.R[1][3] = 0; E.R[2][3] = 0;  E.R[3][3] = 1; E.R[4][3] = vz;
  } else if (DIR == 1) {
    E.L[1][0] = -vx; E.L[1][1] = 1; E.L[1][2] = 0; E.L[1][3] = 0;  E.L[1][4] = 0;
    E.L[3][0] = vz;  E.L[3][1] = 0; E.L[3][2] = 0; E.L[3][3] = -1; E.L[3][4] = 0;
    E.R[0][1] = 0; E.R[1][1] = 1; E.R[2][1] = 0; E.R[3][1] = 0;  E.R[4][1] = vx;
    E.R[0][3] = 0; E.R[1][3] = 0; E.R[2][3] = 0; E.R[3][3] = -1; E.R[4][3] = -vz;
  } else {
    E.L[1][0] = vx;  E.L[1][1] = -1; E.L[1][2] = 0; E.L[1][3] = 0; E.L[1][4] = 0;
    E.L[2][0] = -vy; E.L[2][1] = 0;  E.L[2][2] = 1; E.L[2][3] = 0; E.L[2][4] = 0;
    E.R[0][1] = 0; E.R[1][1] = -1; E.R[2][1] = 0; E.R[3][1] = 0; E.R[4][1] = -vx;
    E.R[0][2] = 0; E.R[1][2] = 0;  E.R[2][2] = 1; E.R[3][2] = 0; E.R[4][2] = vy;
  }
  for (int r = 0; r < NV; r++) E.lam[r] = vn;
  E.lam[m] = vn - a;
  E.lam[4] = vn + a;
}

// |lambda| with Harten's entropy fix, delta = 1e-6
// (NavierStokes3DUpwind.c:95-101)
__device__ __forceinline__ double harten_abs(double l) {
  const double delta = 0.000001, delta2 = delta * delta;
  const double al = l < 0 ? -l : l;
  return al < delta ? (l * l + delta2) / (2 * delta) : al;
}

// WENO5 interpolant.  (m3,m2,m1,p1,p2) are the stencil values the
// smoothness indicators use, (g3,g2,g1,q1,q2) the values that are
// interpolated (they differ only for the solution, whose weights come from u
// but whose interpolant uses the modified solution uC).  `type` selects the
// weights (0 Jiang-Shu, 1 mapped, 2 WENO-Z, 3 Yamaleev-Carpenter;
// include/interpolation.h:265-500; 4 the frozen optimal weights of
// no_limiting): JS, Z and YC are applied in ratio form (one division), the
// mapped weights as the reference forms them.
__device__ __forceinline__ double weno5(double m3, double m2, double m1, double p1, double p2,
                                        double g3, double g2, double g1, double q1, double q2,
                                        double eps, int type = 0) {
  const double c1312 = 13.0 / 12.0, c14 = 1.0 / 4.0, s6 = 1.0 / 6.0;
  double t, r;
  t = m3 - 2 * m2 + m1; r = m3 - 4 * m2 + 3 * m1;
  const double b1 = c1312 * t * t + c14 * r * r;
  t = m2 - 2 * m1 + p1; r = m2 - p1;
  const double b2 = c1312 * t * t + c14 * r * r;
  t = m1 - 2 * p1 + p2; r = 3 * m1 - 4 * p1 + p2;
  const double b3 = c1312 * t * t + c14 * r * r;
  const double f1 = (2 * s6) * g3 + (-7 * s6) * g2 + (11 * s6) * g1;
  const double f2 = (-s6) * g2 + (5 * s6) * g1 + (2 * s6) * q1;
  const double f3 = (2 * s6) * g1 + (5 * s6) * q1 + (-s6) * q2;
  if (type == 4) return 0.1 * f1 + 0.6 * f2 + 0.3 * f3;
  const double s1 = (b1 + eps) * (b1 + eps);
  const double s2 = (b2 + eps) * (b2 + eps);
  const double s3 = (b3 + eps) * (b3 + eps);
  if (type == 1) {
    // mapped (Henrick, Aslam & Powers): JS weights w mapped by
    // w (c + c^2 - 3 c w + w^2) / (c^2 + w (1 - 2c)), normalised again
    const double c1 = 0.1, c2 = 0.6, c3 = 0.3;
    const double a1 = c1 / s1, a2 = c2 / s2, a3 = c3 / s3;
    const double in = 1.0 / (a1 + a2 + a3);
    const double w1 = a1 * in, w2 = a2 * in, w3 = a3 * in;
    const double d1 = w1 * (c1 + c1 * c1 - 3 * c1 * w1 + w1 * w1) / (c1 * c1 + w1 * (1.0 - 2.0 * c1));
    const double d2 = w2 * (c2 + c2 * c2 - 3 * c2 * w2 + w2 * w2) / (c2 * c2 + w2 * (1.0 - 2.0 * c2));
    const double d3 = w3 * (c3 + c3 * c3 - 3 * c3 * w3 + w3 * w3) / (c3 * c3 + w3 * (1.0 - 2.0 * c3));
    return (d1 * f1 + d2 * f2 + d3 * f3) / (d1 + d2 + d3);
  }
  if (type == 2 || type == 3) {
    // a_k = c_k (1 + (tau/(b_k+eps))^2), multiplied through by s1 s2 s3
    double tau;
    if (type == 2) {
      tau = fabs(b3 - b1);
    } else {
      const double d4 = m3 - 4 * m2 + 6 * m1 - 4 * p1 + p2;
      tau = d4 * d4;
    }
    const double t2 = tau * tau;
    const double a1 = 0.1 * ((s1 + t2) * (s2 * s3)), a2 = 0.6 * ((s2 + t2) * (s1 * s3)),
                 a3 = 0.3 * ((s3 + t2) * (s1 * s2));
    return (a1 * f1 + a2 * f2 + a3 * f3) / (a1 + a2 + a3);
  }
  const double a1 = 0.1 * (s2 * s3), a2 = 0.6 * (s1 * s3), a3 = 0.3 * (s1 * s2);
  return (a1 * f1 + a2 * f2 + a3 * f3) / (a1 + a2 + a3);
}

// periodic index, i in [-n, 2n)
__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Remap the hardware block id so that every XCD works through one
// contiguous range of logical blocks: neighbouring tiles, which share
// stencil rows, then hit the same L2.
__device__ __forceinline__ size_t xcd_block() {
  const unsigned b = blockIdx.x, nb = gridDim.x;
  const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
  return (size_t)((x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k);
}

// Persistent blocks striding over nitems work items: XCD x of the 8 takes a
// contiguous range of the items, which its blocks stride through, so items
// that share cache lines (neighbouring grid lines) meet in one L2.  Needs
// gridDim.x >= min(nitems, 8).
struct XcdItems {
  unsigned first, end, step;
};
__device__ __forceinline__ XcdItems xcd_items(unsigned nitems) {
  const unsigned b = blockIdx.x, nb = gridDim.x, x = b & 7u, k = b >> 3;
  const unsigned qi = nitems >> 3, ri = nitems & 7u;
  const unsigned first = x < ri ? x * (qi + 1) : ri * (qi + 1) + (x - ri) * qi;
  const unsigned len = qi + (x < ri ? 1u : 0u);
  const unsigned qb = nb >> 3, rb = nb & 7u;
  return {first + k, first + len, qb + (x < rb ? 1u : 0u)};
}

// 4th-order central difference (FirstDerivativeFourthOrder.c:103)
__device__ __forceinline__ double d4c(double fm2, double fm1, double fp1, double fp2) {
  return (fm2 - 8 * fm1 + 8 * fp1 - fp2) * (1.0 / 12.0);
}

}  // namespace hm
