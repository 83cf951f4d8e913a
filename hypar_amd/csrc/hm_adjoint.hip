// hm_adjoint.hip -- discrete adjoint of the WENO5 + RK4 step (gfx950, FP64).
//
// The reference has no adjoint; this differentiates this library's own
// discretisation of the NavierStokes3D right-hand side by hand, in reverse
// mode: component-wise and characteristic-wise WENO5 (JS, mapped, Z and YC
// weights) with Roe or Rusanov upwinding (the operator of
// src/HyParFunctions/HyperbolicFunction.c:31-222, NavierStokes3DUpwind.c:40-124,
// 349-420), and the viscous term (NavierStokes3DParabolicFunction.c:50-325).
//
// rhs_i = -sum_d dxinv_i (F_{i+1/2} - F_{i-1/2}) along each direction, so the
// adjoint of interface flux F_{i-1/2} is  g = dxinv_i lam_i - dxinv_{i-1} lam_{i-1}.
// The interface-flux VJP spreads g over the 6 stencil cells as a flux-adjoint
// part (later multiplied by the flux Jacobian transpose of the cell, once per
// cell) and a direct solution-adjoint part (WENO of the solution, Rusanov
// speed).  One workgroup owns one grid line along the direction: the line's
// states, fluxes and modified energies live in LDS (sized to the line), each
// thread is one interface, and each thread writes its six stencil
// contributions to LDS slots that the owning cell then sums in a fixed
// order, so the result is deterministic.
//
// Within an RK4 step the stage adjoint mu_s = dt b_s lam + dt a_{s+1,s}
// nu_{s+1} is formed on the fly from lam and the previous stage's nu, and
// the last direction adds the finished nu_s into the step adjoint.
//
// Non-smooth points (the max of the Rusanov speeds, |v|) take the derivative
// of the branch the forward evaluation selects.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

constexpr int VJP_LINE_MAX = 512;  // longest grid line an adjoint workgroup holds
constexpr int VJP_LDS = 23;        // doubles of LDS per cell of the component-wise line VJP

__device__ __forceinline__ int wrapn(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Lines per workgroup of the line VJPs: a block of `threads` threads (a
// multiple of 64, at least n) takes threads / n consecutive lines of length
// n, so that lines shorter than the block do not leave wavefronts idle (two
// 64-cell lines per 128-thread block).
__host__ __device__ __forceinline__ int vjp_lines_per_block(int threads, int n) { return threads / n; }

// the line and interface of this thread for work item `item` (a group of
// vjp_lines_per_block lines) of a direction with nlines lines of length n;
// sub is the thread's line within the group, whose share of the dynamic LDS
// starts sub * (doubles per cell) * n in.  live is false for threads past the
// group's last line.
struct LineSlot {
  size_t line;
  int t, sub;
  bool live;
};
__device__ __forceinline__ LineSlot line_slot(size_t item, int n, size_t nlines) {
  const int lpb = vjp_lines_per_block((int)blockDim.x, n), sub = (int)threadIdx.x / n;
  LineSlot s;
  s.line = item * lpb + sub;
  s.live = sub < lpb && s.line < nlines;
  s.t = (int)threadIdx.x - sub * n;
  s.sub = s.live ? sub : 0;
  if (!s.live) s.line = 0;
  return s;
}

// 1/d, d > 0 and normal: hardware estimate and two Newton-Raphson steps
__device__ __forceinline__ double vjp_rcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
}

// one side of a WENO5 interpolation from its smoothness indicators beta_k
// and candidate values f_k: the interpolated value, the weights, ob times
// d out/d beta_k (g1..g3) and, for the YC weights, ob times d out/d tau (gt)
// with tau the YC indicator passed in.  W selects the weights
// (include/interpolation.h:265-285, 314-341, 426-447, 479-500):
//  * JS:     a_k = c_k/(beta_k + eps)^2;
//            d out/d beta_k = -2 w_k (f_k - out)/(beta_k + eps)
//  * Z, YC:  a_k = c_k (1 + (t/(beta_k + eps))^2), t = |beta_3 - beta_1| (Z)
//            or tau (YC); t's own derivative is folded into g1, g3 (Z) or
//            returned in gt (YC)
//  * mapped: the JS weights u_k mapped by u (c + c^2 - 3 c u + u^2)/
//            (c^2 + u (1 - 2c)) and renormalised; the adjoint goes back
//            through the mapping and the JS normalisation
//  * optimal (no_limiting): the constant weights 0.1, 0.6, 0.3
// The a_k are formed multiplied through by their common denominator; one
// reciprocal of (sum a_k)(prod (beta_k + eps)) gives both 1/sum a_k and the
// 1/(beta_k + eps) (the mapped weights take a second one for the mapping).
struct W5Side {
  double out, w1, w2, w3, g1, g2, g3, gt;
};

template <int W>
__device__ __forceinline__ W5Side w5_side(double b1, double b2, double b3, double f1, double f2, double f3,
                                          double eps, double ob, double tau) {
  const double e1 = b1 + eps, e2 = b2 + eps, e3 = b3 + eps;
  const double s1 = e1 * e1, s2 = e2 * e2, s3 = e3 * e3;
  W5Side r;
  r.gt = 0.0;
  if (W == WENO_OPTIMAL) {
    // no_limiting: constant optimal weights, no dependence on beta_k
    r.w1 = 0.1;
    r.w2 = 0.6;
    r.w3 = 0.3;
    r.out = 0.1 * f1 + 0.6 * f2 + 0.3 * f3;
    r.g1 = r.g2 = r.g3 = 0.0;
  } else if (W == WENO_JS) {
    const double a1 = 0.1 * (s2 * s3), a2 = 0.6 * (s1 * s3), a3 = 0.3 * (s1 * s2);
    const double A = a1 + a2 + a3, e23 = e2 * e3, P = e1 * e23;
    const double q = vjp_rcp(A * P);
    const double inv = q * P, k = -2.0 * ob * (q * A);
    r.w1 = a1 * inv;
    r.w2 = a2 * inv;
    r.w3 = a3 * inv;
    r.out = r.w1 * f1 + r.w2 * f2 + r.w3 * f3;
    r.g1 = k * r.w1 * (f1 - r.out) * e23;
    r.g2 = k * r.w2 * (f2 - r.out) * (e1 * e3);
    r.g3 = k * r.w3 * (f3 - r.out) * (e1 * e2);
  } else if (W == WENO_Z || W == WENO_YC) {
    const double t = W == WENO_Z ? fabs(b3 - b1) : tau, t2 = t * t;
    const double p1 = s2 * s3, p2 = s1 * s3, p3 = s1 * s2;  // products of the other two s_k
    const double a1 = 0.1 * (s1 + t2) * p1, a2 = 0.6 * (s2 + t2) * p2, a3 = 0.3 * (s3 + t2) * p3;
    const double A = a1 + a2 + a3, e23 = e2 * e3, P = e1 * e23;
    const double q = vjp_rcp(A * P);
    const double inv = q * P;
    r.w1 = a1 * inv;
    r.w2 = a2 * inv;
    r.w3 = a3 * inv;
    r.out = r.w1 * f1 + r.w2 * f2 + r.w3 * f3;
    // d out/d a_k (unscaled a_k) = s1 s2 s3 (f_k - out)/A, d a_k/d beta_k =
    // -2 c_k t^2/(beta_k + eps)^3 and d a_k/d t = 2 c_k t/(beta_k + eps)^2
    const double h1 = ob * (f1 - r.out), h2 = ob * (f2 - r.out), h3 = ob * (f3 - r.out);
    r.g1 = (-2.0 * 0.1) * t2 * h1 * p1 * e23 * q;
    r.g2 = (-2.0 * 0.6) * t2 * h2 * p2 * (e1 * e3) * q;
    r.g3 = (-2.0 * 0.3) * t2 * h3 * p3 * (e1 * e2) * q;
    const double gt = 2.0 * t * inv * (0.1 * h1 * p1 + 0.6 * h2 * p2 + 0.3 * h3 * p3);
    if (W == WENO_Z) {
      // t = |beta_3 - beta_1|; at beta_3 = beta_1 the derivative is taken as 0
      const double sg = b3 > b1 ? 1.0 : (b3 < b1 ? -1.0 : 0.0);
      r.g3 += sg * gt;
      r.g1 -= sg * gt;
    } else {
      r.gt = gt;
    }
  } else {
    const double c1 = 0.1, c2 = 0.6, c3 = 0.3;
    const double at1 = c1 * (s2 * s3), at2 = c2 * (s1 * s3), at3 = c3 * (s1 * s2);
    const double At = at1 + at2 + at3, e23 = e2 * e3, P = e1 * e23;
    const double q = vjp_rcp(At * P);
    const double it = q * P;
    const double u1 = at1 * it, u2 = at2 * it, u3 = at3 * it;  // JS weights
    const double N1 = c1 + c1 * c1 - 3 * c1 * u1 + u1 * u1, D1 = c1 * c1 + u1 * (1.0 - 2.0 * c1);
    const double N2 = c2 + c2 * c2 - 3 * c2 * u2 + u2 * u2, D2 = c2 * c2 + u2 * (1.0 - 2.0 * c2);
    const double N3 = c3 + c3 * c3 - 3 * c3 * u3 + u3 * u3, D3 = c3 * c3 + u3 * (1.0 - 2.0 * c3);
    const double o1 = D2 * D3, o2 = D1 * D3, o3 = D1 * D2, DD = D1 * o1;
    const double B1 = u1 * N1 * o1, B2 = u2 * N2 * o2, B3 = u3 * N3 * o3, B = B1 + B2 + B3;
    const double qm = vjp_rcp(B * DD);
    const double im = qm * DD;
    r.w1 = B1 * im;
    r.w2 = B2 * im;
    r.w3 = B3 * im;
    r.out = r.w1 * f1 + r.w2 * f2 + r.w3 * f3;
    // d out/d a_k = DD (f_k - out)/B with a_k = u_k N_k/D_k, whose derivative
    // is ((N_k + u_k N_k') D_k - u_k N_k D_k')/D_k^2; DD/(B D_k^2) = o_k^2 qm
    const double h1 = ob * (f1 - r.out), h2 = ob * (f2 - r.out), h3 = ob * (f3 - r.out);
    const double du1 = h1 * ((N1 + u1 * (2 * u1 - 3 * c1)) * D1 - u1 * N1 * (1.0 - 2.0 * c1)) * (o1 * o1) * qm;
    const double du2 = h2 * ((N2 + u2 * (2 * u2 - 3 * c2)) * D2 - u2 * N2 * (1.0 - 2.0 * c2)) * (o2 * o2) * qm;
    const double du3 = h3 * ((N3 + u3 * (2 * u3 - 3 * c3)) * D3 - u3 * N3 * (1.0 - 2.0 * c3)) * (o3 * o3) * qm;
    // through u_k = at_k/sum at, at_k = c_k/(beta_k + eps)^2
    const double U = du1 * u1 + du2 * u2 + du3 * u3, iP = q * At;
    r.g1 = -2.0 * (du1 - U) * u1 * e23 * iP;
    r.g2 = -2.0 * (du2 - U) * u2 * (e1 * e3) * iP;
    r.g3 = -2.0 * (du3 - U) * u3 * (e1 * e2) * iP;
  }
  return r;
}

// reverse of the left- and right-biased WENO5 of one interface, with
// smoothness indicators from b[0..5] and interpolated values from g[0..5]
// (cells I-3..I+2; g = b except for the modified energy, a function of the
// same state): adds obL d vl/d(b, g) + obR d vr/d(b, g) to bb and returns
// vl, vr.  The second differences the two sides share are formed once and
// their adjoints summed.
template <int W>
__device__ __forceinline__ void weno5_lr_vjp(const double b[6], const double g[6], double eps, double obL,
                                             double obR, double bb[6], double &vl, double &vr) {
  const double c1312 = 13.0 / 12.0, c14 = 0.25, s6 = 1.0 / 6.0;
  const double d1 = b[0] - 2 * b[1] + b[2], d2 = b[1] - 2 * b[2] + b[3];
  const double d3 = b[2] - 2 * b[3] + b[4], d4 = b[3] - 2 * b[4] + b[5];
  const double q2 = c1312 * (d2 * d2), q3 = c1312 * (d3 * d3);
  const double rl1 = b[0] - 4 * b[1] + 3 * b[2], rl2 = b[1] - b[3], rl3 = 3 * b[2] - 4 * b[3] + b[4];
  const double rr1 = b[5] - 4 * b[4] + 3 * b[3], rr2 = b[4] - b[2], rr3 = 3 * b[3] - 4 * b[2] + b[1];
  // YC: undivided 4th differences of the two stencils, tau = their squares
  const double tl = W == WENO_YC ? b[0] - 4 * b[1] + 6 * b[2] - 4 * b[3] + b[4] : 0.0;
  const double tr = W == WENO_YC ? b[5] - 4 * b[4] + 6 * b[3] - 4 * b[2] + b[1] : 0.0;
  // left: second differences d1, d2, d3; right (mirrored): d4, d3, d2
  const W5Side L = w5_side<W>(c1312 * (d1 * d1) + c14 * (rl1 * rl1), q2 + c14 * (rl2 * rl2),
                              q3 + c14 * (rl3 * rl3), (2 * s6) * g[0] + (-7 * s6) * g[1] + (11 * s6) * g[2],
                              (-s6) * g[1] + (5 * s6) * g[2] + (2 * s6) * g[3],
                              (2 * s6) * g[2] + (5 * s6) * g[3] + (-s6) * g[4], eps, obL, tl * tl);
  const W5Side R = w5_side<W>(c1312 * (d4 * d4) + c14 * (rr1 * rr1), q3 + c14 * (rr2 * rr2),
                              q2 + c14 * (rr3 * rr3), (2 * s6) * g[5] + (-7 * s6) * g[4] + (11 * s6) * g[3],
                              (-s6) * g[4] + (5 * s6) * g[3] + (2 * s6) * g[2],
                              (2 * s6) * g[3] + (5 * s6) * g[2] + (-s6) * g[1], eps, obR, tr * tr);
  vl = L.out;
  vr = R.out;
  // interpolated values
  const double wl1 = obL * L.w1, wl2 = obL * L.w2, wl3 = obL * L.w3;
  const double wr1 = obR * R.w1, wr2 = obR * R.w2, wr3 = obR * R.w3;
  // smoothness indicators: second differences and the one-sided differences
  const double D1 = (2 * c1312) * d1 * L.g1, D2 = (2 * c1312) * d2 * (L.g2 + R.g3);
  const double D3 = (2 * c1312) * d3 * (L.g3 + R.g2), D4 = (2 * c1312) * d4 * R.g1;
  const double L1 = (2 * c14) * rl1 * L.g1, L2 = (2 * c14) * rl2 * L.g2, L3 = (2 * c14) * rl3 * L.g3;
  const double R1 = (2 * c14) * rr1 * R.g1, R2 = (2 * c14) * rr2 * R.g2, R3 = (2 * c14) * rr3 * R.g3;
  bb[0] += (2 * s6) * wl1 + D1 + L1;
  bb[1] += (-7 * s6) * wl1 + (-s6) * wl2 + (-s6) * wr3 - 2 * D1 + D2 - 4 * L1 + L2 + R3;
  bb[2] += (11 * s6) * wl1 + (5 * s6) * wl2 + (2 * s6) * wl3 + (2 * s6) * wr2 + (5 * s6) * wr3 + D1 - 2 * D2 + D3 +
           3 * L1 + 3 * L3 - R2 - 4 * R3;
  bb[3] += (2 * s6) * wl2 + (5 * s6) * wl3 + (11 * s6) * wr1 + (5 * s6) * wr2 + (2 * s6) * wr3 + D2 - 2 * D3 + D4 -
           L2 - 4 * L3 + 3 * R1 + 3 * R3;
  bb[4] += (-s6) * wl3 + (-7 * s6) * wr1 + (-s6) * wr2 + D3 - 2 * D4 + L3 - 4 * R1 + R2;
  bb[5] += (2 * s6) * wr1 + D4 + R1;
  if (W == WENO_YC) {
    const double TL = 2.0 * tl * L.gt, TR = 2.0 * tr * R.gt;
    bb[0] += TL;
    bb[1] += -4 * TL + TR;
    bb[2] += 6 * TL - 4 * TR;
    bb[3] += -4 * TL + 6 * TR;
    bb[4] += TL - 4 * TR;
    bb[5] += TR;
  }
}

// gradient of c + |v_d| of one state with respect to its conserved variables
template <int DIR>
__device__ __forceinline__ void speed_vjp(const double *u, double gamma, double ab, double ub[NV]) {
  const double gm1 = gamma - 1.0;
  const double rho = u[0], ir = 1.0 / rho;
  const double v[3] = {u[1] * ir, u[2] * ir, u[3] * ir};
  const double q2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const double Por = gm1 * (u[4] * ir - 0.5 * q2);  // P / rho
  const double c = sqrt(gamma * Por);
  const double k = ab * gamma / (2.0 * c);           // d a / d(P/rho)
  ub[0] += k * gm1 * (-u[4] * ir * ir + q2 * ir);
  for (int j = 0; j < 3; j++) ub[1 + j] += k * gm1 * (-v[j] * ir);
  ub[4] += k * gm1 * ir;
  const double vn = v[DIR];
  const double sg = vn > 0 ? 1.0 : (vn < 0 ? -1.0 : 0.0);
  ub[0] += ab * (-sg * vn * ir);
  ub[1 + DIR] += ab * (sg * ir);
}

// adjoint of the Roe-averaged velocity (gv) and total enthalpy (gH) of
// (uL, uR) onto the two states (navierstokes3d.h:144-169)
__device__ __forceinline__ void roe_avg_vjp(const double *uL, const double *uR, double gamma, const double gv[3],
                                            double gH, double ubL[NV], double ubR[NV]) {
  const double gm1 = gamma - 1.0;
  const double *U[2] = {uL, uR};
  double t[2], v[2][3], H[2], rho[2], q2[2];
  for (int s = 0; s < 2; s++) {
    rho[s] = U[s][0];
    const double ir = 1.0 / rho[s];
    for (int j = 0; j < 3; j++) v[s][j] = U[s][1 + j] * ir;
    q2[s] = v[s][0] * v[s][0] + v[s][1] * v[s][1] + v[s][2] * v[s][2];
    H[s] = gamma * U[s][4] * ir - gm1 * 0.5 * q2[s];
    t[s] = sqrt(rho[s]);
  }
  const double den = 1.0 / (t[0] + t[1]);
  double va[3];
  for (int j = 0; j < 3; j++) va[j] = (t[0] * v[0][j] + t[1] * v[1][j]) * den;
  const double Ha = (t[0] * H[0] + t[1] * H[1]) * den;
  double *UB[2] = {ubL, ubR};
  for (int s = 0; s < 2; s++) {
    const double w = t[s] * den, ir = 1.0 / rho[s];
    double gt = gH * (H[s] - Ha) * den;
    for (int j = 0; j < 3; j++) gt += gv[j] * (v[s][j] - va[j]) * den;
    const double gHs = gH * w;
    double gvs[3];
    for (int j = 0; j < 3; j++) gvs[j] = gv[j] * w;
    UB[s][0] += gt / (2.0 * t[s]) + gHs * (-gamma * U[s][4] * ir * ir + gm1 * q2[s] * ir);
    for (int j = 0; j < 3; j++) {
      UB[s][0] += gvs[j] * (-v[s][j] * ir);
      UB[s][1 + j] += gvs[j] * ir + gHs * (-gm1 * v[s][j] * ir);
    }
    UB[s][4] += gHs * gamma * ir;
  }
}

// gradient of c + |v_d| of the Roe average of (uL, uR)
template <int DIR>
__device__ __forceinline__ void roe_speed_vjp(const double *uL, const double *uR, double gamma, double ab,
                                              double ubL[NV], double ubR[NV]) {
  const double gm1 = gamma - 1.0;
  double va[3], Ha;
  {
    double t[2], v[2][3], H[2];
    const double *U[2] = {uL, uR};
    for (int s = 0; s < 2; s++) {
      const double ir = 1.0 / U[s][0];
      for (int j = 0; j < 3; j++) v[s][j] = U[s][1 + j] * ir;
      H[s] = gamma * U[s][4] * ir - gm1 * 0.5 * (v[s][0] * v[s][0] + v[s][1] * v[s][1] + v[s][2] * v[s][2]);
      t[s] = sqrt(U[s][0]);
    }
    const double den = 1.0 / (t[0] + t[1]);
    for (int j = 0; j < 3; j++) va[j] = (t[0] * v[0][j] + t[1] * v[1][j]) * den;
    Ha = (t[0] * H[0] + t[1] * H[1]) * den;
  }
  const double ca = sqrt(gm1 * (Ha - 0.5 * (va[0] * va[0] + va[1] * va[1] + va[2] * va[2])));
  const double gH = ab * gm1 / (2.0 * ca);
  double gv[3];
  for (int j = 0; j < 3; j++) gv[j] = -ab * gm1 * va[j] / (2.0 * ca);
  const double sg = va[DIR] > 0 ? 1.0 : (va[DIR] < 0 ? -1.0 : 0.0);
  gv[DIR] += ab * sg;
  roe_avg_vjp(uL, uR, gamma, gv, gH, ubL, ubR);
}

// ub += A(u)^T fb, A = d(Euler flux along DIR)/d(conserved variables)
template <int DIR>
__device__ __forceinline__ void flux_jac_T(const double *u, double gamma, const double fb[NV], double ub[NV]) {
  const double gm1 = gamma - 1.0;
  const double rho = u[0], ir = 1.0 / rho;
  const double v[3] = {u[1] * ir, u[2] * ir, u[3] * ir};
  const double q2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const double P = gm1 * (u[4] - 0.5 * rho * q2);
  const double H = (u[4] + P) * ir;
  const double vn = v[DIR];
  double fm = 0.0;  // sum_i fb_{1+i} v_i
  for (int i = 0; i < 3; i++) fm += fb[1 + i] * v[i];
  ub[0] += -vn * fm + fb[1 + DIR] * gm1 * 0.5 * q2 + fb[4] * vn * (gm1 * 0.5 * q2 - H);
  for (int j = 0; j < 3; j++)
    ub[1 + j] += fb[1 + j] * vn - fb[1 + DIR] * gm1 * v[j] - fb[4] * gm1 * v[j] * vn;
  ub[1 + DIR] += fb[0] + fm + fb[4] * H;
  ub[4] += fb[1 + DIR] * gm1 + fb[4] * gamma * vn;
}

// Roe state of an interface for the adjoints of Roe upwinding and of the
// characteristic scheme, and the rows of L with their derivatives
struct RoeAdj {
  double v[3], vn, a, a2, ia2, ek, h0;
};

template <int DIR>
__device__ __forceinline__ RoeAdj roe_adj_state(const double uL[NV], const double uR[NV], double gamma) {
  const double gm1 = gamma - 1.0;
  RoeAdj s;
  const double iL = 1.0 / uL[0], iR = 1.0 / uR[0];
  double vL[3], vR[3];
  for (int j = 0; j < 3; j++) { vL[j] = uL[1 + j] * iL; vR[j] = uR[1 + j] * iR; }
  const double qL = vL[0] * vL[0] + vL[1] * vL[1] + vL[2] * vL[2];
  const double qR = vR[0] * vR[0] + vR[1] * vR[1] + vR[2] * vR[2];
  const double HL = gamma * uL[4] * iL - gm1 * 0.5 * qL, HR = gamma * uR[4] * iR - gm1 * 0.5 * qR;
  const double tL = sqrt(uL[0]), tR = sqrt(uR[0]), den = 1.0 / (tL + tR);
  for (int j = 0; j < 3; j++) s.v[j] = (tL * vL[j] + tR * vR[j]) * den;
  const double H = (tL * HL + tR * HR) * den;
  s.ek = 0.5 * (s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2]);
  s.a2 = (H - s.ek) * gm1;
  s.a = sqrt(s.a2);
  s.ia2 = 1.0 / s.a2;
  s.vn = s.v[DIR];
  s.h0 = s.a2 / gm1 + s.ek;
  return s;
}

// row k of L at the Roe state: k = 0 entropy, 1 u-a, 2 u+a, 3 and 4 the two
// shear rows in row order (navierstokes3d.h:288-374)
template <int DIR>
__device__ __forceinline__ void l_row(const RoeAdj &s, double gm1, int k, double r[NV]) {
  if (k <= 2) {
    const double ia2 = s.ia2;
    if (k == 0) {
      r[0] = (s.a2 - gm1 * s.ek) * ia2;
      for (int j = 0; j < 3; j++) r[1 + j] = gm1 * s.v[j] * ia2;
      r[4] = -gm1 * ia2;
    } else {
      const double h = 0.5 * ia2, sg = k == 1 ? 1.0 : -1.0;
      r[0] = (gm1 * s.ek + sg * s.a * s.vn) * h;
      for (int j = 0; j < 3; j++) r[1 + j] = -gm1 * s.v[j] * h;
      r[1 + DIR] -= sg * s.a * h;
      r[4] = gm1 * h;
    }
    return;
  }
  for (int j = 0; j < NV; j++) r[j] = 0.0;
  if (DIR == 0) {
    if (k == 3) { r[0] = s.v[1]; r[2] = -1.0; } else { r[0] = -s.v[2]; r[3] = 1.0; }
  } else if (DIR == 1) {
    if (k == 3) { r[0] = -s.v[0]; r[1] = 1.0; } else { r[0] = s.v[2]; r[3] = -1.0; }
  } else {
    if (k == 3) { r[0] = s.v[0]; r[1] = -1.0; } else { r[0] = -s.v[1]; r[2] = 1.0; }
  }
}

// adds (d row_k / d(v, a))^T X to (gv, ab): the Roe-state derivative of
// sum_p (adjoint of the projection at p) * (row_k . x_p), X = the adjoint-
// weighted sum of the projected vectors
template <int DIR>
__device__ __forceinline__ void l_row_vjp(const RoeAdj &s, double gm1, int k, const double X[NV], double gv[3],
                                          double &ab) {
  const double a = s.a, ia2 = s.ia2, ia = a * ia2, ia3 = ia2 * ia;
  const double vX = s.v[0] * X[1] + s.v[1] * X[2] + s.v[2] * X[3];
  if (k == 0) {
    for (int j = 0; j < 3; j++) gv[j] += gm1 * ia2 * (X[1 + j] - X[0] * s.v[j]);
    ab += 2.0 * gm1 * ia3 * (X[0] * s.ek - vX + X[4]);
  } else if (k <= 2) {
    const double sg = k == 1 ? 1.0 : -1.0, h = 0.5 * ia2;
    for (int j = 0; j < 3; j++) gv[j] += gm1 * h * (X[0] * s.v[j] - X[1 + j]);
    gv[DIR] += sg * a * h * X[0];
    ab += -X[0] * (2.0 * gm1 * s.ek + sg * a * s.vn) * h * ia + gm1 * ia3 * (vX - X[4]) + sg * h * X[1 + DIR];
  } else if (DIR == 0) {
    if (k == 3) gv[1] += X[0]; else gv[2] -= X[0];
  } else if (DIR == 1) {
    if (k == 3) gv[0] -= X[0]; else gv[2] += X[0];
  } else {
    if (k == 3) gv[0] += X[0]; else gv[1] -= X[0];
  }
}

__device__ __forceinline__ double harten_d(double l) {
  const double delta = 0.000001;
  return fabs(l) < delta ? l / delta : (l > 0 ? 1.0 : (l < 0 ? -1.0 : 0.0));
}

// nu (+)= (d rhs_DIR / d u)^T mu along one direction, one workgroup per line,
// with mu = cb*lam + ca*nu_in (nu_in may be null).  mode VJP_WRITE stores,
// VJP_ADD accumulates into nu, VJP_ADD_L accumulates into nu and adds the
// finished nu into L, VJP_ADD_LSET sets L = lam + the finished nu.  LDS (dynamic, 23 doubles per cell of the line):
// state [5][n], cell flux [5][n], modified energy [n], stencil
// contributions of the flux and solution adjoints [6][n] each.
enum { VJP_WRITE = 0, VJP_ADD = 1, VJP_ADD_L = 2, VJP_ADD_LSET = 3 };

// state, lam and (when given) nu_in of cell c, every load issued before any
// is used: a null check inside the variable loop would put a wait in each
// iteration
__device__ __forceinline__ void vjp_load_cell(const double *__restrict__ u, const double *__restrict__ lam,
                                              const double *__restrict__ nu_in, size_t ncell, size_t c,
                                              double (&us)[NV], double (&lm)[NV], double (&ni)[NV]) {
#pragma unroll
  for (int v = 0; v < NV; v++) {
    us[v] = u[(size_t)v * ncell + c];
    lm[v] = lam[(size_t)v * ncell + c];
    ni[v] = 0.0;
  }
  if (nu_in)
#pragma unroll
    for (int v = 0; v < NV; v++) ni[v] = nu_in[(size_t)v * ncell + c];
}

// weight of F(r+s) in D at reader r, s = -2, -1, +1, +2
__device__ __forceinline__ double d4w(int s) {
  return s == -2 ? 1.0 / 12.0 : (s == -1 ? -8.0 / 12.0 : (s == 1 ? 8.0 / 12.0 : -1.0 / 12.0));
}

// The viscous term's share of the adjoint of cell idx, (d par/d u)^T mu: QB
// plus the transposed gradient stencils over GB, then back through
// (u, v, w, T)(u) at the cell's state us (NavierStokes3DParabolicFunction.c:
// 85-108).  QB and GB are complete once the tile and ghost-layer kernels of
// the viscous adjoint have run; the x line VJP, which runs after them, adds
// this share into nu together with its own.
__device__ __forceinline__ void visc_gather_cell(const double (&us)[NV], const double *__restrict__ QB,
                                                 const double *__restrict__ GB, int nx, int ny, int nz,
                                                 double gamma, size_t idx, double (&ub)[NV]) {
  const size_t ncell = (size_t)nx * ny * nz;
  const int n[3] = {nx, ny, nz};
  const int pos[3] = {(int)(idx % nx), (int)((idx / nx) % ny), (int)(idx / ((size_t)nx * ny))};
  const size_t st[3] = {1, (size_t)nx, (size_t)nx * ny};
  double qb[4];
#pragma unroll
  for (int c = 0; c < 4; c++) qb[c] = QB[(size_t)c * ncell + idx];
  const int offs[4] = {-2, -1, 1, 2};
#pragma unroll
  for (int e = 0; e < 3; e++) {
    const size_t line0 = idx - (size_t)pos[e] * st[e];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      // reader j = m - s reads this cell with weight d4w(s) (periodic)
      const size_t j = line0 + (size_t)wrapn(pos[e] - offs[s], n[e]) * st[e];
      const double w = d4w(offs[s]);
#pragma unroll
      for (int c = 0; c < 4; c++) qb[c] += w * GB[(size_t)(c * 3 + e) * ncell + j];
    }
  }
  const double gm1 = gamma - 1.0;
  const double ir = 1.0 / us[0];
  const double vx = us[1] * ir, vy = us[2] * ir, vz = us[3] * ir;
  const double q2 = vx * vx + vy * vy + vz * vz;
  const double k = gamma * gm1 * qb[3];  // T = gamma (gamma-1) (e/rho - |m|^2/(2 rho^2))
  ub[0] = -(qb[0] * vx + qb[1] * vy + qb[2] * vz) * ir + k * (-us[4] * ir * ir + q2 * ir);
  ub[1] = qb[0] * ir - k * vx * ir;
  ub[2] = qb[1] * ir - k * vy * ir;
  ub[3] = qb[2] * ir - k * vz * ir;
  ub[4] = k * ir;
}

// the finished adjoint acc of cell c into nu (and L) by mode, after the
// viscous share vb when has_vb: nu = (nu + vb) + acc, the order of a
// separate viscous pass that adds vb before this one.  Every load issues
// before the first store, so that the stores (nu, L and lam may alias as
// far as the compiler knows) do not serialise the loads
__device__ __forceinline__ void vjp_store(double *__restrict__ nu, double *__restrict__ L,
                                          const double *__restrict__ lam, int mode, size_t ncell, size_t c,
                                          const double (&acc)[NV], bool has_vb, const double (&vb)[NV]) {
  if (mode == VJP_WRITE) {
#pragma unroll
    for (int v = 0; v < NV; v++) nu[(size_t)v * ncell + c] = has_vb ? vb[v] + acc[v] : acc[v];
    return;
  }
  double r[NV], l[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int v = 0; v < NV; v++) r[v] = nu[(size_t)v * ncell + c];
  if (mode == VJP_ADD_L) {
#pragma unroll
    for (int v = 0; v < NV; v++) l[v] = L[(size_t)v * ncell + c];
  } else if (mode == VJP_ADD_LSET) {
#pragma unroll
    for (int v = 0; v < NV; v++) l[v] = lam[(size_t)v * ncell + c];
  }
  if (has_vb)
#pragma unroll
    for (int v = 0; v < NV; v++) r[v] = r[v] + vb[v];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    r[v] = r[v] + acc[v];
    nu[(size_t)v * ncell + c] = r[v];
  }
  if (mode == VJP_ADD_L || mode == VJP_ADD_LSET)
#pragma unroll
    for (int v = 0; v < NV; v++) L[(size_t)v * ncell + c] = l[v] + r[v];
}

template <int DIR, int UPW, int W, bool VG = false>
__device__ __forceinline__ void rhs_vjp_line(const double *__restrict__ u, const double *__restrict__ lam,
                                             const double *__restrict__ nu_in, double ca, double cb,
                                             double *__restrict__ nu, double *__restrict__ L, int mode, int nx,
                                             int ny, int nz, const double *__restrict__ dxg, double gamma,
                                             double eps, const double *__restrict__ QB,
                                             const double *__restrict__ GB, size_t line, int t, bool active,
                                             double *smem) {
  const size_t ncell = (size_t)nx * ny * nz;
  const int n = DIR == 0 ? nx : (DIR == 1 ? ny : nz);
  size_t base, stride;
  if (DIR == 0) { base = line * nx; stride = 1; }
  else if (DIR == 1) { base = (line % nx) + (line / nx) * (size_t)nx * ny; stride = nx; }
  else { base = line; stride = (size_t)nx * ny; }
  const double *dxd = dxg + 3 + (DIR >= 1 ? nx + 6 : 0) + (DIR == 2 ? ny + 6 : 0);
  double *sU = smem;            // [5][n]
  double *sFl = sU + 5 * n;     // [5][n]
  double *sC4 = sFl + 5 * n;    // [n]
  double *cF = sC4 + n;         // [6][n]
  double *cU = cF + 6 * n;      // [6][n]
  const double inv_gm1 = 1.0 / (gamma - 1.0);
  const int tm = active ? wrapn(t - 1, n) : 0;
  double g[NV] = {0, 0, 0, 0, 0};
  if (active) {
    const size_t c = base + (size_t)t * stride;
    double us[NV], lm[NV], ni[NV];
    vjp_load_cell(u, lam, nu_in, ncell, c, us, lm, ni);
#pragma unroll
    for (int v = 0; v < NV; v++) {
      sU[v * n + t] = us[v];
      // dxinv mu of this cell, staged in the (not yet used) slot buffer
      cF[v * n + t] = dxd[t] * (cb * lm[v] + (nu_in ? ca * ni[v] : 0.0));
    }
    const Prim q = flowvar(us, gamma);
    double fl[NV];
    euler_flux<DIR>(fl, q);
#pragma unroll
    for (int v = 0; v < NV; v++) sFl[v * n + t] = fl[v];
    sC4[t] = modified_energy(q, inv_gm1);
  }
  __syncthreads();
  if (active)
#pragma unroll
    for (int v = 0; v < NV; v++) g[v] = cF[v * n + t] - cF[v * n + tm];
  __syncthreads();
  int cs[6];
  double alpha = 0.0, abar = 0.0;
  int sel = 0;
  // Roe: d = R |Lambda| L du with du = (uR - uL)/2 and fI = fh - d
  // (NavierStokes3DUpwind.c:40-124); wR = A^T lambda_F, A = R |Lambda| L,
  // so that the solution WENO adjoints are -/+ wR/2
  double wR[NV] = {0, 0, 0, 0, 0}, du[NV] = {0, 0, 0, 0, 0};
  if (active && UPW == UPWIND_ROE) {
#pragma unroll
    for (int s = 0; s < 6; s++) cs[s] = wrapn(t + s - 3, n);
    double u2[NV], u3[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) { u2[v] = sU[v * n + cs[2]]; u3[v] = sU[v * n + cs[3]]; }
    const RoeAdj s = roe_adj_state<DIR>(u2, u3, gamma);
    const double gm1 = gamma - 1.0;
    const double eb = g[4];
    const double sumb = g[0] + s.v[0] * g[1] + s.v[1] * g[2] + s.v[2] * g[3];
    const double difb = g[1 + DIR] + s.vn * eb;
    double lg[5];  // (R^T lambda_F)_k |lambda_k|
    lg[0] = (sumb + s.ek * eb) * harten_abs(s.vn);
    lg[1] = (sumb + s.h0 * eb - s.a * difb) * harten_abs(s.vn - s.a);
    lg[2] = (sumb + s.h0 * eb + s.a * difb) * harten_abs(s.vn + s.a);
    if (DIR == 0) { lg[3] = -g[2] - s.v[1] * eb; lg[4] = g[3] + s.v[2] * eb; }
    else if (DIR == 1) { lg[3] = g[1] + s.v[0] * eb; lg[4] = -g[3] - s.v[2] * eb; }
    else { lg[3] = -g[1] - s.v[0] * eb; lg[4] = g[2] + s.v[1] * eb; }
    lg[3] *= harten_abs(s.vn);
    lg[4] *= harten_abs(s.vn);
#pragma unroll 1
    for (int k = 0; k < 5; k++) {
      double row[NV];
      l_row<DIR>(s, gm1, k, row);
      const double c = k == 0 ? lg[0] : (k == 1 ? lg[1] : (k == 2 ? lg[2] : (k == 3 ? lg[3] : lg[4])));
#pragma unroll
      for (int v = 0; v < NV; v++) wR[v] += c * row[v];
    }
  }
  if (active && UPW == UPWIND_RUSANOV) {
#pragma unroll
    for (int s = 0; s < 6; s++) cs[s] = wrapn(t + s - 3, n);
    double u2[NV], u3[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) { u2[v] = sU[v * n + cs[2]]; u3[v] = sU[v * n + cs[3]]; }
    const Prim q2 = flowvar(u2, gamma), q3 = flowvar(u3, gamma);
    const Prim a = roe_average(u2, u3, gamma);
    const double aL = sqrt(gamma * q2.P / q2.rho) + fabs(vel<DIR>(q2));
    const double aR = sqrt(gamma * q3.P / q3.rho) + fabs(vel<DIR>(q3));
    const double aA = sqrt(gamma * a.P / a.rho) + fabs(vel<DIR>(a));
    const double m01 = fmax(aL, aR), m12 = fmax(aR, aA);
    alpha = fmax(m01, m12);
    // alpha = max(max(aL,aR), max(aR,aA)): derivative of the selected term
    sel = m01 >= m12 ? (aL >= aR ? 0 : 1) : (aR >= aA ? 1 : 2);
  }
  double accF[NV], accU[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    if (active) {
      double fb[6] = {0, 0, 0, 0, 0, 0}, ub[6] = {0, 0, 0, 0, 0, 0};
      double F[6], U[6], C[6];
#pragma unroll
      for (int s = 0; s < 6; s++) {
        F[s] = sFl[v * n + cs[s]];
        U[s] = sU[v * n + cs[s]];
        C[s] = v == 4 ? sC4[cs[s]] : U[s];
      }
      double fl, fr, uLv, uRv;
      // fluxes: smoothness and values are the same stencil
      weno5_lr_vjp<W>(F, F, eps, 0.5 * g[v], 0.5 * g[v], fb, fl, fr);
      // solution: weights from u, values from the modified solution (the
      // same function of u, so both adjoints land on u)
      const double ob = UPW == UPWIND_ROE ? 0.5 * wR[v] : 0.5 * alpha * g[v];
      weno5_lr_vjp<W>(U, C, eps, ob, -ob, ub, uLv, uRv);
      if (UPW == UPWIND_ROE) du[v] = 0.5 * (uRv - uLv);
      else abar += -0.5 * (uRv - uLv) * g[v];
      // slot s of interface t belongs to cell t-3+s
#pragma unroll
      for (int s = 0; s < 6; s++) {
        cF[s * n + t] = fb[s];
        cU[s * n + t] = ub[s];
      }
    }
    __syncthreads();
    // cell t sums slot s of interface t+3-s, s = 0..5, in that order
    accF[v] = accU[v] = 0.0;
    if (active) {
#pragma unroll
      for (int s = 0; s < 6; s++) {
        const int i = wrapn(t + 3 - s, n);
        accF[v] += cF[s * n + i];
        accU[v] += cU[s * n + i];
      }
    }
    __syncthreads();
  }
  // adjoint of the upwinding's dependence on the two cells next to interface
  // t, cell t-1 (slot in cF) and cell t (slot in cU): the Rusanov speed, or
  // the Roe state inside R |Lambda| L
  if (active) {
    double u2[NV], u3[NV], ub2[NV] = {0, 0, 0, 0, 0}, ub3[NV] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < NV; v++) { u2[v] = sU[v * n + cs[2]]; u3[v] = sU[v * n + cs[3]]; }
    if (UPW == UPWIND_ROE) {
      const RoeAdj s = roe_adj_state<DIR>(u2, u3, gamma);
      const double gm1 = gamma - 1.0;
      // adjoint of d is -lambda_F; lambda_g = R^T (-lambda_F)
      double ld[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) ld[v] = -g[v];
      const double eb = ld[4];
      const double sumb = ld[0] + s.v[0] * ld[1] + s.v[1] * ld[2] + s.v[2] * ld[3];
      const double difb = ld[1 + DIR] + s.vn * eb;
      double lgk[5], lam_k[5], gk[5], pk[5];
      lgk[0] = sumb + s.ek * eb;
      lgk[1] = sumb + s.h0 * eb - s.a * difb;
      lgk[2] = sumb + s.h0 * eb + s.a * difb;
      if (DIR == 0) { lgk[3] = -ld[2] - s.v[1] * eb; lgk[4] = ld[3] + s.v[2] * eb; }
      else if (DIR == 1) { lgk[3] = ld[1] + s.v[0] * eb; lgk[4] = -ld[3] - s.v[2] * eb; }
      else { lgk[3] = -ld[1] - s.v[0] * eb; lgk[4] = ld[2] + s.v[1] * eb; }
      lam_k[0] = lam_k[3] = lam_k[4] = harten_abs(s.vn);
      lam_k[1] = harten_abs(s.vn - s.a);
      lam_k[2] = harten_abs(s.vn + s.a);
      double gv[3] = {0, 0, 0}, ab = 0.0;
      // rows of L: p_k = l_k . du, g_k = |lambda_k| p_k
#pragma unroll
      for (int k = 0; k < 5; k++) {
        double row[NV];
        l_row<DIR>(s, gm1, k, row);
        pk[k] = row[0] * du[0] + row[1] * du[1] + row[2] * du[2] + row[3] * du[3] + row[4] * du[4];
        gk[k] = lam_k[k] * pk[k];
        double X[NV];
        const double c = lgk[k] * lam_k[k];
#pragma unroll
        for (int v = 0; v < NV; v++) X[v] = c * du[v];
        l_row_vjp<DIR>(s, gm1, k, X, gv, ab);
      }
      // R at the Roe state, with g = |Lambda| L du
      const double sum = gk[0] + gk[1] + gk[2];
      for (int j = 0; j < 3; j++) gv[j] += ld[1 + j] * sum;
      const double h0b = eb * (gk[1] + gk[2]);
      const double ekb = eb * gk[0] + h0b;
      gv[DIR] += eb * s.a * (gk[2] - gk[1]);
      ab += difb * (gk[2] - gk[1]) + h0b * 2.0 * s.a / gm1;
      if (DIR == 0) { gv[1] -= eb * gk[3]; gv[2] += eb * gk[4]; }
      else if (DIR == 1) { gv[0] += eb * gk[3]; gv[2] -= eb * gk[4]; }
      else { gv[1] += eb * gk[4]; gv[0] -= eb * gk[3]; }
      for (int j = 0; j < 3; j++) gv[j] += ekb * s.v[j];
      // |lambda_k| with Harten's fix
      const double hm = harten_d(s.vn - s.a), hp = harten_d(s.vn + s.a);
      gv[DIR] += harten_d(s.vn) * (lgk[0] * pk[0] + lgk[3] * pk[3] + lgk[4] * pk[4]) + hm * lgk[1] * pk[1] +
                 hp * lgk[2] * pk[2];
      ab += hp * lgk[2] * pk[2] - hm * lgk[1] * pk[1];
      const double ka = ab * gm1 / (2.0 * s.a);
      for (int j = 0; j < 3; j++) gv[j] -= ka * s.v[j];
      roe_avg_vjp(u2, u3, gamma, gv, ka, ub2, ub3);
    } else {
      if (sel == 0) speed_vjp<DIR>(u2, gamma, abar, ub2);
      else if (sel == 1) speed_vjp<DIR>(u3, gamma, abar, ub3);
      else roe_speed_vjp<DIR>(u2, u3, gamma, abar, ub2, ub3);
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
      cF[v * n + t] = ub2[v];
      cU[v * n + t] = ub3[v];
    }
  }
  __syncthreads();
  if (active) {
    const int tp = wrapn(t + 1, n);
    double uc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      uc[v] = sU[v * n + t];
      accU[v] += cF[v * n + tp] + cU[v * n + t];
    }
    flux_jac_T<DIR>(uc, gamma, accF, accU);
    const size_t c = base + (size_t)t * stride;
    double vb[NV] = {0, 0, 0, 0, 0};
    if (VG) visc_gather_cell(uc, QB, GB, nx, ny, nz, gamma, c, vb);
    vjp_store(nu, L, lam, mode, ncell, c, accU, VG, vb);
  }
}

// ---------------------------------------------------------------------------
// characteristic-wise WENO5-JS with Roe or Rusanov upwinding
// ---------------------------------------------------------------------------
// The forward flux at an interface is F = R g with, for every characteristic
// field k (rows of L at the Roe average of the two cells next to it),
//   g_k = (fL_k + fR_k)/2 - l_k (uR_k - uL_k)/2,
// fL/fR (uL/uR) the WENO5 values of L f (L u) over the stencil and l_k = |lambda_k|
// with Harten's fix (Roe) or the Rusanov speed; this is the reference's
// R (WENO of L f) - R |Lambda| L R (WENO of L u) with L R = I
// (Interp1PrimFifthOrderWENOChar.c:86-198, NavierStokes3DUpwind.c:40-124).
// The reverse pass runs field by field: WENO adjoints of the projected
// stencils, then back through L onto the cells' fluxes and states, while the
// derivatives of L, R and l_k with respect to the Roe state (v, a) are summed
// and finally taken back through the Roe average onto cells I-1 and I.  The
// interpolated values use the modified solution, which equals u exactly in
// exact arithmetic, so both adjoints land on u.
// longest grid line a characteristic adjoint workgroup holds: at 512 cells
// the block is 8 wavefronts, two per SIMD (the register budget the
// two-wave hint sets for every block size) and 88 KiB of LDS
constexpr int VJP_CHAR_LINE_MAX = 512;
constexpr int VJP_CHAR_CS = 11;   // doubles per cell record: state [5], flux [5], pad
constexpr int VJP_CHAR_ES = 11;   // doubles per exchange record
constexpr int VJP_CHAR_LDS = VJP_CHAR_CS + VJP_CHAR_ES;  // doubles of LDS per cell of the line

// nu (+)= (d hyp_DIR-part of rhs / d u)^T mu along one direction for the
// characteristic scheme; one workgroup per line, one thread per interface,
// modes as k_rhs_vjp.  LDS (dynamic, VJP_CHAR_LDS doubles per cell of the
// line), cell-major so that one address per stencil cell serves all its
// variables through immediate offsets, with odd record lengths so that the
// 64-bit accesses of a wave spread over all banks: a record per cell of its
// state and flux [n][11] and an exchange record per interface [n][11].
// The WENO5 of L f and the WENO5 of L u are taken back in two parts of five
// fields each.  After each field k every interface posts the adjoints of its
// six projected values and its row of L; every cell then takes the six posts
// that touch it back through those rows, in a fixed order, so the sums are
// deterministic.  The Roe-state derivatives of R, L and |lambda_k| are summed
// field by field in registers.
template <int DIR, int UPW, int W, bool VG = false>
__device__ __forceinline__ void rhs_vjp_char_line(
    const double *__restrict__ u, const double *__restrict__ lam, const double *__restrict__ nu_in, double ca,
    double cb, double *__restrict__ nu, double *__restrict__ L, int mode, int nx, int ny, int nz,
    const double *__restrict__ dxg, double gamma, double eps, const double *__restrict__ QB,
    const double *__restrict__ GB, size_t line, int t, bool active, double *smem) {
  const size_t ncell = (size_t)nx * ny * nz;
  const int n = DIR == 0 ? nx : (DIR == 1 ? ny : nz);
  size_t base, stride;
  if (DIR == 0) { base = line * nx; stride = 1; }
  else if (DIR == 1) { base = (line % nx) + (line / nx) * (size_t)nx * ny; stride = nx; }
  else { base = line; stride = (size_t)nx * ny; }
  const double *dxd = dxg + 3 + (DIR >= 1 ? nx + 6 : 0) + (DIR == 2 ? ny + 6 : 0);
  constexpr int CS = VJP_CHAR_CS, ES = VJP_CHAR_ES;
  double *sC = smem;        // [n][CS]: state [5], flux [5]
  double *E = sC + CS * n;  // [n][ES]: six stencil adjoints [6], row of L [5]
  const double gm1 = gamma - 1.0;
  double lF[NV] = {0, 0, 0, 0, 0};  // adjoint of the interface flux
  if (active) {
    const size_t c = base + (size_t)t * stride;
    double us[NV], lm[NV], ni[NV];
    vjp_load_cell(u, lam, nu_in, ncell, c, us, lm, ni);
#pragma unroll
    for (int v = 0; v < NV; v++) {
      sC[t * CS + v] = us[v];
      // dxinv mu of this cell, staged in the (not yet used) exchange buffer
      E[t * ES + v] = dxd[t] * (cb * lm[v] + (nu_in ? ca * ni[v] : 0.0));
    }
    const Prim q = flowvar(us, gamma);
    double fl[NV];
    euler_flux<DIR>(fl, q);
#pragma unroll
    for (int v = 0; v < NV; v++) sC[t * CS + NV + v] = fl[v];
  }
  __syncthreads();
  if (active) {
    const int tm = wrapn(t - 1, n);
#pragma unroll
    for (int v = 0; v < NV; v++) lF[v] = E[t * ES + v] - E[tm * ES + v];
  }
  __syncthreads();
  int cs[6];
#pragma unroll
  for (int p = 0; p < 6; p++) cs[p] = active ? wrapn(t + p - 3, n) : 0;
  RoeAdj s = {};
  double ls = 0.0, lm = 0.0, lp = 0.0, sumb = 0.0, difb = 0.0;
  int sel = 0;
  const double eb = lF[4];
  if (active) {
    double u2[NV], u3[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) { u2[v] = sC[cs[2] * CS + v]; u3[v] = sC[cs[3] * CS + v]; }
    s = roe_adj_state<DIR>(u2, u3, gamma);
    if (UPW == UPWIND_RUSANOV) {
      // max(max(aL, aR), max(aR, aA)), the branch the forward selects
      const Prim q2 = flowvar(u2, gamma), q3 = flowvar(u3, gamma);
      const double aL = sqrt(gamma * q2.P / q2.rho) + fabs(vel<DIR>(q2));
      const double aR = sqrt(gamma * q3.P / q3.rho) + fabs(vel<DIR>(q3));
      const double aA = s.a + fabs(s.vn);
      const double m01 = fmax(aL, aR), m12 = fmax(aR, aA);
      ls = lm = lp = fmax(m01, m12);
      sel = m01 >= m12 ? (aL >= aR ? 0 : 1) : (aR >= aA ? 1 : 2);
    } else {
      ls = harten_abs(s.vn);
      lm = harten_abs(s.vn - s.a);
      lp = harten_abs(s.vn + s.a);
    }
    // lambda_g = R^T lambda_F (columns of R, navierstokes3d.h:384-470)
    sumb = lF[0] + s.v[0] * lF[1] + s.v[1] * lF[2] + s.v[2] * lF[3];
    difb = lF[1 + DIR] + s.vn * eb;
  }
  double gv[3] = {0, 0, 0}, ab = 0.0, alb = 0.0;
  // part 0 takes the WENO5 of L f (left and right biased) back onto the cell
  // fluxes, part 1 the WENO5 of L u onto the cell states; in between, the
  // flux adjoint of every cell goes through its flux Jacobian
  double acc[NV] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int part = 0; part < 2; part++) {
    const int S = part == 0 ? NV : 0;  // flux or state within the cell record
#pragma unroll
    for (int k = 0; k < 5; k++) {
      // the shear rows k = 3, 4 of L have two entries: r0 at 0 and +-1 at J
      const bool shear = k >= 3;
      constexpr int JA = DIR == 0 ? 2 : 1, JB = DIR == 2 ? 2 : 3;
      constexpr double CA = DIR == 1 ? 1.0 : -1.0, CB = DIR == 1 ? -1.0 : 1.0;
      const int J = k == 3 ? JA : JB;
      const double cJ = k == 3 ? CA : CB;
      if (active) {
        double gb;  // (R^T lambda_F)_k
        if (k == 0) gb = sumb + s.ek * eb;
        else if (k == 1) gb = sumb + s.h0 * eb - s.a * difb;
        else if (k == 2) gb = sumb + s.h0 * eb + s.a * difb;
        else if (DIR == 0) gb = k == 3 ? -lF[2] - s.v[1] * eb : lF[3] + s.v[2] * eb;
        else if (DIR == 1) gb = k == 3 ? lF[1] + s.v[0] * eb : -lF[3] - s.v[2] * eb;
        else gb = k == 3 ? -lF[1] - s.v[0] * eb : lF[2] + s.v[1] * eb;
        const double lk = k == 1 ? lm : (k == 2 ? lp : ls);
        double row[NV], pr[6], b[6];
        if (!shear) {
          l_row<DIR>(s, gm1, k, row);
#pragma unroll
          for (int p = 0; p < 6; p++) {
            double x = 0.0;
#pragma unroll
            for (int v = 0; v < NV; v++) x += row[v] * sC[cs[p] * CS + S + v];
            pr[p] = x;
          }
        } else {
          if (DIR == 0) row[0] = k == 3 ? s.v[1] : -s.v[2];
          else if (DIR == 1) row[0] = k == 3 ? -s.v[0] : s.v[2];
          else row[0] = k == 3 ? s.v[0] : -s.v[1];
#pragma unroll
          for (int p = 0; p < 6; p++) pr[p] = row[0] * sC[cs[p] * CS + S] + cJ * sC[cs[p] * CS + S + J];
        }
        const double obL = part == 0 ? 0.5 * gb : 0.5 * lk * gb;
#pragma unroll
        for (int p = 0; p < 6; p++) b[p] = 0.0;
        double wl, wr;
        weno5_lr_vjp<W>(pr, pr, eps, obL, part == 0 ? obL : -obL, b, wl, wr);
        // post the six adjoints and the row; X for the row's derivatives
        double X[NV] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int p = 0; p < 6; p++) E[t * ES + p] = b[p];
        if (!shear) {
#pragma unroll
          for (int p = 0; p < 6; p++)
#pragma unroll
            for (int v = 0; v < NV; v++) X[v] += b[p] * sC[cs[p] * CS + S + v];
#pragma unroll
          for (int v = 0; v < NV; v++) E[t * ES + 6 + v] = row[v];
        } else {
          // only r0 depends on the Roe state
#pragma unroll
          for (int p = 0; p < 6; p++) X[0] += b[p] * sC[cs[p] * CS + S];
          E[t * ES + 6] = row[0];
        }
        // this part's share of the forward g_k = (fL + fR)/2 - l_k (uR - uL)/2
        const double gk = part == 0 ? 0.5 * (wl + wr) : -0.5 * lk * (wr - wl);
        // term k of R g at the Roe state
        if (k < 3) {
#pragma unroll
          for (int j = 0; j < 3; j++) gv[j] += (lF[1 + j] + eb * s.v[j]) * gk;
          if (k > 0) {
            const double sg = k == 2 ? 1.0 : -1.0;
            gv[DIR] += sg * eb * s.a * gk;
            ab += sg * difb * gk + eb * gk * 2.0 * s.a / gm1;
          }
        } else if (DIR == 0) {
          if (k == 3) gv[1] -= eb * gk; else gv[2] += eb * gk;
        } else if (DIR == 1) {
          if (k == 3) gv[0] += eb * gk; else gv[2] -= eb * gk;
        } else {
          if (k == 3) gv[0] -= eb * gk; else gv[1] += eb * gk;
        }
        // l_k: Harten's |lambda| (Roe) or the common Rusanov speed
        if (part == 1) {
          const double lbk = -0.5 * (wr - wl) * gb;  // adjoint of l_k
          if (UPW == UPWIND_ROE) {
            if (k == 1) { const double h = harten_d(s.vn - s.a) * lbk; gv[DIR] += h; ab -= h; }
            else if (k == 2) { const double h = harten_d(s.vn + s.a) * lbk; gv[DIR] += h; ab += h; }
            else gv[DIR] += harten_d(s.vn) * lbk;
          } else {
            alb += lbk;
          }
        }
        l_row_vjp<DIR>(s, gm1, k, X, gv, ab);
      }
      __syncthreads();
      // cell t takes post p of interface t+3-p, p = 0..5, back through that
      // interface's row
      if (active) {
        if (!shear) {
#pragma unroll
          for (int p = 0; p < 6; p++) {
            const double *e = E + wrapn(t + 3 - p, n) * ES;
            const double bp = e[p];
#pragma unroll
            for (int v = 0; v < NV; v++) acc[v] += bp * e[6 + v];
          }
        } else {
          double bsum = 0.0;
#pragma unroll
          for (int p = 0; p < 6; p++) {
            const double *e = E + wrapn(t + 3 - p, n) * ES;
            const double bp = e[p];
            acc[0] += bp * e[6];
            bsum += bp;
          }
          acc[JA] += (k == 3 ? CA : 0.0) * bsum;
          acc[JB] += (k == 4 ? CB : 0.0) * bsum;
        }
      }
      __syncthreads();
    }
    if (part == 0 && active) {
      double uc[NV], ub[NV] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < NV; v++) uc[v] = sC[t * CS + v];
      flux_jac_T<DIR>(uc, gamma, acc, ub);
#pragma unroll
      for (int v = 0; v < NV; v++) acc[v] = ub[v];
    }
  }
  // the Roe state's (and the Rusanov speed's) dependence on cells t-1 and t
  if (active) {
    double u2[NV], u3[NV], ub2[NV] = {0, 0, 0, 0, 0}, ub3[NV] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < NV; v++) { u2[v] = sC[cs[2] * CS + v]; u3[v] = sC[cs[3] * CS + v]; }
    if (UPW == UPWIND_RUSANOV) {
      if (sel == 0) speed_vjp<DIR>(u2, gamma, alb, ub2);
      else if (sel == 1) speed_vjp<DIR>(u3, gamma, alb, ub3);
      else roe_speed_vjp<DIR>(u2, u3, gamma, alb, ub2, ub3);
    }
    // a = sqrt((H - |v|^2/2)(gamma-1)) -> (v, H) of the Roe average -> the two cells
    const double ka = ab * gm1 / (2.0 * s.a);
    for (int j = 0; j < 3; j++) gv[j] -= ka * s.v[j];
    roe_avg_vjp(u2, u3, gamma, gv, ka, ub2, ub3);
#pragma unroll
    for (int v = 0; v < NV; v++) {
      E[t * ES + v] = ub2[v];
      E[t * ES + NV + v] = ub3[v];
    }
  }
  __syncthreads();
  if (active) {
    const int tp = wrapn(t + 1, n);
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] += E[tp * ES + v] + E[t * ES + NV + v];
    const size_t c = base + (size_t)t * stride;
    double vb[NV] = {0, 0, 0, 0, 0};
    if (VG) {
      double uc[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) uc[v] = sC[t * CS + v];
      visc_gather_cell(uc, QB, GB, nx, ny, nz, gamma, c, vb);
    }
    vjp_store(nu, L, lam, mode, ncell, c, acc, VG, vb);
  }
}

// ---------------------------------------------------------------------------
// viscous term
// ---------------------------------------------------------------------------
// par_i = sum_d dxinv_d(i) D_d[FV_d](i), D the 4th-order central difference
// over FV_d at i-2..i+2 along d, where FV_d is the viscous flux of the
// primitive variables Q = (u, v, w, T) and their gradients (4th-order central,
// periodic) inside the grid, and the ghost-layer flux (biased stencil in the
// second layer, zero derivatives across d) at the two positions on each side
// of a line (NavierStokes3DParabolicFunction.c:50-325,
// FirstDerivativeFourthOrder.c:36-129).  The reverse pass:
//  k_visc_vjp_tile   flux adjoints of every cell from its interior readers,
//                    back through the flux onto Q (stored in QB) and onto the
//                    gradients (stored in GB, already times dxinv), by tiles
//                    of VTX x VTY columns marching through VKS z-planes: Q
//                    and the readers' adjoint weights of a plane sit in LDS
//                    with a 2-cell halo in x and y, the z stencil in
//                    registers, so each is read about once per tile;
//  k_visc_vjp_ghost  the same for the ghost-layer fluxes of each line, whose
//                    gradient adjoints are spread over their stencils into QB;
//  visc_gather_cell  QB plus the transposed gradient stencils over GB, then
//                    back through the primitive variables: the x line VJP,
//                    which runs last, adds this onto the nu that the z and y
//                    line VJPs left, before its own share.
struct ViscC {
  double gamma, inv_Re, inv_gm1, inv_Pr;
};

// adjoint of the viscous flux along DIR at a point: fb = adjoint of
// (tau_d1, tau_d2, tau_d3, energy flux); adds to qb (u, v, w, T values) and
// grb (gradient entries, grb[c][e] = d/d(dq_c/dx_e)); returns the adjoint of mu
template <int DIR>
__device__ __forceinline__ double visc_flux_vjp(const double q[4], const double (&gr)[4][3], double mu,
                                                const ViscC &vc, const double fb[4], double qb[4],
                                                double (&grb)[4][3]) {
  const double tt = 2.0 / 3.0;
  const double mr = mu * vc.inv_Re, kq = mr * vc.inv_gm1 * vc.inv_Pr;
  const double ux = gr[0][0], uy = gr[0][1], uz = gr[0][2];
  const double vx = gr[1][0], vy = gr[1][1], vz = gr[1][2];
  const double wx = gr[2][0], wy = gr[2][1], wz = gr[2][2];
  double b1, b2, b3;
  if (DIR == 0) { b1 = tt * (2 * ux - vy - wz); b2 = uy + vx; b3 = uz + wx; }
  else if (DIR == 1) { b1 = uy + vx; b2 = tt * (-ux + 2 * vy - wz); b3 = vz + wy; }
  else { b1 = uz + wx; b2 = vz + wy; b3 = tt * (-ux - vy + 2 * wz); }
  const double gT = gr[3][DIR];
  const double tb1 = fb[0] + q[0] * fb[3], tb2 = fb[1] + q[1] * fb[3], tb3 = fb[2] + q[2] * fb[3];
  qb[0] += fb[3] * (mr * b1);
  qb[1] += fb[3] * (mr * b2);
  qb[2] += fb[3] * (mr * b3);
  const double mub = (tb1 * b1 + tb2 * b2 + tb3 * b3) * vc.inv_Re + fb[3] * gT * (vc.inv_Re * vc.inv_gm1 * vc.inv_Pr);
  const double c1 = mr * tb1, c2 = mr * tb2, c3 = mr * tb3;
  if (DIR == 0) {
    grb[0][0] += 2 * tt * c1; grb[1][1] -= tt * c1; grb[2][2] -= tt * c1;
    grb[0][1] += c2; grb[1][0] += c2;
    grb[0][2] += c3; grb[2][0] += c3;
  } else if (DIR == 1) {
    grb[0][1] += c1; grb[1][0] += c1;
    grb[0][0] -= tt * c2; grb[1][1] += 2 * tt * c2; grb[2][2] -= tt * c2;
    grb[1][2] += c3; grb[2][1] += c3;
  } else {
    grb[0][2] += c1; grb[2][0] += c1;
    grb[1][2] += c2; grb[2][1] += c2;
    grb[0][0] -= tt * c3; grb[1][1] -= tt * c3; grb[2][2] += 2 * tt * c3;
  }
  grb[3][DIR] += kq * fb[3];
  return mub;
}

// mu_c at a cell: momentum and energy components of cb*lam + ca*nu_in
__device__ __forceinline__ void visc_mu(const double *__restrict__ lam, const double *__restrict__ nu_in, double ca,
                                        double cb, size_t ncell, size_t c, double m[4]) {
  // all loads before any use: a null check of nu_in inside the loop would
  // put a wait in each iteration
  double l[4], n[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; k++) l[k] = lam[(size_t)(1 + k) * ncell + c];
  if (nu_in)
#pragma unroll
    for (int k = 0; k < 4; k++) n[k] = nu_in[(size_t)(1 + k) * ncell + c];
#pragma unroll
  for (int k = 0; k < 4; k++) m[k] = cb * l[k] + (nu_in ? ca * n[k] : 0.0);
}

__device__ __forceinline__ void visc_vjp_tile(const double *__restrict__ Q, const double *__restrict__ lam,
                                              const double *__restrict__ nu_in, double ca, double cb,
                                              double *__restrict__ QB, double *__restrict__ GB, int nx, int ny,
                                              int nz, const double *__restrict__ dxg, const ViscC &vc, int b,
                                              double (&sQ)[4][VHY][VHX], double (&sM)[4][VHY][VHX]) {
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t plane = (size_t)nx * ny;
  const int ntx = (nx + VTX - 1) / VTX, nty = (ny + VTY - 1) / VTY;
  const int tx = b % ntx, ty = (b / ntx) % nty, kz = b / (ntx * nty);
  const int x0 = tx * VTX, y0 = ty * VTY, k0 = kz * VKS;
  const int lx = threadIdx.x % VTX, ly = threadIdx.x / VTX;
  const int i = x0 + lx, j = y0 + ly;
  const bool own = i < nx && j < ny;
  const int ic = own ? i : 0, jc = own ? j : 0;
  const size_t col = (size_t)jc * nx + ic;
  const double *dx = dxg + 3, *dy = dxg + (nx + 6) + 3, *dz = dxg + (nx + 6) + (ny + 6) + 3;
  const int k1 = min(k0 + VKS, nz);
  const int offs[4] = {-2, -1, 1, 2};
  // window: Q and the weights m of this column at planes k-2..k+2
  double q[5][4], m[5][4];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const size_t c = col + (size_t)wrapn(k0 - 2 + s, nz) * plane;
#pragma unroll
    for (int v = 0; v < 4; v++) q[s + 1][v] = Q[(size_t)v * ncell + c];
    visc_mu(lam, nu_in, ca, cb, ncell, c, m[s + 1]);
  }
  for (int k = k0; k < k1; k++) {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int v = 0; v < 4; v++) { q[s][v] = q[s + 1][v]; m[s][v] = m[s + 1][v]; }
    {
      const size_t c = col + (size_t)wrapn(k + 2, nz) * plane;
#pragma unroll
      for (int v = 0; v < 4; v++) q[4][v] = Q[(size_t)v * ncell + c];
      visc_mu(lam, nu_in, ca, cb, ncell, c, m[4]);
    }
    // tile of plane k with halo
    const size_t pk = (size_t)k * plane;
    for (int c = threadIdx.x; c < VHX * VHY; c += VTX * VTY) {
      const int hx = c % VHX, hy = c / VHX;
      // x0 - 2 + hx spans [x0 - 2, x0 + VTX + 1] and y0 - 2 + hy spans
      // [y0 - 2, y0 + VTY + 1], which can exceed 2 nx and 2 ny
      const int gi = (x0 - 2 + hx + 2 * nx) % nx, gj = (y0 - 2 + hy + 2 * ny) % ny;
      const size_t h = pk + (size_t)gj * nx + gi;
      double mh[4];
      visc_mu(lam, nu_in, ca, cb, ncell, h, mh);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        sQ[v][hy][hx] = Q[(size_t)v * ncell + h];
        sM[v][hy][hx] = mh[v];
      }
    }
    __syncthreads();
    if (own) {
      double gr[4][3];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const double *r = &sQ[c][ly + 2][0];
        gr[c][0] = d4c(r[lx], r[lx + 1], r[lx + 3], r[lx + 4]) * dx[i];
        gr[c][1] = d4c(sQ[c][ly][lx + 2], sQ[c][ly + 1][lx + 2], sQ[c][ly + 3][lx + 2], sQ[c][ly + 4][lx + 2]) * dy[j];
        gr[c][2] = d4c(q[0][c], q[1][c], q[3][c], q[4][c]) * dz[k];
      }
      const double qc[4] = {q[2][0], q[2][1], q[2][2], q[2][3]};
      const double mu = exp(0.76 * log(qc[3]));
      double qb[4] = {0, 0, 0, 0}, grb[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, mub = 0.0;
      // direction by direction: reader r = pos - s reads this cell's flux
      // with weight d4w(s) dxinv(r), inside the grid only
#pragma unroll
      for (int d = 0; d < 3; d++) {
        double fb[4] = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const int o = offs[s];
          const int r = (d == 0 ? i : (d == 1 ? j : k)) - o;
          if (r >= 0 && r < (d == 0 ? nx : (d == 1 ? ny : nz))) {
            const double w = d4w(o) * (d == 0 ? dx[r] : (d == 1 ? dy[r] : dz[r]));
#pragma unroll
            for (int c = 0; c < 4; c++)
              fb[c] += w * (d == 0 ? sM[c][ly + 2][lx + 2 - o] : (d == 1 ? sM[c][ly + 2 - o][lx + 2] : m[2 - o][c]));
          }
        }
        if (d == 0) mub += visc_flux_vjp<0>(qc, gr, mu, vc, fb, qb, grb);
        else if (d == 1) mub += visc_flux_vjp<1>(qc, gr, mu, vc, fb, qb, grb);
        else mub += visc_flux_vjp<2>(qc, gr, mu, vc, fb, qb, grb);
      }
      qb[3] += mub * (0.76 * mu / qc[3]);
      const size_t idx = pk + col;
      const double dxe[3] = {dx[i], dy[j], dz[k]};
#pragma unroll
      for (int c = 0; c < 4; c++) {
        QB[(size_t)c * ncell + idx] = qb[c];
#pragma unroll
        for (int e = 0; e < 3; e++) GB[(size_t)(c * 3 + e) * ncell + idx] = grb[c][e] * dxe[e];
      }
    }
    __syncthreads();
  }
}

// one thread per line along DIR: the four ghost-layer fluxes -2, -1, n, n+1
template <int DIR>
__device__ __forceinline__ void visc_vjp_ghost_line(const double *__restrict__ Q, const double *__restrict__ lam,
                                                    const double *__restrict__ nu_in, double ca, double cb,
                                                    double *__restrict__ QB, int nx, int ny, int nz,
                                                    const double *__restrict__ dxg, const ViscC &vc, size_t line) {
  const size_t ncell = (size_t)nx * ny * nz;
  const int n = DIR == 0 ? nx : (DIR == 1 ? ny : nz);
  const size_t nplane = ncell / n;
  if (line >= nplane) return;
  size_t base, stride;
  if (DIR == 0) { base = line * nx; stride = 1; }
  else if (DIR == 1) { base = (line % nx) + (line / nx) * (size_t)nx * ny; stride = nx; }
  else { base = line; stride = (size_t)nx * ny; }
  const double *dxd = dxg + (DIR >= 1 ? nx + 6 : 0) + (DIR == 2 ? ny + 6 : 0);  // entry 0: position -3
  // readers 0, 1, n-2, n-1, times their dxinv
  double a[4][4];
  const int rd[4] = {0, 1, n - 2, n - 1};
#pragma unroll
  for (int r = 0; r < 4; r++) {
    visc_mu(lam, nu_in, ca, cb, ncell, base + (size_t)rd[r] * stride, a[r]);
#pragma unroll
    for (int k = 0; k < 4; k++) a[r][k] *= dxd[3 + rd[r]];
  }
  // the seven cells gp-3 .. gp+3 of slot s (positions -2, -1, n, n+1,
  // periodic) are entries s .. s+6 of the window n-5 .. n-1, 0 .. 4 (n >= 5),
  // loaded once; the slots' adjoints onto Q accumulate per window entry and
  // reach QB with one read-modify-write each (entries that are the same cell
  // of a short line add in turn)
  double f[10][4], acc[10][4];
  size_t cw[10];
#pragma unroll
  for (int j = 0; j < 10; j++) {
    cw[j] = base + (size_t)(j < 5 ? n - 5 + j : j - 5) * stride;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      f[j][c] = Q[(size_t)c * ncell + cw[j]];
      acc[j][c] = 0.0;
    }
  }
  const double r12 = 1.0 / 12.0;
#pragma unroll
  for (int slot = 0; slot < 4; slot++) {
    const int gp = slot < 2 ? slot - 2 : n + slot - 2;
    double fb[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (slot == 0) fb[k] = d4w(-2) * a[0][k];
      else if (slot == 1) fb[k] = d4w(-1) * a[0][k] + d4w(-2) * a[1][k];
      else if (slot == 2) fb[k] = d4w(1) * a[3][k] + d4w(2) * a[2][k];
      else fb[k] = d4w(2) * a[3][k];
    }
    // the forward evaluation at this ghost position (k_visc_ghost_x, k_visc_ghost)
    const double(*w)[4] = f + slot;  // positions gp-3 .. gp+3
    // stencil weights over positions gp-3 .. gp+3
    double w7[7] = {0, 0, 0, 0, 0, 0, 0};
    if (slot == 0) { w7[2] = -3 * r12; w7[3] = -10 * r12; w7[4] = 18 * r12; w7[5] = -6 * r12; w7[6] = r12; }
    else if (slot == 3) { w7[0] = -r12; w7[1] = 6 * r12; w7[2] = -18 * r12; w7[3] = 10 * r12; w7[4] = 3 * r12; }
    else { w7[1] = r12; w7[2] = -8 * r12; w7[4] = 8 * r12; w7[5] = -r12; }
    const double dxgh = dxd[3 + gp];
    double gr[4][3];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      double d = 0.0;
      if (slot == 0) d = (-3 * w[2][c] - 10 * w[3][c] + 18 * w[4][c] - 6 * w[5][c] + w[6][c]) * r12;
      else if (slot == 3) d = (-w[0][c] + 6 * w[1][c] - 18 * w[2][c] + 10 * w[3][c] + 3 * w[4][c]) * r12;
      else d = d4c(w[1][c], w[2][c], w[4][c], w[5][c]);
#pragma unroll
      for (int e = 0; e < 3; e++) gr[c][e] = e == DIR ? d * dxgh : 0.0;
    }
    const double q[4] = {w[3][0], w[3][1], w[3][2], w[3][3]};
    const double mu = exp(0.76 * log(q[3]));
    double qb[4] = {0, 0, 0, 0}, grb[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const double mub = visc_flux_vjp<DIR>(q, gr, mu, vc, fb, qb, grb);
    qb[3] += mub * (0.76 * mu / q[3]);
#pragma unroll
    for (int c = 0; c < 4; c++) acc[slot + 3][c] += qb[c];
#pragma unroll
    for (int s = 0; s < 7; s++) {
      if (w7[s] == 0.0) continue;
#pragma unroll
      for (int c = 0; c < 4; c++) acc[slot + s][c] += w7[s] * dxgh * grb[c][DIR];
    }
  }
#pragma unroll
  for (int j = 0; j < 10; j++)
#pragma unroll
    for (int c = 0; c < 4; c++) QB[(size_t)c * ncell + cw[j]] += acc[j][c];
}

// ---------------------------------------------------------------------------
// kernels: one grid
// ---------------------------------------------------------------------------
// neighbouring lines share cache lines along x: keep them on one XCD
template <int DIR, int UPW, int W, bool VG>
__global__ void __launch_bounds__(VJP_LINE_MAX) k_rhs_vjp(const double *__restrict__ u, const double *__restrict__ lam,
                                                      const double *__restrict__ nu_in, double ca, double cb,
                                                      double *__restrict__ nu, double *__restrict__ L, int mode,
                                                      int nx, int ny, int nz, const double *__restrict__ dxg,
                                                      double gamma, double eps, const double *__restrict__ QB,
                                                      const double *__restrict__ GB) {
  extern __shared__ double smem[];
  const int n = DIR == 0 ? nx : (DIR == 1 ? ny : nz);
  const LineSlot s = line_slot(xcd_block(), n, (size_t)nx * ny * nz / n);
  rhs_vjp_line<DIR, UPW, W, VG>(u, lam, nu_in, ca, cb, nu, L, mode, nx, ny, nz, dxg, gamma, eps, QB, GB, s.line,
                                s.t, s.live, smem + (size_t)s.sub * VJP_LDS * n);
}

template <int DIR, int UPW, int W, bool VG>
__global__ void __launch_bounds__(VJP_CHAR_LINE_MAX) __attribute__((amdgpu_waves_per_eu(2))) k_rhs_vjp_char(
    const double *__restrict__ u, const double *__restrict__ lam, const double *__restrict__ nu_in, double ca,
    double cb, double *__restrict__ nu, double *__restrict__ L, int mode, int nx, int ny, int nz,
    const double *__restrict__ dxg, double gamma, double eps, const double *__restrict__ QB,
    const double *__restrict__ GB) {
  extern __shared__ double smem[];
  const int n = DIR == 0 ? nx : (DIR == 1 ? ny : nz);
  const LineSlot s = line_slot(xcd_block(), n, (size_t)nx * ny * nz / n);
  rhs_vjp_char_line<DIR, UPW, W, VG>(u, lam, nu_in, ca, cb, nu, L, mode, nx, ny, nz, dxg, gamma, eps, QB, GB,
                                     s.line, s.t, s.live, smem + (size_t)s.sub * VJP_CHAR_LDS * n);
}

__global__ void __launch_bounds__(VT) k_visc_vjp_tile(const double *__restrict__ Q, const double *__restrict__ lam,
                                                     const double *__restrict__ nu_in, double ca, double cb,
                                                     double *__restrict__ QB, double *__restrict__ GB, int nx, int ny,
                                                     int nz, const double *__restrict__ dxg, ViscC vc) {
  __shared__ double sQ[4][VHY][VHX], sM[4][VHY][VHX];
  visc_vjp_tile(Q, lam, nu_in, ca, cb, QB, GB, nx, ny, nz, dxg, vc, blockIdx.x, sQ, sM);
}

template <int DIR>
__global__ void __launch_bounds__(256) k_visc_vjp_ghost(const double *__restrict__ Q, const double *__restrict__ lam,
                                                        const double *__restrict__ nu_in, double ca, double cb,
                                                        double *__restrict__ QB, int nx, int ny, int nz,
                                                        const double *__restrict__ dxg, ViscC vc) {
  visc_vjp_ghost_line<DIR>(Q, lam, nu_in, ca, cb, QB, nx, ny, nz, dxg, vc,
                           (size_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// kernels: all members of an ensemble
// ---------------------------------------------------------------------------
// Member m's state, adjoints and work space sit at 5*off, 5*off and 20*off;
// its stage coefficients are dt_m * (A, B).  Line work lists hold each
// member's first line along a direction (nmem + 1 entries).
struct EnsVjp {
  const EnsMember *mem;
  int nmem;
  const double *u, *lam, *nu_in;
  double A, B;
  double *nu, *L, *work;
  int mode;
  double gamma, eps;
  ViscC vc;
};

__device__ __forceinline__ int ens_find(const unsigned *first, int nmem, unsigned item) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int ens_find_cell(const EnsMember *mem, int nmem, size_t idx) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mem[mid].off <= idx) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Calls f(member index) for cell idx of a launch with one thread per cell.
// A block of consecutive cells nearly always lies in one member: the member
// is then found once for the block, the same for every lane (scalar loads
// of its table entry); otherwise every lane searches for its own.
template <class F>
__device__ __forceinline__ void with_cell_member(const EnsMember *mem, int nmem, size_t idx, const F &f) {
  const size_t b0 = (size_t)blockIdx.x * blockDim.x;
  const int m0 = ens_find_cell(mem, nmem, b0);
  if (b0 + blockDim.x <= mem[m0].off + mem[m0].ncell) f(m0);
  else f(ens_find_cell(mem, nmem, idx));
}

// with VG (the x lines of a viscous ensemble) also QB and GB of the member's
// work space, whose viscous share the lines add
#define ENS_MEMBER_ARGS(a, M)                                                                              \
  a.u + 5 * M.off, a.lam + 5 * M.off, a.nu_in ? a.nu_in + 5 * M.off : nullptr, M.dt * a.A, M.dt * a.B,     \
      a.nu + 5 * M.off, a.L ? a.L + 5 * M.off : nullptr, a.mode, M.n[0], M.n[1], M.n[2], M.dxg, a.gamma, a.eps, \
      VG ? a.work + 20 * M.off + 4 * M.ncell : nullptr, VG ? a.work + 20 * M.off + 8 * M.ncell : nullptr

template <int DIR, int UPW, int W, bool VG>
__global__ void __launch_bounds__(VJP_LINE_MAX) k_ens_rhs_vjp(EnsVjp a, const unsigned *__restrict__ first,
                                                          unsigned nitems) {
  extern __shared__ double smem[];
  const XcdItems r = xcd_items(nitems);
  for (unsigned item = r.first; item < r.end; item += r.step) {
    const int m = ens_find(first, a.nmem, item);
    const EnsMember M = a.mem[m];
    const LineSlot s = line_slot(item - first[m], M.n[DIR], M.ncell / M.n[DIR]);
    rhs_vjp_line<DIR, UPW, W, VG>(ENS_MEMBER_ARGS(a, M), s.line, s.t, s.live,
                                  smem + (size_t)s.sub * VJP_LDS * M.n[DIR]);
    __syncthreads();
  }
}

template <int DIR, int UPW, int W, bool VG>
__global__ void __launch_bounds__(VJP_CHAR_LINE_MAX) __attribute__((amdgpu_waves_per_eu(2))) k_ens_rhs_vjp_char(EnsVjp a, const unsigned *__restrict__ first,
                                                                    unsigned nitems) {
  extern __shared__ double smem[];
  const XcdItems r = xcd_items(nitems);
  for (unsigned item = r.first; item < r.end; item += r.step) {
    const int m = ens_find(first, a.nmem, item);
    const EnsMember M = a.mem[m];
    const LineSlot s = line_slot(item - first[m], M.n[DIR], M.ncell / M.n[DIR]);
    rhs_vjp_char_line<DIR, UPW, W, VG>(ENS_MEMBER_ARGS(a, M), s.line, s.t, s.live,
                                       smem + (size_t)s.sub * VJP_CHAR_LDS * M.n[DIR]);
    __syncthreads();
  }
}

#undef ENS_MEMBER_ARGS

// primitive variables of every member's state into its work space
__global__ void __launch_bounds__(256) k_ens_vjp_prims(EnsVjp a, size_t ncell) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  with_cell_member(a.mem, a.nmem, idx, [&](int m) {
    const EnsMember &M = a.mem[m];
    const size_t c = idx - M.off;
    double s[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) s[v] = a.u[5 * M.off + (size_t)v * M.ncell + c];
    const Prim q = flowvar(s, a.gamma);
    double *Q = a.work + 20 * M.off;
    Q[c] = q.vx;
    Q[M.ncell + c] = q.vy;
    Q[2 * M.ncell + c] = q.vz;
    Q[3 * M.ncell + c] = a.gamma * q.P / q.rho;
  });
}

// tiles of all members; tiles0 holds each member's first tile
__global__ void __launch_bounds__(VT) k_ens_visc_vjp_tile(EnsVjp a, const unsigned *__restrict__ tiles0,
                                                         unsigned ntiles) {
  __shared__ double sQ[4][VHY][VHX], sM[4][VHY][VHX];
  const unsigned item = blockIdx.x;  // one block per tile (a loop over tiles spills 12 B per lane)
  if (item >= ntiles) return;
  const int m = ens_find(tiles0, a.nmem, item);
  const EnsMember M = a.mem[m];
  const double *Q = a.work + 20 * M.off;
  double *QB = a.work + 20 * M.off + 4 * M.ncell, *GB = a.work + 20 * M.off + 8 * M.ncell;
  visc_vjp_tile(Q, a.lam + 5 * M.off, a.nu_in ? a.nu_in + 5 * M.off : nullptr, M.dt * a.A, M.dt * a.B, QB, GB,
                M.n[0], M.n[1], M.n[2], M.dxg, a.vc, (int)(item - tiles0[m]), sQ, sM);
}

template <int DIR>
__global__ void __launch_bounds__(256) k_ens_visc_vjp_ghost(EnsVjp a, const unsigned *__restrict__ first,
                                                            unsigned nitems) {
  const unsigned item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= nitems) return;
  const int m = ens_find(first, a.nmem, item);
  const EnsMember M = a.mem[m];
  const double *Q = a.work + 20 * M.off;
  double *QB = a.work + 20 * M.off + 4 * M.ncell;
  visc_vjp_ghost_line<DIR>(Q, a.lam + 5 * M.off, a.nu_in ? a.nu_in + 5 * M.off : nullptr, M.dt * a.A,
                           M.dt * a.B, QB, M.n[0], M.n[1], M.n[2], M.dxg, a.vc, item - first[m]);
}

// y = a*x + b*y
__global__ void k_axpby(const double *__restrict__ x, double a, double *__restrict__ y, double b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = (a == 0.0 ? 0.0 : a * x[i]) + b * y[i];  // x unread when a == 0
}

// ---------------------------------------------------------------------------
bool vjp_supported(const Grid &g, const Scheme &sc) {
  const int lmax = sc.interp == INTERP_CHAR ? VJP_CHAR_LINE_MAX : VJP_LINE_MAX;
  return g.n[0] <= lmax && g.n[1] <= lmax && g.n[2] <= lmax;
}

size_t vjp_work_size(const Grid &g, const Scheme &sc) { return sc.Re > 0 ? 20 * g.ncell : 0; }

// dynamic LDS beyond 64 KiB needs the kernel's opt-in (160 KiB per CU on gfx950)
static void allow_lds(const void *kernel, size_t bytes) {
  if (bytes > 65536) (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// calls f(tag) with the scheme's upwinding and WENO weights as compile-time
// constants tag.upw, tag.weno
template <int U, int Wt>
struct SchemeTag {
  static constexpr int upw = U, weno = Wt;
};

template <class F>
static void with_scheme(const Scheme &sc, F &&f) {
  const bool roe = sc.upwind == UPWIND_ROE;
  switch (sc.weno) {
    case WENO_MAPPED: roe ? f(SchemeTag<UPWIND_ROE, WENO_MAPPED>{}) : f(SchemeTag<UPWIND_RUSANOV, WENO_MAPPED>{}); break;
    case WENO_Z: roe ? f(SchemeTag<UPWIND_ROE, WENO_Z>{}) : f(SchemeTag<UPWIND_RUSANOV, WENO_Z>{}); break;
    case WENO_YC: roe ? f(SchemeTag<UPWIND_ROE, WENO_YC>{}) : f(SchemeTag<UPWIND_RUSANOV, WENO_YC>{}); break;
    case WENO_OPTIMAL:
      roe ? f(SchemeTag<UPWIND_ROE, WENO_OPTIMAL>{}) : f(SchemeTag<UPWIND_RUSANOV, WENO_OPTIMAL>{});
      break;
    default: roe ? f(SchemeTag<UPWIND_ROE, WENO_JS>{}) : f(SchemeTag<UPWIND_RUSANOV, WENO_JS>{}); break;
  }
}

template <int DIR>
static void launch_vjp_dir(const Grid &g, const Scheme &sc, const double *u, const double *lam, const double *nu_in,
                           double ca, double cb, double *nu, double *L, int mode, const double *dxg, hipStream_t st,
                           const double *QB = nullptr, const double *GB = nullptr) {
  const unsigned threads = (unsigned)((g.n[DIR] + 63) / 64 * 64);
  const unsigned lpb = (unsigned)vjp_lines_per_block((int)threads, g.n[DIR]);
  const unsigned lines = (unsigned)((g.ncell / g.n[DIR] + lpb - 1) / lpb);  // work items
  with_scheme(sc, [&](auto tag) {
    constexpr int U = decltype(tag)::upw, W = decltype(tag)::weno;
    auto go = [&](auto vg) {
      constexpr bool VG = decltype(vg)::value;
      if (sc.interp == INTERP_COMP) {
        const size_t lds = (size_t)VJP_LDS * lpb * g.n[DIR] * sizeof(double);
        allow_lds((const void *)k_rhs_vjp<DIR, U, W, VG>, lds);
        hipLaunchKernelGGL((k_rhs_vjp<DIR, U, W, VG>), dim3(lines), dim3(threads), lds, st, u, lam, nu_in, ca, cb,
                           nu, L, mode, g.n[0], g.n[1], g.n[2], dxg, sc.gamma, sc.eps, QB, GB);
      } else {
        const size_t lds = (size_t)VJP_CHAR_LDS * lpb * g.n[DIR] * sizeof(double);
        allow_lds((const void *)k_rhs_vjp_char<DIR, U, W, VG>, lds);
        hipLaunchKernelGGL((k_rhs_vjp_char<DIR, U, W, VG>), dim3(lines), dim3(threads), lds, st, u, lam, nu_in, ca,
                           cb, nu, L, mode, g.n[0], g.n[1], g.n[2], dxg, sc.gamma, sc.eps, QB, GB);
      }
    };
    // only the x lines of a viscous run add the viscous share
    if (DIR == 0 && QB) go(std::integral_constant<bool, DIR == 0>{});
    else go(std::false_type{});
  });
}

void launch_rhs_vjp(const Grid &g, const Scheme &sc, const double *u, const double *lam, const double *nu_in,
                    double ca, double cb, double *nu, double *L, bool set_L, const double *dxg, double *work,
                    hipStream_t st) {
  const bool visc = sc.Re > 0;
  // z first and x last: the strided z lines take the plain store; the y
  // lines add into nu; the viscous kernels leave the flux and gradient
  // adjoints in QB and GB, and the contiguous x lines add the viscous share
  // and their own into nu (x also into L)
  const int last = L ? (set_L ? VJP_ADD_LSET : VJP_ADD_L) : VJP_ADD;
  launch_vjp_dir<2>(g, sc, u, lam, nu_in, ca, cb, nu, L, VJP_WRITE, dxg, st);
  launch_vjp_dir<1>(g, sc, u, lam, nu_in, ca, cb, nu, L, VJP_ADD, dxg, st);
  // work: Q [4], QB [4], GB [12] (times ncell)
  double *QB = visc ? work + 4 * g.ncell : nullptr, *GB = visc ? work + 8 * g.ncell : nullptr;
  if (visc) {
    double *Q = work;
    const ViscC vc = {sc.gamma, 1.0 / sc.Re, 1.0 / (sc.gamma - 1.0), 1.0 / sc.Pr};
    launch_prims(g, sc, u, Q, st);
    hipLaunchKernelGGL(k_visc_vjp_tile, dim3(visc_tiles(g.n[0], g.n[1], g.n[2])), dim3(VT), 0, st, Q, lam, nu_in, ca,
                       cb, QB, GB, g.n[0], g.n[1], g.n[2], dxg, vc);
    const unsigned gx = (unsigned)((g.ncell / g.n[0] + 255) / 256), gy = (unsigned)((g.ncell / g.n[1] + 255) / 256),
                   gz = (unsigned)((g.ncell / g.n[2] + 255) / 256);
    hipLaunchKernelGGL(k_visc_vjp_ghost<0>, dim3(gx), dim3(256), 0, st, Q, lam, nu_in, ca, cb, QB, g.n[0], g.n[1],
                       g.n[2], dxg, vc);
    hipLaunchKernelGGL(k_visc_vjp_ghost<1>, dim3(gy), dim3(256), 0, st, Q, lam, nu_in, ca, cb, QB, g.n[0], g.n[1],
                       g.n[2], dxg, vc);
    hipLaunchKernelGGL(k_visc_vjp_ghost<2>, dim3(gz), dim3(256), 0, st, Q, lam, nu_in, ca, cb, QB, g.n[0], g.n[1],
                       g.n[2], dxg, vc);
  }
  launch_vjp_dir<0>(g, sc, u, lam, nu_in, ca, cb, nu, L, last, dxg, st, QB, GB);
}

bool ens_vjp_supported(const Scheme &sc, const int nmax[3]) {
  Grid g;
  for (int d = 0; d < 3; d++) g.n[d] = nmax[d];
  g.ncell = 1;
  return vjp_supported(g, sc);
}

size_t ens_vjp_work_size(const Scheme &sc, size_t ncell) { return sc.Re > 0 ? 20 * ncell : 0; }

int ens_vjp_threads(int nmax) { return (nmax + 63) / 64 * 64; }
int ens_vjp_lines_per_item(int nmax, int n) { return vjp_lines_per_block(ens_vjp_threads(nmax), n); }

template <int DIR>
static void launch_ens_vjp_dir(const Scheme &sc, const EnsVjpPlan &p, const EnsVjp &a, hipStream_t st) {
  // work items: groups of vjp_lines_per_block(threads, n) lines of a member
  const unsigned items = p.nitems[DIR];
  const unsigned grid = items < 4096 ? items : 4096;  // persistent blocks stride over the items
  const unsigned threads = (unsigned)ens_vjp_threads(p.nmax[DIR]);
  with_scheme(sc, [&](auto tag) {
    constexpr int U = decltype(tag)::upw, W = decltype(tag)::weno;
    auto go = [&](auto vg) {
      constexpr bool VG = decltype(vg)::value;
      if (sc.interp == INTERP_COMP) {
        const size_t lds = (size_t)VJP_LDS * threads * sizeof(double);
        allow_lds((const void *)k_ens_rhs_vjp<DIR, U, W, VG>, lds);
        hipLaunchKernelGGL((k_ens_rhs_vjp<DIR, U, W, VG>), dim3(grid), dim3(threads), lds, st, a, p.items0[DIR],
                           items);
      } else {
        const size_t lds = (size_t)VJP_CHAR_LDS * threads * sizeof(double);
        allow_lds((const void *)k_ens_rhs_vjp_char<DIR, U, W, VG>, lds);
        hipLaunchKernelGGL((k_ens_rhs_vjp_char<DIR, U, W, VG>), dim3(grid), dim3(threads), lds, st, a,
                           p.items0[DIR], items);
      }
    };
    // only the x lines of a viscous ensemble add the viscous share (a.work)
    if (DIR == 0 && a.work) go(std::integral_constant<bool, DIR == 0>{});
    else go(std::false_type{});
  });
}

void launch_ens_rhs_vjp(const Scheme &sc, const EnsVjpPlan &p, const double *u, const double *lam,
                        const double *nu_in, double A, double B, double *nu, double *L, bool set_L, double *work,
                        hipStream_t st) {
  EnsVjp a;
  a.mem = p.mem;
  a.nmem = p.nmem;
  a.u = u;
  a.lam = lam;
  a.nu_in = nu_in;
  a.A = A;
  a.B = B;
  a.nu = nu;
  a.L = L;
  a.work = work;
  a.gamma = sc.gamma;
  a.eps = sc.eps;
  a.vc = {sc.gamma, sc.Re > 0 ? 1.0 / sc.Re : 0.0, 1.0 / (sc.gamma - 1.0), 1.0 / sc.Pr};
  const bool visc = sc.Re > 0;
  // the order of the single grid: z (plain store), y, the viscous kernels, x
  a.mode = VJP_WRITE;
  launch_ens_vjp_dir<2>(sc, p, a, st);
  a.mode = VJP_ADD;
  launch_ens_vjp_dir<1>(sc, p, a, st);
  if (visc) {
    const unsigned nb = (unsigned)((p.ncell + 255) / 256);
    hipLaunchKernelGGL(k_ens_vjp_prims, dim3(nb), dim3(256), 0, st, a, p.ncell);
    hipLaunchKernelGGL(k_ens_visc_vjp_tile, dim3(p.ntiles), dim3(VT), 0, st, a, p.tiles0, p.ntiles);
    hipLaunchKernelGGL(k_ens_visc_vjp_ghost<0>, dim3((p.nlines[0] + 255) / 256), dim3(256), 0, st, a, p.lines0[0],
                       p.nlines[0]);
    hipLaunchKernelGGL(k_ens_visc_vjp_ghost<1>, dim3((p.nlines[1] + 255) / 256), dim3(256), 0, st, a, p.lines0[1],
                       p.nlines[1]);
    hipLaunchKernelGGL(k_ens_visc_vjp_ghost<2>, dim3((p.nlines[2] + 255) / 256), dim3(256), 0, st, a, p.lines0[2],
                       p.nlines[2]);
  }
  // the x lines add the viscous share (QB, GB of a.work) before their own
  a.mode = L ? (set_L ? VJP_ADD_LSET : VJP_ADD_L) : VJP_ADD;
  launch_ens_vjp_dir<0>(sc, p, a, st);
}

void launch_axpby(const double *x, double a, double *y, double b, size_t n, hipStream_t st) {
  size_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)nb), dim3(256), 0, st, x, a, y, b, n);
}

}  // namespace hm
