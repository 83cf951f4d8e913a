// hm_pass.hip -- fused direction passes of the NS3D WENO5 right-hand side
// and RK4 stage (gfx950, FP64).
//
// One RK stage is three launches, one per direction (after the viscous-term
// kernel of hm_kernels.hip when the run is viscous):
//   pass x:  R = dxinv (F_{i+1/2} - F_{i-1/2})
//   pass y:  R = R + dyinv (G_{j+1/2} - G_{j-1/2})
//   pass z:  r = -(R + dzinv (H_{k+1/2} - H_{k-1/2})) + par,
//            then the RK4 stage update in 3N form (or r itself for hm_rhs)
// i.e. rhs = -hyp + par of TimeRHSFunctionExplicit.c:88-93 summed in the
// reference's order, with the hyperbolic term of HyperbolicFunction.c:75-101
// accumulated direction by direction and the parabolic term par of
// NavierStokes3DParabolicFunction.c:50-325.
//
// Each pass stages the cells of its pencils in LDS once, evaluates every
// cell's flux and modified energy once, computes the interface fluxes
// (WENO5-JS, component-wise or characteristic, with Roe or Rusanov
// upwinding) from LDS, and differences them in LDS: interface fluxes never
// reach HBM.
//  * x pass: one wavefront per x-line, chunks of 64 interfaces along x.
//  * y, z passes: a block owns 64 consecutive columns (x fastest) of the
//    plane across y (z), which may span several x-rows of that plane, and
//    marches along y (z) four rows per step, with a 9-row ring of cell data
//    and a 5-row ring of interface fluxes in LDS.
// Output of cell c needs the interfaces c-1/2 and c+1/2; a step emits the
// cells one row (lane) behind its interfaces, and the last cell of a line,
// whose right interface is the periodic image of interface 0, is emitted
// after the march.
//
// Restatements relative to the reference arithmetic (equal in exact
// arithmetic; rounding-level differences only):
//  * 1/rho, the Jiang-Shu normalisation and the Roe-average weights use a
//    Newton-Raphson-refined reciprocal (two steps r <- r + r(1 - d r) from
//    the hardware estimate); velocities are m * (1/rho).
//  * The mass flux of a cell is its momentum m_d (the reference forms
//    rho * (m_d / rho)).
//  * The smoothness indicators are evaluated as 4*beta with 4*eps, which
//    scales every (beta_k + eps)^2 by 16 and leaves the weights unchanged;
//    the second differences shared by the left- and right-biased
//    indicators of an interface are formed once.
//  * The characteristic projections use the structure of the left
//    eigenvectors (navierstokes3d.h:288-374): with
//    w = v.x_{1..3} - x_4, rows 0, d and 4 are
//      L00 x0 + (g-1)/a^2 w,  Ld0 x0 - (g-1)/(2a^2) w -/+ a/(2a^2) x_d,
//    and the two shear rows have two non-zero entries; the product with the
//    right eigenvectors (navierstokes3d.h:384-470) is expanded the same
//    way.  The speed of sound of the Roe state is sqrt((g-1)(H - |v|^2/2)),
//    the reference's sqrt(g P / rho) with P from H.
#include <hip/hip_runtime.h>

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

namespace {

constexpr int NQ = 10;  // per-cell LDS data: U0..U4, F1..F4 (F0 = U_d), modified energy


// 1/d, d > 0 and normal: hardware estimate refined by two Newton-Raphson
// steps, each of which squares the relative error.
__device__ __forceinline__ double nr_rcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
}

struct PassConst {
  int n[3];
  size_t ncell;
  double gamma, gm1, inv_gm1, eps4;
  const double *dxg;  // ghosted dxinv x|y|z
};

struct PassIO {
  const double *U;    // stage state
  double *R;          // hyperbolic-term accumulator
  const double *par;  // viscous term, par[c*ncell + cell], c = momentum x, y, z, energy
  double *u, *S, *acc, *out, *Q;
  int visc, mode, stage;
  double ca, cb;
};

// Stage update of the z pass in the separate kernel instantiations RK =
// RK_AFFINE: with r the stage's right-hand side and a, b the operands A, B
// at the cell,
//   O1 = a + c1 r,  O2 = b + c2 r (the next stage state),  O3 = a + c3 r;
// the caller picks operands, outputs and coefficients per stage.  The first
// nout outputs are stored; the caller leaves out only an output that
// repeats O1 (same array, same value), and every pointer is valid whether
// stored or not.  The update has no per-stage branches that choose among
// outputs: the compiler merged the stores of a three-way stage branch into
// one store whose address register one path never set.
enum { RK_CLASSIC = 0, RK_AFFINE = 1 };
struct SspIO {
  const double *A, *B;
  double *O1, *O2, *O3;
  double c1, c2, c3;
  int nout;  // outputs stored: O1, then O2 (nout >= 2), then O3 (nout = 3)
};

// per-cell data of the pass direction (NavierStokes3DFlux.c:23-57 and
// NavierStokes3DModifiedSolution.c:31-70)
template <int DIR>
__device__ __forceinline__ void cell_data(const double u[NV], const PassConst &k, double c[NQ]) {
  const double rho = u[0];
  const double ir = rho == 0.0 ? 0.0 : nr_rcp(rho);
  const double vx = u[1] * ir, vy = u[2] * ir, vz = u[3] * ir;
  const double vsq = vx * vx + vy * vy + vz * vz;
  const double P = (u[4] - 0.5 * rho * vsq) * k.gm1;
  const double vn = DIR == 0 ? vx : (DIR == 1 ? vy : vz);
  const double m = u[DIR + 1];
#pragma unroll
  for (int q = 0; q < NV; q++) c[q] = u[q];
  c[5] = m * vx + (DIR == 0 ? P : 0.0);
  c[6] = m * vy + (DIR == 1 ? P : 0.0);
  c[7] = m * vz + (DIR == 2 ? P : 0.0);
  c[8] = (u[4] + P) * vn;
  c[9] = (P * k.inv_gm1) + (0.5 * rho * vsq);
}

// WENO5 on both sides of one interface.  b[0..5] are the values of cells
// I-3..I+2 the smoothness indicators use, g[0..5] the values that are
// interpolated (include/interpolation.h:265-500,
// Interp1PrimFifthOrderWENO.c:146-158; the right-biased stencil is the
// mirror image, WENOFifthOrderCalculateWeights.c:166-177).  The indicators
// arrive scaled by 4 (l1..l3 = 4 beta_k) together with eps4 = 4 eps; W
// selects the weights:
//  * WENO_JS:     a_k = c_k / (beta_k + eps)^2
//  * WENO_Z:      a_k = c_k (1 + (tau / (beta_k + eps))^2), tau = |beta_1 - beta_3|
//  * WENO_YC:     the same with tau = (m3 - 4 m2 + 6 m1 - 4 p1 + p2)^2
//  * WENO_MAPPED: the JS weights w mapped by
//                 w (c + c^2 - 3 c w + w^2) / (c^2 + w (1 - 2c)) and renormalised
// All are applied in ratio form with one reciprocal (two for the mapped).
template <int W>
__device__ __forceinline__ double weno_combine(double l1, double l2, double l3, double d4, double g3, double g2,
                                               double g1, double q1, double q2, double eps4) {
  const double s6 = 1.0 / 6.0;
  const double s1 = (l1 + eps4) * (l1 + eps4);
  const double s2 = (l2 + eps4) * (l2 + eps4);
  const double s3 = (l3 + eps4) * (l3 + eps4);
  const double f1 = (2 * s6) * g3 + (-7 * s6) * g2 + (11 * s6) * g1;
  const double f2 = (-s6) * g2 + (5 * s6) * g1 + (2 * s6) * q1;
  const double f3 = (2 * s6) * g1 + (5 * s6) * q1 + (-s6) * q2;
  // no_limiting: the weights stay at their optimal initial values
  // (WENOFifthOrderInitializeWeights.c:33-228, HyperbolicFunction.c:62)
  if (W == WENO_OPTIMAL) return 0.1 * f1 + 0.6 * f2 + 0.3 * f3;
  double a1, a2, a3;
  if (W == WENO_Z || W == WENO_YC) {
    // tau scales like beta for WENO-Z; the YC tau does not, hence 4 tau
    const double tau = W == WENO_Z ? fabs(l3 - l1) : 4.0 * (d4 * d4);
    const double t2 = tau * tau;
    a1 = 0.1 * ((s1 + t2) * (s2 * s3));
    a2 = 0.6 * ((s2 + t2) * (s1 * s3));
    a3 = 0.3 * ((s3 + t2) * (s1 * s2));
  } else {
    a1 = 0.1 * (s2 * s3);
    a2 = 0.6 * (s1 * s3);
    a3 = 0.3 * (s1 * s2);
  }
  if (W == WENO_MAPPED) {
    const double c1 = 0.1, c2 = 0.6, c3 = 0.3;
    const double in = nr_rcp(a1 + a2 + a3);
    const double w1 = a1 * in, w2 = a2 * in, w3 = a3 * in;
    const double n1 = w1 * (c1 + c1 * c1 - 3 * c1 * w1 + w1 * w1), e1 = c1 * c1 + w1 * (1.0 - 2.0 * c1);
    const double n2 = w2 * (c2 + c2 * c2 - 3 * c2 * w2 + w2 * w2), e2 = c2 * c2 + w2 * (1.0 - 2.0 * c2);
    const double n3 = w3 * (c3 + c3 * c3 - 3 * c3 * w3 + w3 * w3), e3 = c3 * c3 + w3 * (1.0 - 2.0 * c3);
    // n_k / e_k multiplied through by e1 e2 e3
    a1 = n1 * (e2 * e3);
    a2 = n2 * (e1 * e3);
    a3 = n3 * (e1 * e2);
  }
  return (a1 * f1 + a2 * f2 + a3 * f3) * nr_rcp(a1 + a2 + a3);
}

template <int W>
__device__ __forceinline__ void weno5_lr(const double (&b)[6], const double (&g)[6], double eps4, double &vl,
                                         double &vr) {
  const double c13 = 13.0 / 3.0;  // 4 * 13/12
  const double d1 = b[0] - 2 * b[1] + b[2], d2 = b[1] - 2 * b[2] + b[3];
  const double d3 = b[2] - 2 * b[3] + b[4], d4 = b[3] - 2 * b[4] + b[5];
  const double q2 = c13 * (d2 * d2), q3 = c13 * (d3 * d3);
  double e;
  e = b[0] - 4 * b[1] + 3 * b[2];
  const double l1 = c13 * (d1 * d1) + e * e;
  e = b[1] - b[3];
  const double l2 = q2 + e * e;
  e = 3 * b[2] - 4 * b[3] + b[4];
  const double l3 = q3 + e * e;
  e = b[5] - 4 * b[4] + 3 * b[3];
  const double r1 = c13 * (d4 * d4) + e * e;
  e = b[4] - b[2];
  const double r2 = q3 + e * e;
  e = 3 * b[3] - 4 * b[2] + b[1];
  const double r3 = q2 + e * e;
  // 4th undivided differences of the two stencils (YC weights only)
  const double tl = W == WENO_YC ? b[0] - 4 * b[1] + 6 * b[2] - 4 * b[3] + b[4] : 0.0;
  const double tr = W == WENO_YC ? b[5] - 4 * b[4] + 6 * b[3] - 4 * b[2] + b[1] : 0.0;
  vl = weno_combine<W>(l1, l2, l3, tl, g[0], g[1], g[2], g[3], g[4], eps4);
  vr = weno_combine<W>(r1, r2, r3, tr, g[5], g[4], g[3], g[2], g[1], eps4);
}

// Roe-averaged state of two cells (navierstokes3d.h:144-169) and the
// eigen-structure constants of the DIR flux Jacobian there.
struct RoeState {
  double vx, vy, vz, vn, a, ek, h0;
  double L00, Lm0, L40, k0, kh, aih;  // left-eigenvector entries (see top)
  double cL, cR;                      // sound speeds + |v_d| of the two cells
};

template <int DIR>
__device__ __forceinline__ RoeState roe_state(const double uL[NV], const double uR[NV], const PassConst &k) {
  RoeState s;
  const double rL = uL[0], rR = uR[0];
  const double iL = rL == 0.0 ? 0.0 : nr_rcp(rL), iR = rR == 0.0 ? 0.0 : nr_rcp(rR);
  const double vxL = uL[1] * iL, vyL = uL[2] * iL, vzL = uL[3] * iL;
  const double vxR = uR[1] * iR, vyR = uR[2] * iR, vzR = uR[3] * iR;
  const double qL = vxL * vxL + vyL * vyL + vzL * vzL, qR = vxR * vxR + vyR * vyR + vzR * vzR;
  const double PL = (uL[4] - 0.5 * rL * qL) * k.gm1, PR = (uR[4] - 0.5 * rR * qR) * k.gm1;
  const double c2L = k.gamma * PL * iL, c2R = k.gamma * PR * iR;
  const double HL = 0.5 * qL + c2L * k.inv_gm1, HR = 0.5 * qR + c2R * k.inv_gm1;
  const double tL = sqrt(rL), tR = sqrt(rR);
  const double inv = nr_rcp(tL + tR);
  s.vx = (tL * vxL + tR * vxR) * inv;
  s.vy = (tL * vyL + tR * vyR) * inv;
  s.vz = (tL * vzL + tR * vzR) * inv;
  const double H = (tL * HL + tR * HR) * inv;
  s.ek = 0.5 * (s.vx * s.vx + s.vy * s.vy + s.vz * s.vz);
  const double a2 = (H - s.ek) * k.gm1;
  s.a = sqrt(a2);
  s.vn = DIR == 0 ? s.vx : (DIR == 1 ? s.vy : s.vz);
  s.h0 = a2 * k.inv_gm1 + s.ek;
  const double ia2 = nr_rcp(a2), ih2 = 0.5 * ia2;
  const double gek = k.gm1 * s.ek, avn = s.a * s.vn;
  s.L00 = (a2 - gek) * ia2;
  s.Lm0 = (gek + avn) * ih2;
  s.L40 = (gek - avn) * ih2;
  s.k0 = k.gm1 * ia2;
  s.kh = k.gm1 * ih2;
  s.aih = s.a * ih2;
  const double vnL = DIR == 0 ? vxL : (DIR == 1 ? vyL : vzL);
  const double vnR = DIR == 0 ? vxR : (DIR == 1 ? vyR : vzR);
  s.cL = sqrt(c2L) + fabs(vnL);
  s.cR = sqrt(c2R) + fabs(vnR);
  return s;
}

// |lambda| with Harten's entropy fix, delta = 1e-6 (NavierStokes3DUpwind.c:95-101)
__device__ __forceinline__ double harten(double l) {
  const double delta = 0.000001, delta2 = delta * delta, inv2d = 0.5 / delta;
  const double al = fabs(l);
  return al < delta ? (l * l + delta2) * inv2d : al;
}

// rows 0, d, 4 of L x (u-a wave in row d = DIR+1)
template <int DIR>
__device__ __forceinline__ void proj_acoustic(const RoeState &s, double x0, double x1, double x2, double x3,
                                              double x4, double &p0, double &pm, double &p4) {
  const double w = s.vx * x1 + s.vy * x2 + s.vz * x3 - x4;
  const double xm = DIR == 0 ? x1 : (DIR == 1 ? x2 : x3);
  const double t = s.kh * w;
  p0 = s.L00 * x0 + s.k0 * w;
  pm = s.Lm0 * x0 - (t + s.aih * xm);
  p4 = s.L40 * x0 + (s.aih * xm - t);
}

// the two shear rows of L x, in row order
template <int DIR>
__device__ __forceinline__ void proj_shear(const RoeState &s, double x0, double x1, double x2, double x3,
                                           double &pa, double &pb) {
  if (DIR == 0) {        // rows 2, 3: (vy, 0, -1, 0, 0), (-vz, 0, 0, 1, 0)
    pa = s.vy * x0 - x2;
    pb = x3 - s.vz * x0;
  } else if (DIR == 1) { // rows 1, 3: (-vx, 1, 0, 0, 0), (vz, 0, 0, -1, 0)
    pa = x1 - s.vx * x0;
    pb = s.vz * x0 - x3;
  } else {               // rows 1, 2: (vx, -1, 0, 0, 0), (-vy, 0, 1, 0, 0)
    pa = s.vx * x0 - x1;
    pb = x2 - s.vy * x0;
  }
}

// f = R g, with g0 (entropy), gm (u-a), g4 (u+a), ga, gb (shear rows in
// row order)
template <int DIR>
__device__ __forceinline__ void right_mult(const RoeState &s, double g0, double gm, double g4, double ga,
                                           double gb, double f[NV]) {
  const double sum = g0 + gm + g4;
  const double dif = s.a * (g4 - gm);
  f[0] = sum;
  f[1] = s.vx * sum;
  f[2] = s.vy * sum;
  f[3] = s.vz * sum;
  f[DIR + 1] += dif;
  double e = s.ek * g0 + s.h0 * (gm + g4) + s.vn * dif;
  if (DIR == 0) {        // columns 2, 3: (0,0,-1,0,-vy), (0,0,0,1,vz)
    f[2] -= ga;
    f[3] += gb;
    e += s.vz * gb - s.vy * ga;
  } else if (DIR == 1) { // columns 1, 3: (0,1,0,0,vx), (0,0,0,-1,-vz)
    f[1] += ga;
    f[3] -= gb;
    e += s.vx * ga - s.vz * gb;
  } else {               // columns 1, 2: (0,-1,0,0,-vx), (0,0,1,0,vy)
    f[1] -= ga;
    f[2] += gb;
    e += s.vy * gb - s.vx * ga;
  }
  f[4] = e;
}

// Interface flux at I-1/2 from the LDS data of cells I-3..I+2;
// ld(s, q) returns quantity q of cell I-3+s.
template <int DIR, int INTERP, int UPW, int W, class LD>
__device__ __forceinline__ void iface(const LD &ld, const PassConst &k, double fI[NV]) {
  double uL[NV], uR[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    uL[q] = ld(2, q);
    uR[q] = ld(3, q);
  }
  const RoeState s = roe_state<DIR>(uL, uR, k);
  const double aA = s.a + fabs(s.vn);
  const double alpha = fmax(fmax(s.cL, s.cR), fmax(s.cR, aA));  // NavierStokes3DUpwind.c:349-420

  if (INTERP == INTERP_CHAR) {
    // rows 0, d, 4 (WENOFifthOrderCalculateWeightsChar, Interp1PrimFifthOrderWENOChar.c:86-198)
    double f0[6], fm[6], f4[6], u0[6], um[6], u4[6], c0[6], cm[6], c4[6];
#pragma unroll
    for (int p = 0; p < 6; p++) {
      const double x0 = ld(p, 0), x1 = ld(p, 1), x2 = ld(p, 2), x3 = ld(p, 3), x4 = ld(p, 4);
      const double F0 = ld(p, DIR + 1), F1 = ld(p, 5), F2 = ld(p, 6), F3 = ld(p, 7), F4 = ld(p, 8);
      const double dC = ld(p, 9) - x4;
      proj_acoustic<DIR>(s, F0, F1, F2, F3, F4, f0[p], fm[p], f4[p]);
      proj_acoustic<DIR>(s, x0, x1, x2, x3, x4, u0[p], um[p], u4[p]);
      c0[p] = u0[p] - s.k0 * dC;
      cm[p] = um[p] + s.kh * dC;
      c4[p] = u4[p] + s.kh * dC;
    }
    double fl, fr, ul, ur;
    const double l0 = UPW == UPWIND_ROE ? harten(s.vn) : alpha;
    const double lm = UPW == UPWIND_ROE ? harten(s.vn - s.a) : alpha;
    const double l4 = UPW == UPWIND_ROE ? harten(s.vn + s.a) : alpha;
    weno5_lr<W>(f0, f0, k.eps4, fl, fr);
    weno5_lr<W>(u0, c0, k.eps4, ul, ur);
    const double g0 = 0.5 * (fl + fr) - l0 * (0.5 * (ur - ul));
    weno5_lr<W>(fm, fm, k.eps4, fl, fr);
    weno5_lr<W>(um, cm, k.eps4, ul, ur);
    const double gm = 0.5 * (fl + fr) - lm * (0.5 * (ur - ul));
    weno5_lr<W>(f4, f4, k.eps4, fl, fr);
    weno5_lr<W>(u4, c4, k.eps4, ul, ur);
    const double g4 = 0.5 * (fl + fr) - l4 * (0.5 * (ur - ul));
    // shear rows: the modified energy does not enter them (L[row][4] = 0)
    double fa[6], fb[6], ua[6], ub[6];
#pragma unroll
    for (int p = 0; p < 6; p++) {
      proj_shear<DIR>(s, ld(p, DIR + 1), ld(p, 5), ld(p, 6), ld(p, 7), fa[p], fb[p]);
      proj_shear<DIR>(s, ld(p, 0), ld(p, 1), ld(p, 2), ld(p, 3), ua[p], ub[p]);
    }
    const double ls = UPW == UPWIND_ROE ? harten(s.vn) : alpha;
    weno5_lr<W>(fa, fa, k.eps4, fl, fr);
    weno5_lr<W>(ua, ua, k.eps4, ul, ur);
    const double ga = 0.5 * (fl + fr) - ls * (0.5 * (ur - ul));
    weno5_lr<W>(fb, fb, k.eps4, fl, fr);
    weno5_lr<W>(ub, ub, k.eps4, ul, ur);
    const double gb = 0.5 * (fl + fr) - ls * (0.5 * (ur - ul));
    right_mult<DIR>(s, g0, gm, g4, ga, gb, fI);
  } else {
    // component-wise (Interp1PrimFifthOrderWENO.c:74-168)
    double fh[NV], du[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double b[6], g[6];
#pragma unroll
      for (int p = 0; p < 6; p++) b[p] = ld(p, v == 0 ? DIR + 1 : 4 + v);
      double fl, fr, ul, ur;
      weno5_lr<W>(b, b, k.eps4, fl, fr);
#pragma unroll
      for (int p = 0; p < 6; p++) {
        b[p] = ld(p, v);
        g[p] = v == 4 ? ld(p, 9) : b[p];
      }
      weno5_lr<W>(b, g, k.eps4, ul, ur);
      fh[v] = 0.5 * (fl + fr);
      du[v] = 0.5 * (ur - ul);
    }
    if (UPW == UPWIND_RUSANOV) {
#pragma unroll
      for (int v = 0; v < NV; v++) fI[v] = fh[v] - alpha * du[v];
    } else {
      // R |Lambda| L (uR - uL)/2 (NavierStokes3DUpwind.c:40-124)
      double p0, pm, p4, pa, pb;
      proj_acoustic<DIR>(s, du[0], du[1], du[2], du[3], du[4], p0, pm, p4);
      proj_shear<DIR>(s, du[0], du[1], du[2], du[3], pa, pb);
      const double hs = harten(s.vn);
      double d[NV];
      right_mult<DIR>(s, hs * p0, harten(s.vn - s.a) * pm, harten(s.vn + s.a) * p4, hs * pa, hs * pb, d);
#pragma unroll
      for (int v = 0; v < NV; v++) fI[v] = fh[v] - d[v];
    }
  }
}

// The hyperbolic-term accumulator R, read and written plainly by the
// per-pass kernels.  In the one-launch-per-stage kernel (COH) blocks hand R
// to each other within the launch, so there R goes through agent-scope
// relaxed atomics: coherent loads and write-through stores (sc1) that
// bypass the stale lines of the eight non-coherent XCD L2s.  The stage
// kernel's dependency counters then need no L2 write-back or invalidate.
template <bool COH>
__device__ __forceinline__ double load_r(const double *p) {
  if (COH) return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return *p;
}
template <bool COH>
__device__ __forceinline__ void store_r(double *p, double v) {
  if (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

// Inputs of cell idx's output that come from global memory: its dxinv, the
// hyperbolic term so far (y, z), the viscous term and the RK4 operands (z).
// The marching passes issue these loads in one block with the next ring
// row's, so that their latencies overlap.
struct EmitIn {
  double dxi, par[NV], rin[NV], a[NV], b[NV];
};

template <int DIR, int RK = RK_CLASSIC, bool COH = false>
__device__ __forceinline__ EmitIn emit_load(const PassConst &k, const PassIO &io, size_t idx, int pos,
                                            const SspIO &sp = SspIO{}) {
  EmitIn in;
  in.dxi = k.dxg[(DIR >= 1 ? k.n[0] + 6 : 0) + (DIR == 2 ? k.n[1] + 6 : 0) + 3 + pos];
#pragma unroll
  for (int v = 0; v < NV; v++) in.par[v] = in.rin[v] = in.a[v] = in.b[v] = 0.0;
  if (DIR > 0)
#pragma unroll
    for (int v = 0; v < NV; v++) in.rin[v] = load_r<COH>(io.R + (size_t)v * k.ncell + idx);
  if (DIR == 2 && io.visc && io.mode != OUT_HYP)
#pragma unroll
    for (int c = 0; c < 4; c++) in.par[c + 1] = io.par[(size_t)c * k.ncell + idx];
  // the RK4 operands of the 3N update (u at stage 0, acc and u at stages
  // 1-2, acc at stage 3), one uniform branch per stage kind, each with its
  // loads in one block
  if (RK == RK_AFFINE && DIR == 2) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      in.a[v] = sp.A[(size_t)v * k.ncell + idx];
      in.b[v] = sp.B[(size_t)v * k.ncell + idx];
    }
  } else if (DIR == 2 && io.mode == OUT_STAGE) {
    if (io.stage == 0) {
#pragma unroll
      for (int v = 0; v < NV; v++) in.a[v] = io.u[(size_t)v * k.ncell + idx];
    } else if (io.stage == 3) {
#pragma unroll
      for (int v = 0; v < NV; v++) in.a[v] = io.acc[(size_t)v * k.ncell + idx];
    } else {
#pragma unroll
      for (int v = 0; v < NV; v++) {
        in.a[v] = io.acc[(size_t)v * k.ncell + idx];
        in.b[v] = io.u[(size_t)v * k.ncell + idx];
      }
    }
  }
  return in;
}

// Finish cell idx of the pass: flux difference and the pass's output.
template <int DIR, int RK = RK_CLASSIC, bool COH = false>
__device__ __forceinline__ void emit_store(const PassConst &k, const PassIO &io, size_t idx, const EmitIn &in,
                                           const double (&fm)[NV], const double (&fp)[NV],
                                           const SspIO &sp = SspIO{}) {
  double r[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double h = in.dxi * (fp[v] - fm[v]);
    // hyp = (hx + hy) + hz and rhs = (-hyp) + par (TimeRHSFunctionExplicit.c:88-93)
    if (DIR == 0) r[v] = h;
    else if (DIR == 1 || io.mode == OUT_HYP) r[v] = in.rin[v] + h;
    else r[v] = (-(in.rin[v] + h)) + in.par[v];
  }
  if (DIR < 2) {
#pragma unroll
    for (int v = 0; v < NV; v++) store_r<COH>(io.R + (size_t)v * k.ncell + idx, r[v]);
    return;
  }
  if (io.mode != OUT_STAGE) {
#pragma unroll
    for (int v = 0; v < NV; v++) io.out[(size_t)v * k.ncell + idx] = r[v];
    return;
  }
  double nu[NV];
  if (RK == RK_AFFINE) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const size_t o = (size_t)v * k.ncell + idx;
      nu[v] = in.b[v] + sp.c2 * r[v];
      sp.O1[o] = in.a[v] + sp.c1 * r[v];
      if (sp.nout > 1) sp.O2[o] = nu[v];
      if (sp.nout > 2) sp.O3[o] = in.a[v] + sp.c3 * r[v];
    }
  } else {
  // RK4 stage update in 3N form (TimeRK.c:126-195)
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const size_t o = (size_t)v * k.ncell + idx;
    if (io.stage == 0) {
      io.acc[o] = in.a[v] + io.cb * r[v];
      nu[v] = in.a[v] + io.ca * r[v];
      io.S[o] = nu[v];
    } else if (io.stage < 3) {
      io.acc[o] = in.a[v] + io.cb * r[v];
      nu[v] = in.b[v] + io.ca * r[v];
      io.S[o] = nu[v];
    } else {
      nu[v] = in.a[v] + io.cb * r[v];
      io.u[o] = nu[v];
    }
  }
  }
  if (io.visc) {
    // primitive variables (u, v, w, T) of the new state for the next
    // viscous-term evaluation (NavierStokes3DParabolicFunction.c:85-108)
    const Prim q = flowvar(nu, k.gamma);
    io.Q[idx] = q.vx;
    io.Q[k.ncell + idx] = q.vy;
    io.Q[2 * k.ncell + idx] = q.vz;
    io.Q[3 * k.ncell + idx] = k.gamma * q.P / q.rho;
  }
}

template <int DIR, int RK = RK_CLASSIC, bool COH = false>
__device__ __forceinline__ void emit(const PassConst &k, const PassIO &io, size_t idx, int pos,
                                     const double (&fm)[NV], const double (&fp)[NV],
                                     const SspIO &sp = SspIO{}) {
  emit_store<DIR, RK, COH>(k, io, idx, emit_load<DIR, RK, COH>(k, io, idx, pos, sp), fm, fp, sp);
}

// ---------------------------------------------------------------------------
// x pass: one wavefront per x-line, LPB lines per block
// ---------------------------------------------------------------------------
constexpr int XCH = 64;        // interfaces per chunk
constexpr int XNC = XCH + 6;   // cells per chunk
constexpr int XPAD = 72;

template <int LPB>
struct XShared {
  double sC[LPB][NQ][XPAD];
  double sF[LPB][NV][XCH + 1];  // slot 0: fI[x0-1]; slot 1+l: fI[x0+l]
  double sF0[LPB][NV];          // fI[0] of the line
};

// lines LPB*group .. LPB*group + LPB-1 of the grid (line = j + ny*k)
template <int INTERP, int UPW, int LPB, int W, bool COH = false>
__device__ __forceinline__ void pass_x_body(const PassConst &k, const PassIO &io, size_t group,
                                            XShared<LPB> &sh) {
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nx = k.n[0];
  const size_t nlines = (size_t)k.n[1] * k.n[2];
  const size_t line = group * LPB + w;
  const bool active = line < nlines;
  const size_t base = (active ? line : 0) * (size_t)nx;
  int x0 = 0;
  for (; x0 < nx; x0 += XCH) {
    if (x0 > 0) {
      // carry fI[x0-1] into slot 0 (the previous chunk's output phase is done)
      if (lane < NV) sh.sF[w][lane][0] = sh.sF[w][lane][XCH];
    }
    const int ncells = min(XNC, nx + 6 - x0);
    for (int c = lane; c < ncells; c += 64) {
      const size_t cell = base + wrap(x0 - 3 + c, nx);
      double u[NV], cd[NQ];
#pragma unroll
      for (int v = 0; v < NV; v++) u[v] = io.U[(size_t)v * k.ncell + cell];
      cell_data<0>(u, k, cd);
#pragma unroll
      for (int q = 0; q < NQ; q++) sh.sC[w][q][c] = cd[q];
    }
    __syncthreads();
    if (active && x0 + lane < nx) {
      auto ld = [&](int s, int q) { return sh.sC[w][q][lane + s]; };
      double f[NV];
      iface<0, INTERP, UPW, W>(ld, k, f);
#pragma unroll
      for (int v = 0; v < NV; v++) sh.sF[w][v][lane + 1] = f[v];
      if (x0 == 0 && lane == 0)
#pragma unroll
        for (int v = 0; v < NV; v++) sh.sF0[w][v] = f[v];
    }
    __syncthreads();
    const int cell = x0 - 1 + lane;
    if (active && cell >= 0 && cell <= nx - 2) {
      double fm[NV], fp[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        fm[v] = sh.sF[w][v][lane];
        fp[v] = sh.sF[w][v][lane + 1];
      }
      emit<0, RK_CLASSIC, COH>(k, io, base + cell, cell, fm, fp);
    }
    __syncthreads();
  }
  // last cell: fI[nx-1] of the last chunk, fI[nx] = fI[0]
  x0 -= XCH;
  if (active && lane == 0) {
    double fm[NV], fp[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      fm[v] = sh.sF[w][v][nx - x0];
      fp[v] = sh.sF0[w][v];
    }
    emit<0, RK_CLASSIC, COH>(k, io, base + nx - 1, nx - 1, fm, fp);
  }
}

// ---------------------------------------------------------------------------
// y, z passes: 64 columns marching along DIR, MR rows per step
// ---------------------------------------------------------------------------
constexpr int MW = 64;  // columns

template <int MR>
struct MShared {
  static constexpr int NRING = MR + 5;  // cell rows t0-3 .. t0+MR+1
  static constexpr int NFR = MR + 1;    // interface rows t0-1 .. t0+MR-1
  double sC[NQ][NRING][MW];
  double sF[NV][NFR][MW];
  double sF0[NV][MW];
};

// columns MW*group .. MW*group + MW-1 of the plane across DIR
template <int DIR, int INTERP, int UPW, int MR, int W, int RK = RK_CLASSIC, bool COH = false>
__device__ __forceinline__ void pass_m_body(const PassConst &k, const PassIO &io, size_t group,
                                            MShared<MR> &sh, const SspIO &sp = SspIO{}) {
  constexpr int NRING = MShared<MR>::NRING, NFR = MShared<MR>::NFR;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nx = k.n[0], ny = k.n[1];
  const int n = k.n[DIR];
  // the 64 lanes take 64 consecutive columns c = i + nx*o of the plane
  // across DIR (o = z index in the y pass, y index in the z pass)
  const size_t ncol = k.ncell / n;
  const size_t c = group * MW + lane;
  const bool col = c < ncol;
  const size_t line = col ? c : ncol - 1;
  const size_t o = line / nx;
  const size_t stride = DIR == 1 ? (size_t)nx : (size_t)nx * ny;
  const size_t cbase = DIR == 1 ? o * nx * ny + (line - o * nx) : line;  // cell at position 0

  auto load_row = [&](int r) {  // r in [-3, MR+1]
    const size_t cell = cbase + (size_t)wrap(r, n) * stride;
    double u[NV], cd[NQ];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = io.U[(size_t)v * k.ncell + cell];
    cell_data<DIR>(u, k, cd);
    const int slot = (r + NRING) % NRING;
#pragma unroll
    for (int q = 0; q < NQ; q++) sh.sC[q][slot][lane] = cd[q];
  };

  // ring row r's state, loaded one march step ahead in the same block as
  // the output's operands, so that their latencies overlap
  auto fetch_row = [&](int r, double (&u)[NV]) {  // r in [-3, n+MR]
    const size_t cell = cbase + (size_t)wrap(r, n) * stride;
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = io.U[(size_t)v * k.ncell + cell];
  };
  auto stage_row = [&](int r, const double (&u)[NV]) {
    double cd[NQ];
    cell_data<DIR>(u, k, cd);
    const int slot = (r + NRING) % NRING;
#pragma unroll
    for (int q = 0; q < NQ; q++) sh.sC[q][slot][lane] = cd[q];
  };

  // rows -3..MR+1
  for (int r = w; r < MR + 5; r += MR) load_row(r - 3);
  for (int t0 = 0; t0 < n; t0 += MR) {
    __syncthreads();
    const int t = t0 + w;
    if (t < n) {
      const int sl0 = (t - 3 + NRING) % NRING;
      auto ld = [&](int s, int q) {
        int sl = sl0 + s;
        sl = sl >= NRING ? sl - NRING : sl;
        return sh.sC[q][sl][lane];
      };
      double f[NV];
      iface<DIR, INTERP, UPW, W>(ld, k, f);
      const int fs = t % NFR;
#pragma unroll
      for (int v = 0; v < NV; v++) sh.sF[v][fs][lane] = f[v];
      if (t == 0)
#pragma unroll
        for (int v = 0; v < NV; v++) sh.sF0[v][lane] = f[v];
    }
    const int e = t0 - 1 + w;  // row emitted by this wavefront
    const bool emits = col && e >= 0 && e <= n - 2;
    __syncthreads();
    // the next step's ring row (its slot held row t0-3+w, which no later
    // step reads) and the output's operands, loaded by every lane (a stand-in
    // row 0 after the last step, output rows clamped into the line) so that
    // no divergent branch separates the loads
    const bool more = t0 + MR < n;
    const int rn = more ? t0 + MR + 2 + w : 0;  // at most n+MR
    double un[NV];
    fetch_row(rn, un);
    const int ec = e < 0 ? 0 : (e > n - 2 ? n - 2 : e);
    const EmitIn in = emit_load<DIR, RK, COH>(k, io, cbase + (size_t)ec * stride, ec, sp);
    if (more) stage_row(rn, un);
    if (emits) {
      double fm[NV], fp[NV];
      const int a = e % NFR, b = (e + 1) % NFR;
#pragma unroll
      for (int v = 0; v < NV; v++) {
        fm[v] = sh.sF[v][a][lane];
        fp[v] = sh.sF[v][b][lane];
      }
      emit_store<DIR, RK, COH>(k, io, cbase + (size_t)e * stride, in, fm, fp, sp);
    }
  }
  // last row: fI[n-1] and fI[n] = fI[0]
  if (w == 0 && col) {
    double fm[NV], fp[NV];
    const int a = (n - 1) % NFR;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      fm[v] = sh.sF[v][a][lane];
      fp[v] = sh.sF0[v][lane];
    }
    emit<DIR, RK, COH>(k, io, cbase + (size_t)(n - 1) * stride, n - 1, fm, fp, sp);
  }
}

template <int INTERP, int UPW, int W>
__global__ void __launch_bounds__(256) k_pass_x(PassConst k, PassIO io) {
  __shared__ XShared<4> sh;
  pass_x_body<INTERP, UPW, 4, W>(k, io, blockIdx.x, sh);
}

// at most 256 registers: two waves per SIMD (the characteristic passes sit
// at the limit, and one register more would halve the occupancy)
template <int DIR, int INTERP, int UPW, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_pass_m(PassConst k, PassIO io) {
  __shared__ MShared<4> sh;
  pass_m_body<DIR, INTERP, UPW, 4, W>(k, io, blockIdx.x, sh);
}

// the z pass with the three-output stage update (SSPRK3 stages)
template <int INTERP, int UPW, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_pass_z_affine(PassConst k, PassIO io,
                                                                                             SspIO sp) {
  __shared__ MShared<4> sh;
  pass_m_body<2, INTERP, UPW, 4, W, RK_AFFINE>(k, io, blockIdx.x, sh, sp);
}

// ---------------------------------------------------------------------------
// batched passes: all members of an ensemble in one persistent launch
// ---------------------------------------------------------------------------
struct EnsArgs {
  const EnsMember *mem;    // member table
  int nmem;
  const unsigned *item0;   // first work item of each member, nmem + 1 entries
  unsigned nitems;
  double gamma, gm1, inv_gm1, eps4;
  const double *U;         // stage states of all members (member m at 5*off)
  double *R, *u, *S, *acc;
  int mode, stage, visc;
};

// the member of a work item and the item's index within the member
__device__ __forceinline__ int ens_member(const EnsArgs &a, unsigned item) {
  int lo = 0, hi = a.nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.item0[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void ens_setup(const EnsArgs &a, int m, int dir, PassConst &k, PassIO &io) {
  const EnsMember M = a.mem[m];
  for (int d = 0; d < 3; d++) k.n[d] = M.n[d];
  k.ncell = M.ncell;
  k.gamma = a.gamma;
  k.gm1 = a.gm1;
  k.inv_gm1 = a.inv_gm1;
  k.eps4 = a.eps4;
  k.dxg = M.dxg;
  const size_t o = 5 * M.off;
  io.U = a.U + o;
  io.R = a.R + o;
  io.par = M.par;
  io.u = a.u + o;
  io.S = a.S ? a.S + o : nullptr;
  io.acc = a.acc + o;
  io.out = nullptr;
  io.Q = M.q;
  io.visc = a.visc;
  io.mode = a.mode;
  io.stage = a.stage;
  const double A_sub[3] = {0.5, 0.5, 1.0}, b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  io.ca = a.stage < 3 ? M.dt * A_sub[a.stage] : 0.0;
  io.cb = M.dt * b[a.stage];
}

template <int INTERP, int UPW, int LPB, int W>
__global__ void __launch_bounds__(64 * LPB) k_ens_x(EnsArgs a) {
  __shared__ XShared<LPB> sh;
  for (unsigned item = blockIdx.x; item < a.nitems; item += gridDim.x) {
    const int m = ens_member(a, item);
    PassConst k;
    PassIO io;
    ens_setup(a, m, 0, k, io);
    pass_x_body<INTERP, UPW, LPB, W>(k, io, item - a.item0[m], sh);
    __syncthreads();
  }
}

template <int DIR, int INTERP, int UPW, int MR, int W>
__global__ void __launch_bounds__(64 * MR) __attribute__((amdgpu_waves_per_eu(2))) k_ens_m(EnsArgs a) {
  __shared__ MShared<MR> sh;
  for (unsigned item = blockIdx.x; item < a.nitems; item += gridDim.x) {
    const int m = ens_member(a, item);
    PassConst k;
    PassIO io;
    ens_setup(a, m, DIR, k, io);
    pass_m_body<DIR, INTERP, UPW, MR, W>(k, io, item - a.item0[m], sh);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// one persistent launch per RK stage: the x, y and z work items of all members
// ---------------------------------------------------------------------------
// Blocks take items in ticket order: every x item of every member (xg
// groups of x-lines each), then the y items, then the z items.  A member's y items add to the rhs its x items
// wrote, and its z items read what its y items left, so a y (z) item first
// waits until all x (y) items of its member are done, as counted in ctr.
// Every item a block waits on holds a smaller ticket, taken by a block that
// is running and that waits only on smaller tickets in turn, so the waits
// end whatever the number of resident blocks.  Each wait is bounded: after
// spin_limit polls it sets the error flag ctr[2] and goes on, so the kernel
// always exits, and the host reports the flag.  The last block to leave
// zeroes the ticket and the counters for the next launch, so graph replays
// need no memset.
struct EnsStageArgs {
  EnsArgs a;                 // item0 and nitems of a are not used
  const unsigned *item0[3];  // first item of each member per direction
  unsigned nitems[3];
  const unsigned *xgroups0;  // first x-line group of each member
  unsigned xg;               // x-line groups per x item
  unsigned *ctr;             // [0] ticket, [1] blocks done, [2] error flag, [3] unused,
                             // [4, 4+nmem) x items done, [4+nmem, 4+2nmem) y items done
  unsigned spin_limit;
};

template <int SH>
union StageShared {
  XShared<SH> x;
  MShared<SH> m;
};

// the member of item `item` of a direction whose first items are item0
__device__ __forceinline__ int stage_member(const unsigned *item0, int nmem, unsigned item) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (item0[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Wave 0 acts for the block: all 64 of its lanes take the same branches and
// only lane 0 contributes to the counters (the others add 0).  A branch on
// threadIdx.x == 0 between two barriers of the item loop would let the
// compiler move lane 0's atomics out of the loop (a loop that lanes leave
// one by one), so that the rest of the block passed the next barrier before
// lane 0 had counted its item or fetched the next ticket.
__device__ __forceinline__ bool lead_wave() { return __builtin_amdgcn_readfirstlane(threadIdx.x) < 64u; }

__device__ __forceinline__ unsigned lane0_one() { return (threadIdx.x & 63u) == 0u ? 1u : 0u; }

// The only data blocks of the launch hand to each other is R, and every
// access to it here is an agent-scope atomic (load_r / store_r with COH):
// the stores write through the XCD's L2, the loads bypass stale lines of
// the L1 and the L2.  A count of finished items therefore needs only each
// wave's stores acknowledged before it, and a wait only the count seen
// before the block's first load of R; neither writes back nor invalidates
// an L2 (the fences a release or acquire at agent scope would add, once per
// work item, for every XCD).
//
// Whole block: wait until *c >= need (bounded).  The barrier keeps the
// block's loads after the poll that saw the count.
__device__ __forceinline__ void stage_wait(const unsigned *c, unsigned need, unsigned *err, unsigned limit) {
  if (lead_wave()) {
    for (unsigned polls = 0;; polls++) {
      const unsigned v =
          __builtin_amdgcn_readfirstlane(__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (v >= need) break;
      if (polls >= limit) {
        __hip_atomic_fetch_or(err, lane0_one(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
}

// Whole block: every wave waits until its write-through stores of R are
// acknowledged, then one wave counts the item.
__device__ __forceinline__ void stage_signal(unsigned *c) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (lead_wave()) __hip_atomic_fetch_add(c, lane0_one(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int INTERP, int UPW, int SH, int W>
__global__ void __launch_bounds__(64 * SH) __attribute__((amdgpu_waves_per_eu(2))) k_ens_stage(EnsStageArgs s) {
  __shared__ StageShared<SH> sh;
  __shared__ unsigned s_ticket, s_last;
  const int nmem = s.a.nmem;
  unsigned *const done_x = s.ctr + 4, *const done_y = s.ctr + 4 + nmem;
  const unsigned nx_items = s.nitems[0], nxy_items = s.nitems[0] + s.nitems[1];
  const unsigned total = nxy_items + s.nitems[2];
  for (;;) {
    if (lead_wave()) {
      const unsigned old = __hip_atomic_fetch_add(&s.ctr[0], lane0_one(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_ticket = __builtin_amdgcn_readfirstlane(old);  // lane 0's: the value before its add
    }
    __syncthreads();
    const unsigned t = __builtin_amdgcn_readfirstlane(s_ticket);
    if (t >= total) break;
    const int phase = t < nx_items ? 0 : (t < nxy_items ? 1 : 2);
    const unsigned item = t - (phase == 0 ? 0u : (phase == 1 ? nx_items : nxy_items));
    const unsigned *const item0 = s.item0[phase];
    const int m = stage_member(item0, nmem, item);
    if (phase > 0) {
      const unsigned *before = s.item0[phase - 1];
      stage_wait(phase == 1 ? &done_x[m] : &done_y[m], before[m + 1] - before[m], &s.ctr[2], s.spin_limit);
    }
    const unsigned local = item - item0[m];
    if (phase == 0) {
      PassConst k;
      PassIO io;
      ens_setup(s.a, m, 0, k, io);
      const unsigned g1 = min((local + 1) * s.xg, s.xgroups0[m + 1] - s.xgroups0[m]);
      for (unsigned g = local * s.xg; g < g1; g++) {
        pass_x_body<INTERP, UPW, SH, W, true>(k, io, g, sh.x);
        __syncthreads();
      }
      stage_signal(&done_x[m]);
    } else if (phase == 1) {
      PassConst k;
      PassIO io;
      ens_setup(s.a, m, 1, k, io);
      pass_m_body<1, INTERP, UPW, SH, W, RK_CLASSIC, true>(k, io, local, sh.m);
      stage_signal(&done_y[m]);
    } else {
      PassConst k;
      PassIO io;
      ens_setup(s.a, m, 2, k, io);
      pass_m_body<2, INTERP, UPW, SH, W, RK_CLASSIC, true>(k, io, local, sh.m);
      __syncthreads();
    }
  }
  // the last block out resets the counters (every other block has stopped
  // touching them)
  if (lead_wave()) {
    const unsigned old = __hip_atomic_fetch_add(&s.ctr[1], lane0_one(), __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = __builtin_amdgcn_readfirstlane(old) == gridDim.x - 1 ? 1u : 0u;
  }
  __syncthreads();
  if (__builtin_amdgcn_readfirstlane(s_last)) {
    for (unsigned i = threadIdx.x; i < 2u * (unsigned)nmem; i += blockDim.x) s.ctr[4 + i] = 0u;
    if (threadIdx.x < 2u) s.ctr[threadIdx.x] = 0u;  // ticket, blocks done
  }
}

// primitive variables (u, v, w, T) of a state
__global__ void __launch_bounds__(256) k_prims(const double *__restrict__ u, double *__restrict__ Q, size_t ncell,
                                               double gamma) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  double s[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) s[v] = u[(size_t)v * ncell + idx];
  const Prim q = flowvar(s, gamma);
  Q[idx] = q.vx;
  Q[ncell + idx] = q.vy;
  Q[2 * ncell + idx] = q.vz;
  Q[3 * ncell + idx] = gamma * q.P / q.rho;
}

template <int DIR, int INTERP, int UPW, int W>
void launch_t(const Grid &g, const PassConst &k, const PassIO &io, hipStream_t st) {
  if (DIR == 0) {
    const size_t nlines = (size_t)g.n[1] * g.n[2];
    hipLaunchKernelGGL((k_pass_x<INTERP, UPW, W>), dim3((unsigned)((nlines + 3) / 4)), dim3(256), 0, st, k, io);
  } else {
    const size_t nb = (g.ncell / g.n[DIR] + MW - 1) / MW;
    hipLaunchKernelGGL((k_pass_m<DIR == 0 ? 1 : DIR, INTERP, UPW, W>), dim3((unsigned)nb), dim3(256), 0, st, k, io);
  }
}

template <int DIR, int INTERP, int UPW>
void launch_w(const Grid &g, const Scheme &sc, const PassConst &k, const PassIO &io, hipStream_t st) {
  if (sc.weno == WENO_MAPPED) launch_t<DIR, INTERP, UPW, WENO_MAPPED>(g, k, io, st);
  else if (sc.weno == WENO_Z) launch_t<DIR, INTERP, UPW, WENO_Z>(g, k, io, st);
  else if (sc.weno == WENO_YC) launch_t<DIR, INTERP, UPW, WENO_YC>(g, k, io, st);
  else if (sc.weno == WENO_OPTIMAL) launch_t<DIR, INTERP, UPW, WENO_OPTIMAL>(g, k, io, st);
  else launch_t<DIR, INTERP, UPW, WENO_JS>(g, k, io, st);
}

template <int DIR>
void launch_d(const Grid &g, const Scheme &sc, const PassConst &k, const PassIO &io, hipStream_t st) {
  if (sc.interp == INTERP_CHAR) {
    if (sc.upwind == UPWIND_ROE) launch_w<DIR, INTERP_CHAR, UPWIND_ROE>(g, sc, k, io, st);
    else launch_w<DIR, INTERP_CHAR, UPWIND_RUSANOV>(g, sc, k, io, st);
  } else {
    if (sc.upwind == UPWIND_ROE) launch_w<DIR, INTERP_COMP, UPWIND_ROE>(g, sc, k, io, st);
    else launch_w<DIR, INTERP_COMP, UPWIND_RUSANOV>(g, sc, k, io, st);
  }
}

}  // namespace

namespace {
void pass_args(const Grid &g, const Scheme &sc, const PassBuffers &b, int dir, int mode, int stage, double ca,
               double cb, const double *dxg, PassConst &k, PassIO &io) {
  for (int d = 0; d < 3; d++) k.n[d] = g.n[d];
  k.ncell = g.ncell;
  k.gamma = sc.gamma;
  k.gm1 = sc.gamma - 1.0;
  k.inv_gm1 = 1.0 / (sc.gamma - 1.0);
  k.eps4 = 4.0 * sc.eps;
  k.dxg = dxg;
  io.U = b.U;
  io.R = b.R;
  io.par = b.par;
  io.u = b.u;
  io.S = b.S;
  io.acc = b.acc;
  io.out = b.out;
  io.Q = b.Q;
  io.visc = sc.Re > 0 ? 1 : 0;
  io.mode = mode;
  io.stage = stage;
  io.ca = ca;
  io.cb = cb;
  (void)dir;
}

template <int INTERP, int UPW, int W>
void launch_ssp_t(const Grid &g, const PassConst &k, const PassIO &io, const SspIO &sp, hipStream_t st) {
  const size_t nb = (g.ncell / g.n[2] + MW - 1) / MW;
  hipLaunchKernelGGL((k_pass_z_affine<INTERP, UPW, W>), dim3((unsigned)nb), dim3(256), 0, st, k, io, sp);
}

template <int INTERP, int UPW>
void launch_ssp_w(const Grid &g, const Scheme &sc, const PassConst &k, const PassIO &io, const SspIO &sp,
                  hipStream_t st) {
  if (sc.weno == WENO_MAPPED) launch_ssp_t<INTERP, UPW, WENO_MAPPED>(g, k, io, sp, st);
  else if (sc.weno == WENO_Z) launch_ssp_t<INTERP, UPW, WENO_Z>(g, k, io, sp, st);
  else if (sc.weno == WENO_YC) launch_ssp_t<INTERP, UPW, WENO_YC>(g, k, io, sp, st);
  else if (sc.weno == WENO_OPTIMAL) launch_ssp_t<INTERP, UPW, WENO_OPTIMAL>(g, k, io, sp, st);
  else launch_ssp_t<INTERP, UPW, WENO_JS>(g, k, io, sp, st);
}
}  // namespace

void launch_pass(const Grid &g, const Scheme &sc, int dir, const PassBuffers &b, int mode, int stage, double ca,
                 double cb, const double *dxg, hipStream_t st) {
  PassConst k;
  PassIO io;
  pass_args(g, sc, b, dir, mode, stage, ca, cb, dxg, k, io);
  if (dir == 0) launch_d<0>(g, sc, k, io, st);
  else if (dir == 1) launch_d<1>(g, sc, k, io, st);
  else launch_d<2>(g, sc, k, io, st);
}

void launch_pass_z_affine(const Grid &g, const Scheme &sc, const PassBuffers &b, const double *A, const double *B,
                          double *O1, double *O2, double *O3, double c1, double c2, double c3, int nout,
                          const double *dxg, hipStream_t st) {
  PassConst k;
  PassIO io;
  pass_args(g, sc, b, 2, OUT_STAGE, 0, 0.0, 0.0, dxg, k, io);
  const SspIO sp{A, B, O1, O2, O3, c1, c2, c3, nout};
  if (sc.interp == INTERP_CHAR) {
    if (sc.upwind == UPWIND_ROE) launch_ssp_w<INTERP_CHAR, UPWIND_ROE>(g, sc, k, io, sp, st);
    else launch_ssp_w<INTERP_CHAR, UPWIND_RUSANOV>(g, sc, k, io, sp, st);
  } else {
    if (sc.upwind == UPWIND_ROE) launch_ssp_w<INTERP_COMP, UPWIND_ROE>(g, sc, k, io, sp, st);
    else launch_ssp_w<INTERP_COMP, UPWIND_RUSANOV>(g, sc, k, io, sp, st);
  }
}


namespace {
template <int INTERP, int UPW, int W>
void ens_launch_w(int dir, int shape, const EnsArgs &a, unsigned grid, hipStream_t st) {
  if (dir == 0) {
    if (shape == 0) hipLaunchKernelGGL((k_ens_x<INTERP, UPW, 4, W>), dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_ens_x<INTERP, UPW, 2, W>), dim3(grid), dim3(128), 0, st, a);
  } else if (dir == 1) {
    if (shape == 0) hipLaunchKernelGGL((k_ens_m<1, INTERP, UPW, 4, W>), dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_ens_m<1, INTERP, UPW, 2, W>), dim3(grid), dim3(128), 0, st, a);
  } else {
    if (shape == 0) hipLaunchKernelGGL((k_ens_m<2, INTERP, UPW, 4, W>), dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_ens_m<2, INTERP, UPW, 2, W>), dim3(grid), dim3(128), 0, st, a);
  }
}

template <int INTERP, int UPW>
void ens_launch_t(int weno, int dir, int shape, const EnsArgs &a, unsigned grid, hipStream_t st) {
  if (weno == WENO_MAPPED) ens_launch_w<INTERP, UPW, WENO_MAPPED>(dir, shape, a, grid, st);
  else if (weno == WENO_Z) ens_launch_w<INTERP, UPW, WENO_Z>(dir, shape, a, grid, st);
  else if (weno == WENO_YC) ens_launch_w<INTERP, UPW, WENO_YC>(dir, shape, a, grid, st);
  else if (weno == WENO_OPTIMAL) ens_launch_w<INTERP, UPW, WENO_OPTIMAL>(dir, shape, a, grid, st);
  else ens_launch_w<INTERP, UPW, WENO_JS>(dir, shape, a, grid, st);
}
}  // namespace

unsigned ens_items_per_member(int dir, int shape, const int n[3]) {
  const size_t ncell = (size_t)n[0] * n[1] * n[2];
  if (dir == 0) {
    const size_t lpb = shape == 0 ? 4 : 2;
    return (unsigned)(((size_t)n[1] * n[2] + lpb - 1) / lpb);
  }
  return (unsigned)((ncell / n[dir] + MW - 1) / MW);
}

void launch_ens_pass(const Scheme &sc, int dir, int shape, const EnsMember *mem, int nmem, const unsigned *item0,
                     unsigned nitems, unsigned grid, const double *U, double *R, double *u, double *S, double *acc,
                     int stage, hipStream_t st) {
  EnsArgs a;
  a.mem = mem;
  a.nmem = nmem;
  a.item0 = item0;
  a.nitems = nitems;
  a.gamma = sc.gamma;
  a.gm1 = sc.gamma - 1.0;
  a.inv_gm1 = 1.0 / (sc.gamma - 1.0);
  a.eps4 = 4.0 * sc.eps;
  a.U = U;
  a.R = R;
  a.u = u;
  a.S = S;
  a.acc = acc;
  a.mode = OUT_STAGE;
  a.stage = stage;
  a.visc = sc.Re > 0;
  if (grid == 0 || grid > nitems) grid = nitems;
  if (sc.interp == INTERP_CHAR) {
    if (sc.upwind == UPWIND_ROE) ens_launch_t<INTERP_CHAR, UPWIND_ROE>(sc.weno, dir, shape, a, grid, st);
    else ens_launch_t<INTERP_CHAR, UPWIND_RUSANOV>(sc.weno, dir, shape, a, grid, st);
  } else {
    if (sc.upwind == UPWIND_ROE) ens_launch_t<INTERP_COMP, UPWIND_ROE>(sc.weno, dir, shape, a, grid, st);
    else ens_launch_t<INTERP_COMP, UPWIND_RUSANOV>(sc.weno, dir, shape, a, grid, st);
  }
}

namespace {
// blocks of `kern` that one CU holds at once
template <class K>
unsigned resident_per_cu(K kern, int threads) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, threads, 0) != hipSuccess || nb < 1) nb = 1;
  return (unsigned)nb;
}

template <int INTERP, int UPW, int W>
void stage_launch_w(int rows, const EnsStageArgs &s, unsigned grid, int cus, hipStream_t st) {
  const unsigned total = s.nitems[0] + s.nitems[1] + s.nitems[2];
  if (rows == 4) {
    static const unsigned per_cu = resident_per_cu(k_ens_stage<INTERP, UPW, 4, W>, 256);
    if (grid == 0) grid = per_cu * (unsigned)cus;
    hipLaunchKernelGGL((k_ens_stage<INTERP, UPW, 4, W>), dim3(std::min(grid, total)), dim3(256), 0, st, s);
  } else {
    static const unsigned per_cu = resident_per_cu(k_ens_stage<INTERP, UPW, 2, W>, 128);
    if (grid == 0) grid = per_cu * (unsigned)cus;
    hipLaunchKernelGGL((k_ens_stage<INTERP, UPW, 2, W>), dim3(std::min(grid, total)), dim3(128), 0, st, s);
  }
}

template <int INTERP, int UPW>
void stage_launch_t(int weno, int rows, const EnsStageArgs &s, unsigned grid, int cus, hipStream_t st) {
  if (weno == WENO_MAPPED) stage_launch_w<INTERP, UPW, WENO_MAPPED>(rows, s, grid, cus, st);
  else if (weno == WENO_Z) stage_launch_w<INTERP, UPW, WENO_Z>(rows, s, grid, cus, st);
  else if (weno == WENO_YC) stage_launch_w<INTERP, UPW, WENO_YC>(rows, s, grid, cus, st);
  else if (weno == WENO_OPTIMAL) stage_launch_w<INTERP, UPW, WENO_OPTIMAL>(rows, s, grid, cus, st);
  else stage_launch_w<INTERP, UPW, WENO_JS>(rows, s, grid, cus, st);
}
}  // namespace

void launch_ens_stage(const Scheme &sc, int rows, const EnsMember *mem, int nmem, const unsigned *const item0[3],
                      const unsigned nitems[3], const unsigned *xgroups0, unsigned xg, unsigned grid, int cus,
                      unsigned *ctr, unsigned spin_limit, const double *U, double *R, double *u, double *S,
                      double *acc, int stage, hipStream_t st) {
  EnsStageArgs s;
  EnsArgs &a = s.a;
  a.mem = mem;
  a.nmem = nmem;
  a.item0 = nullptr;
  a.nitems = 0;
  a.gamma = sc.gamma;
  a.gm1 = sc.gamma - 1.0;
  a.inv_gm1 = 1.0 / (sc.gamma - 1.0);
  a.eps4 = 4.0 * sc.eps;
  a.U = U;
  a.R = R;
  a.u = u;
  a.S = S;
  a.acc = acc;
  a.mode = OUT_STAGE;
  a.stage = stage;
  a.visc = sc.Re > 0;
  for (int d = 0; d < 3; d++) {
    s.item0[d] = item0[d];
    s.nitems[d] = nitems[d];
  }
  s.xgroups0 = xgroups0;
  s.xg = xg;
  s.ctr = ctr;
  s.spin_limit = spin_limit;
  if (sc.interp == INTERP_CHAR) {
    if (sc.upwind == UPWIND_ROE) stage_launch_t<INTERP_CHAR, UPWIND_ROE>(sc.weno, rows, s, grid, cus, st);
    else stage_launch_t<INTERP_CHAR, UPWIND_RUSANOV>(sc.weno, rows, s, grid, cus, st);
  } else {
    if (sc.upwind == UPWIND_ROE) stage_launch_t<INTERP_COMP, UPWIND_ROE>(sc.weno, rows, s, grid, cus, st);
    else stage_launch_t<INTERP_COMP, UPWIND_RUSANOV>(sc.weno, rows, s, grid, cus, st);
  }
}

void launch_prims(const Grid &g, const Scheme &sc, const double *u, double *Q, hipStream_t st) {
  hipLaunchKernelGGL(k_prims, dim3((unsigned)((g.ncell + 255) / 256)), dim3(256), 0, st, u, Q, g.ncell, sc.gamma);
}

}  // namespace hm
