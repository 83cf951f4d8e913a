// hm_adjoint.hip -- discrete adjoint of the WENO5 + RK4 step (gfx950, FP64).
//
// The reference has no adjoint; this differentiates this library's own
// discretisation of the inviscid NavierStokes3D right-hand side with
// component-wise WENO5-JS reconstruction and Rusanov upwinding
// (the operator of src/HyParFunctions/HyperbolicFunction.c:31-222 with
// NavierStokes3DUpwindRusanov, NavierStokes3DUpwind.c:349-420), by hand, in
// reverse mode.
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

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

constexpr int VJP_LINE_MAX = 512;  // longest grid line an adjoint workgroup holds

__device__ __forceinline__ int wrapn(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// 1/d, d > 0 and normal: hardware estimate and two Newton-Raphson steps
__device__ __forceinline__ double vjp_rcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
}

// reverse of weno5(m3..p2; g3..q2): adds d out/d m (smoothness stencil) to mb
// and d out/d g (interpolated values) to gb, scaled by ob; returns out
__device__ __forceinline__ double weno5_vjp(const double m[5], const double g[5], double eps, double ob,
                                            double mb[5], double gb[5]) {
  const double c1312 = 13.0 / 12.0, c14 = 0.25, s6 = 1.0 / 6.0;
  const double t1 = m[0] - 2 * m[1] + m[2], r1 = m[0] - 4 * m[1] + 3 * m[2];
  const double t2 = m[1] - 2 * m[2] + m[3], r2 = m[1] - m[3];
  const double t3 = m[2] - 2 * m[3] + m[4], r3 = 3 * m[2] - 4 * m[3] + m[4];
  const double b1 = c1312 * t1 * t1 + c14 * r1 * r1;
  const double b2 = c1312 * t2 * t2 + c14 * r2 * r2;
  const double b3 = c1312 * t3 * t3 + c14 * r3 * r3;
  const double e1 = b1 + eps, e2 = b2 + eps, e3 = b3 + eps;
  const double s1 = e1 * e1, s2 = e2 * e2, s3 = e3 * e3;
  const double a1 = 0.1 * (s2 * s3), a2 = 0.6 * (s1 * s3), a3 = 0.3 * (s1 * s2);
  const double inv = vjp_rcp(a1 + a2 + a3);
  const double w1 = a1 * inv, w2 = a2 * inv, w3 = a3 * inv;
  const double f1 = (2 * s6) * g[0] + (-7 * s6) * g[1] + (11 * s6) * g[2];
  const double f2 = (-s6) * g[1] + (5 * s6) * g[2] + (2 * s6) * g[3];
  const double f3 = (2 * s6) * g[2] + (5 * s6) * g[3] + (-s6) * g[4];
  const double out = w1 * f1 + w2 * f2 + w3 * f3;
  // values
  gb[0] += ob * (w1 * (2 * s6));
  gb[1] += ob * (w1 * (-7 * s6) + w2 * (-s6));
  gb[2] += ob * (w1 * (11 * s6) + w2 * (5 * s6) + w3 * (2 * s6));
  gb[3] += ob * (w2 * (2 * s6) + w3 * (5 * s6));
  gb[4] += ob * (w3 * (-s6));
  // weights: d out/d b_k = -2 w_k (f_k - out) / (b_k + eps), with
  // 1/e_k from one reciprocal of e1 e2 e3
  const double re = vjp_rcp(e1 * e2 * e3);
  const double bb1 = -2.0 * ob * w1 * (f1 - out) * (re * (e2 * e3));
  const double bb2 = -2.0 * ob * w2 * (f2 - out) * (re * (e1 * e3));
  const double bb3 = -2.0 * ob * w3 * (f3 - out) * (re * (e1 * e2));
  const double tt1 = bb1 * (2 * c1312) * t1, rr1 = bb1 * (2 * c14) * r1;
  const double tt2 = bb2 * (2 * c1312) * t2, rr2 = bb2 * (2 * c14) * r2;
  const double tt3 = bb3 * (2 * c1312) * t3, rr3 = bb3 * (2 * c14) * r3;
  mb[0] += tt1 + rr1;
  mb[1] += -2 * tt1 - 4 * rr1 + tt2 + rr2;
  mb[2] += tt1 + 3 * rr1 - 2 * tt2 + tt3 + 3 * rr3;
  mb[3] += tt2 - rr2 - 2 * tt3 - 4 * rr3;
  mb[4] += tt3 + rr3;
  return out;
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

// gradient of c + |v_d| of the Roe average of (uL, uR)
template <int DIR>
__device__ __forceinline__ void roe_speed_vjp(const double *uL, const double *uR, double gamma, double ab,
                                              double ubL[NV], double ubR[NV]) {
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
  const double ca = sqrt(gm1 * (Ha - 0.5 * (va[0] * va[0] + va[1] * va[1] + va[2] * va[2])));
  const double gH = ab * gm1 / (2.0 * ca);
  double gv[3];
  for (int j = 0; j < 3; j++) gv[j] = -ab * gm1 * va[j] / (2.0 * ca);
  const double sg = va[DIR] > 0 ? 1.0 : (va[DIR] < 0 ? -1.0 : 0.0);
  gv[DIR] += ab * sg;
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

// nu (+)= (d rhs_DIR / d u)^T mu along one direction, one workgroup per line,
// with mu = cb*lam + ca*nu_in (nu_in may be null).  mode VJP_WRITE stores,
// VJP_ADD accumulates into nu, VJP_ADD_L accumulates into nu and adds the
// finished nu into L, VJP_ADD_LSET sets L = lam + the finished nu.  LDS (dynamic, 23 doubles per cell of the line):
// state [5][n], cell flux [5][n], modified energy [n], stencil
// contributions of the flux and solution adjoints [6][n] each.
enum { VJP_WRITE = 0, VJP_ADD = 1, VJP_ADD_L = 2, VJP_ADD_LSET = 3 };

template <int DIR>
__global__ void __launch_bounds__(VJP_LINE_MAX) k_rhs_vjp(const double *__restrict__ u, const double *__restrict__ lam,
                                                      const double *__restrict__ nu_in, double ca, double cb,
                                                      double *__restrict__ nu, double *__restrict__ L, int mode,
                                                      int nx, int ny, int nz, const double *__restrict__ dxg,
                                                      double gamma, double eps) {
  extern __shared__ double smem[];
  const size_t ncell = (size_t)nx * ny * nz;
  const int n = DIR == 0 ? nx : (DIR == 1 ? ny : nz);
  // neighbouring lines share cache lines along x: keep them on one XCD
  const size_t line = xcd_block();
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
  const int t = threadIdx.x;
  const bool active = t < n;
  const double inv_gm1 = 1.0 / (gamma - 1.0);
  const int tm = active ? wrapn(t - 1, n) : 0;
  double g[NV] = {0, 0, 0, 0, 0};
  if (active) {
    const size_t c = base + (size_t)t * stride, cm = base + (size_t)tm * stride;
    double us[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      us[v] = u[(size_t)v * ncell + c];
      sU[v * n + t] = us[v];
      const size_t o = (size_t)v * ncell;
      const double mu = cb * lam[o + c] + (nu_in ? ca * nu_in[o + c] : 0.0);
      const double mum = cb * lam[o + cm] + (nu_in ? ca * nu_in[o + cm] : 0.0);
      g[v] = dxd[t] * mu - dxd[tm] * mum;
    }
    const Prim q = flowvar(us, gamma);
    double fl[NV];
    euler_flux<DIR>(fl, q);
#pragma unroll
    for (int v = 0; v < NV; v++) sFl[v * n + t] = fl[v];
    sC4[t] = modified_energy(q, inv_gm1);
  }
  __syncthreads();
  int cs[6];
  double alpha = 0.0, abar = 0.0;
  int sel = 0;
  if (active) {
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
      double mL[5], gL[5], mR[5], gR[5];
      double mbL[5] = {0, 0, 0, 0, 0}, gbL[5] = {0, 0, 0, 0, 0};
      double mbR[5] = {0, 0, 0, 0, 0}, gbR[5] = {0, 0, 0, 0, 0};
      // fluxes: smoothness and values are the same stencil
#pragma unroll
      for (int k = 0; k < 5; k++) { mL[k] = F[k]; mR[k] = F[5 - k]; }
      weno5_vjp(mL, mL, eps, 0.5 * g[v], mbL, gbL);
      weno5_vjp(mR, mR, eps, 0.5 * g[v], mbR, gbR);
#pragma unroll
      for (int k = 0; k < 5; k++) {
        fb[k] += mbL[k] + gbL[k];
        fb[5 - k] += mbR[k] + gbR[k];
        mbL[k] = gbL[k] = mbR[k] = gbR[k] = 0.0;
      }
      // solution: weights from u, values from the modified solution (the
      // same function of u, so both adjoints land on u)
#pragma unroll
      for (int k = 0; k < 5; k++) { mL[k] = U[k]; gL[k] = C[k]; mR[k] = U[5 - k]; gR[k] = C[5 - k]; }
      const double uLv = weno5_vjp(mL, gL, eps, 0.5 * alpha * g[v], mbL, gbL);
      const double uRv = weno5_vjp(mR, gR, eps, -0.5 * alpha * g[v], mbR, gbR);
      abar += -0.5 * (uRv - uLv) * g[v];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        ub[k] += mbL[k] + gbL[k];
        ub[5 - k] += mbR[k] + gbR[k];
      }
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
  // Rusanov speed adjoint onto the two cells next to interface t: cell t-1
  // (slot in cF) and cell t (slot in cU)
  if (active) {
    double u2[NV], u3[NV], ub2[NV] = {0, 0, 0, 0, 0}, ub3[NV] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < NV; v++) { u2[v] = sU[v * n + cs[2]]; u3[v] = sU[v * n + cs[3]]; }
    if (sel == 0) speed_vjp<DIR>(u2, gamma, abar, ub2);
    else if (sel == 1) speed_vjp<DIR>(u3, gamma, abar, ub3);
    else roe_speed_vjp<DIR>(u2, u3, gamma, abar, ub2, ub3);
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
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const size_t o = (size_t)v * ncell + c;
      if (mode == VJP_WRITE) {
        nu[o] = accU[v];
      } else {
        const double r = nu[o] + accU[v];
        nu[o] = r;
        if (mode == VJP_ADD_L) L[o] += r;
        else if (mode == VJP_ADD_LSET) L[o] = lam[o] + r;
      }
    }
  }
}

// y = a*x + b*y
__global__ void k_axpby(const double *__restrict__ x, double a, double *__restrict__ y, double b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = (a == 0.0 ? 0.0 : a * x[i]) + b * y[i];  // x unread when a == 0
}

// ---------------------------------------------------------------------------
bool vjp_supported(const Grid &g, const Scheme &sc) {
  return sc.interp == INTERP_COMP && sc.upwind == UPWIND_RUSANOV && sc.weno == WENO_JS && !(sc.Re > 0) &&
         g.n[0] <= VJP_LINE_MAX &&
         g.n[1] <= VJP_LINE_MAX && g.n[2] <= VJP_LINE_MAX;
}

void launch_rhs_vjp(const Grid &g, const Scheme &sc, const double *u, const double *lam, const double *nu_in,
                    double ca, double cb, double *nu, double *L, bool set_L, const double *dxg, hipStream_t st) {
  for (int d = 0; d < 3; d++) {
    const unsigned lines = (unsigned)(g.ncell / g.n[d]);
    const unsigned threads = (unsigned)((g.n[d] + 63) / 64 * 64);
    const size_t lds = (size_t)23 * g.n[d] * sizeof(double);
    const int mode = d == 0 ? VJP_WRITE : (d == 2 && L ? (set_L ? VJP_ADD_LSET : VJP_ADD_L) : VJP_ADD);
    if (d == 0)
      hipLaunchKernelGGL(k_rhs_vjp<0>, dim3(lines), dim3(threads), lds, st, u, lam, nu_in, ca, cb, nu, L, mode,
                         g.n[0], g.n[1], g.n[2], dxg, sc.gamma, sc.eps);
    else if (d == 1)
      hipLaunchKernelGGL(k_rhs_vjp<1>, dim3(lines), dim3(threads), lds, st, u, lam, nu_in, ca, cb, nu, L, mode,
                         g.n[0], g.n[1], g.n[2], dxg, sc.gamma, sc.eps);
    else
      hipLaunchKernelGGL(k_rhs_vjp<2>, dim3(lines), dim3(threads), lds, st, u, lam, nu_in, ca, cb, nu, L, mode,
                         g.n[0], g.n[1], g.n[2], dxg, sc.gamma, sc.eps);
  }
}

void launch_axpby(const double *x, double a, double *y, double b, size_t n, hipStream_t st) {
  size_t nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)nb), dim3(256), 0, st, x, a, y, b, n);
}

}  // namespace hm
