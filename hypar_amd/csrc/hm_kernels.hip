// hm_kernels.hip -- gfx950 kernels of the NS3D WENO5 + RK4 path.
//
// State layout: structure of arrays without ghosts (see include/hypar_mi355x.h).
// All boundaries are periodic; stencil indices wrap.
//
// Kernels per RK stage:
//   k_iface<DIR,...>   interface fluxes F_{i-1/2} along DIR for every cell i
//                      (Roe/Rusanov upwinding of WENO5 reconstructions);
//                      HyperbolicFunction.c:75-93 with ReconstructHyperbolic
//   k_visc_flux        viscous fluxes of x, y, z at every cell from the
//                      primitive variables the x pass stores
//                      (NavierStokes3DParabolicFunction.c:85-195)
//   k_visc_ghost_x, k_visc_ghost<DIR>  viscous fluxes at the reference's ghost layers
//   k_stage            flux differences in x, y, z, the viscous divergence,
//                      rhs = -hyp + par, and the RK4 stage update in 3N form
//                      (TimeRHSFunctionExplicit.c:88-93, TimeRK.c:126-195)
#include <hip/hip_runtime.h>

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

// ---------------------------------------------------------------------------
// interface flux kernel
// ---------------------------------------------------------------------------
template <int DIR, int INTERP, int UPW>
__device__ __forceinline__ void iface_flux(const double (&U)[6][NV], double gamma, double eps, int weno,
                                           double fI[NV]) {
  const double inv_gm1 = 1.0 / (gamma - 1.0);
  double F[6][NV];
  double UC4[6];
  for (int s = 0; s < 6; s++) {
    Prim q = flowvar(U[s], gamma);
    euler_flux<DIR>(F[s], q);
    UC4[s] = modified_energy(q, inv_gm1);
  }
  // upwinding states are the cells adjacent to the interface (I-1, I)
  Prim avg = roe_average(U[2], U[3], gamma);

  if (INTERP == INTERP_CHAR) {
    Eigen E;
    eigen<DIR>(E, avg, gamma);
    double alpha = 0.0;
    if (UPW == UPWIND_RUSANOV) {
      Prim l = flowvar(U[2], gamma), r = flowvar(U[3], gamma);
      double aL = sqrt(gamma * l.P / l.rho) + fabs(vel<DIR>(l));
      double aR = sqrt(gamma * r.P / r.rho) + fabs(vel<DIR>(r));
      double aA = sqrt(gamma * avg.P / avg.rho) + fabs(vel<DIR>(avg));
      alpha = fmax(fmax(aL, aR), fmax(aR, aA));
    }
    double g[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double pf[6], pu[6], pc[6];
#pragma unroll
      for (int s = 0; s < 6; s++) {
        double af = 0.0, au = 0.0;
#pragma unroll
        for (int k = 0; k < NV; k++) {
          af += E.L[v][k] * F[s][k];
          au += E.L[v][k] * U[s][k];
        }
        pf[s] = af;
        pu[s] = au;
        pc[s] = au + E.L[v][4] * (UC4[s] - U[s][4]);
      }
      double fcL = weno5(pf[0], pf[1], pf[2], pf[3], pf[4], pf[0], pf[1], pf[2], pf[3], pf[4], eps, weno);
      double fcR = weno5(pf[5], pf[4], pf[3], pf[2], pf[1], pf[5], pf[4], pf[3], pf[2], pf[1], eps, weno);
      double ucL = weno5(pu[0], pu[1], pu[2], pu[3], pu[4], pc[0], pc[1], pc[2], pc[3], pc[4], eps, weno);
      double ucR = weno5(pu[5], pu[4], pu[3], pu[2], pu[1], pc[5], pc[4], pc[3], pc[2], pc[1], eps, weno);
      double lam = (UPW == UPWIND_ROE) ? harten_abs(E.lam[v]) : alpha;
      g[v] = 0.5 * (fcL + fcR) - lam * (0.5 * (ucR - ucL));
    }
#pragma unroll
    for (int r = 0; r < NV; r++) {
      double acc = 0.0;
#pragma unroll
      for (int c = 0; c < NV; c++) acc += E.R[r][c] * g[c];
      fI[r] = acc;
    }
  } else {
    double fL[NV], fR[NV], uL[NV], uR[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double c0 = (v == 4) ? UC4[0] : U[0][v], c1 = (v == 4) ? UC4[1] : U[1][v];
      double c2 = (v == 4) ? UC4[2] : U[2][v], c3 = (v == 4) ? UC4[3] : U[3][v];
      double c4 = (v == 4) ? UC4[4] : U[4][v], c5 = (v == 4) ? UC4[5] : U[5][v];
      fL[v] = weno5(F[0][v], F[1][v], F[2][v], F[3][v], F[4][v], F[0][v], F[1][v], F[2][v], F[3][v], F[4][v], eps, weno);
      fR[v] = weno5(F[5][v], F[4][v], F[3][v], F[2][v], F[1][v], F[5][v], F[4][v], F[3][v], F[2][v], F[1][v], eps, weno);
      uL[v] = weno5(U[0][v], U[1][v], U[2][v], U[3][v], U[4][v], c0, c1, c2, c3, c4, eps, weno);
      uR[v] = weno5(U[5][v], U[4][v], U[3][v], U[2][v], U[1][v], c5, c4, c3, c2, c1, eps, weno);
    }
    if (UPW == UPWIND_RUSANOV) {
      Prim l = flowvar(U[2], gamma), r = flowvar(U[3], gamma);
      double aL = sqrt(gamma * l.P / l.rho) + fabs(vel<DIR>(l));
      double aR = sqrt(gamma * r.P / r.rho) + fabs(vel<DIR>(r));
      double aA = sqrt(gamma * avg.P / avg.rho) + fabs(vel<DIR>(avg));
      double alpha = fmax(fmax(aL, aR), fmax(aR, aA));
#pragma unroll
      for (int v = 0; v < NV; v++) fI[v] = 0.5 * (fL[v] + fR[v]) - alpha * (0.5 * (uR[v] - uL[v]));
    } else {
      Eigen E;
      eigen<DIR>(E, avg, gamma);
      double c[NV];
#pragma unroll
      for (int r = 0; r < NV; r++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < NV; k++) acc += E.L[r][k] * (0.5 * (uR[k] - uL[k]));
        c[r] = harten_abs(E.lam[r]) * acc;
      }
#pragma unroll
      for (int r = 0; r < NV; r++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < NV; k++) acc += E.R[r][k] * c[k];
        fI[r] = 0.5 * (fL[r] + fR[r]) - acc;
      }
    }
  }
}

// one thread per interface i-1/2 (stored at the index of cell i).  The
// x-direction pass also stores the primitive variables (u, v, w, T) of its
// cell for the viscous kernels (Q != nullptr).
template <int DIR, int INTERP, int UPW>
__global__ void __launch_bounds__(256) k_iface(const double *__restrict__ u, double *__restrict__ fI,
                                               double *__restrict__ Q, int nx, int ny, int nz,
                                               double gamma, double eps, int weno) {
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t idx = xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  const int i = (int)(idx % nx);
  const int j = (int)((idx / nx) % ny);
  const int k = (int)(idx / ((size_t)nx * ny));
  const int n = DIR == 0 ? nx : (DIR == 1 ? ny : nz);
  const int pos = DIR == 0 ? i : (DIR == 1 ? j : k);
  const size_t stride = DIR == 0 ? 1 : (DIR == 1 ? (size_t)nx : (size_t)nx * ny);
  const size_t base = idx - (size_t)pos * stride;
  double U[6][NV];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const size_t c = base + (size_t)wrap(pos + s - 3, n) * stride;
#pragma unroll
    for (int v = 0; v < NV; v++) U[s][v] = u[(size_t)v * ncell + c];
  }
  if (Q) {
    Prim q = flowvar(U[3], gamma);
    Q[idx] = q.vx;
    Q[ncell + idx] = q.vy;
    Q[2 * ncell + idx] = q.vz;
    Q[3 * ncell + idx] = gamma * q.P / q.rho;
  }
  double f[NV];
  iface_flux<DIR, INTERP, UPW>(U, gamma, eps, weno, f);
#pragma unroll
  for (int v = 0; v < NV; v++) fI[(size_t)v * ncell + idx] = f[v];
}

// ---------------------------------------------------------------------------
// viscous terms
// ---------------------------------------------------------------------------
// Primitive variables used by the viscous term: (u, v, w, T) with
// T = gamma*P/rho (NavierStokes3DParabolicFunction.c:85-108), stored as
// four arrays Q[c*ncell + cell].
struct QT {
  double u, v, w, T;
};

__device__ __forceinline__ QT load_q(const double *__restrict__ Q, size_t ncell, size_t c) {
  QT q;
  q.u = Q[c];
  q.v = Q[ncell + c];
  q.w = Q[2 * ncell + c];
  q.T = Q[3 * ncell + c];
  return q;
}

struct ViscParams {
  double gamma, inv_Re, inv_gm1, inv_Pr;
};

// Viscous flux along DIR at a point from the primitive values and the
// gradient components (derivatives already scaled by dxinv).
// grad[c][e] = d(q_c)/dx_e for c in (u,v,w,T).
template <int DIR>
__device__ __forceinline__ void visc_flux(double fv[4], const QT &q, const double (&gr)[4][3],
                                          double mu, const ViscParams &vp) {
  const double two_third = 2.0 / 3.0;
  const double mr = mu * vp.inv_Re;
  const double kq = mu * vp.inv_Re * vp.inv_gm1 * vp.inv_Pr;
  const double ux = gr[0][0], uy = gr[0][1], uz = gr[0][2];
  const double vx = gr[1][0], vy = gr[1][1], vz = gr[1][2];
  const double wx = gr[2][0], wy = gr[2][1], wz = gr[2][2];
  double t1, t2, t3, qd;
  if (DIR == 0) {
    t1 = two_third * mr * (2 * ux - vy - wz);
    t2 = mr * (uy + vx);
    t3 = mr * (uz + wx);
    qd = kq * gr[3][0];
  } else if (DIR == 1) {
    t1 = mr * (uy + vx);
    t2 = two_third * mr * (-ux + 2 * vy - wz);
    t3 = mr * (vz + wy);
    qd = kq * gr[3][1];
  } else {
    t1 = mr * (uz + wx);
    t2 = mr * (vz + wy);
    t3 = two_third * mr * (-ux - vy + 2 * wz);
    qd = kq * gr[3][2];
  }
  fv[0] = t1; fv[1] = t2; fv[2] = t3;
  fv[3] = q.u * t1 + q.v * t2 + q.w * t3 + qd;
}

// Viscous fluxes of x, y, z at every cell: FV[(d*4+c)*ncell + cell],
// c = momentum x, y, z, energy.  Gradients use the 4th-order central
// difference with periodic neighbours, scaled by the cell's dxinv.
//
// A block owns a VTX x VTY tile of (x, y) columns and marches through VKS
// z-planes.  Each thread keeps its column's (u, v, w, T) at k-2..k+2 in
// registers (the z-derivative stencil, one new plane loaded per step); the
// tile's plane k with a 2-cell halo in x and y is staged in LDS for the x-
// and y-derivatives.  Every plane of Q is thus read from HBM about once per
// tile, instead of once per stencil point.
// (tile constants VTX, VTY, VKS and visc_tiles: hm_internal.h)

// tile b of the grid; every thread of the block takes part
__device__ __forceinline__ void visc_flux_tile(const double *__restrict__ Q, double *__restrict__ FV, int nx,
                                               int ny, int nz,
                                               const double *__restrict__ dxg,  // ghosted dxinv x|y|z
                                               const ViscParams &vp, int b, double (&sQ)[4][VHY][VHX]) {
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
  const double dxi = dxg[3 + ic], dyi = dxg[(nx + 6) + 3 + jc];
  const double *dzi = dxg + (nx + 6) + (ny + 6) + 3;
  const int k1 = min(k0 + VKS, nz);
  // window: planes k-2..k+2 of this column
  QT q[5];
#pragma unroll
  for (int s = 0; s < 4; s++) q[s + 1] = load_q(Q, ncell, col + (size_t)wrap(k0 - 2 + s, nz) * plane);
  for (int k = k0; k < k1; k++) {
#pragma unroll
    for (int s = 0; s < 4; s++) q[s] = q[s + 1];
    q[4] = load_q(Q, ncell, col + (size_t)wrap(k + 2, nz) * plane);
    // tile of plane k with halo
    const size_t pk = (size_t)k * plane;
    for (int c = threadIdx.x; c < VHX * VHY; c += VTX * VTY) {
      const int hx = c % VHX, hy = c / VHX;
      // x0 - 2 + hx spans [x0 - 2, x0 + VTX + 1] and y0 - 2 + hy spans
      // [y0 - 2, y0 + VTY + 1], which can exceed 2 nx and 2 ny
      const int gi = (x0 - 2 + hx + 2 * nx) % nx, gj = (y0 - 2 + hy + 2 * ny) % ny;
      const QT h = load_q(Q, ncell, pk + (size_t)gj * nx + gi);
      sQ[0][hy][hx] = h.u;
      sQ[1][hy][hx] = h.v;
      sQ[2][hy][hx] = h.w;
      sQ[3][hy][hx] = h.T;
    }
    __syncthreads();
    if (own) {
      double gr[4][3];
      const double dzk = dzi[k];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const double *r = &sQ[c][ly + 2][0];
        gr[c][0] = d4c(r[lx], r[lx + 1], r[lx + 3], r[lx + 4]) * dxi;
        gr[c][1] = d4c(sQ[c][ly][lx + 2], sQ[c][ly + 1][lx + 2], sQ[c][ly + 3][lx + 2], sQ[c][ly + 4][lx + 2]) * dyi;
      }
      gr[0][2] = d4c(q[0].u, q[1].u, q[3].u, q[4].u) * dzk;
      gr[1][2] = d4c(q[0].v, q[1].v, q[3].v, q[4].v) * dzk;
      gr[2][2] = d4c(q[0].w, q[1].w, q[3].w, q[4].w) * dzk;
      gr[3][2] = d4c(q[0].T, q[1].T, q[3].T, q[4].T) * dzk;
      const QT &q0 = q[2];
      const double mu = exp(0.76 * log(q0.T));
      const size_t idx = pk + col;
      double fv[4];
      visc_flux<0>(fv, q0, gr, mu, vp);
      for (int c = 0; c < 4; c++) FV[(size_t)(0 * 4 + c) * ncell + idx] = fv[c];
      visc_flux<1>(fv, q0, gr, mu, vp);
      for (int c = 0; c < 4; c++) FV[(size_t)(1 * 4 + c) * ncell + idx] = fv[c];
      visc_flux<2>(fv, q0, gr, mu, vp);
      for (int c = 0; c < 4; c++) FV[(size_t)(2 * 4 + c) * ncell + idx] = fv[c];
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(VTX * VTY) k_visc_flux(const double *__restrict__ Q, double *__restrict__ FV,
                                                         int nx, int ny, int nz, const double *__restrict__ dxg,
                                                         ViscParams vp) {
  __shared__ double sQ[4][VHY][VHX];
  visc_flux_tile(Q, FV, nx, ny, nz, dxg, vp, blockIdx.x, sQ);
}

// Viscous flux along DIR at the ghost positions gp in {-2,-1,n,n+1} of every
// line along DIR (other indices interior), as the reference evaluates it
// from its ghost layers: primitive values are the periodic copies, the
// derivative along DIR uses the biased stencils of
// FirstDerivativeFourthOrder.c:74-93/106-125 in the second-outermost layer
// and the central one otherwise, scaled by the ghost dxinv (a copy of the
// nearest interior value, InitialSolution.c:101-113), and the derivatives
// across DIR are zero because the reference computes them only on lines whose
// other indices are interior.
// Output FVg[(slot*4 + c)*nplane + line], slot 0..3 for gp = -2,-1,n,n+1.
template <int DIR>
__device__ __forceinline__ void visc_ghost_one(const double *__restrict__ Q, double *__restrict__ FVg, int nx,
                                               int ny, int nz, const double *__restrict__ dxg, const ViscParams &vp,
                                               size_t t) {
  const size_t ncell = (size_t)nx * ny * nz;
  const int n = DIR == 0 ? nx : (DIR == 1 ? ny : nz);
  const size_t nplane = ncell / n;
  if (t >= 4 * nplane) return;
  const int slot = (int)(t / nplane);
  const size_t line = t % nplane;
  size_t base, stride;
  if (DIR == 0) { base = line * nx; stride = 1; }                          // line = j + ny*k
  else if (DIR == 1) { base = (line % nx) + (line / nx) * (size_t)nx * ny; stride = nx; }  // i + nx*k
  else { base = line; stride = (size_t)nx * ny; }                           // i + nx*j
  const int gp = slot < 2 ? slot - 2 : n + slot - 2;
  const double *dxd = dxg + (DIR >= 1 ? nx + 6 : 0) + (DIR == 2 ? ny + 6 : 0);
  QT f[7];  // positions gp-3 .. gp+3 (periodic copies)
#pragma unroll
  for (int s = 0; s < 7; s++) f[s] = load_q(Q, ncell, base + (size_t)wrap(gp - 3 + s, n) * stride);
  double d[4];
  const double r12 = 1.0 / 12.0;
  if (gp == -2) {
    // (-3 f(-3) - 10 f(-2) + 18 f(-1) - 6 f(0) + f(1)) / 12
    d[0] = (-3 * f[2].u - 10 * f[3].u + 18 * f[4].u - 6 * f[5].u + f[6].u) * r12;
    d[1] = (-3 * f[2].v - 10 * f[3].v + 18 * f[4].v - 6 * f[5].v + f[6].v) * r12;
    d[2] = (-3 * f[2].w - 10 * f[3].w + 18 * f[4].w - 6 * f[5].w + f[6].w) * r12;
    d[3] = (-3 * f[2].T - 10 * f[3].T + 18 * f[4].T - 6 * f[5].T + f[6].T) * r12;
  } else if (gp == n + 1) {
    // (-f(n-2) + 6 f(n-1) - 18 f(n) + 10 f(n+1) + 3 f(n+2)) / 12
    d[0] = (-f[0].u + 6 * f[1].u - 18 * f[2].u + 10 * f[3].u + 3 * f[4].u) * r12;
    d[1] = (-f[0].v + 6 * f[1].v - 18 * f[2].v + 10 * f[3].v + 3 * f[4].v) * r12;
    d[2] = (-f[0].w + 6 * f[1].w - 18 * f[2].w + 10 * f[3].w + 3 * f[4].w) * r12;
    d[3] = (-f[0].T + 6 * f[1].T - 18 * f[2].T + 10 * f[3].T + 3 * f[4].T) * r12;
  } else {
    d[0] = d4c(f[1].u, f[2].u, f[4].u, f[5].u);
    d[1] = d4c(f[1].v, f[2].v, f[4].v, f[5].v);
    d[2] = d4c(f[1].w, f[2].w, f[4].w, f[5].w);
    d[3] = d4c(f[1].T, f[2].T, f[4].T, f[5].T);
  }
  const double dxgh = dxd[3 + gp];  // entry 0 is position -3
  double gr[4][3];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int e = 0; e < 3; e++) gr[c][e] = (e == DIR) ? d[c] * dxgh : 0.0;
  const QT &q = f[3];
  const double mu = exp(0.76 * log(q.T));
  double fv[4];
  visc_flux<DIR>(fv, q, gr, mu, vp);
#pragma unroll
  for (int c = 0; c < 4; c++) FVg[(size_t)(slot * 4 + c) * nplane + line] = fv[c];
}

// The four ghost-layer fluxes of x-line `line` (line = j + ny*k) from one
// thread.  Slot s (positions -2, -1, n, n+1) reads the seven cells
// gp-3 .. gp+3, periodic; together the four read the ten cells n-5 .. n-1,
// 0 .. 4, loaded once, with slot s taking window entries s .. s+6.  The x
// lines lie along the contiguous direction, so the one-thread-per-(slot,
// line) form of visc_ghost_one reads each cell's cache line from four
// threads far apart in the grid.
__device__ __forceinline__ void visc_ghost_xline(const double *__restrict__ Q, double *__restrict__ FVg, int nx,
                                                 int ny, int nz, const double *__restrict__ dxg,
                                                 const ViscParams &vp, size_t line) {
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t nplane = ncell / nx;
  if (line >= nplane) return;
  const size_t base = line * nx;
  QT f[10];  // cells n-5 .. n-1, 0 .. 4 (n >= 5)
#pragma unroll
  for (int j = 0; j < 5; j++) {
    f[j] = load_q(Q, ncell, base + (size_t)(nx - 5 + j));
    f[5 + j] = load_q(Q, ncell, base + (size_t)j);
  }
  const double r12 = 1.0 / 12.0;
#pragma unroll
  for (int slot = 0; slot < 4; slot++) {
    const QT *w = f + slot;  // positions gp-3 .. gp+3
    const int gp = slot < 2 ? slot - 2 : nx + slot - 2;
    double d[4];
    if (slot == 0) {
      // (-3 f(-3) - 10 f(-2) + 18 f(-1) - 6 f(0) + f(1)) / 12
      d[0] = (-3 * w[2].u - 10 * w[3].u + 18 * w[4].u - 6 * w[5].u + w[6].u) * r12;
      d[1] = (-3 * w[2].v - 10 * w[3].v + 18 * w[4].v - 6 * w[5].v + w[6].v) * r12;
      d[2] = (-3 * w[2].w - 10 * w[3].w + 18 * w[4].w - 6 * w[5].w + w[6].w) * r12;
      d[3] = (-3 * w[2].T - 10 * w[3].T + 18 * w[4].T - 6 * w[5].T + w[6].T) * r12;
    } else if (slot == 3) {
      // (-f(n-2) + 6 f(n-1) - 18 f(n) + 10 f(n+1) + 3 f(n+2)) / 12
      d[0] = (-w[0].u + 6 * w[1].u - 18 * w[2].u + 10 * w[3].u + 3 * w[4].u) * r12;
      d[1] = (-w[0].v + 6 * w[1].v - 18 * w[2].v + 10 * w[3].v + 3 * w[4].v) * r12;
      d[2] = (-w[0].w + 6 * w[1].w - 18 * w[2].w + 10 * w[3].w + 3 * w[4].w) * r12;
      d[3] = (-w[0].T + 6 * w[1].T - 18 * w[2].T + 10 * w[3].T + 3 * w[4].T) * r12;
    } else {
      d[0] = d4c(w[1].u, w[2].u, w[4].u, w[5].u);
      d[1] = d4c(w[1].v, w[2].v, w[4].v, w[5].v);
      d[2] = d4c(w[1].w, w[2].w, w[4].w, w[5].w);
      d[3] = d4c(w[1].T, w[2].T, w[4].T, w[5].T);
    }
    const double dxgh = dxg[3 + gp];  // entry 0 is position -3
    double gr[4][3];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int e = 0; e < 3; e++) gr[c][e] = e == 0 ? d[c] * dxgh : 0.0;
    const QT &q = w[3];
    const double mu = exp(0.76 * log(q.T));
    double fv[4];
    visc_flux<0>(fv, q, gr, mu, vp);
#pragma unroll
    for (int c = 0; c < 4; c++) FVg[(size_t)(slot * 4 + c) * nplane + line] = fv[c];
  }
}

__global__ void __launch_bounds__(256) k_visc_ghost_x(const double *__restrict__ Q, double *__restrict__ FVg, int nx,
                                                      int ny, int nz, const double *__restrict__ dxg,
                                                      ViscParams vp) {
  visc_ghost_xline(Q, FVg, nx, ny, nz, dxg, vp, (size_t)blockIdx.x * blockDim.x + threadIdx.x);
}

template <int DIR>
__global__ void __launch_bounds__(256) k_visc_ghost(const double *__restrict__ Q, double *__restrict__ FVg,
                                                    int nx, int ny, int nz, const double *__restrict__ dxg,
                                                    ViscParams vp) {
  visc_ghost_one<DIR>(Q, FVg, nx, ny, nz, dxg, vp, (size_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Viscous fluxes of all members of an ensemble: one persistent launch over
// work items (member, tile) and (member, VT ghost positions of one
// direction); item0v holds each member's first item.
struct EnsViscArgs {
  const EnsMember *mem;
  int nmem;
  const unsigned *item0v;
  unsigned nitems;
  ViscParams vp;
};

// work items of VT threads for the ghost-layer fluxes of one direction: one
// thread per x-line (all four positions), one per (position, line) along y
// and z
__host__ __device__ __forceinline__ unsigned ghost_blocks(int dir, size_t ncell, int n) {
  const size_t threads = dir == 0 ? ncell / n : 4 * (ncell / n);
  return (unsigned)((threads + VT - 1) / VT);
}

__global__ void __launch_bounds__(VTX * VTY) k_ens_visc(EnsViscArgs a) {
  __shared__ double sQ[4][VHY][VHX];
  for (unsigned item = blockIdx.x; item < a.nitems; item += gridDim.x) {
    int lo = 0, hi = a.nmem - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.item0v[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const EnsMember &M = a.mem[lo];
    const int nx = M.n[0], ny = M.n[1], nz = M.n[2];
    unsigned r = item - a.item0v[lo];
    const unsigned tiles = visc_tiles(nx, ny, nz);
    if (r < tiles) {
      visc_flux_tile(M.q, M.fv, nx, ny, nz, M.dxg, a.vp, (int)r, sQ);
      continue;
    }
    r -= tiles;
    const unsigned gx = ghost_blocks(0, M.ncell, nx), gy = ghost_blocks(1, M.ncell, ny);
    const size_t t = threadIdx.x;
    if (r < gx) visc_ghost_xline(M.q, M.fvg[0], nx, ny, nz, M.dxg, a.vp, (size_t)r * VT + t);
    else if (r < gx + gy) visc_ghost_one<1>(M.q, M.fvg[1], nx, ny, nz, M.dxg, a.vp, (size_t)(r - gx) * VT + t);
    else visc_ghost_one<2>(M.q, M.fvg[2], nx, ny, nz, M.dxg, a.vp, (size_t)(r - gx - gy) * VT + t);
  }
}

// primitive variables of every member's state
__global__ void __launch_bounds__(256) k_ens_prims(const EnsMember *__restrict__ mem, int nmem,
                                                   const double *__restrict__ u, size_t ncell, double gamma) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mem[mid].off <= idx) lo = mid;
    else hi = mid - 1;
  }
  const EnsMember &M = mem[lo];
  const size_t c = idx - M.off;
  const double *um = u + 5 * M.off;
  double s[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) s[v] = um[(size_t)v * M.ncell + c];
  const Prim q = flowvar(s, gamma);
  M.q[c] = q.vx;
  M.q[M.ncell + c] = q.vy;
  M.q[2 * M.ncell + c] = q.vz;
  M.q[3 * M.ncell + c] = gamma * q.P / q.rho;
}

// ---------------------------------------------------------------------------
// flux differences + viscous divergence + RK4 stage update (3N form)
// ---------------------------------------------------------------------------
// mode OUT_STAGE:
// stage 0: acc = u + cb*k;   S = u + ca*k
// stage 1,2: acc = acc + cb*k; S = u + ca*k
// stage 3: u = acc + cb*k
// with k = -hyp + par and hyp = (hx + hy) + hz.  Other modes store k, hyp
// or par in `out` instead.  FVg holds the ghost viscous fluxes of x, y, z
// (4 slots x 4 components x plane each).
__global__ void __launch_bounds__(256) k_stage(
    const double *__restrict__ fIx, const double *__restrict__ fIy, const double *__restrict__ fIz,
    const double *__restrict__ FV, const double *__restrict__ FVgx, const double *__restrict__ FVgy,
    const double *__restrict__ FVgz, double *__restrict__ u, double *__restrict__ S,
    double *__restrict__ acc, double *__restrict__ out, int mode, int nx, int ny, int nz,
    const double *__restrict__ dxg, int viscous, int stage, double ca, double cb) {
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t idx = xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  const int i = (int)(idx % nx);
  const int j = (int)((idx / nx) % ny);
  const int k = (int)(idx / ((size_t)nx * ny));
  const double *dxi = dxg + 3, *dyi = dxg + (nx + 6) + 3, *dzi = dxg + (nx + 6) + (ny + 6) + 3;
  const size_t ip = idx - i + wrap(i + 1, nx);
  const size_t jp = idx - (size_t)j * nx + (size_t)wrap(j + 1, ny) * nx;
  const size_t kp = idx - (size_t)k * nx * ny + (size_t)wrap(k + 1, nz) * nx * ny;
  double par[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (viscous) {
    double dv[4];
    visc_div_dir<0>(dv, FV, FVgx, ncell, idx, i, nx, 1, (size_t)j + (size_t)ny * k);
    for (int v = 0; v < 4; v++) par[v + 1] = dxi[i] * dv[v];
    visc_div_dir<1>(dv, FV, FVgy, ncell, idx, j, ny, (size_t)nx, (size_t)i + (size_t)nx * k);
    for (int v = 0; v < 4; v++) par[v + 1] += dyi[j] * dv[v];
    visc_div_dir<2>(dv, FV, FVgz, ncell, idx, k, nz, (size_t)nx * ny, (size_t)i + (size_t)nx * j);
    for (int v = 0; v < 4; v++) par[v + 1] += dzi[k] * dv[v];
  }
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const size_t o = (size_t)v * ncell;
    double hyp = dxi[i] * (fIx[o + ip] - fIx[o + idx]);
    hyp += dyi[j] * (fIy[o + jp] - fIy[o + idx]);
    hyp += dzi[k] * (fIz[o + kp] - fIz[o + idx]);
    const double r = (-hyp) + par[v];
    if (mode != OUT_STAGE) {
      out[o + idx] = mode == OUT_RHS ? r : (mode == OUT_HYP ? hyp : par[v]);
      continue;
    }
    if (stage == 0) {
      const double u0 = u[o + idx];
      acc[o + idx] = u0 + cb * r;
      S[o + idx] = u0 + ca * r;
    } else if (stage < 3) {
      acc[o + idx] = acc[o + idx] + cb * r;
      S[o + idx] = u[o + idx] + ca * r;
    } else {
      u[o + idx] = acc[o + idx] + cb * r;
    }
  }
}

// ---------------------------------------------------------------------------
// reductions: CFL max and squared-difference sum
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// per-block partial maxima of the local CFL; NaN propagates as +Inf flag
__global__ void __launch_bounds__(256) k_cfl(const double *__restrict__ u, double *__restrict__ part,
                                             int nx, int ny, int nz, const double *__restrict__ dxg,
                                             double gamma, double dt) {
  const size_t ncell = (size_t)nx * ny * nz;
  __shared__ double red[4];
  double m = 0.0;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < ncell;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx % nx);
    const int j = (int)((idx / nx) % ny);
    const int k = (int)(idx / ((size_t)nx * ny));
    double s[NV];
    for (int v = 0; v < NV; v++) s[v] = u[(size_t)v * ncell + idx];
    Prim q = flowvar(s, gamma);
    const double c = sqrt(gamma * q.P / q.rho);
    const double lx = (fabs(q.vx) + c) * dt * dxg[3 + i];
    const double ly = (fabs(q.vy) + c) * dt * dxg[(nx + 6) + 3 + j];
    const double lz = (fabs(q.vz) + c) * dt * dxg[(nx + 6) + (ny + 6) + 3 + k];
    double l = fmax(fmax(lx, ly), lz);
    if (!(l == l) || l > 1.0e300) l = __builtin_inf();
    m = fmax(m, l);
  }
  m = wave_max(m);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) r = fmax(r, red[q]);
    part[blockIdx.x] = r;
  }
}

__global__ void __launch_bounds__(256) k_sqdiff(const double *__restrict__ a, const double *__restrict__ b,
                                                double *__restrict__ part, size_t n) {
  __shared__ double red[4];
  double s = 0.0;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (size_t)gridDim.x * blockDim.x) {
    const double d = a[idx] - b[idx];
    s += d * d;
  }
  s = wave_sum(s);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) r += red[q];
    part[blockIdx.x] = r;
  }
}

// ---------------------------------------------------------------------------
// layout conversion (reference AoS with ghosts <-> hm SoA)
// ---------------------------------------------------------------------------
__global__ void k_aos_to_soa(const double *__restrict__ aos, double *__restrict__ soa, int nx, int ny,
                             int nz, int g) {
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  const int i = (int)(idx % nx);
  const int j = (int)((idx / nx) % ny);
  const int k = (int)(idx / ((size_t)nx * ny));
  const size_t p = ((size_t)(k + g) * (ny + 2 * g) + (j + g)) * (nx + 2 * g) + (i + g);
  for (int v = 0; v < NV; v++) soa[(size_t)v * ncell + idx] = aos[p * NV + v];
}

__global__ void k_soa_to_aos(const double *__restrict__ soa, double *__restrict__ aos, int nx, int ny,
                             int nz, int g) {
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  const int i = (int)(idx % nx);
  const int j = (int)((idx / nx) % ny);
  const int k = (int)(idx / ((size_t)nx * ny));
  const size_t p = ((size_t)(k + g) * (ny + 2 * g) + (j + g)) * (nx + 2 * g) + (i + g);
  for (int v = 0; v < NV; v++) aos[p * NV + v] = soa[(size_t)v * ncell + idx];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblocks(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

template <int DIR, int INTERP, int UPW>
static void launch_iface_t(const Grid &g, const double *u, double *fI, double *Q, double gamma,
                           double eps, int weno, hipStream_t st) {
  hipLaunchKernelGGL((k_iface<DIR, INTERP, UPW>), dim3(nblocks(g.ncell, 256)), dim3(256), 0, st, u, fI, Q,
                     g.n[0], g.n[1], g.n[2], gamma, eps, weno);
}

template <int DIR>
static void launch_iface_d(const Grid &g, const Scheme &sc, const double *u, double *fI, double *Q,
                           hipStream_t st) {
  if (sc.interp == INTERP_CHAR) {
    if (sc.upwind == UPWIND_ROE) launch_iface_t<DIR, INTERP_CHAR, UPWIND_ROE>(g, u, fI, Q, sc.gamma, sc.eps, sc.weno, st);
    else launch_iface_t<DIR, INTERP_CHAR, UPWIND_RUSANOV>(g, u, fI, Q, sc.gamma, sc.eps, sc.weno, st);
  } else {
    if (sc.upwind == UPWIND_ROE) launch_iface_t<DIR, INTERP_COMP, UPWIND_ROE>(g, u, fI, Q, sc.gamma, sc.eps, sc.weno, st);
    else launch_iface_t<DIR, INTERP_COMP, UPWIND_RUSANOV>(g, u, fI, Q, sc.gamma, sc.eps, sc.weno, st);
  }
}

void launch_iface(const Grid &g, const Scheme &sc, int dir, const double *u, double *fI, double *Q,
                  hipStream_t st) {
  if (dir == 0) launch_iface_d<0>(g, sc, u, fI, Q, st);
  else if (dir == 1) launch_iface_d<1>(g, sc, u, fI, nullptr, st);
  else launch_iface_d<2>(g, sc, u, fI, nullptr, st);
}

static ViscParams visc_params(const Scheme &sc) {
  ViscParams vp;
  vp.gamma = sc.gamma;
  vp.inv_Re = sc.Re > 0 ? 1.0 / sc.Re : 0.0;
  vp.inv_gm1 = 1.0 / (sc.gamma - 1.0);
  vp.inv_Pr = 1.0 / sc.Pr;
  return vp;
}

size_t visc_ghost_size(const Grid &g, int dir) { return 16 * (g.ncell / g.n[dir]); }

void launch_visc_flux(const Grid &g, const Scheme &sc, const double *Q, double *FV, double *const FVg[3],
                      const double *dxg, hipStream_t st) {
  const ViscParams vp = visc_params(sc);
  const unsigned nb = visc_tiles(g.n[0], g.n[1], g.n[2]);
  hipLaunchKernelGGL(k_visc_flux, dim3(nb), dim3(VTX * VTY), 0, st, Q, FV, g.n[0], g.n[1], g.n[2], dxg, vp);
  hipLaunchKernelGGL(k_visc_ghost_x, dim3(nblocks(g.ncell / g.n[0], 256)), dim3(256), 0, st, Q, FVg[0], g.n[0],
                     g.n[1], g.n[2], dxg, vp);
  hipLaunchKernelGGL(k_visc_ghost<1>, dim3(nblocks(4 * (g.ncell / g.n[1]), 256)), dim3(256), 0, st, Q,
                     FVg[1], g.n[0], g.n[1], g.n[2], dxg, vp);
  hipLaunchKernelGGL(k_visc_ghost<2>, dim3(nblocks(4 * (g.ncell / g.n[2]), 256)), dim3(256), 0, st, Q,
                     FVg[2], g.n[0], g.n[1], g.n[2], dxg, vp);
}

unsigned ens_visc_items_per_member(const int n[3]) {
  const size_t ncell = (size_t)n[0] * n[1] * n[2];
  return visc_tiles(n[0], n[1], n[2]) + ghost_blocks(0, ncell, n[0]) + ghost_blocks(1, ncell, n[1]) +
         ghost_blocks(2, ncell, n[2]);
}

void launch_ens_visc(const Scheme &sc, const EnsMember *mem, int nmem, const unsigned *item0v, unsigned nitems,
                     hipStream_t st) {
  EnsViscArgs a;
  a.mem = mem;
  a.nmem = nmem;
  a.item0v = item0v;
  a.nitems = nitems;
  a.vp = visc_params(sc);
  hipLaunchKernelGGL(k_ens_visc, dim3(nitems), dim3(VTX * VTY), 0, st, a);
}

void launch_ens_prims(const Scheme &sc, const EnsMember *mem, int nmem, const double *u, size_t ncell,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_ens_prims, dim3(nblocks(ncell, 256)), dim3(256), 0, st, mem, nmem, u, ncell, sc.gamma);
}

void launch_stage(const Grid &g, const Scheme &sc, const double *fIx, const double *fIy,
                  const double *fIz, const double *FV, double *const FVg[3], double *u, double *S,
                  double *acc, double *out, int mode, const double *dxg, int stage, double ca,
                  double cb, hipStream_t st) {
  hipLaunchKernelGGL(k_stage, dim3(nblocks(g.ncell, 256)), dim3(256), 0, st, fIx, fIy, fIz, FV,
                     FVg ? FVg[0] : nullptr, FVg ? FVg[1] : nullptr, FVg ? FVg[2] : nullptr, u, S, acc,
                     out, mode, g.n[0], g.n[1], g.n[2], dxg, sc.Re > 0 ? 1 : 0, stage, ca, cb);
}

// Volume integral of every conserved variable: sum of u dV over the
// interior cells with dV = 1/dxinv_x 1/dxinv_y 1/dxinv_z
// (src/HyParFunctions/VolumeIntegral.c:14-50); one partial per block and
// variable, part[v*gridDim.x + block].
__global__ void __launch_bounds__(256) k_volint(const double *__restrict__ u, double *__restrict__ part, int nx,
                                                int ny, int nz, const double *__restrict__ dxg) {
  const size_t ncell = (size_t)nx * ny * nz;
  __shared__ double red[NV][4];
  double acc[NV] = {0, 0, 0, 0, 0};
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < ncell;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx % nx);
    const int j = (int)((idx / nx) % ny);
    const int k = (int)(idx / ((size_t)nx * ny));
    double dV = 1.0;
    dV *= 1.0 / dxg[3 + i];
    dV *= 1.0 / dxg[(nx + 6) + 3 + j];
    dV *= 1.0 / dxg[(nx + 6) + (ny + 6) + 3 + k];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] += u[(size_t)v * ncell + idx] * dV;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double t = wave_sum(acc[v]);
    if (lane == 0) red[v][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double r = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) r += red[threadIdx.x][q];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = r;
  }
}

unsigned reduction_blocks(size_t n) {
  unsigned b = nblocks(n, 256);
  return b > 2048 ? 2048 : (b == 0 ? 1 : b);
}

void launch_volint(const Grid &g, const double *u, double *part, unsigned nb, const double *dxg, hipStream_t st) {
  hipLaunchKernelGGL(k_volint, dim3(nb), dim3(256), 0, st, u, part, g.n[0], g.n[1], g.n[2], dxg);
}

void launch_cfl(const Grid &g, const Scheme &sc, const double *u, double *part, unsigned nb,
                const double *dxg, double dt, hipStream_t st) {
  hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(256), 0, st, u, part, g.n[0], g.n[1], g.n[2], dxg, sc.gamma, dt);
}

void launch_sqdiff(const double *a, const double *b, double *part, unsigned nb, size_t n,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_sqdiff, dim3(nb), dim3(256), 0, st, a, b, part, n);
}

void launch_aos_to_soa(const Grid &g, const double *aos, double *soa, int ghosts, hipStream_t st) {
  hipLaunchKernelGGL(k_aos_to_soa, dim3(nblocks(g.ncell, 256)), dim3(256), 0, st, aos, soa, g.n[0], g.n[1],
                     g.n[2], ghosts);
}

void launch_soa_to_aos(const Grid &g, const double *soa, double *aos, int ghosts, hipStream_t st) {
  hipLaunchKernelGGL(k_soa_to_aos, dim3(nblocks(g.ncell, 256)), dim3(256), 0, st, soa, aos, g.n[0], g.n[1],
                     g.n[2], ghosts);
}

}  // namespace hm
