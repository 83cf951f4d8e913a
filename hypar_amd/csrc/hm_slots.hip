// hm_slots.hip -- the reference's per-slot operators as device kernels.
//
// HyPar composes its hyperbolic term from function-pointer slots
// (ReconstructHyperbolic, src/HyParFunctions/HyperbolicFunction.c:167-222):
// FFunction -> SetInterpLimiterVar -> UFunction -> 4x InterpolateInterfacesHyp
// -> Upwind -> flux difference.  These kernels implement each slot on
// hm-layout device arrays so that a host solver object can keep that
// composition (and swap any slot for its own device function); the fused
// kernels in hm_kernels.hip compute the same operator in one pass per
// direction.
//
// Interface arrays hold, at the index of cell i, the interface i-1/2 along
// the direction (periodic: n interfaces per line, interface n-1/2 is -1/2).
#include <hip/hip_runtime.h>

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

struct LineIdx {
  size_t idx, base, stride;
  int pos, n;
};

template <int DIR>
__device__ __forceinline__ LineIdx line_index(size_t idx, int nx, int ny, int nz) {
  LineIdx L;
  L.idx = idx;
  const int i = (int)(idx % nx), j = (int)((idx / nx) % ny), k = (int)(idx / ((size_t)nx * ny));
  L.n = DIR == 0 ? nx : (DIR == 1 ? ny : nz);
  L.pos = DIR == 0 ? i : (DIR == 1 ? j : k);
  L.stride = DIR == 0 ? 1 : (DIR == 1 ? (size_t)nx : (size_t)nx * ny);
  L.base = idx - (size_t)L.pos * L.stride;
  return L;
}

__device__ __forceinline__ size_t at(const LineIdx &L, int off) {
  return L.base + (size_t)wrapi(L.pos + off, L.n) * L.stride;
}

// FFunction = NavierStokes3DFlux (NavierStokes3DFlux.c:23-57)
template <int DIR>
__global__ void __launch_bounds__(256) k_flux(const double *__restrict__ u, double *__restrict__ f,
                                              size_t ncell, double gamma) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  double s[NV], fl[NV];
  for (int v = 0; v < NV; v++) s[v] = u[(size_t)v * ncell + idx];
  euler_flux<DIR>(fl, flowvar(s, gamma));
  for (int v = 0; v < NV; v++) f[(size_t)v * ncell + idx] = fl[v];
}

// UFunction = NavierStokes3DModifiedSolution without gravity
// (NavierStokes3DModifiedSolution.c:31-70)
__global__ void __launch_bounds__(256) k_modsol(const double *__restrict__ u, double *__restrict__ uC,
                                                size_t ncell, double gamma) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  double s[NV];
  for (int v = 0; v < NV; v++) s[v] = u[(size_t)v * ncell + idx];
  Prim q = flowvar(s, gamma);
  for (int v = 0; v < 4; v++) uC[(size_t)v * ncell + idx] = s[v];
  uC[4 * ncell + idx] = modified_energy(q, 1.0 / (gamma - 1.0));
}

// InterpolateInterfacesHyp = Interp1PrimFifthOrderWENO[Char] with the
// weights of SetInterpLimiterVar = WENOFifthOrderCalculateWeights[Char]:
// the smoothness indicators use fC (uflag = 0) or the raw solution u
// (uflag = 1, fC is then the modified solution), the characteristic
// variables use the Roe average of u at the cells left and right of the
// interface.
template <int DIR, int CHAR>
__global__ void __launch_bounds__(256) k_interp(double *__restrict__ fI, const double *__restrict__ fC,
                                                const double *__restrict__ u, int upw, int uflag,
                                                int nx, int ny, int nz, double gamma, double eps, int weno) {
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  const LineIdx L = line_index<DIR>(idx, nx, ny, nz);
  // stencil cells m3, m2, m1, p1, p2 of the biased interpolant
  int off[5];
  if (upw > 0) { off[0] = -3; off[1] = -2; off[2] = -1; off[3] = 0; off[4] = 1; }
  else         { off[0] = 2; off[1] = 1; off[2] = 0; off[3] = -1; off[4] = -2; }
  double val[5][NV], sm[5][NV];
  for (int s = 0; s < 5; s++) {
    const size_t c = at(L, off[s]);
    for (int v = 0; v < NV; v++) {
      val[s][v] = fC[(size_t)v * ncell + c];
      sm[s][v] = uflag ? u[(size_t)v * ncell + c] : val[s][v];
    }
  }
  if (!CHAR) {
    for (int v = 0; v < NV; v++)
      fI[(size_t)v * ncell + idx] = weno5(sm[0][v], sm[1][v], sm[2][v], sm[3][v], sm[4][v], val[0][v],
                                          val[1][v], val[2][v], val[3][v], val[4][v], eps, weno);
    return;
  }
  double ul[NV], ur[NV];
  for (int v = 0; v < NV; v++) {
    ul[v] = u[(size_t)v * ncell + at(L, -1)];
    ur[v] = u[(size_t)v * ncell + at(L, 0)];
  }
  Eigen E;
  eigen<DIR>(E, roe_average(ul, ur, gamma), gamma);
  double fc[NV];
  for (int r = 0; r < NV; r++) {
    double ps[5], pv[5];
    for (int s = 0; s < 5; s++) {
      double a = 0.0, b = 0.0;
      for (int k = 0; k < NV; k++) {
        a += E.L[r][k] * sm[s][k];
        b += E.L[r][k] * val[s][k];
      }
      ps[s] = a;
      pv[s] = b;
    }
    fc[r] = weno5(ps[0], ps[1], ps[2], ps[3], ps[4], pv[0], pv[1], pv[2], pv[3], pv[4], eps, weno);
  }
  for (int r = 0; r < NV; r++) {
    double a = 0.0;
    for (int c = 0; c < NV; c++) a += E.R[r][c] * fc[c];
    fI[(size_t)r * ncell + idx] = a;
  }
}

// Upwind = NavierStokes3DUpwindRoe / NavierStokes3DUpwindRusanov
// (NavierStokes3DUpwind.c:40-124, 349-420)
template <int DIR, int UPW>
__global__ void __launch_bounds__(256) k_upwind(double *__restrict__ fI, const double *__restrict__ fL,
                                                const double *__restrict__ fR, const double *__restrict__ uL,
                                                const double *__restrict__ uR, const double *__restrict__ u,
                                                int nx, int ny, int nz, double gamma) {
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  const LineIdx L = line_index<DIR>(idx, nx, ny, nz);
  double ul[NV], ur[NV], du[NV];
  for (int v = 0; v < NV; v++) {
    ul[v] = u[(size_t)v * ncell + at(L, -1)];
    ur[v] = u[(size_t)v * ncell + at(L, 0)];
    du[v] = 0.5 * (uR[(size_t)v * ncell + idx] - uL[(size_t)v * ncell + idx]);
  }
  Prim avg = roe_average(ul, ur, gamma);
  double diss[NV];
  if (UPW == UPWIND_RUSANOV) {
    Prim l = flowvar(ul, gamma), r = flowvar(ur, gamma);
    const double aL = sqrt(gamma * l.P / l.rho) + fabs(vel<DIR>(l));
    const double aR = sqrt(gamma * r.P / r.rho) + fabs(vel<DIR>(r));
    const double aA = sqrt(gamma * avg.P / avg.rho) + fabs(vel<DIR>(avg));
    const double alpha = fmax(fmax(aL, aR), fmax(aR, aA));
    for (int v = 0; v < NV; v++) diss[v] = alpha * du[v];
  } else {
    Eigen E;
    eigen<DIR>(E, avg, gamma);
    double c[NV];
    for (int r = 0; r < NV; r++) {
      double a = 0.0;
      for (int k = 0; k < NV; k++) a += E.L[r][k] * du[k];
      c[r] = harten_abs(E.lam[r]) * a;
    }
    for (int r = 0; r < NV; r++) {
      double a = 0.0;
      for (int k = 0; k < NV; k++) a += E.R[r][k] * c[k];
      diss[r] = a;
    }
  }
  for (int v = 0; v < NV; v++)
    fI[(size_t)v * ncell + idx] = 0.5 * (fL[(size_t)v * ncell + idx] + fR[(size_t)v * ncell + idx]) - diss[v];
}

// hyp (+)= dxinv * (F_{i+1/2} - F_{i-1/2}) along DIR (HyperbolicFunction.c:94-101)
template <int DIR>
__global__ void __launch_bounds__(256) k_fluxdiff(double *__restrict__ hyp, const double *__restrict__ fI,
                                                  int nx, int ny, int nz, const double *__restrict__ dxg,
                                                  int accumulate) {
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  const LineIdx L = line_index<DIR>(idx, nx, ny, nz);
  const double *dxi = dxg + 3 + (DIR >= 1 ? nx + 6 : 0) + (DIR == 2 ? ny + 6 : 0);
  const size_t ip = at(L, 1);
  for (int v = 0; v < NV; v++) {
    const size_t o = (size_t)v * ncell;
    const double d = dxi[L.pos] * (fI[o + ip] - fI[o + idx]);
    hyp[o + idx] = accumulate ? hyp[o + idx] + d : d;
  }
}

__global__ void k_axpy(const double *__restrict__ x, double a, double *__restrict__ y, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

__global__ void k_fill(double *__restrict__ x, double value, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = value;
}

// ---------------------------------------------------------------------------
static inline unsigned nb(size_t n) { return (unsigned)((n + 255) / 256); }
static inline unsigned nb_grid(size_t n) {
  size_t b = (n + 255) / 256;
  return (unsigned)(b > 65536 ? 65536 : (b == 0 ? 1 : b));
}

void launch_flux(const Grid &g, const Scheme &sc, int dir, const double *u, double *f, hipStream_t st) {
  if (dir == 0) hipLaunchKernelGGL(k_flux<0>, dim3(nb(g.ncell)), dim3(256), 0, st, u, f, g.ncell, sc.gamma);
  else if (dir == 1) hipLaunchKernelGGL(k_flux<1>, dim3(nb(g.ncell)), dim3(256), 0, st, u, f, g.ncell, sc.gamma);
  else hipLaunchKernelGGL(k_flux<2>, dim3(nb(g.ncell)), dim3(256), 0, st, u, f, g.ncell, sc.gamma);
}

void launch_modsol(const Grid &g, const Scheme &sc, const double *u, double *uC, hipStream_t st) {
  hipLaunchKernelGGL(k_modsol, dim3(nb(g.ncell)), dim3(256), 0, st, u, uC, g.ncell, sc.gamma);
}

template <int DIR>
static void interp_d(const Grid &g, const Scheme &sc, double *fI, const double *fC, const double *u,
                     int upw, int uflag, hipStream_t st) {
  if (sc.interp == INTERP_CHAR)
    hipLaunchKernelGGL((k_interp<DIR, 1>), dim3(nb(g.ncell)), dim3(256), 0, st, fI, fC, u, upw, uflag,
                       g.n[0], g.n[1], g.n[2], sc.gamma, sc.eps, sc.weno);
  else
    hipLaunchKernelGGL((k_interp<DIR, 0>), dim3(nb(g.ncell)), dim3(256), 0, st, fI, fC, u, upw, uflag,
                       g.n[0], g.n[1], g.n[2], sc.gamma, sc.eps, sc.weno);
}

void launch_interp(const Grid &g, const Scheme &sc, int dir, double *fI, const double *fC,
                   const double *u, int upw, int uflag, hipStream_t st) {
  if (dir == 0) interp_d<0>(g, sc, fI, fC, u, upw, uflag, st);
  else if (dir == 1) interp_d<1>(g, sc, fI, fC, u, upw, uflag, st);
  else interp_d<2>(g, sc, fI, fC, u, upw, uflag, st);
}

template <int DIR>
static void upwind_d(const Grid &g, const Scheme &sc, double *fI, const double *fL, const double *fR,
                     const double *uL, const double *uR, const double *u, hipStream_t st) {
  if (sc.upwind == UPWIND_ROE)
    hipLaunchKernelGGL((k_upwind<DIR, UPWIND_ROE>), dim3(nb(g.ncell)), dim3(256), 0, st, fI, fL, fR, uL, uR,
                       u, g.n[0], g.n[1], g.n[2], sc.gamma);
  else
    hipLaunchKernelGGL((k_upwind<DIR, UPWIND_RUSANOV>), dim3(nb(g.ncell)), dim3(256), 0, st, fI, fL, fR,
                       uL, uR, u, g.n[0], g.n[1], g.n[2], sc.gamma);
}

void launch_upwind(const Grid &g, const Scheme &sc, int dir, double *fI, const double *fL,
                   const double *fR, const double *uL, const double *uR, const double *u,
                   hipStream_t st) {
  if (dir == 0) upwind_d<0>(g, sc, fI, fL, fR, uL, uR, u, st);
  else if (dir == 1) upwind_d<1>(g, sc, fI, fL, fR, uL, uR, u, st);
  else upwind_d<2>(g, sc, fI, fL, fR, uL, uR, u, st);
}

void launch_fluxdiff(const Grid &g, int dir, double *hyp, const double *fI, const double *dxg,
                     int accumulate, hipStream_t st) {
  if (dir == 0)
    hipLaunchKernelGGL(k_fluxdiff<0>, dim3(nb(g.ncell)), dim3(256), 0, st, hyp, fI, g.n[0], g.n[1], g.n[2], dxg, accumulate);
  else if (dir == 1)
    hipLaunchKernelGGL(k_fluxdiff<1>, dim3(nb(g.ncell)), dim3(256), 0, st, hyp, fI, g.n[0], g.n[1], g.n[2], dxg, accumulate);
  else
    hipLaunchKernelGGL(k_fluxdiff<2>, dim3(nb(g.ncell)), dim3(256), 0, st, hyp, fI, g.n[0], g.n[1], g.n[2], dxg, accumulate);
}

void launch_axpy(const double *x, double a, double *y, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy, dim3(nb_grid(n)), dim3(256), 0, st, x, a, y, n);
}

void launch_fill(double *x, double value, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(nb_grid(n)), dim3(256), 0, st, x, value, n);
}

}  // namespace hm
