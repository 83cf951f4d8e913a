// hm_kernels.hip -- gfx950 kernels of the NS3D WENO5 + RK4 path besides the
// direction passes (hm_pass.hip).
//
// State layout: structure of arrays without ghosts (see include/hypar_mi355x.h).
// All boundaries are periodic; stencil indices wrap.
//
//   k_visc_par, k_ens_visc  the viscous term of every cell from the primitive
//                           variables the previous z pass stored
//                           (NavierStokes3DParabolicFunction.c:50-325), with
//                           the reference's ghost-layer fluxes at the faces
//   k_par_out               the viscous term as a five-variable array
//   k_cfl, k_sqdiff, k_volint  reductions: CFL number, step norm, volume
//                           integrals (NavierStokes3DComputeCFL.c:16-55,
//                           TimePostStep.c:43-63, VolumeIntegral.c:14-50)
//   k_aos_to_soa, k_soa_to_aos  layout conversion at the host transfers
#include <hip/hip_runtime.h>

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

// ---------------------------------------------------------------------------
// viscous terms
// ---------------------------------------------------------------------------
// Primitive variables used by the viscous term: (u, v, w, T) with
// T = gamma*P/rho (NavierStokes3DParabolicFunction.c:85-108).
struct QT {
  double u, v, w, T;
};

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

// Viscous term of the right-hand side (NavierStokes3DParabolicFunction.c:50-325):
//   par = (dxinv Dx FVx + dyinv Dy FVy) + dzinv Dz FVz,
// summed in the reference's order, with FVd the viscous flux along d of the
// primitive variables (u, v, w, T) and their gradient (4th-order central
// differences with periodic neighbours, scaled by the point's dxinv;
// FirstDerivativeFourthOrder.c:103) and Dd the same central difference of
// the flux.  par[c*ncell + cell], c = momentum x, y, z, energy (the mass
// component is zero).  Q holds u, v, w, T of every cell, Q[c*ncell + cell];
// the viscosity is mu = T^0.76 = exp(0.76 log T) (raiseto,
// include/math_ops.h).
//
// Where the stencil of Dd leaves the grid (positions -2, -1, n, n+1 along
// d) the reference reads the fluxes of its ghost layers: the primitive
// values there are the periodic copies, the derivative along d uses the
// biased stencils of FirstDerivativeFourthOrder.c:74-93/106-125 in the outer
// layer and the central one in the inner layer, scaled by the ghost dxinv
// (a copy of the nearest interior value, InitialSolution.c:101-113), and the
// derivatives across d are zero, because the reference computes those only
// on lines whose other indices are interior.  The kernel evaluates these
// ghost-layer fluxes where its stencils need them.
//
// A block owns a TX x TY tile of (x, y) columns and marches through the
// planes of one z segment [k0, k1) (pks planes).  Plane q of Q with a
// 4-cell halo in x and y is staged in LDS.  Every thread evaluates the
// gradient at its own column and at most one halo point (the rows y0-2,
// y0-1, y0+TY, y0+TY+1 and the columns x0-2, x0-1, x0+TX, x0+TX+1 of the
// tile), taking the z derivative from a ring of that point's values on
// planes q-2..q+2 held in registers (in a thread-private part of LDS for
// the halo point).  The next plane's loads are issued before the current
// plane's arithmetic and stored to LDS after it.  The x and y fluxes of plane q go to
// LDS, where their x and y differences are formed.  The z fluxes of the
// column stay in a five-plane register ring, so the term of plane q-2 is
// complete once plane q is done.  The z differences of the segment need FVz
// on planes k0-2..k1+1, so the march covers them; on the ghost planes (-2,
// -1, nz, nz+1) the z flux comes from direct loads of the column.
template <int TX, int TY>
struct ViscParShared {
  double q[4][TY + 8][TX + 8];  // plane q of Q: x0-4..x0+TX+3, y0-4..y0+TY+3
  double fx[4][TY][TX + 4];     // x fluxes of the tile's rows at x0-2..x0+TX+1
  double fy[4][TY + 4][TX];     // y fluxes of the tile's columns at y0-2..y0+TY+1
  // z ring of each halo point, planes by p mod 5, and the x + y part of
  // the own column's term on planes q-2..q, by p mod 3 (private to the thread)
  double hr[5][4][4 * TX + 4 * TY];
  double pxy[3][4][TX * TY];
};

__device__ __forceinline__ int wrapi(int p, int n) {
  const int r = p % n;
  return r < 0 ? r + n : r;
}

// d/dx_d of one primitive at the ghost position of slot s (0..3 for the
// positions -2, -1, n, n+1) from f(o), its values at offsets o from that
// position along d (periodic copies); not yet scaled by dxinv.  Only the
// offsets a slot's stencil uses are read: -1..3 (slot 0), -3..1 (slot 3),
// -2..2 (slots 1, 2).
template <class F>
__device__ __forceinline__ double ghost_deriv(int slot, const F &f) {
  const double r12 = 1.0 / 12.0;
  if (slot == 0) return (-3 * f(-1) - 10 * f(0) + 18 * f(1) - 6 * f(2) + f(3)) * r12;
  if (slot == 3) return (-f(-3) + 6 * f(-2) - 18 * f(-1) + 10 * f(0) + 3 * f(1)) * r12;
  return d4c(f(-2), f(-1), f(1), f(2));
}

// Viscous flux along DIR at a ghost position: derivative along DIR only
// (d[c], already scaled), primitive values q.
template <int DIR>
__device__ __forceinline__ void ghost_flux(double fv[4], const QT &q, const double (&d)[4], const ViscParams &vp) {
  const double mu = exp(0.76 * log(q.T));
  double gr[4][3];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int e = 0; e < 3; e++) gr[c][e] = e == DIR ? d[c] : 0.0;
  visc_flux<DIR>(fv, q, gr, mu, vp);
}

// tile b of the grid; every thread of the block takes part
template <int TX, int TY>
__device__ __forceinline__ void visc_par_tile(const double *__restrict__ Q, double *__restrict__ par, int nx,
                                              int ny, int nz,
                                              const double *__restrict__ dxg,  // ghosted dxinv x|y|z
                                              const ViscParams &vp, int b, int pks, ViscParShared<TX, TY> &sh) {
  constexpr int NT = TX * TY;
  static_assert(4 * TX + 4 * TY <= NT, "every halo point needs a thread");
  const size_t ncell = (size_t)nx * ny * nz;
  const size_t plane = (size_t)nx * ny;
  const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
  const int tx = b % ntx, ty = (b / ntx) % nty, kz = b / (ntx * nty);
  const int x0 = tx * TX, y0 = ty * TY, k0 = kz * pks;
  const int k1 = min(k0 + pks, nz);
  const double *dxa = dxg, *dya = dxg + (nx + 6), *dza = dxg + (nx + 6) + (ny + 6);  // entry 3+p: position p
  const int t = threadIdx.x;
  const int lx = t % TX, ly = t / TX;
  const int i = x0 + lx, j = y0 + ly;
  const bool own = i < nx && j < ny;
  const size_t col = (size_t)wrapi(j, ny) * nx + wrapi(i, nx);
  // halo point: y fluxes of rows y0-2, y0-1, y0+TY, y0+TY+1 (threads
  // 0..4*TX-1), x fluxes of columns x0-2, x0-1, x0+TX, x0+TX+1 (the next
  // 4*TY threads)
  int hkind = -1, hx = 0, hy = 0;
  if (t < 4 * TX) {
    const int r = t / TX;
    hkind = 1;
    hx = t % TX;
    hy = r < 2 ? r - 2 : TY + r - 2;
  } else if (t < 4 * TX + 4 * TY) {
    const int u = t - 4 * TX, r = u % 4;
    hkind = 0;
    hy = u / 4;
    hx = r < 2 ? r - 2 : TX + r - 2;
  }
  const int hi = x0 + hx, hj = y0 + hy;
  const size_t hcol = (size_t)wrapi(hj, ny) * nx + wrapi(hi, nx);
  // does the halo point carry a flux the tile needs, and is it interior?
  const bool hneed = hkind == 1 ? (hi < nx && hj >= -2 && hj < ny + 2)
                                : (hkind == 0 && hj < ny && hi >= -2 && hi < nx + 2);
  const bool hint = hkind == 1 ? (hj >= 0 && hj < ny) : (hi >= 0 && hi < nx);

  auto ldq = [&](size_t c, int k, double (&r)[4]) {
    const size_t o = (size_t)wrapi(k, nz) * plane + c;
#pragma unroll
    for (int v = 0; v < 4; v++) r[v] = Q[(size_t)v * ncell + o];
  };
  // Q on LDS at tile offsets (ox, oy) (-4 <= ox < TX + 4, -4 <= oy < TY + 4)
  auto sq = [&](int v, int oy, int ox) { return sh.q[v][oy + 4][ox + 4]; };
  auto qat = [&](int ox, int oy) {
    QT q;
    q.u = sq(0, oy, ox);
    q.v = sq(1, oy, ox);
    q.w = sq(2, oy, ox);
    q.T = sq(3, oy, ox);
    return q;
  };
  // gradient at tile offset (ox, oy) of interior plane q; zr(s, v) is
  // component v on plane q-2+s of the point
  auto grad = [&](int ox, int oy, int q, const auto &zr, double (&gr)[4][3]) {
    const double dxi = dxa[3 + x0 + ox], dyi = dya[3 + y0 + oy], dzi = dza[3 + q];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      gr[v][0] = d4c(sq(v, oy, ox - 2), sq(v, oy, ox - 1), sq(v, oy, ox + 1), sq(v, oy, ox + 2)) * dxi;
      gr[v][1] = d4c(sq(v, oy - 2, ox), sq(v, oy - 1, ox), sq(v, oy + 1, ox), sq(v, oy + 2, ox)) * dyi;
      gr[v][2] = d4c(zr(0, v), zr(1, v), zr(3, v), zr(4, v)) * dzi;
    }
  };
  // ghost-layer x / y flux at tile offset (ox, oy)
  auto ghost_x = [&](int ox, int oy) {
    const int gx = x0 + ox, slot = gx < 0 ? gx + 2 : gx - nx + 2;
    double d[4];
#pragma unroll
    for (int v = 0; v < 4; v++) d[v] = ghost_deriv(slot, [&](int o) { return sq(v, oy, ox + o); }) * dxa[3 + gx];
    double fv[4];
    ghost_flux<0>(fv, qat(ox, oy), d, vp);
#pragma unroll
    for (int v = 0; v < 4; v++) sh.fx[v][oy][ox + 2] = fv[v];
  };
  auto ghost_y = [&](int ox, int oy) {
    const int gy = y0 + oy, slot = gy < 0 ? gy + 2 : gy - ny + 2;
    double d[4];
#pragma unroll
    for (int v = 0; v < 4; v++) d[v] = ghost_deriv(slot, [&](int o) { return sq(v, oy + o, ox); }) * dya[3 + gy];
    double fv[4];
    ghost_flux<1>(fv, qat(ox, oy), d, vp);
#pragma unroll
    for (int v = 0; v < 4; v++) sh.fy[v][oy + 2][ox] = fv[v];
  };

  // staging of plane q into LDS: every thread takes NR points of the
  // (TX + 8) x (TY + 8) region; the loads are issued a plane ahead into
  // registers and stored once the previous plane's reads are done
  constexpr int NQP = (TX + 8) * (TY + 8), NR = (NQP + NT - 1) / NT;
  double pre[NR][4];
  auto fetch_plane = [&](int q) {
    const size_t pk = (size_t)q * plane;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int c = t + r * NT;
      if (c < NQP) {
        const int ox = c % (TX + 8), oy = c / (TX + 8);
        const size_t o = pk + (size_t)wrapi(y0 - 4 + oy, ny) * nx + wrapi(x0 - 4 + ox, nx);
#pragma unroll
        for (int v = 0; v < 4; v++) pre[r][v] = Q[(size_t)v * ncell + o];
      }
    }
  };
  auto store_plane = [&]() {
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int c = t + r * NT;
      if (c < NQP) {
        const int ox = c % (TX + 8), oy = c / (TX + 8);
#pragma unroll
        for (int v = 0; v < 4; v++) sh.q[v][oy][ox] = pre[r][v];
      }
    }
  };

  // rings of planes q-2..q+2 of the own column (registers) and the halo
  // point (LDS, plane p in slot p mod 5), and the next plane of each, loaded
  // a step ahead
  double rq[5][4], nq[4], nh[4];
  double fz[5][4];
  auto hslot = [&](int p, int v) -> double & { return sh.hr[wrapi(p, 5)][v][t]; };
#pragma unroll
  for (int s = 0; s < 5; s++)
#pragma unroll
    for (int v = 0; v < 4; v++) rq[s][v] = fz[s][v] = 0.0;
#pragma unroll
  for (int v = 0; v < 4; v++) nh[v] = 0.0;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    ldq(col, k0 - 4 + s, rq[s + 1]);
    if (hkind >= 0) {
      double r[4];
      ldq(hcol, k0 - 4 + s, r);
#pragma unroll
      for (int v = 0; v < 4; v++) hslot(k0 - 4 + s, v) = r[v];
    }
  }
  ldq(col, k0, nq);
  if (hkind >= 0) ldq(hcol, k0, nh);
  if (k0 - 2 >= 0) {  // the first plane of the march, unless a ghost plane
    fetch_plane(k0 - 2);
    store_plane();
  }
  for (int q = k0 - 2; q < k1 + 2; q++) {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int v = 0; v < 4; v++) rq[s][v] = rq[s + 1][v];
#pragma unroll
    for (int v = 0; v < 4; v++) rq[4][v] = nq[v];
    if (hkind >= 0)
#pragma unroll
      for (int v = 0; v < 4; v++) hslot(q + 2, v) = nh[v];
    const bool gz = q < 0 || q >= nz;     // a ghost plane along z (uniform)
    const bool outq = q >= k0 && q < k1;  // a plane whose x and y terms are formed
    const bool next = q + 1 < k1 + 2 && q + 1 >= 0 && q + 1 < nz;  // plane q+1 is staged
    __syncthreads();
    if (next) fetch_plane(q + 1);
    if (q + 1 < k1 + 2) {
      ldq(col, q + 3, nq);
      if (hkind >= 0) ldq(hcol, q + 3, nh);
    }
    double fzq[4] = {0.0, 0.0, 0.0, 0.0};
    if (gz) {
      if (own) {
        const int slot = q < 0 ? q + 2 : q - nz + 2;
        double d[4];
#pragma unroll
        for (int v = 0; v < 4; v++)
          d[v] = ghost_deriv(slot, [&](int o) { return Q[(size_t)v * ncell + (size_t)wrapi(q + o, nz) * plane + col]; }) *
                 dza[3 + q];
        const size_t o = (size_t)wrapi(q, nz) * plane + col;
        QT qv;
        qv.u = Q[o];
        qv.v = Q[ncell + o];
        qv.w = Q[2 * ncell + o];
        qv.T = Q[3 * ncell + o];
        ghost_flux<2>(fzq, qv, d, vp);
      }
    } else {
      if (own) {
        double gr[4][3];
        grad(lx, ly, q, [&](int s2, int v) { return rq[s2][v]; }, gr);
        const QT qv = qat(lx, ly);
        const double mu = exp(0.76 * log(qv.T));
        visc_flux<2>(fzq, qv, gr, mu, vp);
        if (outq) {
          double fv[4];
          visc_flux<0>(fv, qv, gr, mu, vp);
#pragma unroll
          for (int v = 0; v < 4; v++) sh.fx[v][ly][lx + 2] = fv[v];
          visc_flux<1>(fv, qv, gr, mu, vp);
#pragma unroll
          for (int v = 0; v < 4; v++) sh.fy[v][ly + 2][lx] = fv[v];
        }
      } else if (outq) {
        // the column lies past the grid: ghost-layer x flux at i = nx, nx+1
        // (rows inside), ghost-layer y flux at j = ny, ny+1 (columns inside)
        if (j < ny && i < nx + 2) ghost_x(lx, ly);
        if (i < nx && j < ny + 2) ghost_y(lx, ly);
      }
      if (outq && hneed) {
        if (hint) {
          double gr[4][3];
          grad(hx, hy, q, [&](int s2, int v) { return hslot(q - 2 + s2, v); }, gr);
          const QT qv = qat(hx, hy);
          const double mu = exp(0.76 * log(qv.T));
          double fv[4];
          if (hkind == 0) {
            visc_flux<0>(fv, qv, gr, mu, vp);
#pragma unroll
            for (int v = 0; v < 4; v++) sh.fx[v][hy][hx + 2] = fv[v];
          } else {
            visc_flux<1>(fv, qv, gr, mu, vp);
#pragma unroll
            for (int v = 0; v < 4; v++) sh.fy[v][hy + 2][hx] = fv[v];
          }
        } else if (hkind == 0) {
          ghost_x(hx, hy);
        } else {
          ghost_y(hx, hy);
        }
      }
    }
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int v = 0; v < 4; v++) fz[s][v] = fz[s + 1][v];
#pragma unroll
    for (int v = 0; v < 4; v++) fz[4][v] = fzq[v];
    __syncthreads();
    if (next) store_plane();  // nothing reads plane q from LDS any more
    if (outq && own) {
      const double dxi = dxa[3 + i], dyi = dya[3 + j];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const double px = dxi * d4c(sh.fx[v][ly][lx], sh.fx[v][ly][lx + 1], sh.fx[v][ly][lx + 3], sh.fx[v][ly][lx + 4]);
        const double py = dyi * d4c(sh.fy[v][ly][lx], sh.fy[v][ly + 1][lx], sh.fy[v][ly + 3][lx], sh.fy[v][ly + 4][lx]);
        sh.pxy[wrapi(q, 3)][v][t] = px + py;
      }
    }
    const int p = q - 2;  // plane whose z difference is now complete
    if (p >= k0 && own) {
      const double dzi = dza[3 + p];
      const size_t o = (size_t)p * plane + col;
#pragma unroll
      for (int v = 0; v < 4; v++)
        par[(size_t)v * ncell + o] = sh.pxy[wrapi(p, 3)][v][t] + dzi * d4c(fz[0][v], fz[1][v], fz[3][v], fz[4][v]);
    }
  }
}

__global__ void __launch_bounds__(PTX * PTY) k_visc_par(const double *__restrict__ Q, double *__restrict__ par,
                                                         int nx, int ny, int nz, int pks,
                                                         const double *__restrict__ dxg, ViscParams vp) {
  __shared__ ViscParShared<PTX, PTY> sh;
  visc_par_tile<PTX, PTY>(Q, par, nx, ny, nz, dxg, vp, blockIdx.x, pks, sh);
}

// Viscous terms of all members of an ensemble: one launch with a block per
// (member, tile) work item; ptiles0 holds each member's first item.
struct EnsViscArgs {
  const EnsMember *mem;
  int nmem;
  const unsigned *ptiles0;
  unsigned nitems;
  ViscParams vp;
};

__global__ void __launch_bounds__(PTX * PTY) k_ens_visc(EnsViscArgs a) {
  __shared__ ViscParShared<PTX, PTY> sh;
  const unsigned item = blockIdx.x;  // one block per item (a loop over items spills 8 B per lane)
  if (item >= a.nitems) return;
  int lo = 0, hi = a.nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.ptiles0[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  const EnsMember &M = a.mem[lo];
  visc_par_tile<PTX, PTY>(M.q, M.par, M.n[0], M.n[1], M.n[2], M.dxg, a.vp, (int)(item - a.ptiles0[lo]), PKS, sh);
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
// the parabolic term on its own (hm_parabolic): par, with the zero mass
// component, into a five-variable array
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_par_out(const double *__restrict__ par, double *__restrict__ out,
                                                 size_t ncell) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell) return;
  out[idx] = 0.0;
#pragma unroll
  for (int v = 0; v < 4; v++) out[(size_t)(v + 1) * ncell + idx] = par[(size_t)v * ncell + idx];
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

static ViscParams visc_params(const Scheme &sc) {
  ViscParams vp;
  vp.gamma = sc.gamma;
  vp.inv_Re = sc.Re > 0 ? 1.0 / sc.Re : 0.0;
  vp.inv_gm1 = 1.0 / (sc.gamma - 1.0);
  vp.inv_Pr = 1.0 / sc.Pr;
  return vp;
}

void launch_visc_par(const Grid &g, const Scheme &sc, const double *Q, double *par, const double *dxg,
                     hipStream_t st) {
  const int pks = visc_pks(g.n[0], g.n[1], g.n[2]);
  const unsigned nb = par_tiles(g.n[0], g.n[1], g.n[2], pks);
  hipLaunchKernelGGL(k_visc_par, dim3(nb), dim3(PTX * PTY), 0, st, Q, par, g.n[0], g.n[1], g.n[2], pks, dxg,
                     visc_params(sc));
}

void launch_ens_visc(const Scheme &sc, const EnsMember *mem, int nmem, const unsigned *ptiles0, unsigned nitems,
                     hipStream_t st) {
  EnsViscArgs a;
  a.mem = mem;
  a.nmem = nmem;
  a.ptiles0 = ptiles0;
  a.nitems = nitems;
  a.vp = visc_params(sc);
  hipLaunchKernelGGL(k_ens_visc, dim3(nitems), dim3(PTX * PTY), 0, st, a);
}

void launch_ens_prims(const Scheme &sc, const EnsMember *mem, int nmem, const double *u, size_t ncell,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_ens_prims, dim3(nblocks(ncell, 256)), dim3(256), 0, st, mem, nmem, u, ncell, sc.gamma);
}

void launch_par_out(const Grid &g, const double *par, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_par_out, dim3(nblocks(g.ncell, 256)), dim3(256), 0, st, par, out, g.ncell);
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
