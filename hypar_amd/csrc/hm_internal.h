// hm_internal.h -- declarations shared by the C-ABI and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include <cmath>
#include <string>

struct hm_solver;

namespace hm {

// drops every slot binding of s (hm_bind.hip); hm_destroy calls it
void drop_bindings(const ::hm_solver *s);

enum { INTERP_COMP = 0, INTERP_CHAR = 1 };
enum { UPWIND_ROE = 0, UPWIND_RUSANOV = 1 };
enum { OUT_STAGE = 0, OUT_RHS = 1, OUT_HYP = 2, OUT_PAR = 3 };
enum { WENO_JS = 0, WENO_MAPPED = 1, WENO_Z = 2, WENO_YC = 3, WENO_OPTIMAL = 4 };

struct Grid {
  int n[3];
  size_t ncell;
};

struct Scheme {
  int interp, upwind;
  double gamma;
  double Re;  // as the solver uses it (physics.inp Re / Minf); <= 0 inviscid
  double Pr;
  double eps;
  int weno;   // WENO_*
};

// records msg as the calling thread's hm_last_error() and returns code
int set_error(int code, const std::string &msg);

// Metrics of the uniform periodic grid x_i = i*(2*pi/n) with g ghost points
// per side: ghost coordinates extrapolated linearly (ReadArray.c:59-90),
// dxinv = 2/(x[i+1]-x[i-1]) inside and copies of the edge value in the
// ghosts (InitialSolution.c:74-113).  X is caller scratch of n + 2g doubles;
// out receives n + 2g doubles.  Host code (compiled without FMA contraction
// on x86-64), so the values are those of the reference's own host code.
inline void uniform_dxinv(int n, int g, double *X, double *out) {
  const double h = 2.0 * M_PI / n;
  for (int i = 0; i < n; i++) X[g + i] = (double)i * h;
  for (int i = 0; i < g; i++) {
    int delta = g - i;
    X[i] = X[g] + ((double)delta) * (X[g] - X[g + 1]);
  }
  for (int i = n + g; i < n + 2 * g; i++) {
    int delta = i - (n + g - 1);
    X[i] = X[n + g - 1] + ((double)delta) * (X[n + g - 1] - X[n + g - 2]);
  }
  for (int i = 0; i < n; i++) out[i + g] = 2.0 / (X[i + 1 + g] - X[i - 1 + g]);
  for (int i = 0; i < g; i++) out[i] = out[g];
  for (int i = n + g; i < n + 2 * g; i++) out[i] = out[n + g - 1];
}

// the viscous term par (4*ncell: momentum x, y, z, energy) of the state
// whose primitive variables (u, v, w, T) are in Q (4*ncell)
void launch_visc_par(const Grid &g, const Scheme &sc, const double *Q, double *par, const double *dxg,
                     hipStream_t st);
// par as a five-variable array with a zero mass component
void launch_par_out(const Grid &g, const double *par, double *out, hipStream_t st);
unsigned reduction_blocks(size_t n);
// volume integral partials: part[v*nb + block], 5*nb entries
void launch_volint(const Grid &g, const double *u, double *part, unsigned nb, const double *dxg, hipStream_t st);
void launch_cfl(const Grid &g, const Scheme &sc, const double *u, double *part, unsigned nb,
                const double *dxg, double dt, hipStream_t st);
void launch_sqdiff(const double *a, const double *b, double *part, unsigned nb, size_t n,
                   hipStream_t st);
void launch_aos_to_soa(const Grid &g, const double *aos, double *soa, int ghosts, hipStream_t st);
void launch_soa_to_aos(const Grid &g, const double *soa, double *aos, int ghosts, hipStream_t st);

// per-slot operators (hm_slots.hip)
void launch_flux(const Grid &g, const Scheme &sc, int dir, const double *u, double *f, hipStream_t st);
void launch_modsol(const Grid &g, const Scheme &sc, const double *u, double *uC, hipStream_t st);
void launch_interp(const Grid &g, const Scheme &sc, int dir, double *fI, const double *fC,
                   const double *u, int upw, int uflag, hipStream_t st);
void launch_upwind(const Grid &g, const Scheme &sc, int dir, double *fI, const double *fL,
                   const double *fR, const double *uL, const double *uR, const double *u,
                   hipStream_t st);
void launch_fluxdiff(const Grid &g, int dir, double *hyp, const double *fI, const double *dxg,
                     int accumulate, hipStream_t st);
void launch_axpy(const double *x, double a, double *y, size_t n, hipStream_t st);
void launch_fill(double *x, double value, size_t n, hipStream_t st);

// fused direction passes (hm_pass.hip)
struct PassBuffers {
  const double *U;        // stage state
  double *R;              // hyperbolic-term accumulator, 5*ncell
  const double *par;      // viscous term, 4*ncell (viscous runs; read by the z pass)
  double *u, *S, *acc;    // RK4 state, next stage state, accumulator (z pass, OUT_STAGE)
  double *out;            // right-hand side or hyperbolic term (z pass, OUT_RHS / OUT_HYP)
  double *Q;              // u, v, w, T of the new state (z pass, viscous OUT_STAGE)
};
// pass along dir: x stores its flux differences in R, y adds its own, z adds
// its own and forms rhs = -hyp + par, then finishes the stage (mode
// OUT_STAGE with stage, ca = dt*A[s+1][s], cb = dt*b[s]) or writes rhs
// (OUT_RHS) or hyp (OUT_HYP)
void launch_pass(const Grid &g, const Scheme &sc, int dir, const PassBuffers &b, int mode, int stage, double ca,
                 double cb, const double *dxg, hipStream_t st);
// the z pass with a three-output stage update: with r the right-hand side
// and a, b the operands A, B at each cell, O1 = a + c1 r, O2 = b + c2 r (the
// next stage state; its primitive variables go to b.Q when viscous),
// O3 = a + c3 r.  The first nout (1-3) outputs are stored; an output left
// out must repeat O1 (same array and value).  O2 may be B.  The x and y
// passes of the stage are launch_pass's.
void launch_pass_z_affine(const Grid &g, const Scheme &sc, const PassBuffers &b, const double *A, const double *B,
                          double *O1, double *O2, double *O3, double c1, double c2, double c3, int nout,
                          const double *dxg, hipStream_t st);
// primitive variables (u, v, w, T) of u into Q
void launch_prims(const Grid &g, const Scheme &sc, const double *u, double *Q, hipStream_t st);

// tiles of the viscous kernels (hm_kernels.hip, hm_adjoint.hip): a block
// owns VTX x VTY (x, y) columns and marches through VKS z-planes
constexpr int VTX = 64, VTY = 8, VKS = 16;
constexpr int VT = VTX * VTY;  // threads of a viscous block
constexpr int VHX = VTX + 4, VHY = VTY + 4;

// tiles (x, y, z-segment) of a grid
__host__ __device__ inline unsigned visc_tiles(int nx, int ny, int nz) {
  return (unsigned)(((nx + VTX - 1) / VTX) * ((ny + VTY - 1) / VTY) * ((nz + VKS - 1) / VKS));
}

// tiles of the viscous-term kernel (hm_kernels.hip): a block owns PTX x PTY
// (x, y) columns and marches through a z segment of pks planes: PKS in the
// batched ensemble, visc_pks() of the grid for one grid
#ifndef HM_PTX
#define HM_PTX 32
#endif
#ifndef HM_PTY
#define HM_PTY 16
#endif
#ifndef HM_PKS
#define HM_PKS 16
#endif
constexpr int PTX = HM_PTX, PTY = HM_PTY, PKS = HM_PKS;
__host__ __device__ inline unsigned par_tiles(int nx, int ny, int nz, int pks = PKS) {
  return (unsigned)(((nx + PTX - 1) / PTX) * ((ny + PTY - 1) / PTY) * ((nz + pks - 1) / pks));
}
// Segment length of the one-grid viscous kernel: the longest of 128, 64, 32
// planes that still gives at least VISC_MIN_TILES tiles (the kernel holds
// one 512-thread block per CU, 256 CUs), else PKS.  Longer segments read
// fewer halo planes: a segment stages 8 planes besides its own.
#ifndef HM_VISC_MIN_TILES
#define HM_VISC_MIN_TILES 512
#endif
inline int visc_pks(int nx, int ny, int nz) {
  for (int p = 128; p > PKS; p >>= 1)
    if (par_tiles(nx, ny, nz, p) >= (unsigned)HM_VISC_MIN_TILES) return p;
  return PKS;
}

// batched ensemble passes (hm_pass.hip): members of one scheme
struct EnsMember {
  int n[3];
  size_t ncell;
  size_t off;         // first cell of the member; its state is u[5*off .. 5*(off+ncell))
  const double *dxg;  // ghosted dxinv x|y|z of the member (device)
  double dt;          // the member's time step
  // viscous members: viscous term (4*ncell) and primitive variables
  // (4*ncell); null when inviscid
  double *par, *q;
};
// work items of one member in pass dir with tile shape `shape` (0: 4
// wavefronts per block, 1: 2 wavefronts per block)
unsigned ens_items_per_member(int dir, int shape, const int n[3]);
// one pass of RK stage `stage` over all members; grid = number of
// persistent blocks (0: one block per work item)
void launch_ens_pass(const Scheme &sc, int dir, int shape, const EnsMember *mem, int nmem, const unsigned *item0,
                     unsigned nitems, unsigned grid, const double *U, double *R, double *u, double *S, double *acc,
                     int stage, hipStream_t st);
// all three passes of RK stage `stage` over all members in one persistent
// launch (rows = 4 or 2 wavefronts per block, the tile shapes of the split
// passes; grid 0 = the blocks the GPU holds at once, cus CUs).  An x item
// is ENS_STAGE_XG groups of x-lines: item0[0] lists each member's first x
// item, xgroups0 its first x-line group (the split x pass's items).  ctr
// holds 4 + 2*nmem zeroed counters and is zero again when the launch ends;
// a dependency wait that polls more than spin_limit times sets ctr[2].
#ifndef HM_ENS_XG
#define HM_ENS_XG 64
#endif
constexpr unsigned ENS_STAGE_XG = HM_ENS_XG;
void launch_ens_stage(const Scheme &sc, int rows, const EnsMember *mem, int nmem, const unsigned *const item0[3],
                      const unsigned nitems[3], const unsigned *xgroups0, unsigned xg, unsigned grid, int cus,
                      unsigned *ctr, unsigned spin_limit, const double *U, double *R, double *u, double *S,
                      double *acc, int stage, hipStream_t st);

// viscous terms of all members (hm_kernels.hip): one launch over the
// (member, par_tiles) items, ptiles0 = each member's first item (nmem + 1
// entries); primitive variables of all members
void launch_ens_visc(const Scheme &sc, const EnsMember *mem, int nmem, const unsigned *ptiles0, unsigned nitems,
                     hipStream_t st);
void launch_ens_prims(const Scheme &sc, const EnsMember *mem, int nmem, const double *u, size_t ncell,
                      hipStream_t st);

// discrete adjoint (hm_adjoint.hip)
bool vjp_supported(const Grid &g, const Scheme &sc);
// doubles of device work space launch_rhs_vjp needs (viscous schemes)
size_t vjp_work_size(const Grid &g, const Scheme &sc);
// nu = (d rhs/d u)^T mu at u with mu = cb*lam + ca*nu_in (nu_in may be
// null); when L is not null the finished nu is also added into L (set_L:
// L = lam + nu); work holds vjp_work_size doubles
void launch_rhs_vjp(const Grid &g, const Scheme &sc, const double *u, const double *lam, const double *nu_in,
                    double ca, double cb, double *nu, double *L, bool set_L, const double *dxg, double *work,
                    hipStream_t st);
void launch_axpby(const double *x, double a, double *y, double b, size_t n, hipStream_t st);

// the same for all members of an ensemble at once: member m's stage
// coefficients are dt_m*A and dt_m*B; lines0[d] lists each member's first
// line along d (nmem + 1 entries), nmax[d] the longest line along d, and
// items0[d] each member's first work item of the line VJPs along d, an item
// being ens_vjp_lines_per_item(nmax[d], n[d]) consecutive lines of the member
struct EnsVjpPlan {
  const EnsMember *mem;  // device member table
  int nmem;
  size_t ncell;
  const unsigned *lines0[3];
  unsigned nlines[3];
  const unsigned *items0[3];
  unsigned nitems[3];
  int nmax[3];
  const unsigned *tiles0;  // viscous members: each member's first viscous tile (nmem + 1 entries)
  unsigned ntiles;
};
bool ens_vjp_supported(const Scheme &sc, const int nmax[3]);
int ens_vjp_threads(int nmax);                 // block size of the line VJPs along a direction
int ens_vjp_lines_per_item(int nmax, int n);  // lines of length n per work item
size_t ens_vjp_work_size(const Scheme &sc, size_t ncell);
void launch_ens_rhs_vjp(const Scheme &sc, const EnsVjpPlan &p, const double *u, const double *lam,
                        const double *nu_in, double A, double B, double *nu, double *L, bool set_L, double *work,
                        hipStream_t st);

}  // namespace hm
