// hm_api.hip -- C-ABI of libhypar_mi355x.so (declared in include/hypar_mi355x.h).
//
// The solver object owns the per-stage scratch in HBM; the caller owns the
// state arrays.  All launches go to the solver's stream and are
// asynchronous unless the entry point returns a host value.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hypar_mi355x.h"
#include "hm_internal.h"

// The adjoint tape keeps a checkpoint every K steps (K = checkpoint_every)
// and, for `sets` segments of K steps at a time, either the four RK stage
// states of every step (stage level) or only the state each step starts
// from (step level: the reverse sweep runs each step forward again into
// three scratch states before taking it back).  The forward sweep keeps the
// last `sets` segments; the reverse sweep recomputes every earlier segment
// from its checkpoint into the set that segment's later neighbour freed.
// The forward sweep takes the cheapest schedule whose states fit the tape
// budget: stage level with 3, 2 or 1 sets (nck-3, nck-2 or nck-1 segments
// recomputed, nck = number of checkpoints), then step level with 3, 2 or 1
// sets (the same, plus one forward step per step).  A third stage-level
// set costs 4K states (128^3, K = 10: 3.4 GB) and saves K recomputed steps.
constexpr int TAPE_SETS = 3;  // most sets a schedule uses

struct TapePlan {
  int sets = TAPE_SETS;
  bool step_level = false;
};

namespace {
// states in the sets of a plan
size_t tape_slots(const TapePlan &p, int nsteps, int every) {
  return (size_t)(p.step_level ? 1 : 4) * (size_t)std::min(nsteps, p.sets * every);
}
// every state the tape and the reverse sweep hold: the checkpoints, the
// sets, 4 scratch states, and 3 stage states at step level
size_t tape_total_states(const TapePlan &p, int nsteps, int every) {
  const size_t nck = (size_t)((nsteps + every - 1) / every);
  return nck + tape_slots(p, nsteps, every) + 4 + (p.step_level ? 3 : 0);
}
// the cheapest plan whose states (state_bytes each, plus fixed_bytes of
// work space) fit in budget bytes
bool choose_tape(size_t budget, size_t state_bytes, size_t fixed_bytes, int nsteps, int every, TapePlan *out) {
  for (int level = 0; level < 2; level++)
    for (int sets = TAPE_SETS; sets >= 1; sets--) {
      TapePlan p;
      p.sets = sets;
      p.step_level = level == 1;
      if (fixed_bytes + tape_total_states(p, nsteps, every) * state_bytes <= budget) {
        *out = p;
        return true;
      }
    }
  return false;
}
// The budget when the caller gave none: the device's free memory plus what
// the tape already holds (it is reused or freed), less a reserve of 2 GiB or
// 2% of the device, whichever is larger.
size_t auto_tape_budget(size_t held_bytes) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return SIZE_MAX;
  const size_t reserve = std::max<size_t>((size_t)2 << 30, total_b / 50);
  const size_t avail = free_b + held_bytes;
  return avail > reserve ? avail - reserve : 0;
}
}  // namespace

struct hm_solver {
  hm_solver() : g(), sc(), device(0), stream(nullptr), dxg(nullptr), R(nullptr), par(nullptr), Q(nullptr),
                S{nullptr, nullptr},
                acc(nullptr), part(nullptr), part_host(nullptr), npart(0), tmp(nullptr), tmp_elems(0),
                tape_nsteps(0), tape_every(0), tape_dt(0.0) {
    for (int i = 0; i < TAPE_SETS; i++) tape_holds[i] = -1;
  }
  hm::Grid g;
  hm::Scheme sc;
  int device;
  hipStream_t stream;
  double *dxg;        // ghosted dxinv x|y|z on the device
  double *R;          // hyperbolic-term accumulator of the passes, 5*ncell
  double *par;        // viscous term, 4*ncell (viscous runs only)
  double *Q;          // primitive variables u, v, w, T, 4*ncell (viscous only)
  double *S[2];       // RK stage states (double-buffered)
  double *acc;        // RK accumulator
  double *part;       // reduction partials (device)
  double *part_host;  // reduction partials (pinned host)
  unsigned npart;
  double *tmp;        // layout-conversion staging
  size_t tmp_elems;
  // adjoint tape: checkpoints, segment stage states, scratch
  std::vector<double *> ckpt, stages, vscratch;
  std::vector<double *> vwork;  // viscous adjoint work space (one buffer)
  int tape_nsteps, tape_every;
  double tape_dt;
  TapePlan tape_plan;            // schedule of the recorded trajectory
  size_t tape_budget = 0;        // bytes the tape may hold (0: from the device's free memory)
  std::vector<double *> sscratch;  // step level: the stage states of the step being taken back
  // segment whose RK stage states set i holds (-1: none).  A reverse sweep
  // recomputes every segment its set does not hold, so a later sweep over
  // the same forward (or one after hm_rk4_vjp reused set 0) stays exact.
  int tape_holds[TAPE_SETS];
  // reverse sweep: a second stream recomputes the next segment while the
  // solver's stream runs the adjoint of the current one
  hipStream_t aux = nullptr;
  hipEvent_t ev_ready[TAPE_SETS] = {}, ev_done[TAPE_SETS] = {}, ev_start = nullptr;
};

namespace {

thread_local std::string g_err;

hm_status fail(hm_status code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HM_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(HM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

hm_status check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(HM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return HM_OK;
}

void uniform_dxinv(int n, int g, double *out) {
  std::vector<double> X(n + 2 * g);
  hm::uniform_dxinv(n, g, X.data(), out);
}

hm_status dmalloc(double **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, n * sizeof(double));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(HM_ERR_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(double)) + " bytes: " +
                                  hipGetErrorString(e));
  }
  return HM_OK;
}

bool is_device_ptr(const void *p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

hm_status check_state(const hm_solver *s, const void *p, const char *name) {
  if (!s) return fail(HM_ERR_ARG, "solver is NULL");
  if (!p) return fail(HM_ERR_ARG, std::string(name) + " is NULL");
  if (!is_device_ptr(p)) return fail(HM_ERR_ARG, std::string(name) + " is not a device pointer");
  return HM_OK;
}

// "rk 44" tableau (TimeExplicitRKInitialize.c:58-68)
const double RK_A_SUB[3] = {0.5, 0.5, 1.0};  // A[s+1][s]
const double RK_B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
// "rk ssprk3" tableau (TimeExplicitRKInitialize.c:69-79): A[1][0] = 1,
// A[2][0] = A[2][1] = 1/4, b = 1/6, 1/6, 2/3
const double SSP_A10 = 1.0, SSP_A20 = 0.25, SSP_A21 = 0.25;
const double SSP_B[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};

hm::PassBuffers stage_buffers(hm_solver *s, double *u, const double *U, double *Sn) {
  hm::PassBuffers b;
  b.U = U;
  b.R = s->R;
  b.par = s->par;
  b.u = u;
  b.S = Sn;
  b.acc = s->acc;
  b.out = nullptr;
  b.Q = s->Q;
  return b;
}

// One RK4 step of u in place through the fused direction passes.  Stage s
// reads Sx[s-1] (u for s = 0) and writes Sx[s]; the last stage writes u.
// Viscous runs need the primitive variables of u in Q on entry and leave
// those of the new u there.
hm_status rk4_step(hm_solver *s, double *u, double dt, double *const Sx[3]) {
  const bool visc = s->sc.Re > 0;
  for (int stage = 0; stage < 4; stage++) {
    const hm::PassBuffers b = stage_buffers(s, u, stage == 0 ? u : Sx[stage - 1], stage < 3 ? Sx[stage] : nullptr);
    if (visc) hm::launch_visc_par(s->g, s->sc, s->Q, s->par, s->dxg, s->stream);
    const double ca = stage < 3 ? dt * RK_A_SUB[stage] : 0.0, cb = dt * RK_B[stage];
    for (int d = 0; d < 3; d++) hm::launch_pass(s->g, s->sc, d, b, hm::OUT_STAGE, stage, ca, cb, s->dxg, s->stream);
    hm_status st = check_launch("RK4 stage passes");
    if (st != HM_OK) return st;
  }
  return HM_OK;
}

// One SSPRK3 step of u in place (TimeRK.c:126-195 with the "ssprk3"
// tableau):  U1 = u + dt k0,  U2 = (u + dt/4 k0) + dt/4 k1,
// u_new = ((u + dt/6 k0) + dt/6 k1) + 2dt/3 k2, in the reference's order.
// With the z pass's three-output update (O1 = a + c1 r, O2 = b + c2 r,
// O3 = a + c3 r):
//   stage 0 (from u):    acc = u + dt/6 k0, S0 = u + dt k0, S1 = u + dt/4 k0
//   stage 1 (from S0):   acc = acc + dt/6 k1, S1 = S1 + dt/4 k1 (in place), acc again
//   stage 2 (from S1):   u = acc + 2dt/3 k2 (all three outputs)
hm_status ssprk3_step(hm_solver *s, double *u, double dt) {
  const bool visc = s->sc.Re > 0;
  double *acc = s->acc, *S0 = s->S[0], *S1 = s->S[1];
  const double *Uin[3] = {u, S0, S1};
  const double *A[3] = {u, acc, acc}, *B[3] = {u, S1, acc};
  double *O1[3] = {acc, acc, u}, *O2[3] = {S0, S1, u}, *O3[3] = {S1, acc, u};
  const double c1[3] = {dt * SSP_B[0], dt * SSP_B[1], dt * SSP_B[2]};
  const double c2[3] = {dt * SSP_A10, dt * SSP_A21, dt * SSP_B[2]};
  const double c3[3] = {dt * SSP_A20, dt * SSP_B[1], dt * SSP_B[2]};
  // distinct outputs per stage: stage 1 stores acc twice (O3 = O1), stage 2
  // stores u three times (O1 = O2 = O3, with B = A)
  const int nout[3] = {3, 2, 1};
  for (int stage = 0; stage < 3; stage++) {
    const hm::PassBuffers b = stage_buffers(s, u, Uin[stage], nullptr);
    if (visc) hm::launch_visc_par(s->g, s->sc, s->Q, s->par, s->dxg, s->stream);
    for (int d = 0; d < 2; d++) hm::launch_pass(s->g, s->sc, d, b, hm::OUT_STAGE, stage, 0.0, 0.0, s->dxg, s->stream);
    hm::launch_pass_z_affine(s->g, s->sc, b, A[stage], B[stage], O1[stage], O2[stage], O3[stage], c1[stage],
                             c2[stage], c3[stage], nout[stage], s->dxg, s->stream);
    hm_status st = check_launch("SSPRK3 stage passes");
    if (st != HM_OK) return st;
  }
  return HM_OK;
}

hm_status ensure_tmp(hm_solver *s, size_t n) {
  if (s->tmp_elems >= n) return HM_OK;
  if (s->tmp) (void)hipFree(s->tmp);
  s->tmp = nullptr;
  s->tmp_elems = 0;
  hm_status st = dmalloc(&s->tmp, n);
  if (st == HM_OK) s->tmp_elems = n;
  return st;
}

}  // namespace

int hm::set_error(int code, const std::string &msg) { return (int)fail((hm_status)code, msg); }

extern "C" {

const char *hm_last_error(void) { return g_err.c_str(); }
const char *hm_version(void) { return "hypar-mi355x 0.1 (gfx950, fp64)"; }

void hm_destroy(hm_solver *s) {
  if (!s) return;
  hm::drop_bindings(s);
  double *bufs[] = {s->dxg, s->R, s->par, s->Q, s->S[0], s->S[1], s->acc, s->part, s->tmp};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  if (s->part_host) (void)hipHostFree(s->part_host);
  for (auto *v : {&s->ckpt, &s->stages, &s->vscratch, &s->vwork, &s->sscratch})
    for (double *b : *v) (void)hipFree(b);
  for (int i = 0; i < TAPE_SETS; i++)
    for (hipEvent_t e : {s->ev_ready[i], s->ev_done[i]})
      if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : {s->ev_start})
    if (e) (void)hipEventDestroy(e);
  if (s->aux) (void)hipStreamDestroy(s->aux);
  delete s;
}

hm_status hm_create(const hm_config *cfg, hm_solver **out) {
  if (!cfg || !out) return fail(HM_ERR_ARG, "hm_create: NULL argument");
  *out = nullptr;
  for (int d = 0; d < 3; d++)
    if (cfg->dim[d] < 5) return fail(HM_ERR_ARG, "hm_create: every grid dimension must be >= 5");
  if (cfg->interp != HM_INTERP_COMPONENTS && cfg->interp != HM_INTERP_CHARACTERISTIC)
    return fail(HM_ERR_UNSUPPORTED, "hm_create: unknown interpolation type");
  if (cfg->upwind != HM_UPWIND_ROE && cfg->upwind != HM_UPWIND_RUSANOV)
    return fail(HM_ERR_UNSUPPORTED, "hm_create: unknown upwinding scheme");
  if (!(cfg->gamma > 1.0)) return fail(HM_ERR_ARG, "hm_create: gamma must exceed 1");
  if (!(cfg->weno_eps > 0.0)) return fail(HM_ERR_ARG, "hm_create: weno_eps must be positive");
  if (cfg->weno_type < HM_WENO_JS || cfg->weno_type > HM_WENO_OPTIMAL)
    return fail(HM_ERR_UNSUPPORTED, "hm_create: unknown WENO weight formulation");
  if (cfg->Re > 0 && !(cfg->Minf > 0 && cfg->Pr > 0))
    return fail(HM_ERR_ARG, "hm_create: viscous runs need Minf > 0 and Pr > 0");

  hm_solver *s = new hm_solver();
  for (int d = 0; d < 3; d++) s->g.n[d] = cfg->dim[d];
  s->g.ncell = (size_t)cfg->dim[0] * cfg->dim[1] * cfg->dim[2];
  s->sc.interp = cfg->interp == HM_INTERP_CHARACTERISTIC ? hm::INTERP_CHAR : hm::INTERP_COMP;
  s->sc.upwind = cfg->upwind == HM_UPWIND_ROE ? hm::UPWIND_ROE : hm::UPWIND_RUSANOV;
  s->sc.gamma = cfg->gamma;
  s->sc.Re = cfg->Re > 0 ? cfg->Re / cfg->Minf : -1.0;
  s->sc.Pr = cfg->Pr;
  s->sc.eps = cfg->weno_eps;
  s->sc.weno = cfg->weno_type;
  if (hipGetDevice(&s->device) != hipSuccess) {
    hm_destroy(s);
    return fail(HM_ERR_HIP, "hm_create: no HIP device");
  }
  s->stream = nullptr;

  const int g = 3;
  size_t ndx = (size_t)(cfg->dim[0] + 2 * g) + (cfg->dim[1] + 2 * g) + (cfg->dim[2] + 2 * g);
  std::vector<double> dxh(ndx);
  if (cfg->dxinv) {
    std::memcpy(dxh.data(), cfg->dxinv, ndx * sizeof(double));
  } else {
    size_t off = 0;
    for (int d = 0; d < 3; d++) {
      uniform_dxinv(cfg->dim[d], g, dxh.data() + off);
      off += cfg->dim[d] + 2 * g;
    }
  }
  const size_t n5 = 5 * s->g.ncell;
  hm_status st = HM_OK;
  if ((st = dmalloc(&s->dxg, ndx)) != HM_OK) goto bad;
  if ((st = dmalloc(&s->R, n5)) != HM_OK) goto bad;
  if (s->sc.Re > 0) {
    if ((st = dmalloc(&s->par, 4 * s->g.ncell)) != HM_OK) goto bad;
    if ((st = dmalloc(&s->Q, 4 * s->g.ncell)) != HM_OK) goto bad;
  }
  if ((st = dmalloc(&s->S[0], n5)) != HM_OK) goto bad;
  if ((st = dmalloc(&s->S[1], n5)) != HM_OK) goto bad;
  if ((st = dmalloc(&s->acc, n5)) != HM_OK) goto bad;
  s->npart = hm::reduction_blocks(n5);
  if ((st = dmalloc(&s->part, s->npart)) != HM_OK) goto bad;
  if (hipHostMalloc((void **)&s->part_host, s->npart * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    st = fail(HM_ERR_NOMEM, "hm_create: pinned host allocation failed");
    goto bad;
  }
  if (hipMemcpy(s->dxg, dxh.data(), ndx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    st = fail(HM_ERR_HIP, "hm_create: copy of dxinv failed");
    goto bad;
  }
  *out = s;
  return HM_OK;
bad:
  hm_destroy(s);
  return st;
}

hm_status hm_set_stream(hm_solver *s, void *stream) {
  if (!s) return fail(HM_ERR_ARG, "hm_set_stream: solver is NULL");
  s->stream = (hipStream_t)stream;
  return HM_OK;
}

size_t hm_ncell(const hm_solver *s) { return s ? s->g.ncell : 0; }

hm_status hm_upload(hm_solver *s, double *dev, const double *host, int layout) {
  hm_status st = check_state(s, dev, "hm_upload: dev");
  if (st != HM_OK) return st;
  if (!host) return fail(HM_ERR_ARG, "hm_upload: host is NULL");
  const size_t n5 = 5 * s->g.ncell;
  if (layout == HM_LAYOUT_SOA_INTERIOR) {
    HM_HIP(hipMemcpyAsync(dev, host, n5 * sizeof(double), hipMemcpyHostToDevice, s->stream));
  } else if (layout == HM_LAYOUT_AOS_INTERIOR || layout == HM_LAYOUT_REFERENCE) {
    const int g = layout == HM_LAYOUT_REFERENCE ? 3 : 0;
    const size_t n = 5 * (size_t)(s->g.n[0] + 2 * g) * (s->g.n[1] + 2 * g) * (s->g.n[2] + 2 * g);
    if ((st = ensure_tmp(s, n)) != HM_OK) return st;
    HM_HIP(hipMemcpyAsync(s->tmp, host, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    hm::launch_aos_to_soa(s->g, s->tmp, dev, g, s->stream);
    if ((st = check_launch("hm_upload")) != HM_OK) return st;
  } else {
    return fail(HM_ERR_ARG, "hm_upload: unknown layout");
  }
  HM_HIP(hipStreamSynchronize(s->stream));
  return HM_OK;
}

hm_status hm_download(hm_solver *s, double *host, const double *dev, int layout) {
  hm_status st = check_state(s, dev, "hm_download: dev");
  if (st != HM_OK) return st;
  if (!host) return fail(HM_ERR_ARG, "hm_download: host is NULL");
  const size_t n5 = 5 * s->g.ncell;
  if (layout == HM_LAYOUT_SOA_INTERIOR) {
    HM_HIP(hipMemcpyAsync(host, dev, n5 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  } else if (layout == HM_LAYOUT_AOS_INTERIOR || layout == HM_LAYOUT_REFERENCE) {
    const int g = layout == HM_LAYOUT_REFERENCE ? 3 : 0;
    const size_t n = 5 * (size_t)(s->g.n[0] + 2 * g) * (s->g.n[1] + 2 * g) * (s->g.n[2] + 2 * g);
    if ((st = ensure_tmp(s, n)) != HM_OK) return st;
    // ghost entries of the reference layout come back as zeros
    HM_HIP(hipMemsetAsync(s->tmp, 0, n * sizeof(double), s->stream));
    hm::launch_soa_to_aos(s->g, dev, s->tmp, g, s->stream);
    if ((st = check_launch("hm_download")) != HM_OK) return st;
    HM_HIP(hipMemcpyAsync(host, s->tmp, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  } else {
    return fail(HM_ERR_ARG, "hm_download: unknown layout");
  }
  HM_HIP(hipStreamSynchronize(s->stream));
  return HM_OK;
}

hm_status hm_rk4(hm_solver *s, double *u, double dt, int nsteps) {
  hm_status st = check_state(s, u, "hm_rk4: u");
  if (st != HM_OK) return st;
  if (nsteps < 0) return fail(HM_ERR_ARG, "hm_rk4: nsteps < 0");
  if (!std::isfinite(dt)) return fail(HM_ERR_ARG, "hm_rk4: dt is not finite");
  if (nsteps == 0) return HM_OK;
  if (s->sc.Re > 0) hm::launch_prims(s->g, s->sc, u, s->Q, s->stream);
  double *const Sx[3] = {s->S[0], s->S[1], s->S[0]};
  for (int it = 0; it < nsteps; it++)
    if ((st = rk4_step(s, u, dt, Sx)) != HM_OK) return st;
  return HM_OK;
}

hm_status hm_ssprk3(hm_solver *s, double *u, double dt, int nsteps) {
  hm_status st = check_state(s, u, "hm_ssprk3: u");
  if (st != HM_OK) return st;
  if (nsteps < 0) return fail(HM_ERR_ARG, "hm_ssprk3: nsteps < 0");
  if (!std::isfinite(dt)) return fail(HM_ERR_ARG, "hm_ssprk3: dt is not finite");
  if (nsteps == 0) return HM_OK;
  if (s->sc.Re > 0) hm::launch_prims(s->g, s->sc, u, s->Q, s->stream);
  for (int it = 0; it < nsteps; it++)
    if ((st = ssprk3_step(s, u, dt)) != HM_OK) return st;
  return HM_OK;
}

// -hyp + par (OUT_RHS) or hyp (OUT_HYP) of u into out through the passes
static hm_status rhs_passes(hm_solver *s, double *out, const double *u, int mode, const char *name) {
  hm_status st = check_state(s, u, name);
  if (st != HM_OK) return st;
  if ((st = check_state(s, out, name)) != HM_OK) return st;
  if (out == u) return fail(HM_ERR_ARG, std::string(name) + ": the output must not alias u");
  if (s->sc.Re > 0 && mode == hm::OUT_RHS) {
    hm::launch_prims(s->g, s->sc, u, s->Q, s->stream);
    hm::launch_visc_par(s->g, s->sc, s->Q, s->par, s->dxg, s->stream);
  }
  hm::PassBuffers b;
  b.U = u;
  b.R = s->R;
  b.par = s->par;
  b.u = b.S = b.acc = b.Q = nullptr;
  b.out = out;
  for (int d = 0; d < 3; d++) hm::launch_pass(s->g, s->sc, d, b, mode, 0, 0.0, 0.0, s->dxg, s->stream);
  return check_launch(name);
}

hm_status hm_rhs(hm_solver *s, double *rhs, const double *u) {
  return rhs_passes(s, rhs, u, hm::OUT_RHS, "hm_rhs");
}
hm_status hm_hyperbolic(hm_solver *s, double *hyp, const double *u) {
  return rhs_passes(s, hyp, u, hm::OUT_HYP, "hm_hyperbolic");
}
hm_status hm_parabolic(hm_solver *s, double *par, const double *u) {
  hm_status st = check_state(s, u, "hm_parabolic: u");
  if (st != HM_OK) return st;
  if ((st = check_state(s, par, "hm_parabolic: par")) != HM_OK) return st;
  if (s->sc.Re > 0) {
    hm::launch_prims(s->g, s->sc, u, s->Q, s->stream);
    hm::launch_visc_par(s->g, s->sc, s->Q, s->par, s->dxg, s->stream);
    hm::launch_par_out(s->g, s->par, par, s->stream);
  } else {
    hm::launch_fill(par, 0.0, 5 * s->g.ncell, s->stream);
  }
  return check_launch("hm_parabolic");
}

hm_status hm_cfl(hm_solver *s, const double *u, double dt, double *cfl) {
  hm_status st = check_state(s, u, "hm_cfl: u");
  if (st != HM_OK) return st;
  if (!cfl) return fail(HM_ERR_ARG, "hm_cfl: cfl is NULL");
  unsigned nb = hm::reduction_blocks(s->g.ncell);
  hm::launch_cfl(s->g, s->sc, u, s->part, nb, s->dxg, dt, s->stream);
  if ((st = check_launch("hm_cfl")) != HM_OK) return st;
  HM_HIP(hipMemcpyAsync(s->part_host, s->part, nb * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HM_HIP(hipStreamSynchronize(s->stream));
  double m = 0.0;
  for (unsigned i = 0; i < nb; i++) m = s->part_host[i] > m ? s->part_host[i] : m;
  *cfl = m;
  if (!std::isfinite(m)) return fail(HM_ERR_NONFINITE, "hm_cfl: state holds NaN/Inf");
  return HM_OK;
}

hm_status hm_volume_integral(hm_solver *s, const double *u, double *out) {
  hm_status st = check_state(s, u, "hm_volume_integral: u");
  if (st != HM_OK) return st;
  if (!out) return fail(HM_ERR_ARG, "hm_volume_integral: out is NULL");
  unsigned nb = hm::reduction_blocks(s->g.ncell);
  if (5 * nb > s->npart) nb = s->npart / 5 > 0 ? s->npart / 5 : 1;
  hm::launch_volint(s->g, u, s->part, nb, s->dxg, s->stream);
  if ((st = check_launch("hm_volume_integral")) != HM_OK) return st;
  HM_HIP(hipMemcpyAsync(s->part_host, s->part, 5 * nb * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HM_HIP(hipStreamSynchronize(s->stream));
  for (int v = 0; v < 5; v++) {
    double sum = 0.0;
    for (unsigned b = 0; b < nb; b++) sum += s->part_host[(size_t)v * nb + b];
    out[v] = sum;
    if (!std::isfinite(sum)) return fail(HM_ERR_NONFINITE, "hm_volume_integral: state holds NaN/Inf");
  }
  return HM_OK;
}

hm_status hm_step_norm(hm_solver *s, const double *a, const double *b, double *norm) {
  hm_status st = check_state(s, a, "hm_step_norm: a");
  if (st != HM_OK) return st;
  if ((st = check_state(s, b, "hm_step_norm: b")) != HM_OK) return st;
  if (!norm) return fail(HM_ERR_ARG, "hm_step_norm: norm is NULL");
  const size_t n5 = 5 * s->g.ncell;
  unsigned nb = hm::reduction_blocks(n5);
  hm::launch_sqdiff(a, b, s->part, nb, n5, s->stream);
  if ((st = check_launch("hm_step_norm")) != HM_OK) return st;
  HM_HIP(hipMemcpyAsync(s->part_host, s->part, nb * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HM_HIP(hipStreamSynchronize(s->stream));
  double sum = 0.0;
  for (unsigned i = 0; i < nb; i++) sum += s->part_host[i];
  *norm = std::sqrt(sum / (double)s->g.ncell);
  if (!std::isfinite(*norm)) return fail(HM_ERR_NONFINITE, "hm_step_norm: NaN/Inf");
  return HM_OK;
}

hm_status hm_malloc(double **dev, size_t n) {
  if (!dev) return fail(HM_ERR_ARG, "hm_malloc: NULL output");
  return dmalloc(dev, n ? n : 1);
}

void hm_free(double *dev) {
  if (dev) (void)hipFree(dev);
}

hm_status hm_copy_to_device(double *dev, const double *host, size_t n) {
  if (!dev || !host) return fail(HM_ERR_ARG, "hm_copy_to_device: NULL pointer");
  HM_HIP(hipMemcpy(dev, host, n * sizeof(double), hipMemcpyHostToDevice));
  return HM_OK;
}

hm_status hm_copy_to_host(double *host, const double *dev, size_t n) {
  if (!dev || !host) return fail(HM_ERR_ARG, "hm_copy_to_host: NULL pointer");
  HM_HIP(hipMemcpy(host, dev, n * sizeof(double), hipMemcpyDeviceToHost));
  return HM_OK;
}

hm_status hm_copy_device(hm_solver *s, double *dst, const double *src, size_t n) {
  hm_status st = check_state(s, dst, "hm_copy_device: dst");
  if (st != HM_OK) return st;
  if ((st = check_state(s, src, "hm_copy_device: src")) != HM_OK) return st;
  HM_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  return HM_OK;
}

hm_status hm_synchronize(hm_solver *s) {
  if (!s) return fail(HM_ERR_ARG, "hm_synchronize: solver is NULL");
  HM_HIP(hipStreamSynchronize(s->stream));
  return HM_OK;
}

static hm_status check_dir(int dir, const char *name) {
  if (dir < 0 || dir > 2) return fail(HM_ERR_ARG, std::string(name) + ": dir must be 0, 1 or 2");
  return HM_OK;
}

hm_status hm_flux(hm_solver *s, double *f, const double *u, int dir) {
  hm_status st;
  if ((st = check_dir(dir, "hm_flux")) != HM_OK) return st;
  if ((st = check_state(s, f, "hm_flux: f")) != HM_OK) return st;
  if ((st = check_state(s, u, "hm_flux: u")) != HM_OK) return st;
  hm::launch_flux(s->g, s->sc, dir, u, f, s->stream);
  return check_launch("hm_flux");
}

hm_status hm_modified_solution(hm_solver *s, double *uC, const double *u) {
  hm_status st;
  if ((st = check_state(s, uC, "hm_modified_solution: uC")) != HM_OK) return st;
  if ((st = check_state(s, u, "hm_modified_solution: u")) != HM_OK) return st;
  hm::launch_modsol(s->g, s->sc, u, uC, s->stream);
  return check_launch("hm_modified_solution");
}

hm_status hm_interpolate(hm_solver *s, double *fI, const double *fC, const double *u, int upw,
                         int dir, int uflag) {
  hm_status st;
  if ((st = check_dir(dir, "hm_interpolate")) != HM_OK) return st;
  if (upw == 0) return fail(HM_ERR_ARG, "hm_interpolate: upw must be nonzero");
  if ((st = check_state(s, fI, "hm_interpolate: fI")) != HM_OK) return st;
  if ((st = check_state(s, fC, "hm_interpolate: fC")) != HM_OK) return st;
  if ((st = check_state(s, u, "hm_interpolate: u")) != HM_OK) return st;
  hm::launch_interp(s->g, s->sc, dir, fI, fC, u, upw, uflag ? 1 : 0, s->stream);
  return check_launch("hm_interpolate");
}

hm_status hm_upwind(hm_solver *s, double *fI, const double *fL, const double *fR,
                    const double *uL, const double *uR, const double *u, int dir) {
  hm_status st;
  if ((st = check_dir(dir, "hm_upwind")) != HM_OK) return st;
  const void *ptrs[] = {fI, fL, fR, uL, uR, u};
  for (const void *p : ptrs)
    if ((st = check_state(s, p, "hm_upwind: array")) != HM_OK) return st;
  hm::launch_upwind(s->g, s->sc, dir, fI, fL, fR, uL, uR, u, s->stream);
  return check_launch("hm_upwind");
}

hm_status hm_flux_difference(hm_solver *s, double *hyp, const double *fI, int dir, int accumulate) {
  hm_status st;
  if ((st = check_dir(dir, "hm_flux_difference")) != HM_OK) return st;
  if ((st = check_state(s, hyp, "hm_flux_difference: hyp")) != HM_OK) return st;
  if ((st = check_state(s, fI, "hm_flux_difference: fI")) != HM_OK) return st;
  hm::launch_fluxdiff(s->g, dir, hyp, fI, s->dxg, accumulate ? 1 : 0, s->stream);
  return check_launch("hm_flux_difference");
}

hm_status hm_axpy(hm_solver *s, const double *x, double a, double *y, size_t n) {
  hm_status st;
  if ((st = check_state(s, x, "hm_axpy: x")) != HM_OK) return st;
  if ((st = check_state(s, y, "hm_axpy: y")) != HM_OK) return st;
  hm::launch_axpy(x, a, y, n, s->stream);
  return check_launch("hm_axpy");
}

hm_status hm_fill(hm_solver *s, double *x, double value, size_t n) {
  hm_status st;
  if ((st = check_state(s, x, "hm_fill: x")) != HM_OK) return st;
  hm::launch_fill(x, value, n, s->stream);
  return check_launch("hm_fill");
}

// ---------------------------------------------------------------------------
// discrete adjoint
// ---------------------------------------------------------------------------

static hm_status ensure_buffers(std::vector<double *> &v, size_t count, size_t n) {
  while (v.size() < count) {
    double *p = nullptr;
    hm_status st = dmalloc(&p, n);
    if (st != HM_OK) return st;
    v.push_back(p);
  }
  return HM_OK;
}

// exactly `count` buffers of n doubles: frees the ones beyond count, so the
// tape holds what its schedule needs and no more
static hm_status fit_buffers(std::vector<double *> &v, size_t count, size_t n) {
  while (v.size() > count) {
    (void)hipFree(v.back());
    v.pop_back();
  }
  return ensure_buffers(v, count, n);
}

static size_t held_bytes(const std::vector<double *> &v, size_t n) { return v.size() * n * sizeof(double); }

// one RK4 step of u in place keeping the stage states U1, U2, U3
static hm_status rk4_step_store(hm_solver *s, double *u, double dt, double *const U[3]) {
  if (s->sc.Re > 0) hm::launch_prims(s->g, s->sc, u, s->Q, s->stream);
  return rk4_step(s, u, dt, U);
}

// Step adjoint given the stage states U[0..3] (U[0] = u_n):
// lam_out = lam + sum_s nu_s, nu_s = J(U_s)^T mu_s,
// mu_s = dt b_s lam + dt a_{s+1,s} nu_{s+1} (formed inside the VJP kernels);
// lam_out must differ from lam
static hm_status vjp_work(hm_solver *s, double **w) {
  const size_t n = hm::vjp_work_size(s->g, s->sc);
  *w = nullptr;
  if (n == 0) return HM_OK;
  hm_status st = ensure_buffers(s->vwork, 1, n);
  if (st == HM_OK) *w = s->vwork[0];
  return st;
}

static hm_status rk4_step_vjp(hm_solver *s, double *const U[4], double dt, const double *lam, double *lam_out) {
  const size_t n5 = 5 * s->g.ncell;
  hm_status st = ensure_buffers(s->vscratch, 3, n5);
  if (st != HM_OK) return st;
  double *work;
  if ((st = vjp_work(s, &work)) != HM_OK) return st;
  double *nuA = s->vscratch[0], *nuB = s->vscratch[1];
  const double *nu_in = nullptr;
  double *nu_out = nuA;
  for (int stage = 3; stage >= 0; stage--) {
    const double ca = stage < 3 ? dt * RK_A_SUB[stage] : 0.0, cb = dt * RK_B[stage];
    hm::launch_rhs_vjp(s->g, s->sc, U[stage], lam, nu_in, ca, cb, nu_out, lam_out, stage == 3, s->dxg, work,
                       s->stream);
    if ((st = check_launch("rk4_step_vjp")) != HM_OK) return st;
    nu_in = nu_out;
    nu_out = nu_out == nuA ? nuB : nuA;
  }
  return HM_OK;
}

static hm_status check_vjp(hm_solver *s, const char *name) {
  if (!s) return fail(HM_ERR_ARG, std::string(name) + ": solver is NULL");
  if (!hm::vjp_supported(s->g, s->sc))
    return fail(HM_ERR_UNSUPPORTED, std::string(name) + ": the adjoint holds grid lines of at most 512 cells");
  return HM_OK;
}

hm_status hm_rhs_vjp(hm_solver *s, double *out, const double *u, const double *lam) {
  hm_status st;
  if ((st = check_vjp(s, "hm_rhs_vjp")) != HM_OK) return st;
  if ((st = check_state(s, out, "hm_rhs_vjp: out")) != HM_OK) return st;
  if ((st = check_state(s, u, "hm_rhs_vjp: u")) != HM_OK) return st;
  if ((st = check_state(s, lam, "hm_rhs_vjp: lam")) != HM_OK) return st;
  if (out == u || out == lam) return fail(HM_ERR_ARG, "hm_rhs_vjp: out must not alias u or lam");
  double *work;
  if ((st = vjp_work(s, &work)) != HM_OK) return st;
  hm::launch_rhs_vjp(s->g, s->sc, u, lam, nullptr, 0.0, 1.0, out, nullptr, false, s->dxg, work, s->stream);
  return check_launch("hm_rhs_vjp");
}

hm_status hm_rk4_vjp(hm_solver *s, const double *u_n, double dt, double *lam) {
  hm_status st;
  if ((st = check_vjp(s, "hm_rk4_vjp")) != HM_OK) return st;
  if ((st = check_state(s, u_n, "hm_rk4_vjp: u_n")) != HM_OK) return st;
  if ((st = check_state(s, lam, "hm_rk4_vjp: lam")) != HM_OK) return st;
  const size_t n5 = 5 * s->g.ncell;
  if ((st = ensure_buffers(s->stages, 4, n5)) != HM_OK) return st;
  double *U[4] = {s->stages[0], s->stages[1], s->stages[2], s->stages[3]};
  HM_HIP(hipMemcpyAsync(U[0], u_n, n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  // a throw-away copy advanced through the step yields the stage states
  if ((st = ensure_buffers(s->vscratch, 4, n5)) != HM_OK) return st;
  double *w = s->vscratch[3];
  HM_HIP(hipMemcpyAsync(w, u_n, n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  double *const Us[3] = {U[1], U[2], U[3]};
  if ((st = rk4_step_store(s, w, dt, Us)) != HM_OK) return st;
  // U[0..3] are the first tape states: a recorded trajectory no longer finds
  // the segments there and recomputes them
  for (int i = 0; i < TAPE_SETS; i++) s->tape_holds[i] = -1;
  double *L = s->vscratch[2];
  if ((st = rk4_step_vjp(s, U, dt, lam, L)) != HM_OK) return st;
  HM_HIP(hipMemcpyAsync(lam, L, n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  return HM_OK;
}

// Segment seg's set is seg % sets; within it step m keeps u_n, U1, U2, U3
// at slots[4m .. 4m+3] (stage level) or u_n at slots[m] (step level).
static double **segment_slots(hm_solver *s, int seg, int K) {
  const TapePlan &p = s->tape_plan;
  return &s->stages[(size_t)(p.step_level ? 1 : 4) * K * (seg % p.sets)];
}

hm_status hm_set_tape_budget(hm_solver *s, size_t bytes) {
  if (!s) return fail(HM_ERR_ARG, "hm_set_tape_budget: solver is NULL");
  s->tape_budget = bytes;
  return HM_OK;
}

hm_status hm_tape_info(const hm_solver *s, int *sets, int *step_level, size_t *bytes) {
  if (!s) return fail(HM_ERR_ARG, "hm_tape_info: solver is NULL");
  if (s->tape_nsteps < 1) return fail(HM_ERR_ARG, "hm_tape_info: no forward sweep recorded");
  const size_t n5 = 5 * s->g.ncell;
  if (sets) *sets = s->tape_plan.sets;
  if (step_level) *step_level = s->tape_plan.step_level ? 1 : 0;
  if (bytes)
    *bytes = hm::vjp_work_size(s->g, s->sc) * sizeof(double) +
             tape_total_states(s->tape_plan, s->tape_nsteps, s->tape_every) * n5 * sizeof(double);
  return HM_OK;
}

hm_status hm_adjoint_forward(hm_solver *s, double *u, double dt, int nsteps, int every) {
  hm_status st;
  if ((st = check_vjp(s, "hm_adjoint_forward")) != HM_OK) return st;
  if ((st = check_state(s, u, "hm_adjoint_forward: u")) != HM_OK) return st;
  if (nsteps < 1 || every < 1) return fail(HM_ERR_ARG, "hm_adjoint_forward: nsteps and checkpoint_every must be >= 1");
  // a failed or partial sweep leaves no tape behind
  s->tape_nsteps = 0;
  for (int i = 0; i < TAPE_SETS; i++) s->tape_holds[i] = -1;
  const size_t n5 = 5 * s->g.ncell, state_bytes = n5 * sizeof(double);
  const size_t work_bytes = hm::vjp_work_size(s->g, s->sc) * sizeof(double);
  const size_t held = held_bytes(s->ckpt, n5) + held_bytes(s->stages, n5) + held_bytes(s->vscratch, n5) +
                      held_bytes(s->sscratch, n5) + (s->vwork.empty() ? 0 : work_bytes);
  const size_t budget = s->tape_budget ? s->tape_budget : auto_tape_budget(held);
  TapePlan plan;
  if (!choose_tape(budget, state_bytes, work_bytes, nsteps, every, &plan)) {
    TapePlan least;
    least.sets = 1;
    least.step_level = true;
    return fail(HM_ERR_NOMEM, "hm_adjoint_forward: the smallest tape (one checkpoint every " + std::to_string(every) +
                                  " steps, one segment of step states) needs " +
                                  std::to_string(work_bytes + tape_total_states(least, nsteps, every) * state_bytes) +
                                  " bytes; the budget is " + std::to_string(budget) +
                                  " (hm_set_tape_budget, or a longer checkpoint interval)");
  }
  s->tape_plan = plan;
  const int nck = (nsteps + every - 1) / every;
  if ((st = fit_buffers(s->ckpt, (size_t)nck, n5)) != HM_OK) return st;
  if ((st = fit_buffers(s->stages, tape_slots(plan, nsteps, every), n5)) != HM_OK) return st;
  if ((st = fit_buffers(s->sscratch, plan.step_level ? 3 : 0, n5)) != HM_OK) return st;
  if ((st = ensure_buffers(s->vscratch, 4, n5)) != HM_OK) return st;
  // the last `sets` segments keep their states, one in each set: the
  // reverse sweep starts on them without recomputation
  const int nkeep = (nck >= plan.sets ? nck - plan.sets : 0) * every;
  for (int n = 0; n < nsteps; n++) {
    if (n % every == 0)
      HM_HIP(hipMemcpyAsync(s->ckpt[n / every], u, n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    if (n < nkeep) {
      if ((st = hm_rk4(s, u, dt, 1)) != HM_OK) return st;
      continue;
    }
    const int seg = n / every, m = n - seg * every;
    double **slots = segment_slots(s, seg, every);
    if (plan.step_level) {
      HM_HIP(hipMemcpyAsync(slots[m], u, n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
      if ((st = hm_rk4(s, u, dt, 1)) != HM_OK) return st;
    } else {
      double **U = slots + (size_t)4 * m;
      HM_HIP(hipMemcpyAsync(U[0], u, n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
      double *const Us[3] = {U[1], U[2], U[3]};
      if ((st = rk4_step_store(s, u, dt, Us)) != HM_OK) return st;
    }
    s->tape_holds[seg % plan.sets] = seg;
  }
  s->tape_nsteps = nsteps;
  s->tape_every = every;
  s->tape_dt = dt;
  return HM_OK;
}

// recompute segment seg from its checkpoint, keeping every step's stage
// states (stage level) or starting states (step level)
static hm_status recompute_segment(hm_solver *s, int seg, int n0, int n1, double dt) {
  const size_t n5 = 5 * s->g.ncell;
  const int K = s->tape_every;
  double **S = segment_slots(s, seg, K);
  hm_status st;
  if (s->tape_plan.step_level) {
    HM_HIP(hipMemcpyAsync(S[0], s->ckpt[seg], n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    for (int m = 1; m < n1 - n0; m++) {
      HM_HIP(hipMemcpyAsync(S[m], S[m - 1], n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
      if ((st = hm_rk4(s, S[m], dt, 1)) != HM_OK) return st;
    }
    return HM_OK;
  }
  if ((st = ensure_buffers(s->vscratch, 4, n5)) != HM_OK) return st;
  for (int n = n0; n < n1; n++) {
    double **U = S + (size_t)4 * (n - n0);
    // U[0] of the first step is the checkpoint; later steps find u_n in
    // their slot, advanced there by the step before
    if (n == n0) HM_HIP(hipMemcpyAsync(U[0], s->ckpt[seg], n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    double *next = n + 1 < n1 ? S[(size_t)4 * (n - n0 + 1)] : s->vscratch[3];
    HM_HIP(hipMemcpyAsync(next, U[0], n5 * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    double *const Us[3] = {U[1], U[2], U[3]};
    if ((st = rk4_step_store(s, next, dt, Us)) != HM_OK) return st;
  }
  return HM_OK;
}

static hm_status adjoint_reverse_sweep(hm_solver *s, double *lam) {
  hm_status st;
  const int N = s->tape_nsteps, K = s->tape_every;
  const double dt = s->tape_dt;
  const TapePlan plan = s->tape_plan;
  const int sets = plan.sets;
  const size_t n5 = 5 * s->g.ncell;
  if ((st = ensure_buffers(s->stages, tape_slots(plan, N, K), n5)) != HM_OK) return st;
  if ((st = ensure_buffers(s->sscratch, plan.step_level ? 3 : 0, n5)) != HM_OK) return st;
  if ((st = ensure_buffers(s->vscratch, 4, n5)) != HM_OK) return st;
  // With two or more stage-level sets, a second stream recomputes the
  // segment before the current one while the solver's stream runs the
  // current one's adjoint.  With one set, or at step level (whose adjoint
  // runs forward steps on the solver's scratch), everything runs in order
  // on the solver's stream.
  const bool overlap = sets >= 2 && !plan.step_level;
  if (overlap) {
    if (!s->aux) HM_HIP(hipStreamCreate(&s->aux));  // default priority
    for (int i = 0; i < TAPE_SETS; i++)
      for (hipEvent_t *e : {&s->ev_ready[i], &s->ev_done[i]})
        if (!*e) HM_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    if (!s->ev_start) HM_HIP(hipEventCreateWithFlags(&s->ev_start, hipEventDisableTiming));
  }
  const int nck = (N + K - 1) / K;
  hipStream_t main = s->stream;
  double *cur = lam;
  if (overlap) {
    // the recomputation stream starts after everything queued so far
    HM_HIP(hipEventRecord(s->ev_start, main));
    HM_HIP(hipStreamWaitEvent(s->aux, s->ev_start, 0));
  }
  // After a forward sweep the sets hold the last `sets` segments; every
  // other segment, and any whose set was reused since (an earlier reverse
  // sweep, hm_rk4_vjp), is recomputed from its checkpoint.
  bool on_aux[TAPE_SETS] = {};
  for (int seg = nck - 1; seg >= 0; seg--) {
    const int n0 = seg * K, n1 = (n0 + K < N) ? n0 + K : N;
    const int o = seg % sets;
    if (s->tape_holds[o] != seg) {
      // the last segment, or (without overlap) any segment: on the solver's
      // stream, after the adjoint of the segment that used the set before
      s->tape_holds[o] = -1;
      if ((st = recompute_segment(s, seg, n0, n1, dt)) != HM_OK) return st;
      s->tape_holds[o] = seg;
      if (overlap) {
        // a recomputation writes the solver's step buffers (primitive
        // variables, viscous terms, right-hand side, accumulator, the last
        // step's scratch): the second stream's next one starts after this one
        HM_HIP(hipEventRecord(s->ev_start, main));
        HM_HIP(hipStreamWaitEvent(s->aux, s->ev_start, 0));
      }
    }
    if (overlap && seg > 0 && s->tape_holds[(seg - 1) % sets] != seg - 1) {
      // recompute the previous segment into its set on the second stream,
      // once the adjoint that last read that set (segment seg-1+sets) is done
      const int op = (seg - 1) % sets;
      s->tape_holds[op] = -1;
      if (seg - 1 + sets < nck) HM_HIP(hipStreamWaitEvent(s->aux, s->ev_done[op], 0));
      s->stream = s->aux;
      st = recompute_segment(s, seg - 1, n0 - K, n0, dt);
      s->stream = main;
      if (st != HM_OK) return st;
      HM_HIP(hipEventRecord(s->ev_ready[op], s->aux));
      s->tape_holds[op] = seg - 1;
      on_aux[op] = true;
    }
    if (on_aux[o]) {
      HM_HIP(hipStreamWaitEvent(main, s->ev_ready[o], 0));
      on_aux[o] = false;
    }
    double **S = segment_slots(s, seg, K);
    for (int n = n1 - 1; n >= n0; n--) {
      // the step adjoint alternates between the caller's array and scratch
      double *out = cur == lam ? s->vscratch[2] : lam;
      if (plan.step_level) {
        // the step's stage states, from its starting state
        double *w = s->vscratch[3];
        HM_HIP(hipMemcpyAsync(w, S[n - n0], n5 * sizeof(double), hipMemcpyDeviceToDevice, main));
        double *const Us[3] = {s->sscratch[0], s->sscratch[1], s->sscratch[2]};
        if ((st = rk4_step_store(s, w, dt, Us)) != HM_OK) return st;
        double *const U[4] = {S[n - n0], Us[0], Us[1], Us[2]};
        if ((st = rk4_step_vjp(s, U, dt, cur, out)) != HM_OK) return st;
      } else if ((st = rk4_step_vjp(s, S + (size_t)4 * (n - n0), dt, cur, out)) != HM_OK) {
        return st;
      }
      cur = out;
    }
    if (overlap) HM_HIP(hipEventRecord(s->ev_done[o], main));
  }
  if (cur != lam) HM_HIP(hipMemcpyAsync(lam, cur, n5 * sizeof(double), hipMemcpyDeviceToDevice, main));
  return HM_OK;
}

hm_status hm_adjoint_reverse(hm_solver *s, double *lam) {
  hm_status st;
  if ((st = check_vjp(s, "hm_adjoint_reverse")) != HM_OK) return st;
  if ((st = check_state(s, lam, "hm_adjoint_reverse: lam")) != HM_OK) return st;
  if (s->tape_nsteps < 1) return fail(HM_ERR_ARG, "hm_adjoint_reverse: no forward sweep recorded");
  hipStream_t main = s->stream;
  st = adjoint_reverse_sweep(s, lam);
  s->stream = main;
  if (st != HM_OK)
    for (int i = 0; i < TAPE_SETS; i++) s->tape_holds[i] = -1;  // contents unknown after a failure
  return st;
}

hm_status hm_profile_kernels(hm_solver *s, const double *u, int reps, double *ms) {
  hm_status st = check_state(s, u, "hm_profile_kernels: u");
  if (st != HM_OK) return st;
  if (!ms || reps < 1) return fail(HM_ERR_ARG, "hm_profile_kernels: bad arguments");
  hipEvent_t e0, e1;
  HM_HIP(hipEventCreate(&e0));
  HM_HIP(hipEventCreate(&e1));
  auto timed = [&](int slot, auto &&launch) -> hm_status {
    launch();  // warm
    if (hipEventRecord(e0, s->stream) != hipSuccess) return fail(HM_ERR_HIP, "event record");
    for (int r = 0; r < reps; r++) launch();
    if (hipEventRecord(e1, s->stream) != hipSuccess) return fail(HM_ERR_HIP, "event record");
    if (hipEventSynchronize(e1) != hipSuccess) return fail(HM_ERR_HIP, "event sync");
    float t = 0.f;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) return fail(HM_ERR_HIP, "elapsed time");
    ms[slot] = (double)t / reps;
    return check_launch("hm_profile_kernels");
  };
  const bool visc = s->sc.Re > 0;
  if (visc) {
    hm::launch_prims(s->g, s->sc, u, s->Q, s->stream);
    hm::launch_visc_par(s->g, s->sc, s->Q, s->par, s->dxg, s->stream);
  }
  // the passes as stage 1 of an RK step runs them: the z pass reads u and
  // acc and writes acc, the next stage state (and Q); u is not written
  HM_HIP(hipMemsetAsync(s->acc, 0, 5 * s->g.ncell * sizeof(double), s->stream));
  hm::PassBuffers b;
  b.U = u;
  b.R = s->R;
  b.par = s->par;
  b.u = const_cast<double *>(u);
  b.S = s->S[1];
  b.acc = s->acc;
  b.out = nullptr;
  b.Q = s->Q;  // primitive variables of the stage state (solver scratch)
  for (int d = 0; d < 3 && st == HM_OK; d++)
    st = timed(d, [&] { hm::launch_pass(s->g, s->sc, d, b, hm::OUT_STAGE, 1, 0.0, 0.0, s->dxg, s->stream); });
  if (st == HM_OK) {
    if (visc) {
      st = timed(3, [&] { hm::launch_visc_par(s->g, s->sc, s->Q, s->par, s->dxg, s->stream); });
      if (st == HM_OK) st = timed(4, [&] { hm::launch_prims(s->g, s->sc, u, s->Q, s->stream); });
    } else {
      ms[3] = ms[4] = 0.0;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}


// ---------------------------------------------------------------------------
// batched ensemble
// ---------------------------------------------------------------------------
}  // extern "C"

struct hm_ensemble {
  hm::Scheme sc;
  int device = 0;
  hipStream_t stream = nullptr;  // caller's stream (launches)
  hipStream_t cap = nullptr;     // private stream for graph capture
  int nmem = 0;
  std::vector<hm::EnsMember> mem;
  hm::EnsMember *mem_dev = nullptr;
  double *dxg = nullptr;
  size_t ncell = 0;
  double *R = nullptr, *S[2] = {nullptr, nullptr}, *acc = nullptr;
  // viscous members: viscous terms and primitive variables of all members,
  // each member's first tile of the viscous-term launch, and (adjoint) of
  // the viscous VJP launches
  double *par = nullptr, *Q = nullptr;
  unsigned *ptiles0 = nullptr, *tiles0 = nullptr;
  unsigned nptiles = 0, ntiles = 0;
  unsigned *item0[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  unsigned nitems[2][3] = {{0, 0, 0}, {0, 0, 0}};
  // x items of the one-launch-per-stage kernel (ENS_STAGE_XG line groups each)
  unsigned *xitem0[2] = {nullptr, nullptr};
  unsigned nxitems[2] = {0, 0};
  // adjoint: first line of every member along each direction, longest
  // lines, and the tape (checkpoints, stage states of one segment, scratch)
  unsigned *lines0[3] = {nullptr, nullptr, nullptr};
  unsigned nlines[3] = {0, 0, 0};
  // work items of the line VJPs (groups of lines, hm::ens_vjp_lines_per_item)
  unsigned *vitems0[3] = {nullptr, nullptr, nullptr};
  unsigned nvitems[3] = {0, 0, 0};
  int nmax[3] = {0, 0, 0};
  std::vector<double *> ckpt, stages, vscratch, vwork;
  int tape_nsteps = 0, tape_every = 0;
  std::vector<double> tape_dt;        // the members' dt of the recorded trajectory
  int tape_holds[TAPE_SETS] = {-1, -1, -1};  // segment held by each set (-1: none)
  TapePlan tape_plan;                 // schedule of the recorded trajectory
  size_t tape_budget = 0;             // bytes the tape may hold (0: from the device's free memory)
  std::vector<double *> sscratch;     // step level: the stage states of the step being taken back
  int shape = 0;
  unsigned grid[3] = {0, 0, 0};
  bool dt_set = false;
  // one-launch-per-stage shapes: dependency counters (4 + 2*nmem, zero
  // between launches), the wait bound, the CU count for the grid
  unsigned *ctr = nullptr;
  unsigned spin_limit = 1u << 21;
  int cus = 256;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const double *graph_u = nullptr;
  int graph_shape = -1;
  unsigned graph_grid[3] = {0, 0, 0};
  unsigned graph_spin = 0;
};

namespace {

void ens_drop_graph(hm_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->graph_u = nullptr;
}

// the launches of RK stage `stage` of all members on stage state U: the
// viscous terms of U (viscous members; the previous stage's z pass left its
// primitive variables in Q), then the three direction passes, as three
// launches (shapes 0, 1) or one (shapes 2, 3)
void ens_stage_launches(hm_ensemble *e, int stage, const double *U, double *u, double *Sn, hipStream_t st) {
  if (e->sc.Re > 0) hm::launch_ens_visc(e->sc, e->mem_dev, e->nmem, e->ptiles0, e->nptiles, st);
  const int ts = e->shape & 1;  // 0: 4 wavefronts per block, 1: 2
  if (e->shape >= 2) {
    const unsigned *const item0[3] = {e->xitem0[ts], e->item0[ts][1], e->item0[ts][2]};
    const unsigned nitems[3] = {e->nxitems[ts], e->nitems[ts][1], e->nitems[ts][2]};
    hm::launch_ens_stage(e->sc, ts == 0 ? 4 : 2, e->mem_dev, e->nmem, item0, nitems, e->item0[ts][0], hm::ENS_STAGE_XG,
                         e->grid[0], e->cus, e->ctr, e->spin_limit, U, e->R, u, Sn, e->acc, stage, st);
    return;
  }
  for (int d = 0; d < 3; d++)
    hm::launch_ens_pass(e->sc, d, ts, e->mem_dev, e->nmem, e->item0[ts][d], e->nitems[ts][d], e->grid[d], U, e->R, u,
                        Sn, e->acc, stage, st);
}

// the launches of one RK4 step of all members (stage s reads S[(s-1)%2],
// writes S[s%2]; the last stage writes u)
void ens_step_launches(hm_ensemble *e, double *u, hipStream_t st) {
  for (int stage = 0; stage < 4; stage++)
    ens_stage_launches(e, stage, stage == 0 ? u : e->S[(stage - 1) % 2], u, stage < 3 ? e->S[stage % 2] : nullptr, st);
}

// After launches of the one-launch-per-stage shapes: wait for the stream and
// report (and clear) the flag of a dependency wait that gave up.
hm_status ens_check_waits(hm_ensemble *e, const char *name) {
  if (e->shape < 2) return HM_OK;
  unsigned flag = 0;
  HM_HIP(hipMemcpyAsync(&flag, e->ctr + 2, sizeof(unsigned), hipMemcpyDeviceToHost, e->stream));
  HM_HIP(hipStreamSynchronize(e->stream));
  if (!flag) return HM_OK;
  const unsigned zero = 0;
  HM_HIP(hipMemcpy(e->ctr + 2, &zero, sizeof(unsigned), hipMemcpyHostToDevice));
  return fail(HM_ERR_TIMEOUT, std::string(name) + ": a dependency wait of the stage kernel gave up after " +
                                  std::to_string(e->spin_limit) + " polls; the state is not valid");
}

hm_status ens_capture(hm_ensemble *e, double *u) {
  ens_drop_graph(e);
  HM_HIP(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal));
  ens_step_launches(e, u, e->cap);
  hipError_t err = hipStreamEndCapture(e->cap, &e->graph);
  if (err != hipSuccess) return fail(HM_ERR_HIP, std::string("hm_ensemble: graph capture: ") + hipGetErrorString(err));
  HM_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  e->graph_u = u;
  e->graph_shape = e->shape;
  e->graph_spin = e->spin_limit;
  for (int d = 0; d < 3; d++) e->graph_grid[d] = e->grid[d];
  return HM_OK;
}

// one RK4 step of all members in place on u, keeping the stage states
// U1, U2, U3 (eager launches; the adjoint's forward and recomputation)
hm_status ens_step_store(hm_ensemble *e, double *u, double *const U[3]) {
  hipStream_t st = e->stream;
  if (e->sc.Re > 0) hm::launch_ens_prims(e->sc, e->mem_dev, e->nmem, u, e->ncell, st);
  for (int stage = 0; stage < 4; stage++)
    ens_stage_launches(e, stage, stage == 0 ? u : U[stage - 1], u, stage < 3 ? U[stage] : nullptr, st);
  return check_launch("hm_ensemble: step");
}

hm_status ens_check(const hm_ensemble *e, const void *u, const char *name) {
  if (!e) return fail(HM_ERR_ARG, std::string(name) + ": ensemble is NULL");
  if (!u) return fail(HM_ERR_ARG, std::string(name) + ": state is NULL");
  if (!is_device_ptr(u)) return fail(HM_ERR_ARG, std::string(name) + ": state is not a device pointer");
  return HM_OK;
}

}  // namespace

extern "C" {

void hm_ensemble_destroy(hm_ensemble *e) {
  if (!e) return;
  ens_drop_graph(e);
  if (e->cap) (void)hipStreamDestroy(e->cap);
  double *bufs[] = {e->dxg, e->R, e->S[0], e->S[1], e->acc, e->par, e->Q};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  if (e->ctr) (void)hipFree(e->ctr);
  if (e->mem_dev) (void)hipFree(e->mem_dev);
  if (e->tiles0) (void)hipFree(e->tiles0);
  if (e->ptiles0) (void)hipFree(e->ptiles0);
  for (int sh = 0; sh < 2; sh++) {
    for (int d = 0; d < 3; d++)
      if (e->item0[sh][d]) (void)hipFree(e->item0[sh][d]);
    if (e->xitem0[sh]) (void)hipFree(e->xitem0[sh]);
  }
  for (int d = 0; d < 3; d++)
    for (unsigned *p : {e->lines0[d], e->vitems0[d]})
      if (p) (void)hipFree(p);
  for (auto *v : {&e->ckpt, &e->stages, &e->vscratch, &e->vwork, &e->sscratch})
    for (double *b : *v) (void)hipFree(b);
  delete e;
}

hm_status hm_ensemble_create(const hm_config *members, int nmembers, hm_ensemble **out) {
  if (!members || !out || nmembers < 1) return fail(HM_ERR_ARG, "hm_ensemble_create: bad arguments");
  *out = nullptr;
  const hm_config &c0 = members[0];
  for (int m = 0; m < nmembers; m++) {
    const hm_config &c = members[m];
    for (int d = 0; d < 3; d++)
      if (c.dim[d] < 5) return fail(HM_ERR_ARG, "hm_ensemble_create: every grid dimension must be >= 5");
    if (c.interp != c0.interp || c.upwind != c0.upwind || c.gamma != c0.gamma || c.weno_eps != c0.weno_eps ||
        c.weno_type != c0.weno_type)
      return fail(HM_ERR_ARG,
                  "hm_ensemble_create: members must share interpolation, upwinding, WENO weights, gamma and eps");
    if ((c.Re > 0) != (c0.Re > 0) || (c0.Re > 0 && (c.Re != c0.Re || c.Minf != c0.Minf || c.Pr != c0.Pr)))
      return fail(HM_ERR_ARG, "hm_ensemble_create: members must share Re, Minf and Pr");
    if (c.Re > 0 && !(c.Minf > 0 && c.Pr > 0))
      return fail(HM_ERR_ARG, "hm_ensemble_create: viscous members need Minf > 0 and Pr > 0");
  }
  if (c0.interp != HM_INTERP_COMPONENTS && c0.interp != HM_INTERP_CHARACTERISTIC)
    return fail(HM_ERR_UNSUPPORTED, "hm_ensemble_create: unknown interpolation type");
  if (c0.upwind != HM_UPWIND_ROE && c0.upwind != HM_UPWIND_RUSANOV)
    return fail(HM_ERR_UNSUPPORTED, "hm_ensemble_create: unknown upwinding scheme");
  if (!(c0.gamma > 1.0)) return fail(HM_ERR_ARG, "hm_ensemble_create: gamma must exceed 1");
  if (!(c0.weno_eps > 0.0)) return fail(HM_ERR_ARG, "hm_ensemble_create: weno_eps must be positive");
  if (c0.weno_type < HM_WENO_JS || c0.weno_type > HM_WENO_OPTIMAL)
    return fail(HM_ERR_UNSUPPORTED, "hm_ensemble_create: unknown WENO weight formulation");

  hm_ensemble *e = new hm_ensemble();
  e->sc.interp = c0.interp == HM_INTERP_CHARACTERISTIC ? hm::INTERP_CHAR : hm::INTERP_COMP;
  e->sc.upwind = c0.upwind == HM_UPWIND_ROE ? hm::UPWIND_ROE : hm::UPWIND_RUSANOV;
  e->sc.gamma = c0.gamma;
  e->sc.Re = c0.Re > 0 ? c0.Re / c0.Minf : -1.0;
  e->sc.Pr = c0.Pr;
  e->sc.eps = c0.weno_eps;
  e->sc.weno = c0.weno_type;
  e->nmem = nmembers;
  hm_status st = HM_OK;
  std::vector<double> dxh;
  std::vector<size_t> dxoff(nmembers);
  e->mem.resize(nmembers);
  for (int m = 0; m < nmembers; m++) {
    const hm_config &c = members[m];
    hm::EnsMember &M = e->mem[m];
    for (int d = 0; d < 3; d++) M.n[d] = c.dim[d];
    M.ncell = (size_t)c.dim[0] * c.dim[1] * c.dim[2];
    M.off = e->ncell;
    M.dt = 0.0;
    M.par = M.q = nullptr;
    e->ncell += M.ncell;
    const size_t ndx = (size_t)(c.dim[0] + 6) + (c.dim[1] + 6) + (c.dim[2] + 6);
    dxoff[m] = dxh.size();
    dxh.resize(dxh.size() + ndx);
    if (c.dxinv) {
      std::memcpy(dxh.data() + dxoff[m], c.dxinv, ndx * sizeof(double));
    } else {
      size_t o = dxoff[m];
      for (int d = 0; d < 3; d++) {
        uniform_dxinv(c.dim[d], 3, dxh.data() + o);
        o += c.dim[d] + 6;
      }
    }
  }
  const size_t n5 = 5 * e->ncell;
  if (hipGetDevice(&e->device) != hipSuccess) {
    st = fail(HM_ERR_HIP, "hm_ensemble_create: no HIP device");
    goto bad;
  }
  if ((st = dmalloc(&e->dxg, dxh.size())) != HM_OK) goto bad;
  if ((st = dmalloc(&e->R, n5)) != HM_OK) goto bad;
  if ((st = dmalloc(&e->S[0], n5)) != HM_OK) goto bad;
  if ((st = dmalloc(&e->S[1], n5)) != HM_OK) goto bad;
  if ((st = dmalloc(&e->acc, n5)) != HM_OK) goto bad;
  if (hipMemcpy(e->dxg, dxh.data(), dxh.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    st = fail(HM_ERR_HIP, "hm_ensemble_create: copy of dxinv failed");
    goto bad;
  }
  for (int m = 0; m < nmembers; m++) e->mem[m].dxg = e->dxg + dxoff[m];
  if (e->sc.Re > 0) {
    if ((st = dmalloc(&e->par, 4 * e->ncell)) != HM_OK) goto bad;
    if ((st = dmalloc(&e->Q, 4 * e->ncell)) != HM_OK) goto bad;
    std::vector<unsigned> pre(nmembers + 1, 0);
    for (int m = 0; m < nmembers; m++) {
      hm::EnsMember &M = e->mem[m];
      M.par = e->par + 4 * M.off;
      M.q = e->Q + 4 * M.off;
      pre[m + 1] = pre[m] + hm::par_tiles(M.n[0], M.n[1], M.n[2]);
    }
    e->nptiles = pre[nmembers];
    if (hipMalloc((void **)&e->ptiles0, pre.size() * sizeof(unsigned)) != hipSuccess ||
        hipMemcpy(e->ptiles0, pre.data(), pre.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) {
      st = fail(HM_ERR_NOMEM, "hm_ensemble_create: viscous-term tile list");
      goto bad;
    }
    for (int m = 0; m < nmembers; m++) {
      const int *n = e->mem[m].n;
      pre[m + 1] = pre[m] + hm::visc_tiles(n[0], n[1], n[2]);
    }
    e->ntiles = pre[nmembers];
    if (hipMalloc((void **)&e->tiles0, pre.size() * sizeof(unsigned)) != hipSuccess ||
        hipMemcpy(e->tiles0, pre.data(), pre.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) {
      st = fail(HM_ERR_NOMEM, "hm_ensemble_create: viscous tile list");
      goto bad;
    }
  }
  if (hipMalloc((void **)&e->mem_dev, nmembers * sizeof(hm::EnsMember)) != hipSuccess) {
    st = fail(HM_ERR_NOMEM, "hm_ensemble_create: member table");
    goto bad;
  }
  if (hipMemcpy(e->mem_dev, e->mem.data(), nmembers * sizeof(hm::EnsMember), hipMemcpyHostToDevice) != hipSuccess) {
    st = fail(HM_ERR_HIP, "hm_ensemble_create: copy of the member table failed");
    goto bad;
  }
  for (int sh = 0; sh < 2; sh++) {
    for (int d = 0; d < 3; d++) {
      std::vector<unsigned> pre(nmembers + 1, 0);
      for (int m = 0; m < nmembers; m++) pre[m + 1] = pre[m] + hm::ens_items_per_member(d, sh, e->mem[m].n);
      e->nitems[sh][d] = pre[nmembers];
      if (hipMalloc((void **)&e->item0[sh][d], pre.size() * sizeof(unsigned)) != hipSuccess ||
          hipMemcpy(e->item0[sh][d], pre.data(), pre.size() * sizeof(unsigned), hipMemcpyHostToDevice) !=
              hipSuccess) {
        st = fail(HM_ERR_NOMEM, "hm_ensemble_create: work lists");
        goto bad;
      }
    }
    std::vector<unsigned> pre(nmembers + 1, 0);
    for (int m = 0; m < nmembers; m++) {
      const unsigned groups = hm::ens_items_per_member(0, sh, e->mem[m].n);
      pre[m + 1] = pre[m] + (groups + hm::ENS_STAGE_XG - 1) / hm::ENS_STAGE_XG;
    }
    e->nxitems[sh] = pre[nmembers];
    if (hipMalloc((void **)&e->xitem0[sh], pre.size() * sizeof(unsigned)) != hipSuccess ||
        hipMemcpy(e->xitem0[sh], pre.data(), pre.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) {
      st = fail(HM_ERR_NOMEM, "hm_ensemble_create: work lists");
      goto bad;
    }
  }
  for (int d = 0; d < 3; d++) {
    std::vector<unsigned> pre(nmembers + 1, 0);
    for (int m = 0; m < nmembers; m++) {
      pre[m + 1] = pre[m] + (unsigned)(e->mem[m].ncell / e->mem[m].n[d]);
      e->nmax[d] = std::max(e->nmax[d], e->mem[m].n[d]);
    }
    e->nlines[d] = pre[nmembers];
    std::vector<unsigned> vpre(nmembers + 1, 0);
    for (int m = 0; m < nmembers; m++) {
      const unsigned lpi = (unsigned)hm::ens_vjp_lines_per_item(e->nmax[d], e->mem[m].n[d]);
      vpre[m + 1] = vpre[m] + (pre[m + 1] - pre[m] + lpi - 1) / lpi;
    }
    e->nvitems[d] = vpre[nmembers];
    if (hipMalloc((void **)&e->lines0[d], pre.size() * sizeof(unsigned)) != hipSuccess ||
        hipMemcpy(e->lines0[d], pre.data(), pre.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc((void **)&e->vitems0[d], vpre.size() * sizeof(unsigned)) != hipSuccess ||
        hipMemcpy(e->vitems0[d], vpre.data(), vpre.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) {
      st = fail(HM_ERR_NOMEM, "hm_ensemble_create: line lists");
      goto bad;
    }
  }
  if (hipMalloc((void **)&e->ctr, (4 + 2 * (size_t)nmembers) * sizeof(unsigned)) != hipSuccess ||
      hipMemset(e->ctr, 0, (4 + 2 * (size_t)nmembers) * sizeof(unsigned)) != hipSuccess) {
    st = fail(HM_ERR_NOMEM, "hm_ensemble_create: stage counters");
    goto bad;
  }
  if (hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, e->device) != hipSuccess || e->cus < 1)
    e->cus = 256;
  if (hipStreamCreate(&e->cap) != hipSuccess) {
    st = fail(HM_ERR_HIP, "hm_ensemble_create: stream creation failed");
    goto bad;
  }
  *out = e;
  return HM_OK;
bad:
  hm_ensemble_destroy(e);
  return st;
}

size_t hm_ensemble_ncell(const hm_ensemble *e) { return e ? e->ncell : 0; }

hm_status hm_ensemble_member(const hm_ensemble *e, int m, size_t *offset, size_t *ncell) {
  if (!e || m < 0 || m >= e->nmem) return fail(HM_ERR_ARG, "hm_ensemble_member: bad member index");
  if (offset) *offset = e->mem[m].off;
  if (ncell) *ncell = e->mem[m].ncell;
  return HM_OK;
}

hm_status hm_ensemble_set_stream(hm_ensemble *e, void *stream) {
  if (!e) return fail(HM_ERR_ARG, "hm_ensemble_set_stream: ensemble is NULL");
  e->stream = (hipStream_t)stream;
  return HM_OK;
}

hm_status hm_ensemble_set_dt(hm_ensemble *e, const double *dt) {
  if (!e || !dt) return fail(HM_ERR_ARG, "hm_ensemble_set_dt: NULL argument");
  for (int m = 0; m < e->nmem; m++) {
    if (!std::isfinite(dt[m])) return fail(HM_ERR_ARG, "hm_ensemble_set_dt: dt is not finite");
    e->mem[m].dt = dt[m];
  }
  // ordered on the caller's stream before any later step
  HM_HIP(hipMemcpyAsync(e->mem_dev, e->mem.data(), e->nmem * sizeof(hm::EnsMember), hipMemcpyHostToDevice,
                        e->stream));
  HM_HIP(hipStreamSynchronize(e->stream));
  e->dt_set = true;
  return HM_OK;
}

hm_status hm_ensemble_set_tiles(hm_ensemble *e, int shape, const unsigned *grid) {
  if (!e || shape < 0 || shape > 3) return fail(HM_ERR_ARG, "hm_ensemble_set_tiles: shape must be 0, 1, 2 or 3");
  e->shape = shape;
  for (int d = 0; d < 3; d++) e->grid[d] = grid ? grid[d] : 0;
  return HM_OK;
}

hm_status hm_ensemble_get_tiles(const hm_ensemble *e, int *shape, unsigned *grid) {
  if (!e) return fail(HM_ERR_ARG, "hm_ensemble_get_tiles: ensemble is NULL");
  if (shape) *shape = e->shape;
  if (grid)
    for (int d = 0; d < 3; d++) grid[d] = e->grid[d];
  return HM_OK;
}

hm_status hm_ensemble_rk4(hm_ensemble *e, double *u, int nsteps) {
  hm_status st = ens_check(e, u, "hm_ensemble_rk4");
  if (st != HM_OK) return st;
  if (nsteps < 0) return fail(HM_ERR_ARG, "hm_ensemble_rk4: nsteps < 0");
  if (!e->dt_set) return fail(HM_ERR_ARG, "hm_ensemble_rk4: set the members' dt first (hm_ensemble_set_dt)");
  if (nsteps == 0) return HM_OK;
  bool stale = !e->exec || e->graph_u != u || e->graph_shape != e->shape || e->graph_spin != e->spin_limit;
  for (int d = 0; d < 3; d++) stale = stale || e->graph_grid[d] != e->grid[d];
  if (stale && (st = ens_capture(e, u)) != HM_OK) return st;
  // primitive variables of the current state for the first stage's viscous
  // fluxes; every later stage finds them written by the z pass before it
  if (e->sc.Re > 0) hm::launch_ens_prims(e->sc, e->mem_dev, e->nmem, u, e->ncell, e->stream);
  for (int it = 0; it < nsteps; it++) HM_HIP(hipGraphLaunch(e->exec, e->stream));
  return ens_check_waits(e, "hm_ensemble_rk4");
}

hm_status hm_ensemble_set_wait_limit(hm_ensemble *e, unsigned polls) {
  if (!e) return fail(HM_ERR_ARG, "hm_ensemble_set_wait_limit: ensemble is NULL");
  e->spin_limit = polls;
  return HM_OK;
}

hm_status hm_ensemble_tune(hm_ensemble *e, const double *u, int reps, int *shape, unsigned *grid) {
  hm_status st = ens_check(e, u, "hm_ensemble_tune");
  if (st != HM_OK) return st;
  if (!e->dt_set) return fail(HM_ERR_ARG, "hm_ensemble_tune: set the members' dt first (hm_ensemble_set_dt)");
  if (reps < 1) reps = 1;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device) != hipSuccess || cus < 1)
    cus = 256;
  const size_t n5 = 5 * e->ncell;
  double *w = nullptr;
  if ((st = dmalloc(&w, n5)) != HM_OK) return st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best = 1e300;
  int best_shape = 0;
  unsigned best_grid[3] = {0, 0, 0};
  const unsigned mults[4] = {0, 1, 2, 4};  // 0: one block per work item
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    st = fail(HM_ERR_HIP, "hm_ensemble_tune: event creation failed");
  }
  // shapes 0, 1: one launch per pass, four grid sizes each; shapes 2, 3:
  // one launch per stage with as many blocks as the GPU holds at once
  for (int sh = 0; sh < 4 && st == HM_OK; sh++)
    for (unsigned mi = 0; mi < (sh < 2 ? 4u : 1u) && st == HM_OK; mi++) {
      const unsigned per_cu = (sh & 1) == 0 ? 2 : 4;  // resident blocks per CU at 2 waves per SIMD
      e->shape = sh;
      for (int d = 0; d < 3; d++) e->grid[d] = sh < 2 ? mults[mi] * per_cu * (unsigned)cus : 0u;
      // candidates time `reps` graph replays of one step on a scratch copy
      if (hipMemcpyAsync(w, u, n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream) != hipSuccess) {
        st = fail(HM_ERR_HIP, "hm_ensemble_tune: copy failed");
        break;
      }
      if ((st = hm_ensemble_rk4(e, w, 1)) != HM_OK) break;  // capture + warm
      (void)hipEventRecord(e0, e->stream);
      if ((st = hm_ensemble_rk4(e, w, reps)) != HM_OK) break;
      (void)hipEventRecord(e1, e->stream);
      if (hipEventSynchronize(e1) != hipSuccess) {
        st = fail(HM_ERR_HIP, "hm_ensemble_tune: event sync failed");
        break;
      }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        best_shape = sh;
        for (int d = 0; d < 3; d++) best_grid[d] = e->grid[d];
      }
    }
  ens_drop_graph(e);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(w);
  e->shape = best_shape;
  for (int d = 0; d < 3; d++) e->grid[d] = best_grid[d];
  if (st != HM_OK) return st;
  if (shape) *shape = best_shape;
  if (grid)
    for (int d = 0; d < 3; d++) grid[d] = best_grid[d];
  return HM_OK;
}

// ---- batched discrete adjoint ----------------------------------------
static hm_status ens_check_vjp(hm_ensemble *e, const void *u, const char *name) {
  hm_status st = ens_check(e, u, name);
  if (st != HM_OK) return st;
  if (!hm::ens_vjp_supported(e->sc, e->nmax))
    return fail(HM_ERR_UNSUPPORTED, std::string(name) + ": the adjoint holds member lines of at most 512 cells");
  if (!e->dt_set) return fail(HM_ERR_ARG, std::string(name) + ": set the members' dt first (hm_ensemble_set_dt)");
  return HM_OK;
}

static hm_status ens_buffers(std::vector<double *> &v, size_t count, size_t n) {
  while (v.size() < count) {
    double *p = nullptr;
    hm_status st = dmalloc(&p, n);
    if (st != HM_OK) return st;
    v.push_back(p);
  }
  return HM_OK;
}

// Segment seg's set is seg % sets; within it step m keeps u_n, U1, U2, U3
// at slots[4m .. 4m+3] (stage level) or u_n at slots[m] (step level).
static double **ens_segment_slots(hm_ensemble *e, int seg, int K) {
  const TapePlan &p = e->tape_plan;
  return &e->stages[(size_t)(p.step_level ? 1 : 4) * K * (seg % p.sets)];
}

hm_status hm_ensemble_set_tape_budget(hm_ensemble *e, size_t bytes) {
  if (!e) return fail(HM_ERR_ARG, "hm_ensemble_set_tape_budget: ensemble is NULL");
  e->tape_budget = bytes;
  return HM_OK;
}

hm_status hm_ensemble_tape_info(const hm_ensemble *e, int *sets, int *step_level, size_t *bytes) {
  if (!e) return fail(HM_ERR_ARG, "hm_ensemble_tape_info: ensemble is NULL");
  if (e->tape_nsteps < 1) return fail(HM_ERR_ARG, "hm_ensemble_tape_info: no forward sweep recorded");
  const size_t n5 = 5 * e->ncell;
  if (sets) *sets = e->tape_plan.sets;
  if (step_level) *step_level = e->tape_plan.step_level ? 1 : 0;
  if (bytes)
    *bytes = hm::ens_vjp_work_size(e->sc, e->ncell) * sizeof(double) +
             tape_total_states(e->tape_plan, e->tape_nsteps, e->tape_every) * n5 * sizeof(double);
  return HM_OK;
}

hm_status hm_ensemble_adjoint_forward(hm_ensemble *e, double *u, int nsteps, int every) {
  hm_status st = ens_check_vjp(e, u, "hm_ensemble_adjoint_forward");
  if (st != HM_OK) return st;
  if (nsteps < 1 || every < 1)
    return fail(HM_ERR_ARG, "hm_ensemble_adjoint_forward: nsteps and checkpoint_every must be >= 1");
  e->tape_nsteps = 0;
  for (int i = 0; i < TAPE_SETS; i++) e->tape_holds[i] = -1;
  const size_t n5 = 5 * e->ncell, state_bytes = n5 * sizeof(double);
  const size_t work_bytes = hm::ens_vjp_work_size(e->sc, e->ncell) * sizeof(double);
  const size_t held = held_bytes(e->ckpt, n5) + held_bytes(e->stages, n5) + held_bytes(e->vscratch, n5) +
                      held_bytes(e->sscratch, n5) + (e->vwork.empty() ? 0 : work_bytes);
  const size_t budget = e->tape_budget ? e->tape_budget : auto_tape_budget(held);
  TapePlan plan;
  if (!choose_tape(budget, state_bytes, work_bytes, nsteps, every, &plan)) {
    TapePlan least;
    least.sets = 1;
    least.step_level = true;
    return fail(HM_ERR_NOMEM, "hm_ensemble_adjoint_forward: the smallest tape (one checkpoint every " +
                                  std::to_string(every) + " steps, one segment of step states) needs " +
                                  std::to_string(work_bytes + tape_total_states(least, nsteps, every) * state_bytes) +
                                  " bytes; the budget is " + std::to_string(budget) +
                                  " (hm_ensemble_set_tape_budget, or a longer checkpoint interval)");
  }
  e->tape_plan = plan;
  const int nck = (nsteps + every - 1) / every;
  if ((st = fit_buffers(e->ckpt, (size_t)nck, n5)) != HM_OK) return st;
  if ((st = fit_buffers(e->stages, tape_slots(plan, nsteps, every), n5)) != HM_OK) return st;
  if ((st = fit_buffers(e->sscratch, plan.step_level ? 3 : 0, n5)) != HM_OK) return st;
  if ((st = ens_buffers(e->vscratch, 4, n5)) != HM_OK) return st;
  // the last `sets` segments keep their states (segment seg in set seg % sets)
  const int nkeep = (nck >= plan.sets ? nck - plan.sets : 0) * every;
  for (int n = 0; n < nsteps; n++) {
    if (n % every == 0)
      HM_HIP(hipMemcpyAsync(e->ckpt[n / every], u, n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    if (n < nkeep) {
      if ((st = hm_ensemble_rk4(e, u, 1)) != HM_OK) return st;
      continue;
    }
    const int seg = n / every, m = n - seg * every;
    double **slots = ens_segment_slots(e, seg, every);
    if (plan.step_level) {
      HM_HIP(hipMemcpyAsync(slots[m], u, n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
      if ((st = hm_ensemble_rk4(e, u, 1)) != HM_OK) return st;
    } else {
      double **U = slots + (size_t)4 * m;
      HM_HIP(hipMemcpyAsync(U[0], u, n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
      double *const Us[3] = {U[1], U[2], U[3]};
      if ((st = ens_step_store(e, u, Us)) != HM_OK) return st;
    }
    e->tape_holds[seg % plan.sets] = seg;
  }
  e->tape_dt.resize(e->nmem);
  for (int m = 0; m < e->nmem; m++) e->tape_dt[m] = e->mem[m].dt;
  e->tape_nsteps = nsteps;
  e->tape_every = every;
  return ens_check_waits(e, "hm_ensemble_adjoint_forward");
}

static hm_status ens_reverse_sweep(hm_ensemble *e, double *lam) {
  hm_status st;
  const int N = e->tape_nsteps, K = e->tape_every;
  const TapePlan plan = e->tape_plan;
  const size_t n5 = 5 * e->ncell;
  if ((st = ens_buffers(e->stages, tape_slots(plan, N, K), n5)) != HM_OK) return st;
  if ((st = ens_buffers(e->sscratch, plan.step_level ? 3 : 0, n5)) != HM_OK) return st;
  if ((st = ens_buffers(e->vscratch, 4, n5)) != HM_OK) return st;
  double *work = nullptr;
  const size_t nw = hm::ens_vjp_work_size(e->sc, e->ncell);
  if (nw) {
    if ((st = ens_buffers(e->vwork, 1, nw)) != HM_OK) return st;
    work = e->vwork[0];
  }
  hm::EnsVjpPlan vp;
  vp.mem = e->mem_dev;
  vp.nmem = e->nmem;
  vp.ncell = e->ncell;
  for (int d = 0; d < 3; d++) {
    vp.lines0[d] = e->lines0[d];
    vp.nlines[d] = e->nlines[d];
    vp.items0[d] = e->vitems0[d];
    vp.nitems[d] = e->nvitems[d];
    vp.nmax[d] = e->nmax[d];
  }
  vp.tiles0 = e->tiles0;
  vp.ntiles = e->ntiles;
  const int nck = (N + K - 1) / K;
  double *cur = lam;
  for (int seg = nck - 1; seg >= 0; seg--) {
    const int n0 = seg * K, n1 = (n0 + K < N) ? n0 + K : N;
    const int o = seg % plan.sets;
    double **S = ens_segment_slots(e, seg, K);
    if (e->tape_holds[o] != seg) {
      // recompute the segment from its checkpoint into its set
      e->tape_holds[o] = -1;
      if (plan.step_level) {
        HM_HIP(hipMemcpyAsync(S[0], e->ckpt[seg], n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
        for (int m = 1; m < n1 - n0; m++) {
          HM_HIP(hipMemcpyAsync(S[m], S[m - 1], n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
          double *const Us[3] = {e->sscratch[0], e->sscratch[1], e->sscratch[2]};
          if ((st = ens_step_store(e, S[m], Us)) != HM_OK) return st;
        }
      } else {
        for (int n = n0; n < n1; n++) {
          double **U = S + (size_t)4 * (n - n0);
          if (n == n0)
            HM_HIP(hipMemcpyAsync(U[0], e->ckpt[seg], n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
          double *next = n + 1 < n1 ? S[(size_t)4 * (n - n0 + 1)] : e->vscratch[3];
          HM_HIP(hipMemcpyAsync(next, U[0], n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
          double *const Us[3] = {U[1], U[2], U[3]};
          if ((st = ens_step_store(e, next, Us)) != HM_OK) return st;
        }
      }
      e->tape_holds[o] = seg;
    }
    for (int n = n1 - 1; n >= n0; n--) {
      double *Ust[4];
      if (plan.step_level) {
        // the step's stage states, from its starting state
        double *w = e->vscratch[3];
        HM_HIP(hipMemcpyAsync(w, S[n - n0], n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
        double *const Us[3] = {e->sscratch[0], e->sscratch[1], e->sscratch[2]};
        if ((st = ens_step_store(e, w, Us)) != HM_OK) return st;
        Ust[0] = S[n - n0];
        for (int i = 0; i < 3; i++) Ust[i + 1] = Us[i];
      } else {
        for (int i = 0; i < 4; i++) Ust[i] = S[(size_t)4 * (n - n0) + i];
      }
      double *out = cur == lam ? e->vscratch[2] : lam;
      // step adjoint: out = cur + sum_s nu_s (stage coefficients formed per member)
      const double *nu_in = nullptr;
      double *nu_out = e->vscratch[0];
      for (int stage = 3; stage >= 0; stage--) {
        const double A = stage < 3 ? RK_A_SUB[stage] : 0.0, B = RK_B[stage];
        hm::launch_ens_rhs_vjp(e->sc, vp, Ust[stage], cur, nu_in, A, B, nu_out, out, stage == 3, work, e->stream);
        if ((st = check_launch("hm_ensemble_adjoint_reverse")) != HM_OK) return st;
        nu_in = nu_out;
        nu_out = nu_out == e->vscratch[0] ? e->vscratch[1] : e->vscratch[0];
      }
      cur = out;
    }
  }
  if (cur != lam) HM_HIP(hipMemcpyAsync(lam, cur, n5 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
  return ens_check_waits(e, "hm_ensemble_adjoint_reverse");
}

hm_status hm_ensemble_adjoint_reverse(hm_ensemble *e, double *lam) {
  hm_status st = ens_check_vjp(e, lam, "hm_ensemble_adjoint_reverse");
  if (st != HM_OK) return st;
  if (e->tape_nsteps < 1) return fail(HM_ERR_ARG, "hm_ensemble_adjoint_reverse: no forward sweep recorded");
  // the sweep differentiates the recorded trajectory: its recomputed steps
  // and VJP kernels read each member's dt from the device member table, so
  // a dt set after the forward sweep is swapped out for the sweep
  bool swap = false;
  for (int m = 0; m < e->nmem; m++) swap = swap || e->mem[m].dt != e->tape_dt[m];
  std::vector<hm::EnsMember> saved;
  if (swap) {
    saved = e->mem;
    for (int m = 0; m < e->nmem; m++) e->mem[m].dt = e->tape_dt[m];
    HM_HIP(hipMemcpyAsync(e->mem_dev, e->mem.data(), e->nmem * sizeof(hm::EnsMember), hipMemcpyHostToDevice,
                          e->stream));
    HM_HIP(hipStreamSynchronize(e->stream));
  }
  st = ens_reverse_sweep(e, lam);
  if (st != HM_OK)
    for (int i = 0; i < TAPE_SETS; i++) e->tape_holds[i] = -1;
  if (swap) {
    e->mem = saved;
    HM_HIP(hipMemcpyAsync(e->mem_dev, e->mem.data(), e->nmem * sizeof(hm::EnsMember), hipMemcpyHostToDevice,
                          e->stream));
    HM_HIP(hipStreamSynchronize(e->stream));
  }
  return st;
}

}  // extern "C"
