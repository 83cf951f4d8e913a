// hm_bind.hip -- HyPar's function-pointer slots with their own signatures
// (include/hypar_mi355x_slots.h).
//
// HyPar hands every slot its solver object as `void *s` and reads its fields
// by casting (e.g. TimeRHSFunctionExplicit.c:38, NavierStokes3DComputeCFL.c:
// 25-30).  The library cannot see HyPar's struct, so hm_bind and
// hm_bind_state record, per object pointer, what a slot would read there:
// the device solver, the state array (solver->u) and how the slots' arrays
// are laid out.  hm_bind_time records the members a time-integration object
// advances (the member loop of TimeRK.c:130-193).  Host code only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hypar_mi355x_slots.h"
#include "hm_internal.h"

namespace {

struct Binding {
  hm_solver *s = nullptr;
  double *u = nullptr;  // solver->u
  int arrays = HM_SLOT_ARRAYS_DEVICE;
  double *stage[2] = {nullptr, nullptr};  // device staging for host arrays
};

struct Member {
  void *obj;
  double dt;
  int scheme;
};

std::mutex g_mu;
std::unordered_map<void *, Binding> g_obj;
std::unordered_map<void *, std::vector<Member>> g_time;

int fail(int code, const std::string &msg) { return hm::set_error(code, msg); }

std::string ptr_str(const void *p) {
  char b[32];
  snprintf(b, sizeof(b), "%p", p);
  return b;
}

void free_stage(Binding &b) {
  for (double *&p : b.stage) {
    hm_free(p);
    p = nullptr;
  }
}

// Copy of the binding of `obj`, with the staging arrays allocated when its
// arrays are host arrays.  Returns 0 or an hm_status.
int lookup(void *obj, const char *slot, Binding &out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_obj.find(obj);
  if (!obj || it == g_obj.end() || !it->second.s)
    return fail(HM_ERR_ARG, std::string(slot) + ": solver object " + ptr_str(obj) + " is not bound (hm_bind)");
  Binding &b = it->second;
  if (b.arrays == HM_SLOT_ARRAYS_REFERENCE_HOST) {
    const size_t n = 5 * hm_ncell(b.s);
    for (double *&p : b.stage)
      if (!p) {
        const hm_status st = hm_malloc(&p, n);
        if (st != HM_OK) return st;
      }
  }
  out = b;
  return HM_OK;
}

// A term slot: out = op(u) on the bound solver, staged through the device
// for host arrays.
template <class Op>
int term_slot(const char *slot, double *out, double *u, void *obj, Op op) {
  Binding b;
  int st = lookup(obj, slot, b);
  if (st != HM_OK) return st;
  if (!out || !u) return fail(HM_ERR_ARG, std::string(slot) + ": NULL array");
  if (b.arrays == HM_SLOT_ARRAYS_DEVICE) return op(b.s, out, u);
  if ((st = hm_upload(b.s, b.stage[0], u, HM_LAYOUT_REFERENCE)) != HM_OK) return st;
  if ((st = op(b.s, b.stage[1], b.stage[0])) != HM_OK) return st;
  return hm_download(b.s, out, b.stage[1], HM_LAYOUT_REFERENCE);
}

hm_status step(hm_solver *s, double *u, double dt, int scheme) {
  return scheme == HM_TIME_SSPRK3 ? hm_ssprk3(s, u, dt, 1) : hm_rk4(s, u, dt, 1);
}

}  // namespace

namespace hm {
// called by hm_destroy: no binding may outlive its solver
void drop_bindings(const hm_solver *s) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto it = g_obj.begin(); it != g_obj.end();) {
    if (it->second.s == s) {
      free_stage(it->second);
      it = g_obj.erase(it);
    } else {
      ++it;
    }
  }
}
}  // namespace hm

extern "C" {

hm_status hm_bind(void *hypar, hm_solver *s) {
  if (!hypar || !s) return (hm_status)fail(HM_ERR_ARG, "hm_bind: NULL solver object or solver");
  std::lock_guard<std::mutex> lk(g_mu);
  Binding &b = g_obj[hypar];
  if (b.s != s) free_stage(b);  // staging is sized for the old solver
  b.s = s;
  return HM_OK;
}

hm_status hm_bind_state(void *hypar, double *u, int arrays) {
  if (arrays != HM_SLOT_ARRAYS_DEVICE && arrays != HM_SLOT_ARRAYS_REFERENCE_HOST)
    return (hm_status)fail(HM_ERR_ARG, "hm_bind_state: arrays must be HM_SLOT_ARRAYS_DEVICE or "
                                       "HM_SLOT_ARRAYS_REFERENCE_HOST");
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_obj.find(hypar);
  if (!hypar || it == g_obj.end())
    return (hm_status)fail(HM_ERR_ARG, "hm_bind_state: solver object " + ptr_str(hypar) + " is not bound (hm_bind)");
  it->second.u = u;
  it->second.arrays = arrays;
  return HM_OK;
}

hm_status hm_bind_time(void *ts, void *hypar, double dt, int scheme) {
  if (!ts) return (hm_status)fail(HM_ERR_ARG, "hm_bind_time: NULL time-integration object");
  if (scheme != HM_TIME_RK4 && scheme != HM_TIME_SSPRK3)
    return (hm_status)fail(HM_ERR_UNSUPPORTED, "hm_bind_time: scheme must be HM_TIME_RK4 or HM_TIME_SSPRK3");
  if (!(dt > 0.0) || !std::isfinite(dt)) return (hm_status)fail(HM_ERR_ARG, "hm_bind_time: dt must be positive");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!hypar || g_obj.find(hypar) == g_obj.end())
    return (hm_status)fail(HM_ERR_ARG, "hm_bind_time: solver object " + ptr_str(hypar) + " is not bound (hm_bind)");
  std::vector<Member> &m = g_time[ts];
  auto it = std::find_if(m.begin(), m.end(), [&](const Member &x) { return x.obj == hypar; });
  if (it != m.end()) {
    it->dt = dt;
    it->scheme = scheme;
  } else {
    m.push_back({hypar, dt, scheme});
  }
  return HM_OK;
}

void hm_unbind(void *obj) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_obj.find(obj);
  if (it != g_obj.end()) {
    free_stage(it->second);
    g_obj.erase(it);
  }
  g_time.erase(obj);
}

int hm_slot_TimeIntegrate(void *ts) {
  std::vector<Member> members;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_time.find(ts);
    if (!ts || it == g_time.end() || it->second.empty())
      return fail(HM_ERR_ARG, "TimeIntegrate: time-integration object " + ptr_str(ts) +
                                  " has no members (hm_bind_time)");
    members = it->second;
  }
  for (const Member &m : members) {
    Binding b;
    int st = lookup(m.obj, "TimeIntegrate", b);
    if (st != HM_OK) return st;
    if (!b.u) return fail(HM_ERR_ARG, "TimeIntegrate: member " + ptr_str(m.obj) + " has no state (hm_bind_state)");
    if (b.arrays == HM_SLOT_ARRAYS_DEVICE) {
      if ((st = step(b.s, b.u, m.dt, m.scheme)) != HM_OK) return st;
      continue;
    }
    if ((st = hm_upload(b.s, b.stage[0], b.u, HM_LAYOUT_REFERENCE)) != HM_OK) return st;
    if ((st = step(b.s, b.stage[0], m.dt, m.scheme)) != HM_OK) return st;
    if ((st = hm_download(b.s, b.u, b.stage[0], HM_LAYOUT_REFERENCE)) != HM_OK) return st;
  }
  return HM_OK;
}

int hm_slot_RHSFunction(double *rhs, double *u, void *s, void *m, double t) {
  (void)m;
  (void)t;
  return term_slot("RHSFunction", rhs, u, s, [](hm_solver *h, double *o, double *x) { return hm_rhs(h, o, x); });
}

int hm_slot_HyperbolicFunction(double *hyp, double *u, void *s, void *m, double t, int LimFlag,
                               int (*FFunction)(double *, double *, int, void *, double),
                               int (*UpwindFunction)(double *, double *, double *, double *, double *, double *,
                                                     int, void *, double)) {
  (void)m;
  (void)t;
  (void)UpwindFunction;
  if (!FFunction)  // zero hyperbolic term (HyperbolicFunction.c:65-67)
    return term_slot("HyperbolicFunction", hyp, u, s, [](hm_solver *h, double *o, double *) {
      return hm_fill(h, o, 0.0, 5 * hm_ncell(h));
    });
  if (LimFlag != 1)
    return fail(HM_ERR_UNSUPPORTED, "HyperbolicFunction: LimFlag must be 1 (the weights are evaluated from u at "
                                    "every call)");
  return term_slot("HyperbolicFunction", hyp, u, s,
                   [](hm_solver *h, double *o, double *x) { return hm_hyperbolic(h, o, x); });
}

int hm_slot_ParabolicFunction(double *par, double *u, void *s, void *m, double t) {
  (void)m;
  (void)t;
  return term_slot("ParabolicFunction", par, u, s,
                   [](hm_solver *h, double *o, double *x) { return hm_parabolic(h, o, x); });
}

double hm_slot_ComputeCFL(void *s, void *m, double dt, double t) {
  (void)m;
  (void)t;
  Binding b;
  if (lookup(s, "ComputeCFL", b) != HM_OK) return -1.0;
  if (!b.u) {
    fail(HM_ERR_ARG, "ComputeCFL: solver object " + ptr_str(s) + " has no state (hm_bind_state)");
    return -1.0;
  }
  const double *u = b.u;
  if (b.arrays == HM_SLOT_ARRAYS_REFERENCE_HOST) {
    if (hm_upload(b.s, b.stage[0], b.u, HM_LAYOUT_REFERENCE) != HM_OK) return -1.0;
    u = b.stage[0];
  }
  double cfl = -1.0;
  if (hm_cfl(b.s, u, dt, &cfl) != HM_OK) return -1.0;
  return cfl;
}

}  // extern "C"
