// hm_noalloc.hip -- hm_cfl_noalloc: the CFL number of hm_cfl without a
// solver object and without any device allocation.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The grid
// metrics and the reduction partials live in `scratch`, a file-scope device
// array of HM_NOALLOC_SCRATCH_DOUBLES doubles that every GPU the library runs
// on holds until the process ends.  Host code never passes `scratch` itself
// to a kernel or a copy: the first call on a device looks up that device's
// address of the array with hipGetSymbolAddress (with the device current)
// and keeps it per device.
//
// Every call on one GPU shares that GPU's array, so calls that overlap on
// one GPU (from several host threads or streams) race; the array keeps the
// last call's contents.  Nothing in it needs to start at zero: each call
// writes the metrics, every partial it reads back and the result before
// reading them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/hypar_mi355x.h"
#include "hm_internal.h"

namespace {

constexpr int NV = 5;
constexpr int GHOSTS = 3;
// scratch layout: ghosted dxinv x|y|z, then the CFL partials, then the result
constexpr size_t DX_OFF = 0;
constexpr size_t DX_MAX = 3 * (size_t)(HM_NOALLOC_MAX_SIDE + 2 * GHOSTS);
constexpr size_t PART_OFF = DX_OFF + DX_MAX;
constexpr size_t PART_MAX = 2048;  // hm::reduction_blocks never exceeds this
constexpr size_t RES_OFF = PART_OFF + PART_MAX;
static_assert(RES_OFF + 1 <= HM_NOALLOC_SCRATCH_DOUBLES, "hm_cfl_noalloc scratch too small");

__device__ __align__(256) double scratch[HM_NOALLOC_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
double *g_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per
// device, with that device current).
hm_status scratch_ptr(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0)
    return (hm_status)hm::set_error(HM_ERR_HIP, "hm_cfl_noalloc: no HIP device");
  if (dev >= MAX_DEVICES) return (hm_status)hm::set_error(HM_ERR_UNSUPPORTED, "hm_cfl_noalloc: device index >= 64");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return (hm_status)hm::set_error(HM_ERR_HIP, std::string("hm_cfl_noalloc: hipGetSymbolAddress: ") +
                                                      hipGetErrorString(e));
    g_ptr[dev] = (double *)p;
  }
  *out = g_ptr[dev];
  return HM_OK;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// max of the nb partials into *out (one block of 256 threads)
__global__ void __launch_bounds__(256) k_max_final(const double *__restrict__ part, unsigned nb,
                                                   double *__restrict__ out) {
  __shared__ double red[4];
  double m = 0.0;
  for (unsigned i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, part[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

bool device_ptr(const void *p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace

extern "C" hm_status hm_cfl_noalloc(const hm_config *cfg, const double *u, double dt, double *cfl, void *stream) {
  if (!cfg || !u || !cfl) return (hm_status)hm::set_error(HM_ERR_ARG, "hm_cfl_noalloc: NULL argument");
  for (int d = 0; d < 3; d++) {
    if (cfg->dim[d] < 5) return (hm_status)hm::set_error(HM_ERR_ARG, "hm_cfl_noalloc: every grid dimension must be >= 5");
    if (cfg->dim[d] > HM_NOALLOC_MAX_SIDE)
      return (hm_status)hm::set_error(HM_ERR_ARG, "hm_cfl_noalloc: grid side " + std::to_string(cfg->dim[d]) +
                                                      " exceeds HM_NOALLOC_MAX_SIDE (" +
                                                      std::to_string(HM_NOALLOC_MAX_SIDE) + ")");
  }
  if (!(cfg->gamma > 1.0)) return (hm_status)hm::set_error(HM_ERR_ARG, "hm_cfl_noalloc: gamma must exceed 1");
  if (!std::isfinite(dt)) return (hm_status)hm::set_error(HM_ERR_ARG, "hm_cfl_noalloc: dt is not finite");
  if (!device_ptr(u)) return (hm_status)hm::set_error(HM_ERR_ARG, "hm_cfl_noalloc: u is not a device pointer");
  double *ptr = nullptr;
  hm_status st = scratch_ptr(&ptr);
  if (st != HM_OK) return st;
  hipStream_t s = (hipStream_t)stream;

  hm::Grid g;
  for (int d = 0; d < 3; d++) g.n[d] = cfg->dim[d];
  g.ncell = (size_t)g.n[0] * g.n[1] * g.n[2];
  hm::Scheme sc{};
  sc.gamma = cfg->gamma;

  // metrics: the caller's HyPar::dxinv, or those of the uniform periodic
  // grid, computed on the host into stack buffers (no allocation)
  double X[HM_NOALLOC_MAX_SIDE + 2 * GHOSTS], dx[3][HM_NOALLOC_MAX_SIDE + 2 * GHOSTS];
  size_t off = 0;
  for (int d = 0; d < 3; d++) {
    const size_t nd = (size_t)g.n[d] + 2 * GHOSTS;
    const double *src = cfg->dxinv ? cfg->dxinv + off : dx[d];
    if (!cfg->dxinv) hm::uniform_dxinv(g.n[d], GHOSTS, X, dx[d]);
    hipError_t e = hipMemcpyAsync(ptr + DX_OFF + off, src, nd * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess)
      return (hm_status)hm::set_error(HM_ERR_HIP, std::string("hm_cfl_noalloc: metrics copy: ") + hipGetErrorString(e));
    off += nd;
  }
  const unsigned nb = hm::reduction_blocks(g.ncell);
  if (nb > PART_MAX) return (hm_status)hm::set_error(HM_ERR_ARG, "hm_cfl_noalloc: too many reduction blocks");
  hm::launch_cfl(g, sc, u, ptr + PART_OFF, nb, ptr + DX_OFF, dt, s);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, ptr + PART_OFF, nb, ptr + RES_OFF);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return (hm_status)hm::set_error(HM_ERR_HIP, std::string("hm_cfl_noalloc: launch: ") + hipGetErrorString(e));
  double m = 0.0;
  e = hipMemcpyAsync(&m, ptr + RES_OFF, sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return (hm_status)hm::set_error(HM_ERR_HIP, std::string("hm_cfl_noalloc: ") + hipGetErrorString(e));
  *cfl = m;
  if (!std::isfinite(m)) return (hm_status)hm::set_error(HM_ERR_NONFINITE, "hm_cfl_noalloc: state holds NaN/Inf");
  return HM_OK;
}
