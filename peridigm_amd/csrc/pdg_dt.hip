// Stable explicit time step on the device
// (src/core/Peridigm_CriticalTimeStep.cpp:54-126, constant horizon).
//
// Per node i: den_i = sum_j (V_j * spring) / |x_j - x_i| with
// spring = 18 K / (pi delta^4), and dt_i = sqrt(2 rho / den_i) (1e50 for a
// node without bonds); the result is min_i dt_i. One thread per node sums its
// CSR row in FP64 with the reference's operations (squared differences added
// x, y, z; volume times spring, then the division); the row order is the
// internal one, so a sum may differ from the reference's list-order sum in
// its last bits. Each workgroup writes the minimum of its nodes to a partial
// array; one more workgroup reduces the partials.
//
// pdg_critical_time_step allocates the partial array for the call;
// pdg_critical_time_step_noalloc uses `scratch`, a file-scope device array of
// fixed size (one copy per device, held until the process ends), and
// allocates nothing.

#include <cmath>
#include <mutex>

#include "pdg_ctx.hpp"

namespace pdg {

namespace {

constexpr int kDtThreads = 256;

// Partial minima of the allocation-free entry point: one per workgroup of
// kDtThreads nodes, so it serves blocks of up to kDtScratchMax * kDtThreads
// = 4,194,304 nodes (128 KiB per device).
constexpr int kDtScratchMax = 16384;
static __device__ __align__(256) double scratch[kDtScratchMax];

// Device address of each device's copy of `scratch`, from
// hipGetSymbolAddress (the array's name is not a device address in host
// code).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

double* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch[device];
}

__device__ __forceinline__ double min_keep(double a, double b) { return b < a ? b : a; }

// Minimum over the 256 threads of a workgroup (result valid in thread 0).
__device__ __forceinline__ double block_min(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min_keep(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kDtThreads / 64; w++) v = min_keep(v, sh[w]);
  }
  return v;
}

__global__ __launch_bounds__(kDtThreads) void k_dt_partial(int n, const int* __restrict__ row,
                                                           const int* __restrict__ col,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ vol,
                                                           double spring, double density,
                                                           double* __restrict__ partial) {
  __shared__ double sh[kDtThreads / 64];
  const int k = blockIdx.x * kDtThreads + threadIdx.x;
  double dt = 1.0e50;
  if (k < n) {
    const int r0 = row[k], r1 = row[k + 1];
    const double X0 = x[3 * k], X1 = x[3 * k + 1], X2 = x[3 * k + 2];
    double den = 0.0;
    for (int b = r0; b < r1; b++) {
      const int j = col[b];
      const double d0 = X0 - x[3 * j];
      const double d1 = X1 - x[3 * j + 1];
      const double d2 = X2 - x[3 * j + 2];
      const double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      den += vol[j] * spring / dist;
    }
    if (r1 > r0) dt = sqrt(2.0 * density / den);
  }
  dt = block_min(dt, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = dt;
}

__global__ __launch_bounds__(kDtThreads) void k_dt_final(int nparts,
                                                         const double* __restrict__ partial,
                                                         double* __restrict__ out) {
  __shared__ double sh[kDtThreads / 64];
  double v = 1.0e50;
  for (int q = threadIdx.x; q < nparts; q += kDtThreads) v = min_keep(v, partial[q]);
  v = block_min(v, sh);
  if (threadIdx.x == 0) *out = v;
}

inline int dt_parts(int n) { return (n + kDtThreads - 1) / kDtThreads; }

double spring_constant(const Ctx* c) {
  const double pi = 3.14159265358979323846;
  const double hz = c->horizon;
  return 18.0 * c->K / (pi * hz * hz * hz * hz);
}

// Enqueues both kernels on the context's stream; partial holds dt_parts(N)
// doubles, out one.
hipError_t launch_dt(const Ctx* c, double density, double* partial, double* out) {
  const int n = c->N, np = std::max(dt_parts(n), 1);
  if (n > 0)
    hipLaunchKernelGGL(k_dt_partial, dim3(np), dim3(kDtThreads), 0, c->stream, n, c->d_row,
                       c->d_col, c->d_x, c->d_vol, spring_constant(c), density, partial);
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(kDtThreads), 0, c->stream, n > 0 ? np : 0, partial,
                     out);
  return hipGetLastError();
}

}  // namespace

// Called by pdg_ctx_create with `device` current, once per device.
hipError_t dt_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = static_cast<double*>(p);
  return hipSuccess;
}

}  // namespace pdg

using namespace pdg;

extern "C" {

// The partials and the result live in device memory allocated for the call.
int pdg_critical_time_step(pdg_ctx* h, double density, double* dtOut) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !dtOut) return PDG_E_ARG;
  if (!(density > 0.0)) return c->fail(PDG_E_ARG, "density must be positive");
  (void)hipSetDevice(c->device);
  double* buf = nullptr;
  const size_t np = (size_t)std::max(dt_parts(c->N), 1);
  int rc = c->hip(hipMalloc((void**)&buf, 8 * (np + 1)), "alloc dt partials");
  if (rc) return rc;
  rc = c->hip(launch_dt(c, density, buf, buf + np), "critical time step");
  double v = 0.0;
  if (!rc) rc = c->hip(hipMemcpyAsync(&v, buf + np, 8, hipMemcpyDeviceToHost, c->stream), "d2h dt");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  (void)hipFree(buf);
  if (rc) return rc;
  *dtOut = v;
  return PDG_OK;
}

// Same result as pdg_critical_time_step, written to device memory, with the
// partials in this device's `scratch`: no allocation. Calls that overlap on
// one device share the array and race.
int pdg_critical_time_step_noalloc(pdg_ctx* h, double density, double* dtOutDevice) {
  Ctx* c = reinterpret_cast<Ctx*>(h);
  if (!c || !c->have_consts || !dtOutDevice) return PDG_E_ARG;
  if (!(density > 0.0)) return c->fail(PDG_E_ARG, "density must be positive");
  if ((long long)c->N > (long long)kDtScratchMax * kDtThreads)
    return c->fail(PDG_E_UNSUPPORTED, "pdg_critical_time_step_noalloc: more than 4194304 nodes "
                                      "(its scratch holds 16384 partial minima of 256 nodes)");
  double* part = scratch_ptr(c->device);
  if (!part) return c->fail(PDG_E_HIP, "scratch address of this device unknown");
  (void)hipSetDevice(c->device);
  int rc = c->hip(launch_dt(c, density, part, dtOutDevice), "critical time step");
  if (!rc) rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  return rc;
}

}  // extern "C"
