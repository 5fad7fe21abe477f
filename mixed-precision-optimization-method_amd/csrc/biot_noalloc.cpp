// biot_misfit_l2 without device allocation, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).  biot_misfit_l2 stages
// the observed traces and the residual in hipMalloc'd buffers and reduces J into a third one;
// here all three live in one file-scope device array, `scratch`, of a fixed size.  Every call
// on one GPU shares that GPU's copy of the array and leaves the last call's contents in it:
// calls that overlap on one GPU (two host threads, or two handles' streams) race.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "../../include/biot.h"
#include "biot_adjoint.hpp"
#include "biot_handle.hpp"

namespace {

// the FP64 accumulator in the first 256 bytes, then the observed and residual traces
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;
constexpr size_t J_BYTES = 256;
static_assert(J_BYTES + (size_t)2 * sizeof(float) * BIOT_MISFIT_NOALLOC_MAX_SAMPLES <= SCRATCH_BYTES,
              "the header's maximum must fit the scratch array");
static __device__ __align__(256) char scratch[SCRATCH_BYTES];

// the array's address on each device: every GPU holds its own copy at its own address, taken
// once per device with that device current (the array's host-side name is no device address)
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char *g_scratch[MAX_DEVICES] = {};

hipError_t scratch_on(int dev, char **out)
{
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[dev] = static_cast<char *>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

__global__ void k_clear_j(double *J)
{
  if (threadIdx.x == 0) *J = 0.0;
}

int nfail(biot_handle *h, int code, const char *fmt, ...)
{
  char buf[256];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  h->err = buf;
  return code;
}

}  // namespace

extern "C" int biot_misfit_l2_noalloc(biot_handle *h, const float *observed, float *residual, double *J)
{
  if (!h || !observed || !J) return BIOT_ERR_ARG;
  const int64_t m = (int64_t)h->ncomp * h->nt_max * h->nrec;
  if (m == 0) return nfail(h, BIOT_ERR_STATE, "no receivers");
  if (m > BIOT_MISFIT_NOALLOC_MAX_SAMPLES)
    return nfail(h, BIOT_ERR_ARG, "the trace layout holds %lld samples; the allocation-free misfit takes at most %lld",
                 (long long)m, (long long)BIOT_MISFIT_NOALLOC_MAX_SAMPLES);
  const int dev = h->g.device;
  if (dev < 0 || dev >= MAX_DEVICES) return nfail(h, BIOT_ERR_ARG, "device %d beyond the scratch table", dev);
  if (hipSetDevice(dev) != hipSuccess) return nfail(h, BIOT_ERR_HIP, "hipSetDevice(%d) failed", dev);
  char *base = nullptr;
  const hipError_t e = scratch_on(dev, &base);
  if (e != hipSuccess)
    return nfail(h, BIOT_ERR_HIP, "hipGetSymbolAddress of the scratch array failed: %s", hipGetErrorString(e));
  double *d_J = reinterpret_cast<double *>(base);
  float *d_obs = reinterpret_cast<float *>(base + J_BYTES);
  float *d_res = d_obs + m;
  const void *d = h->d_traces64 ? (const void *)h->d_traces64 : h->d_traces;
  hipLaunchKernelGGL(k_clear_j, dim3(1), dim3(64), 0, h->stream, d_J);  // a kernel, not a memset node
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(d_obs, observed, (size_t)m * sizeof(float), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      biot::launch_misfit(d, h->d_traces64 != nullptr, d_obs, d_res, m, d_J, h->stream) != hipSuccess ||
      hipMemcpyAsync(J, d_J, sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      (residual &&
       hipMemcpyAsync(residual, d_res, (size_t)m * sizeof(float), hipMemcpyDeviceToHost, h->stream) != hipSuccess) ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return nfail(h, BIOT_ERR_HIP, "misfit reduction failed");
  return BIOT_OK;
}
