// A plain streaming copy on the device, timed with HIP events: the HBM rate (bytes read +
// bytes written per second) that a kernel doing nothing but 16-byte coalesced loads and
// stores reaches on this GPU.  The stencil kernels' achieved rates are read against it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/biot.h"

namespace {

typedef uint32_t dw4 __attribute__((ext_vector_type(4)));

// one 16-byte element per thread; NT: nontemporal loads and stores (no cache allocation)
template <bool NT>
__global__ void __launch_bounds__(256) k_copy16(const dw4 *__restrict__ src, dw4 *__restrict__ dst, int64_t n16)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n16) return;
  if constexpr (NT) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
  else dst[i] = src[i];
}

}  // namespace

extern "C" int biot_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, int32_t variant, double *GBs)
{
  if (!GBs || bytes < 16 || bytes % 16 || bytes > ((int64_t)16 << 30) || reps < 1 || reps > 1000 ||
      (variant != 0 && variant != 1))
    return BIOT_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return BIOT_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return BIOT_ERR_HIP;
  dw4 *src = nullptr, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = BIOT_OK;
  if (hipMalloc((void **)&src, (size_t)bytes) != hipSuccess || hipMalloc((void **)&dst, (size_t)bytes) != hipSuccess)
    rc = BIOT_ERR_NOMEM;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = BIOT_ERR_HIP;
  if (!rc && hipMemset(src, 1, (size_t)bytes) != hipSuccess) rc = BIOT_ERR_HIP;
  const int64_t n16 = bytes / 16;
  const dim3 grid((unsigned)((n16 + 255) / 256)), block(256);
  auto launch = [&]() {
    if (variant == 1) hipLaunchKernelGGL(k_copy16<true>, grid, block, 0, 0, src, dst, n16);
    else hipLaunchKernelGGL(k_copy16<false>, grid, block, 0, 0, src, dst, n16);
  };
  if (!rc) {
    for (int r = 0; r < 3; ++r) launch();  // warm-up
    if (hipEventRecord(e0, 0) != hipSuccess) rc = BIOT_ERR_HIP;
    for (int r = 0; r < reps && !rc; ++r) launch();
    if (!rc && (hipGetLastError() != hipSuccess || hipEventRecord(e1, 0) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess))
      rc = BIOT_ERR_HIP;
  }
  if (!rc) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess || !(ms > 0)) rc = BIOT_ERR_HIP;
    else *GBs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}
