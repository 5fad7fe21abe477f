// Streaming kernels on the device, timed with HIP events: the HBM rate (bytes read + bytes
// written per second) of a plain 16-byte copy, and of the velocity kernel's access pattern
// without its stencil, all at the cell's own index: FP32 (13 float fields read, the 56-byte
// packed coefficient block of 3 x 16 + 8 bytes read, 6 float fields written) and FP16 (13
// binary16 fields, a 16 + 12-byte block, 6 binary16 fields).  The stencil kernels' achieved
// rates are read against these.
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

typedef uint32_t dw3 __attribute__((ext_vector_type(3), aligned(4)));
struct VPattern {
  const float *r[13];
  float *w[6];
  const dw4 *k4[3];
  const uint64_t *k2;
  int64_t n;
};
// FP16 form: 13 binary16 fields, the 28-byte packed block (a 16-byte and a 12-byte record)
// read, 6 binary16 fields written (66 B per cell)
struct HPattern {
  const _Float16 *r[13];
  _Float16 *w[6];
  const dw4 *k4;
  const dw3 *k3;
  int64_t n;
};
template <bool NT>
__global__ void __launch_bounds__(256) k_hpattern(const HPattern a)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 13; ++k) s += (float)a.r[k][i];
  const dw4 v = __builtin_nontemporal_load(a.k4 + i);
  // 12-byte records: addressed in bytes (a 3-element vector type has the size of 4 elements)
  const dw3 t = __builtin_nontemporal_load(
      reinterpret_cast<const dw3 *>(reinterpret_cast<const char *>(a.k3) + (size_t)i * 12));
  s += __builtin_bit_cast(float, v.x ^ v.y ^ v.z ^ v.w ^ t.x ^ t.y ^ t.z);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const _Float16 h = (_Float16)(s + (float)k);
    if constexpr (NT) __builtin_nontemporal_store(h, a.w[k] + i);
    else a.w[k][i] = h;
  }
}
// the velocity kernel's streams at one cell per thread; the loaded values feed the stores
template <bool NT>
__global__ void __launch_bounds__(256) k_vpattern(const VPattern a)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 13; ++k) s += a.r[k][i];
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    const dw4 v = __builtin_nontemporal_load(a.k4[g] + i);
    s += __builtin_bit_cast(float, v.x ^ v.y ^ v.z ^ v.w);
  }
  const uint64_t t = __builtin_nontemporal_load(a.k2 + i);
  s += __builtin_bit_cast(float, (uint32_t)t ^ (uint32_t)(t >> 32));
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if constexpr (NT) __builtin_nontemporal_store(s + (float)k, a.w[k] + i);
    else a.w[k][i] = s + (float)k;
  }
}

}  // namespace

extern "C" int biot_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, int32_t variant, double *GBs)
{
  if (!GBs || bytes < 16 || bytes % 16 || bytes > ((int64_t)16 << 30) || reps < 1 || reps > 1000 ||
      variant < 0 || variant > 5)
    return BIOT_ERR_ARG;
  if (variant >= 2 && bytes > ((int64_t)1 << 30)) return BIOT_ERR_ARG;  // at most 33 x bytes allocated
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return BIOT_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return BIOT_ERR_HIP;
  // copies: source and destination of `bytes`; the velocity pattern: 19 float fields of
  // `bytes` each and a packed block of 14 floats per cell (one allocation)
  const bool pattern = variant >= 2, half = variant >= 4;
  const int64_t n = pattern ? bytes / 4 : bytes / 16;
  // FP16 pattern: the two coefficient record arrays start on 16-byte boundaries
  const size_t h4 = ((size_t)n * 38 + 15) & ~(size_t)15, h3 = (h4 + (size_t)n * 16 + 15) & ~(size_t)15;
  const size_t total = half ? h3 + (size_t)n * 12 : pattern ? (size_t)n * (19 * 4 + 56) : (size_t)bytes * 2;
  char *mem = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = BIOT_OK;
  if (hipMalloc((void **)&mem, total) != hipSuccess) rc = BIOT_ERR_NOMEM;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = BIOT_ERR_HIP;
  if (!rc && hipMemset(mem, 1, total) != hipSuccess) rc = BIOT_ERR_HIP;
  const dw4 *src = (const dw4 *)mem;
  dw4 *dst = (dw4 *)(mem + (pattern ? 0 : bytes));
  VPattern vp{};
  HPattern hp{};
  if (half) {
    for (int k = 0; k < 13; ++k) hp.r[k] = (const _Float16 *)(mem + (size_t)k * n * 2);
    for (int k = 0; k < 6; ++k) hp.w[k] = (_Float16 *)(mem + (size_t)(13 + k) * n * 2);
    hp.k4 = (const dw4 *)(mem + h4);
    hp.k3 = (const dw3 *)(mem + h3);
    hp.n = n;
  } else if (pattern) {
    for (int k = 0; k < 13; ++k) vp.r[k] = (const float *)(mem + (size_t)k * n * 4);
    for (int k = 0; k < 6; ++k) vp.w[k] = (float *)(mem + (size_t)(13 + k) * n * 4);
    for (int g = 0; g < 3; ++g) vp.k4[g] = (const dw4 *)(mem + (size_t)19 * n * 4 + (size_t)g * n * 16);
    vp.k2 = (const uint64_t *)(mem + (size_t)19 * n * 4 + (size_t)3 * n * 16);
    vp.n = n;
  }
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  auto launch = [&]() {
    switch (variant) {
      case 0: hipLaunchKernelGGL(k_copy16<false>, grid, block, 0, 0, src, dst, n); break;
      case 1: hipLaunchKernelGGL(k_copy16<true>, grid, block, 0, 0, src, dst, n); break;
      case 2: hipLaunchKernelGGL(k_vpattern<false>, grid, block, 0, 0, vp); break;
      case 3: hipLaunchKernelGGL(k_vpattern<true>, grid, block, 0, 0, vp); break;
      case 4: hipLaunchKernelGGL(k_hpattern<false>, grid, block, 0, 0, hp); break;
      default: hipLaunchKernelGGL(k_hpattern<true>, grid, block, 0, 0, hp); break;
    }
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
    else *GBs = (half ? 66.0 * (double)n : pattern ? 132.0 * (double)n : 2.0 * (double)bytes) * reps / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (mem) (void)hipFree(mem);
  return rc;
}
