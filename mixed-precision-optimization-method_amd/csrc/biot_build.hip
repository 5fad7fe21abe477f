// GPU-side set-up: Biot moduli (outputfunc.cpp:205-217), staggered averages (main.cpp:322-358),
// the 23 coefficient arrays (main.cpp:431-494) and their conversion to the kernels' scaled
// coefficient arrays, evaluated with the host builder's float/double promotions so the
// results are bit-identical to it (csrc/biot_setup.cpp).  Float division is done as the
// correctly rounded quotient (exact in double, rounded once to float) and contraction into
// fused multiply-adds is off, as on the host.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/biot.h"
#include "biot_device.hpp"

#pragma clang fp contract(off)

namespace biot {
namespace {

// correctly rounded float quotient (a double quotient of two floats rounds to it exactly)
__device__ __forceinline__ float fdiv(float a, float b) { return (float)((double)a / (double)b); }

__device__ __forceinline__ float harmonic4(float a, float b, float c, float d)
{
  if (a == 0 || b == 0 || c == 0 || d == 0) return 0.0f;
  const float s = fdiv(1.0f, a) + fdiv(1.0f, b) + fdiv(1.0f, c) + fdiv(1.0f, d);
  return (float)(1 / (0.25 * s));
}
__device__ __forceinline__ float mean2(float a, float b) { return (float)(0.5 * (a + b)); }

// outputfunc.cpp:205-217 per cell; out: M HH H2u C C1 C2 rho rhof mu
__global__ void __launch_bounds__(256) k_moduli(MediumArgs m, int64_t n)
{
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float *const *o = m.out;
  const double tortuosity = (1 + 2 / m.porousm) * m.tao;
  const float rhos = (float)m.rhos[c];
  const float rhof = (float)m.rhof[c];
  const double phi = m.phi[c];
  const float mu = (float)m.mu[c];
  o[7][c] = rhof;
  o[8][c] = mu;
  if (m.mu[c] == 0.0 && m.Kb[c] == 0.0) {  // fluid cell (outputfunc.cpp:176-193)
    const float kf = (float)m.Kf[c];
    o[0][c] = o[1][c] = o[2][c] = o[3][c] = kf;
    o[4][c] = 0.0f;
    o[5][c] = rhof;
    o[6][c] = rhof;
    return;
  }
  o[6][c] = (float)((1 - phi) * rhos + phi * rhof);
  const double alpha = 1 - m.Kb[c] / m.Ks[c];
  const float M = (float)(m.Kf[c] * m.Ks[c] / (phi * m.Ks[c] + (alpha - phi) * m.Kf[c]));
  const float HH = (float)(alpha * alpha * M + m.Kb[c] + mu * 4.0 / 3.0);
  o[0][c] = M;
  o[1][c] = HH;
  o[2][c] = HH - 2 * mu;
  o[3][c] = (float)(M * alpha);
  o[4][c] = (float)(m.eta[c] / m.perm[c]);
  o[5][c] = (float)(tortuosity * rhof / phi);
}

// main.cpp:322-358 averages (inline) and main.cpp:431-494 coefficients, per cell
__global__ void __launch_bounds__(256) k_coef_ref(CoefRefArgs a)
{
  const int64_t n = (int64_t)a.nx * a.ny * a.nz;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / sz);
  const bool interior = i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1;
  const float *M = a.mod[0], *HH = a.mod[1], *H2u = a.mod[2], *C = a.mod[3], *C1 = a.mod[4];
  const float *C2 = a.mod[5], *rho = a.mod[6], *rhof = a.mod[7], *mu = a.mod[8];
  float avrho[3] = {0, 0, 0}, avrhof[3] = {0, 0, 0}, avC1[3] = {0, 0, 0}, avC2[3] = {0, 0, 0};
  float muxz = 0, muxy = 0, muyz = 0;
  if (interior) {
    muxz = harmonic4(mu[c], mu[c + 1], mu[c + sz], mu[c + sz + 1]);
    muyz = harmonic4(mu[c], mu[c + sy], mu[c + sz], mu[c + sz + sy]);
    muxy = harmonic4(mu[c], mu[c + 1], mu[c + sy], mu[c + sy + 1]);
    const int64_t step[3] = {1, sy, sz};
    for (int d = 0; d < 3; ++d) {
      const int64_t s = step[d];
      avrhof[d] = mean2(rhof[c], rhof[c + s]);
      avrho[d] = mean2(rho[c], rho[c + s]);
      avC1[d] = mean2(C1[c], C1[c + s]);
      avC2[d] = mean2(C2[c], C2[c + s]);
    }
  }
  const float DT = a.DT, Cv = a.Cv, Cp = a.Cp;
  const float Cw1 = 1.0f, Cw2 = 1.0f, Cw3 = 1.0f, Cu1 = 1.0f, Cu2 = 1.0f;
  const float Cpp = 100, Csp = 100;
  float *const *K = a.coef;
  if (avC1[0] == 0 || avC1[1] == 0 || avC1[2] == 0) {
    for (int q = BIOT_W1X; q <= BIOT_W3Z; ++q) K[q][c] = 0;
  } else {
    for (int d = 0; d < 3; ++d) {
      const float rf = avrhof[d], r = avrho[d], c1 = avC1[d], c2 = avC2[d];
      const float mp = c2 - fdiv(rf * rf, r);  // float, as (c2 - (rf * rf / r)) on the host
      K[BIOT_W1X + d][c] = (float)(Cw1 * (mp - 0.5 * DT * c1) / (mp + 0.5 * DT * c1));
      K[BIOT_W2X + d][c] = (float)(Cv * DT * Cw2 * fdiv(rf, r) / (Cp * (mp + 0.5 * DT * c1)));
      K[BIOT_W3X + d][c] = (float)(Cv * DT * Cw3 / (Cp * (mp + 0.5 * DT * c1)));
    }
  }
  if (avrho[0] == 0 || avrho[1] == 0 || avrho[2] == 0) {
    for (int q = BIOT_U1X; q <= BIOT_U2Z; ++q) K[q][c] = 0;
  } else {
    for (int d = 0; d < 3; ++d) {
      K[BIOT_U1X + d][c] = fdiv(Cv * DT * Cu1, avrho[d] * Cp);
      K[BIOT_U2X + d][c] = fdiv(avrhof[d] * Cu2, avrho[d]);
    }
  }
  const float Mi = interior ? M[c] : 0.0f, Ci = interior ? C[c] : 0.0f;
  const float HHi = interior ? HH[c] : 0.0f, H2ui = interior ? H2u[c] : 0.0f;
  K[BIOT_PP1][c] = fdiv(Ci * DT * Cp * Cpp, Cv);
  K[BIOT_PP2][c] = fdiv(Mi * DT * Cp * Cpp, Cv);
  K[BIOT_SP1][c] = fdiv(Cp * DT * H2ui * Csp, Cv);
  K[BIOT_SP2][c] = fdiv(Cp * DT * HHi * Csp, Cv);
  K[BIOT_SP3][c] = fdiv(Cp * DT * Ci * Csp, Cv);
  K[BIOT_SPXY][c] = fdiv(Cp * DT * muxy * Csp, Cv);
  K[BIOT_SPXZ][c] = fdiv(Cp * DT * muxz * Csp, Cv);
  K[BIOT_SPYZ][c] = fdiv(Cp * DT * muyz * Csp, Cv);
}

// reference array of kernel coefficient k and the factor/kind that turns it into K:
// kind 0: K = R, kind 1: K = R * H, kind 2: K = 0.01 * R * H (left to right, in double)
__device__ __forceinline__ int ref_of(int k, int &kind)
{
  if (k <= K_U1Z) return kind = 1, BIOT_U1X + (k - K_U1X);
  if (k <= K_U2Z) return kind = 0, BIOT_U2X + (k - K_U2X);
  if (k <= K_W1Z) return kind = 0, BIOT_W1X + (k - K_W1X);
  if (k <= K_W2Z) return kind = 1, BIOT_W2X + (k - K_W2X);
  if (k == K_W3X) return kind = 1, BIOT_W3X;
  if (k == K_W3Y) return kind = 1, BIOT_W3Y;
  kind = 2;
  switch (k) {
    case K_P1: return BIOT_PP1;
    case K_P2: return BIOT_PP2;
    case K_S1: return BIOT_SP1;
    case K_S2: return BIOT_SP2;
    case K_S3: return BIOT_SP3;
    case K_SXY: return BIOT_SPXY;
    case K_SXZ: return BIOT_SPXZ;
    default: return BIOT_SPYZ;
  }
}
__device__ __forceinline__ double k_value(const float *const *R, int k, int64_t c, double H)
{
  int kind;
  const int r = ref_of(k, kind);
  if (kind == 0) return (double)R[r][c];
  if (kind == 1) return (double)R[r][c] * H;
  return 0.01 * R[r][c] * H;
}

// per-array max |R| (float bits, non-negative) and the forced q response
// max over cells of (|W2| + |W3|) / (1 - W1) of every direction (double bits)
__global__ void __launch_bounds__(256) k_ref_stats(RefArgs a)
{
  const int64_t n = a.n;
  uint32_t amax[BIOT_NCOEF];
  for (int q = 0; q < BIOT_NCOEF; ++q) amax[q] = 0;
  double resp = 0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    for (int q = 0; q < BIOT_NCOEF; ++q) amax[q] = max(amax[q], __float_as_uint(fabsf(a.ref[q][c])));
    for (int d = 0; d < 3; ++d) {
      const double w1 = (double)a.ref[BIOT_W1X + d][c];
      const double w2 = (double)a.ref[BIOT_W2X + d][c] * a.H;
      const double w3 = (double)a.ref[d == 2 ? BIOT_W3Y : BIOT_W3X + d][c] * a.H;
      const double den = 1.0 - w1;
      if (den > 0) resp = fmax(resp, (fabs(w2) + fabs(w3)) / den);
    }
  }
  // wave-level maxima, then one atomic per wave on the orderable bit patterns
  for (int q = 0; q < BIOT_NCOEF; ++q) {
    uint32_t v = amax[q];
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    amax[q] = v;
  }
  unsigned long long rb = (unsigned long long)__double_as_longlong(resp);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = (unsigned long long)__shfl_xor((long long)rb, off);
    rb = o > rb ? o : rb;
  }
  if ((threadIdx.x & 63) == 0) {
    for (int q = 0; q < BIOT_NCOEF; ++q)
      if (amax[q]) atomicMax(&a.stats_u32[q], amax[q]);
    if (rb) atomicMax(a.stats_u64, rb);
  }
}

template <typename S>
__global__ void __launch_bounds__(256) k_ref_to_kernel(RefArgs a)
{
  const int64_t n = a.n;
  const int k = blockIdx.y;
  S *out = (S *)a.kout + (int64_t)k * n;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    double x = k_value(a.ref, k, c, a.H);
    if (a.scaled) {
      x *= a.ratio[k];
      x *= a.inv_ks[k];
    }
    if constexpr (std::is_same<S, double>::value) out[c] = x;
    else if constexpr (std::is_same<S, float>::value) out[c] = (float)x;
    else out[c] = (_Float16)(float)x;
  }
}

}  // namespace

hipError_t launch_moduli(const MediumArgs &m, int64_t n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_moduli, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m, n);
  return hipGetLastError();
}

hipError_t launch_coef_ref(const CoefRefArgs &a, hipStream_t s)
{
  const int64_t n = (int64_t)a.nx * a.ny * a.nz;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_coef_ref, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ref_stats(const RefArgs &a, hipStream_t s)
{
  const unsigned blocks = (unsigned)std::min<int64_t>((a.n + 255) / 256, 2048);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ref_stats, dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ref_to_kernel(int prec, const RefArgs &a, hipStream_t s)
{
  const unsigned blocks = (unsigned)std::min<int64_t>((a.n + 255) / 256, 1024);
  if (blocks == 0) return hipSuccess;
  dim3 grid(blocks, NKC);
  switch (prec) {
    case PREC_FP32: hipLaunchKernelGGL(k_ref_to_kernel<float>, grid, dim3(256), 0, s, a); break;
    case PREC_FP64: hipLaunchKernelGGL(k_ref_to_kernel<double>, grid, dim3(256), 0, s, a); break;
    case PREC_FP16:
    case PREC_FP16C:
    case PREC_FP16T: hipLaunchKernelGGL(k_ref_to_kernel<_Float16>, grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace biot
