// Discrete adjoint of the Biot time step (HIP, gfx950), FP32 or FP64 storage.
//
// One forward step n maps u^n (after the source of step n) to u^{n+1}:
//   velocity half step  D = delta(tau, p); D' = (1+a) D + c_V; c_V <- e (c_V + 2 a D)
//                       q' = W1 q - W2 sum D' - W3 D's;  v' = v + U1 sum D' - U2 (q' - q)
//   stress half step    E = delta(v', q'); E' = (1+a) E + c_T; c_T <- e (c_T + 2 a E)
//                       p' = p - P1 div_u' - P2 div_w';  tau' = tau + S1 .. + S2 .. + S3 div_w'
// (the folded C-PML carry, kernel.cu:188-196 algebraically).  Its transpose, applied
// backward, is split into two passes per half step:
//   local  : adjoints of the twelve post-PML derivatives from the adjoint state at the cell,
//            C-PML carry adjoints  lambda_E = (1+a) lambda_E' + 2 a e lambda_c,
//            lambda_c <- lambda_E' + e lambda_c, the cell-local field adjoints, and the FP64
//            sums of dJ/dK = lambda(output) x (forward term) for every coefficient K, with the
//            forward derivatives recomputed from the stored states;
//   gather : lambda_f(c) += sum_m w_m G(c - o_m), the transposed staggered differences,
//            written per cell so the result is deterministic (no atomics).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "biot_adjoint.hpp"

namespace biot {
namespace {

using u32 = uint32_t;

template <typename T>
__device__ __forceinline__ T &at(void *base, u32 idx)
{
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (u32)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T &at(const void *base, u32 idx)
{
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (u32)sizeof(T));
}

// derivative tables: field, direction (= damped axis), forward(1)/backward(0) difference,
// carry slot on that axis, half-node (1) / integer-node (0) profile
struct Der {
  int field, dir, fwd, slot, half;
};
// velocity half step: x1 x2 x3 s1 y1 y2 y3 s2 z1 z2 z3 s3 (kernel.cu:172-185, 188-216)
__constant__ Der VDER[12] = {{6, 0, 1, 0, 1},  {9, 1, 0, 0, 0},  {10, 2, 0, 0, 0}, {12, 0, 1, 1, 1},
                             {7, 1, 1, 1, 1},  {9, 0, 0, 2, 0},  {11, 2, 0, 1, 0}, {12, 1, 1, 2, 1},
                             {8, 2, 1, 2, 1},  {10, 0, 0, 3, 0}, {11, 1, 0, 3, 0}, {12, 2, 1, 3, 1}};
// stress half step: uxx uyy uzz wx wy wz uxy uyx uxz uzx uyz uzy (kernel.cu:430-482)
__constant__ Der TDER[12] = {{0, 0, 0, 4, 0}, {1, 1, 0, 4, 0}, {2, 2, 0, 4, 0}, {3, 0, 0, 5, 0},
                             {4, 1, 0, 5, 0}, {5, 2, 0, 5, 0}, {0, 1, 1, 6, 1}, {1, 0, 1, 6, 1},
                             {0, 2, 1, 6, 1}, {2, 0, 1, 7, 1}, {1, 2, 1, 7, 1}, {2, 1, 1, 7, 1}};

template <typename S>
struct AdjT {
  using A = typename std::conditional<std::is_same<S, double>::value, double, float>::type;
};

struct CellIdx {
  int i, j, k;
  u32 c;
};

// update-box cell t (x fastest)
__device__ __forceinline__ bool cell_of(const AdjArgs &a, int h, u32 t, CellIdx &ci)
{
  const u32 ni = (u32)(a.nx - 2 * h), nj = (u32)(a.ny - 2 * h);
  ci.i = h + (int)(t % ni);
  const u32 r = t / ni;
  ci.j = h + (int)(r % nj);
  ci.k = h + (int)(r / nj);
  ci.c = (u32)ci.i + (u32)ci.j * (u32)a.sy + (u32)ci.k * (u32)a.sz;
  return ci.k < a.nz - h;
}

template <typename S, int ORDER>
__device__ __forceinline__ typename AdjT<S>::A diff(const void *f, u32 c, u32 s, int fwd)
{
  using A = typename AdjT<S>::A;
  auto L = [&](u32 idx) { return (A)at<S>(f, idx); };
  if (fwd) {
    if constexpr (ORDER == 2) return L(c + s) - L(c);
    return A(9.0 / 8.0) * (L(c + s) - L(c)) + A(-1.0 / 24.0) * (L(c + 2 * s) - L(c - s));
  } else {
    if constexpr (ORDER == 2) return L(c) - L(c - s);
    return A(9.0 / 8.0) * (L(c) - L(c - s)) + A(-1.0 / 24.0) * (L(c + s) - L(c - 2 * s));
  }
}

// transposed difference gathered at c from the derivative-adjoint array G
template <typename A, int ORDER>
__device__ __forceinline__ A diffT(const void *G, u32 c, u32 s, int fwd)
{
  auto L = [&](u32 idx) { return at<A>(G, idx); };
  if (fwd) {
    if constexpr (ORDER == 2) return L(c - s) - L(c);
    return A(9.0 / 8.0) * (L(c - s) - L(c)) + A(-1.0 / 24.0) * (L(c - 2 * s) - L(c + s));
  } else {
    if constexpr (ORDER == 2) return L(c) - L(c + s);
    return A(9.0 / 8.0) * (L(c) - L(c + s)) + A(-1.0 / 24.0) * (L(c - s) - L(c + 2 * s));
  }
}

__device__ __forceinline__ int slab_pos(int i, int n, int Lw, int Rw)
{
  if (i < Lw) return i;
  if (i >= n - Rw) return Lw + i - (n - Rw);
  return -1;
}

// carry element index of slot `slot` on axis `ax` at cell ci, or -1 outside that slab
__device__ __forceinline__ long long carry_idx(const AdjArgs &a, int ax, int slot, const CellIdx &ci)
{
  const int pos[3] = {ci.i, ci.j, ci.k};
  const int nn[3] = {a.nx, a.ny, a.nz};
  const int p = slab_pos(pos[ax], nn[ax], a.L[ax], a.R[ax]);
  if (p < 0) return -1;
  const u32 w = (u32)(a.L[ax] + a.R[ax]);
  u32 o;
  if (ax == 0) o = (u32)p + w * ((u32)ci.j + (u32)a.ny * (u32)ci.k);
  else if (ax == 1) o = (u32)ci.i + (u32)a.nx * ((u32)p + w * (u32)ci.k);
  else o = (u32)ci.i + (u32)a.nx * ((u32)ci.j + (u32)a.ny * (u32)p);
  return (long long)((u32)slot * (u32)a.slab[ax] + o);
}

template <typename A>
__device__ __forceinline__ A profv(const AdjArgs &a, int ax, int q, const CellIdx &ci)
{
  const int pos[3] = {ci.i, ci.j, ci.k};
  return at<A>(a.prof[ax][q], (u32)pos[ax]);
}

template <typename S>
__device__ __forceinline__ typename AdjT<S>::A kcoef(const AdjArgs &a, int kidx, u32 c)
{
  return (typename AdjT<S>::A)at<S>(a.k[kidx], c);
}

// post-PML forward derivative d (table entry D) and the carry adjoint update
template <typename S, int ORDER>
__device__ __forceinline__ void pml_adjoint(const AdjArgs &a, const Der &D, const CellIdx &ci,
                                            const void *const *fwd_fields, const void *const *fwd_carry,
                                            typename AdjT<S>::A lamEp, typename AdjT<S>::A &fwdEp,
                                            typename AdjT<S>::A &G)
{
  using A = typename AdjT<S>::A;
  using SC = S;
  const u32 st[3] = {1u, (u32)a.sy, (u32)a.sz};
  const A E = diff<S, ORDER>(fwd_fields[D.field], ci.c, st[D.dir], D.fwd);
  const long long ix = carry_idx(a, D.dir, D.slot, ci);
  if (ix < 0) {
    fwdEp = E;
    G = lamEp;
    return;
  }
  const A e = profv<A>(a, D.dir, D.half ? 2 : 0, ci);
  const A aa = profv<A>(a, D.dir, D.half ? 3 : 1, ci);
  const A cf = (A)at<SC>(fwd_carry[D.dir], (u32)ix);
  fwdEp = (A(1) + aa) * E + cf;
  A &lc = at<A>(a.lcarry[D.dir], (u32)ix);
  G = (A(1) + aa) * lamEp + A(2) * aa * e * lc;
  lc = lamEp + e * lc;
}

template <typename S, int ORDER>
__global__ void __launch_bounds__(256) k_adj_stress_local(const AdjArgs a)
{
  using A = typename AdjT<S>::A;
  constexpr int h = ORDER / 2;
  CellIdx ci;
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nupd || !cell_of(a, h, t, ci)) return;
  const u32 c = ci.c;
  const A ltxx = at<A>(a.lam[6], c), ltyy = at<A>(a.lam[7], c), ltzz = at<A>(a.lam[8], c);
  const A ltxy = at<A>(a.lam[9], c), ltxz = at<A>(a.lam[10], c), ltyz = at<A>(a.lam[11], c);
  const A lp = at<A>(a.lam[12], c);
  const A S1 = kcoef<S>(a, K_S1, c), S2 = kcoef<S>(a, K_S2, c), S3 = kcoef<S>(a, K_S3, c);
  const A P1 = kcoef<S>(a, K_P1, c), P2 = kcoef<S>(a, K_P2, c);
  const A Sxy = kcoef<S>(a, K_SXY, c), Sxz = kcoef<S>(a, K_SXZ, c), Syz = kcoef<S>(a, K_SYZ, c);
  A lE[12];
  lE[0] = S2 * ltxx + S1 * (ltyy + ltzz) - P1 * lp;
  lE[1] = S2 * ltyy + S1 * (ltxx + ltzz) - P1 * lp;
  lE[2] = S2 * ltzz + S1 * (ltxx + ltyy) - P1 * lp;
  lE[3] = lE[4] = lE[5] = S3 * (ltxx + ltyy + ltzz) - P2 * lp;
  lE[6] = lE[7] = Sxy * ltxy;
  lE[8] = lE[9] = Sxz * ltxz;
  lE[10] = lE[11] = Syz * ltyz;
  A E[12];
#pragma unroll
  for (int d = 0; d < 12; ++d) {
    A G;
    pml_adjoint<S, ORDER>(a, TDER[d], ci, a.fw1, a.fcarry, lE[d], E[d], G);
    at<A>(a.G[d], c) = G;
  }
  // FP64 gradient sums: dJ/dK += lambda(output) x (forward term), kernel.cu:484-490
  const double divu = (double)E[0] + E[1] + E[2], divw = (double)E[3] + E[4] + E[5];
  a.grad[K_P1][c] += -(double)lp * divu;
  a.grad[K_P2][c] += -(double)lp * divw;
  a.grad[K_S1][c] += (double)ltxx * (E[1] + E[2]) + (double)ltyy * (E[0] + E[2]) + (double)ltzz * (E[0] + E[1]);
  a.grad[K_S2][c] += (double)ltxx * E[0] + (double)ltyy * E[1] + (double)ltzz * E[2];
  a.grad[K_S3][c] += ((double)ltxx + ltyy + ltzz) * divw;
  a.grad[K_SXY][c] += (double)ltxy * (E[6] + E[7]);
  a.grad[K_SXZ][c] += (double)ltxz * (E[8] + E[9]);
  a.grad[K_SYZ][c] += (double)ltyz * (E[10] + E[11]);
}

// lambda(v', q') += transposed differences of the stress-side derivative adjoints
template <typename S, int ORDER>
__global__ void __launch_bounds__(256) k_adj_stress_gather(const AdjArgs a)
{
  using A = typename AdjT<S>::A;
  constexpr int h = ORDER / 2;
  CellIdx ci;
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nupd || !cell_of(a, h, t, ci)) return;
  const u32 c = ci.c, sx = 1, sy = (u32)a.sy, sz = (u32)a.sz;
  at<A>(a.lam[0], c) += diffT<A, ORDER>(a.G[0], c, sx, 0) + diffT<A, ORDER>(a.G[6], c, sy, 1) +
                        diffT<A, ORDER>(a.G[8], c, sz, 1);
  at<A>(a.lam[1], c) += diffT<A, ORDER>(a.G[1], c, sy, 0) + diffT<A, ORDER>(a.G[7], c, sx, 1) +
                        diffT<A, ORDER>(a.G[10], c, sz, 1);
  at<A>(a.lam[2], c) += diffT<A, ORDER>(a.G[2], c, sz, 0) + diffT<A, ORDER>(a.G[9], c, sx, 1) +
                        diffT<A, ORDER>(a.G[11], c, sy, 1);
  at<A>(a.lam[3], c) += diffT<A, ORDER>(a.G[3], c, sx, 0);
  at<A>(a.lam[4], c) += diffT<A, ORDER>(a.G[4], c, sy, 0);
  at<A>(a.lam[5], c) += diffT<A, ORDER>(a.G[5], c, sz, 0);
}

template <typename S, int ORDER>
__global__ void __launch_bounds__(256) k_adj_velocity_local(const AdjArgs a)
{
  using A = typename AdjT<S>::A;
  constexpr int h = ORDER / 2;
  CellIdx ci;
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nupd || !cell_of(a, h, t, ci)) return;
  const u32 c = ci.c;
  // per component: lambda_qe = lambda_q' - U2 lambda_v'; lambda_S = U1 lambda_v' - W2 lambda_qe;
  // lambda_Ds = -W3 lambda_qe (qz takes W3y, kernel.cu:220)
  A lS[3], lDs[3], lv[3], lqe[3];
  const int kU1[3] = {K_U1X, K_U1Y, K_U1Z}, kU2[3] = {K_U2X, K_U2Y, K_U2Z};
  const int kW1[3] = {K_W1X, K_W1Y, K_W1Z}, kW2[3] = {K_W2X, K_W2Y, K_W2Z};
  const int kW3[3] = {K_W3X, K_W3Y, K_W3Y};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    lv[q] = at<A>(a.lam[q], c);
    const A lq = at<A>(a.lam[3 + q], c);
    const A U1 = kcoef<S>(a, kU1[q], c), U2 = kcoef<S>(a, kU2[q], c);
    const A W1 = kcoef<S>(a, kW1[q], c), W2 = kcoef<S>(a, kW2[q], c), W3 = kcoef<S>(a, kW3[q], c);
    lqe[q] = lq - U2 * lv[q];
    lS[q] = U1 * lv[q] - W2 * lqe[q];
    lDs[q] = -W3 * lqe[q];
    at<A>(a.lam[3 + q], c) = U2 * lv[q] + W1 * lqe[q];  // lambda_q at step n (lambda_v is unchanged)
  }
  // derivative order x1 x2 x3 s1 y1 y2 y3 s2 z1 z2 z3 s3
  const A lD[12] = {lS[0], lS[0], lS[0], lDs[0], lS[1], lS[1], lS[1], lDs[1],
                    lS[2], lS[2], lS[2], lDs[2]};
  A D[12];
#pragma unroll
  for (int d = 0; d < 12; ++d) {
    A G;
    pml_adjoint<S, ORDER>(a, VDER[d], ci, a.fw, a.fcarry, lD[d], D[d], G);
    at<A>(a.G[d], c) = G;
  }
  // FP64 gradient sums with the forward terms of kernel.cu:218-225
  const double Ssum[3] = {(double)D[0] + D[1] + D[2], (double)D[5] + D[4] + D[6],
                          (double)D[9] + D[10] + D[8]};
  const double Ds[3] = {(double)D[3], (double)D[7], (double)D[11]};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double q0 = (double)at<S>(a.fw[3 + q], c), q1 = (double)at<S>(a.fw1[3 + q], c);
    a.grad[kU1[q]][c] += (double)lv[q] * Ssum[q];
    a.grad[kU2[q]][c] += -(double)lv[q] * (q1 - q0);
    a.grad[kW1[q]][c] += (double)lqe[q] * q0;
    a.grad[kW2[q]][c] += -(double)lqe[q] * Ssum[q];
    a.grad[kW3[q]][c] += -(double)lqe[q] * Ds[q];
  }
}

// lambda(tau, p) += transposed differences of the velocity-side derivative adjoints
template <typename S, int ORDER>
__global__ void __launch_bounds__(256) k_adj_velocity_gather(const AdjArgs a)
{
  using A = typename AdjT<S>::A;
  constexpr int h = ORDER / 2;
  CellIdx ci;
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nupd || !cell_of(a, h, t, ci)) return;
  const u32 c = ci.c, sx = 1, sy = (u32)a.sy, sz = (u32)a.sz;
  at<A>(a.lam[6], c) += diffT<A, ORDER>(a.G[0], c, sx, 1);
  at<A>(a.lam[9], c) += diffT<A, ORDER>(a.G[1], c, sy, 0) + diffT<A, ORDER>(a.G[5], c, sx, 0);
  at<A>(a.lam[10], c) += diffT<A, ORDER>(a.G[2], c, sz, 0) + diffT<A, ORDER>(a.G[9], c, sx, 0);
  at<A>(a.lam[7], c) += diffT<A, ORDER>(a.G[4], c, sy, 1);
  at<A>(a.lam[11], c) += diffT<A, ORDER>(a.G[6], c, sz, 0) + diffT<A, ORDER>(a.G[10], c, sy, 0);
  at<A>(a.lam[8], c) += diffT<A, ORDER>(a.G[8], c, sz, 1);
  at<A>(a.lam[12], c) += diffT<A, ORDER>(a.G[3], c, sx, 1) + diffT<A, ORDER>(a.G[7], c, sy, 1) +
                         diffT<A, ORDER>(a.G[11], c, sz, 1);
}

// lambda(recorded field at receiver) += adjoint source sample of step it
template <typename A>
__global__ void k_adj_inject(const InjectArgs a)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ci = blockIdx.y;
  if (r >= a.nrec || ci >= a.ncomp) return;
  const float v = a.src[((int64_t)ci * a.nt + a.it) * a.nrec + r] * a.scale[ci];
  atomicAdd(&((A *)a.lam[a.comp[ci]])[a.rec[r]], (A)v);
}

// J = 1/2 sum (d - obs)^2 in FP64, residual = d - obs
template <typename T>
__global__ void __launch_bounds__(256) k_misfit(const T *d, const float *obs, float *res, int64_t n, double *J)
{
  __shared__ double part[256];
  double acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double r = (double)d[i] - (double)obs[i];
    if (res) res[i] = (float)r;
    acc += 0.5 * r * r;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(J, part[0]);
}

template <typename S, int ORDER>
hipError_t adj_launch(int which, const AdjArgs &a, hipStream_t s)
{
  const unsigned blocks = (unsigned)((a.nupd + 255) / 256);
  switch (which) {
    case 0: hipLaunchKernelGGL((k_adj_stress_local<S, ORDER>), dim3(blocks), dim3(256), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_adj_stress_gather<S, ORDER>), dim3(blocks), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_adj_velocity_local<S, ORDER>), dim3(blocks), dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_adj_velocity_gather<S, ORDER>), dim3(blocks), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_adjoint_step(int prec, int order, const AdjArgs &a, hipStream_t s)
{
  if (a.nupd == 0) return hipSuccess;
  for (int which = 0; which < 4; ++which) {
    hipError_t e;
    if (prec == PREC_FP32) e = order == 2 ? adj_launch<float, 2>(which, a, s) : adj_launch<float, 4>(which, a, s);
    else if (prec == PREC_FP64) e = order == 2 ? adj_launch<double, 2>(which, a, s) : adj_launch<double, 4>(which, a, s);
    else return hipErrorInvalidValue;
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_adjoint_inject(int prec, const InjectArgs &a, hipStream_t s)
{
  if (a.nrec <= 0 || a.ncomp <= 0) return hipSuccess;
  dim3 grid((unsigned)((a.nrec + 255) / 256), (unsigned)a.ncomp);
  if (prec == PREC_FP32) hipLaunchKernelGGL(k_adj_inject<float>, grid, dim3(256), 0, s, a);
  else if (prec == PREC_FP64) hipLaunchKernelGGL(k_adj_inject<double>, grid, dim3(256), 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_misfit(const void *d, bool d_double, const float *obs, float *res, int64_t n,
                         double *J, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  if (d_double)
    hipLaunchKernelGGL(k_misfit<double>, dim3(blocks), dim3(256), 0, s, (const double *)d, obs, res, n, J);
  else
    hipLaunchKernelGGL(k_misfit<float>, dim3(blocks), dim3(256), 0, s, (const float *)d, obs, res, n, J);
  return hipGetLastError();
}

}  // namespace biot
