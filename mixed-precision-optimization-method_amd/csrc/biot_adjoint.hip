// Discrete adjoint of the Biot time step (HIP, gfx950): FP32, FP64 or FP16 forward states,
// adjoint wavefields in the arithmetic type (float, double for FP64), FP64 gradient sums.
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
// With 16-bit storage everything runs in storage units (fields divided by their scales,
// coefficients as the kernels apply them); the host converts dJ/dK with the per-array factor.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "biot_adjoint.hpp"
#include "biot_packed.hpp"

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

// storage type S -> arithmetic type A (also the C-PML carry type)
template <typename S>
struct AdjT {
  using A = typename std::conditional<std::is_same<S, double>::value, double, float>::type;
  using C = typename std::conditional<std::is_same<S, float>::value, float, S>::type;  // forward carry
  using T = S;  // stored element of fields and coefficients
  static constexpr bool scaled = std::is_same<S, _Float16>::value;
  static constexpr bool res = false;
  static constexpr bool pk = false;
};
// FP16T / FP16C states: binary16 fields and coefficients, an int8 residual on tau and p
// (FP16C: also on v; value = h + r ulp(h)/256, as the forward kernels store it), float C-PML
// carries
struct H16R {};
template <>
struct AdjT<H16R> {
  using A = float;
  using C = float;
  using T = _Float16;
  static constexpr bool scaled = true;
  static constexpr bool res = true;
  static constexpr bool pk = false;
};
// FP16M states: binary16 fields; tau and p completed by the 4-bit residual of the cell's
// 4-byte record (value = h + r ulp(h)/16, as the forward kernels store it); float C-PML
// carries.  Its coefficients are read as FP16T stores them (binary16, packed per cell).
struct H16P {};
template <>
struct AdjT<H16P> {
  using A = float;
  using C = float;
  using T = _Float16;
  static constexpr bool scaled = true;
  static constexpr bool res = false;
  static constexpr bool pk = true;
};
// the residual unit ulp(h)/256 = 2^(e-33), e the biased exponent of h (subnormals e = 1)
__device__ __forceinline__ float res_unit(_Float16 h)
{
  const uint32_t u = __builtin_bit_cast(uint16_t, h);
  const int e = max((int)((u >> 10) & 31u), 1);
  return __builtin_bit_cast(float, (uint32_t)(e + 94) << 23);
}

struct CellIdx {
  int i, j, k;
  u32 c;
};

// update-box cell t (x fastest)
template <class AT>
__device__ __forceinline__ bool cell_of(const AT &a, int h, u32 t, CellIdx &ci)
{
  const u32 ni = (u32)(a.nx - 2 * h), nj = (u32)(a.ny - 2 * h);
  ci.i = h + (int)(t % ni);
  const u32 r = t / ni;
  ci.j = h + (int)(r % nj);
  ci.k = h + (int)(r / nj);
  ci.c = (u32)ci.i + (u32)ci.j * (u32)a.sy + (u32)ci.k * (u32)a.sz;
  return ci.k < a.nz - h;
}

template <class AT>
__device__ __forceinline__ const void *fw_field(const AT &a, int q, int f)
{
  return a.slots + (uint64_t)q * a.sb + (uint64_t)f * a.fbytes;
}
template <class AT>
__device__ __forceinline__ const void *fw_carry(const AT &a, int q, int ax)
{
  return a.slots + (uint64_t)q * a.sb + a.coff[ax];
}
// the residuals of field f in state q (FP16T tau and p, FP16C also v), or null; for FP16M
// the state's record array (256-byte aligned) with the field's nibble in the low 3 bits
template <typename S, class AT>
__device__ __forceinline__ const int8_t *fw_lo(const AT &a, int q, int f)
{
  if (!a.loff || !((a.lomask >> f) & 1u)) return nullptr;
  if constexpr (AdjT<S>::pk)
    return reinterpret_cast<const int8_t *>(a.slots + (uint64_t)q * a.sb + a.loff + (uint64_t)pk_slot(f));
  return reinterpret_cast<const int8_t *>(a.slots + (uint64_t)q * a.sb + a.loff +
                                          (uint64_t)f * (uint64_t)a.nx * a.ny * a.nz);
}

template <typename S, int ORDER>
__device__ __forceinline__ typename AdjT<S>::A diff(const void *f, const int8_t *lo, u32 c, u32 s, int fwd)
{
  using A = typename AdjT<S>::A;
  using T = typename AdjT<S>::T;
  auto L = [&](u32 idx) {
    const T h = at<T>(f, idx);
    if constexpr (AdjT<S>::res) {
      if (lo) return (A)h + (A)lo[idx] * res_unit(h);
    }
    if constexpr (AdjT<S>::pk) {
      if (lo) {
        const uintptr_t t = reinterpret_cast<uintptr_t>(lo);
        const int sl = (int)(t & 7u);
        return (A)pk_join((float)h, pk_nib(at<u32>(reinterpret_cast<const void *>(t - sl), idx), sl));
      }
    }
    return (A)h;
  };
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

// carry element index of slot `slot` on axis `ax` at cell ci, or -1 outside that slab (records
// of four carries per slab cell: slots 0..3, then 4..7)
template <class AT>
__device__ __forceinline__ long long carry_idx(const AT &a, int ax, int slot, const CellIdx &ci)
{
  const int pos = ax == 0 ? ci.i : (ax == 1 ? ci.j : ci.k);
  const int nn = ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz);
  const int p = slab_pos(pos, nn, a.L[ax], a.R[ax]);
  if (p < 0) return -1;
  const u32 w = (u32)(a.L[ax] + a.R[ax]);
  u32 o;
  if (ax == 0) o = (u32)p + w * ((u32)ci.j + (u32)a.ny * (u32)ci.k);
  else if (ax == 1) o = (u32)ci.i + (u32)a.nx * ((u32)p + w * (u32)ci.k);
  else o = (u32)ci.i + (u32)a.nx * ((u32)ci.j + (u32)a.ny * (u32)p);
  return (long long)((u32)(slot >> 2) * 4u * (u32)a.slab[ax] + 4u * o + (u32)(slot & 3));
}

template <typename A, class AT>
__device__ __forceinline__ A profv(const AT &a, int ax, int q, const CellIdx &ci)
{
  const int pos = ax == 0 ? ci.i : (ax == 1 ? ci.j : ci.k);
  return at<A>(a.prof[ax][q], (u32)pos);
}

// the N coefficients K[k0..k0+N) of cell c from the packed copy (storage units, times the
// per-array scale for 16-bit storage)
template <typename S, int N, class AT>
__device__ __forceinline__ void kblock(const AT &a, const void *base, u32 c, int k0,
                                       typename AdjT<S>::A (&out)[N])
{
  using A = typename AdjT<S>::A;
  using T = typename AdjT<S>::T;
  uint32_t d[packed::block_dwords<T, N>()];
  packed::load<T, N, true>(base, (uint64_t)a.nx * a.ny * a.nz, c, d);
#pragma unroll
  for (int e = 0; e < N; ++e) {
    if constexpr (AdjT<S>::scaled) out[e] = (A)packed::value<T>(d, e) * (A)a.ks[k0 + e];
    else out[e] = (A)packed::value<T>(d, e);
  }
}
// the N gradient sums of K[k0..k0+N) at cell c: float partial sums of the packed array p on
// FP32/FP16 handles, the FP64 sums themselves on FP64 handles
template <typename A, int N, class AT>
__device__ __forceinline__ void gload(const AT &a, const void *p, u32 c, int k0, double (&g)[N])
{
  if constexpr (std::is_same<A, double>::value) {
#pragma unroll
    for (int e = 0; e < N; ++e) g[e] = at<double>(a.gacc[k0 + e], c);
  } else {
    uint32_t d[packed::block_dwords<float, N>()];
    packed::load<float, N, false>(p, (uint64_t)a.nx * a.ny * a.nz, c, d);
#pragma unroll
    for (int e = 0; e < N; ++e) g[e] = (double)packed::value<float>(d, e);
  }
}
template <typename A, int N, class AT>
__device__ __forceinline__ void gstore(const AT &a, void *p, u32 c, int k0, const double (&g)[N])
{
  if constexpr (std::is_same<A, double>::value) {
#pragma unroll
    for (int e = 0; e < N; ++e) at<double>(a.gacc[k0 + e], c) = g[e];
  } else {
    uint32_t d[N];
#pragma unroll
    for (int e = 0; e < N; ++e) d[e] = __builtin_bit_cast(uint32_t, (float)g[e]);
    packed::store<float, N>(p, (uint64_t)a.nx * a.ny * a.nz, c, d);
  }
}

// Loads of one derivative-table entry: the forward derivative E (fields of state qf) and, in
// the entry's absorbing slab, the profile values, the forward carry (state qc) and the carry
// adjoint.  Every kernel below issues all of its loads before its first store: the compiler
// cannot prove the arrays distinct and never moves a load above a store, so this order keeps
// a cell's loads in flight together (one memory round trip instead of one per store).
template <typename A>
struct PmlIn {
  A E, e, aa, cf, lc;
  long long ix;
};
template <typename S, int ORDER, class AT>
__device__ __forceinline__ void pml_load(const AT &a, const Der &D, const CellIdx &ci, int qf, int qc,
                                         PmlIn<typename AdjT<S>::A> &r)
{
  using A = typename AdjT<S>::A;
  const u32 st[3] = {1u, (u32)a.sy, (u32)a.sz};
  r.E = diff<S, ORDER>(fw_field(a, qf, D.field), fw_lo<S>(a, qf, D.field), ci.c, st[D.dir], D.fwd);
  r.ix = carry_idx(a, D.dir, D.slot, ci);
  r.e = r.aa = r.cf = r.lc = A(0);
  if (r.ix >= 0) {
    r.e = profv<A>(a, D.dir, D.half ? 2 : 0, ci);
    r.aa = profv<A>(a, D.dir, D.half ? 3 : 1, ci);
    r.cf = (A)at<typename AdjT<S>::C>(fw_carry(a, qc, D.dir), (u32)r.ix);
    r.lc = at<A>(a.lcarry[D.dir], (u32)r.ix);
  }
}
// post-PML forward derivative fwdEp, derivative adjoint G and the updated carry adjoint r.lc
template <typename A>
__device__ __forceinline__ void pml_apply(PmlIn<A> &r, A lamEp, A &fwdEp, A &G)
{
  if (r.ix < 0) {
    fwdEp = r.E;
    G = lamEp;
    return;
  }
  fwdEp = (A(1) + r.aa) * r.E + r.cf;
  G = (A(1) + r.aa) * lamEp + A(2) * r.aa * r.e * r.lc;
  r.lc = lamEp + r.e * r.lc;
}
template <typename A, class AT>
__device__ __forceinline__ void pml_store(const AT &a, const Der &D, const PmlIn<A> &r)
{
  if (r.ix >= 0) at<A>(a.lcarry[D.dir], (u32)r.ix) = r.lc;
}

template <typename S, int ORDER, class AT>
__device__ __forceinline__ void adj_stress_local(const AT &a, int q, u32 t)
{
  using A = typename AdjT<S>::A;
  constexpr int h = ORDER / 2;
  CellIdx ci;
  if (t >= a.nupd || !cell_of(a, h, t, ci)) return;
  const u32 c = ci.c;
  const A ltxx = at<A>(a.lam[6], c), ltyy = at<A>(a.lam[7], c), ltzz = at<A>(a.lam[8], c);
  const A ltxy = at<A>(a.lam[9], c), ltxz = at<A>(a.lam[10], c), ltyz = at<A>(a.lam[11], c);
  const A lp = at<A>(a.lam[12], c);
  A kt[8];  // K[14..21]: P1 P2 S1 S2 S3 Sxy Sxz Syz
  kblock<S, 8>(a, a.kt, c, K_P1, kt);
  const A P1 = kt[0], P2 = kt[1], S1 = kt[2], S2 = kt[3], S3 = kt[4];
  const A Sxy = kt[5], Sxz = kt[6], Syz = kt[7];
  PmlIn<A> in[12];
#pragma unroll
  for (int d = 0; d < 12; ++d)  // velocities of state q+1 (after the velocity half step), carries of state q
    pml_load<S, ORDER>(a, TDER[d], ci, q + 1, q, in[d]);
  double g[8];  // sums of K[14..21]
  gload<A, 8>(a, a.gpt, c, K_P1, g);

  A lE[12];
  lE[0] = S2 * ltxx + S1 * (ltyy + ltzz) - P1 * lp;
  lE[1] = S2 * ltyy + S1 * (ltxx + ltzz) - P1 * lp;
  lE[2] = S2 * ltzz + S1 * (ltxx + ltyy) - P1 * lp;
  lE[3] = lE[4] = lE[5] = S3 * (ltxx + ltyy + ltzz) - P2 * lp;
  lE[6] = lE[7] = Sxy * ltxy;
  lE[8] = lE[9] = Sxz * ltxz;
  lE[10] = lE[11] = Syz * ltyz;
  A E[12], G[12];
#pragma unroll
  for (int d = 0; d < 12; ++d) pml_apply(in[d], lE[d], E[d], G[d]);
  // FP64 gradient sums: dJ/dK += lambda(output) x (forward term), kernel.cu:484-490
  const double divu = (double)E[0] + E[1] + E[2], divw = (double)E[3] + E[4] + E[5];
  g[0] += -(double)lp * divu;
  g[1] += -(double)lp * divw;
  g[2] += (double)ltxx * (E[1] + E[2]) + (double)ltyy * (E[0] + E[2]) + (double)ltzz * (E[0] + E[1]);
  g[3] += (double)ltxx * E[0] + (double)ltyy * E[1] + (double)ltzz * E[2];
  g[4] += ((double)ltxx + ltyy + ltzz) * divw;
  g[5] += (double)ltxy * (E[6] + E[7]);
  g[6] += (double)ltxz * (E[8] + E[9]);
  g[7] += (double)ltyz * (E[10] + E[11]);
#pragma unroll
  for (int d = 0; d < 12; ++d) {
    at<A>(a.G[d], c) = G[d];
    pml_store(a, TDER[d], in[d]);
  }
  gstore<A, 8>(a, a.gpt, c, K_P1, g);
}

// lambda(v', q') += transposed differences of the stress-side derivative adjoints
template <typename S, int ORDER, class AT>
__device__ __forceinline__ void adj_stress_gather(const AT &a, int, u32 t)
{
  using A = typename AdjT<S>::A;
  constexpr int h = ORDER / 2;
  CellIdx ci;
  if (t >= a.nupd || !cell_of(a, h, t, ci)) return;
  const u32 c = ci.c, sx = 1, sy = (u32)a.sy, sz = (u32)a.sz;
  const A l0 = at<A>(a.lam[0], c) + (diffT<A, ORDER>(a.G[0], c, sx, 0) + diffT<A, ORDER>(a.G[6], c, sy, 1) +
                                     diffT<A, ORDER>(a.G[8], c, sz, 1));
  const A l1 = at<A>(a.lam[1], c) + (diffT<A, ORDER>(a.G[1], c, sy, 0) + diffT<A, ORDER>(a.G[7], c, sx, 1) +
                                     diffT<A, ORDER>(a.G[10], c, sz, 1));
  const A l2 = at<A>(a.lam[2], c) + (diffT<A, ORDER>(a.G[2], c, sz, 0) + diffT<A, ORDER>(a.G[9], c, sx, 1) +
                                     diffT<A, ORDER>(a.G[11], c, sy, 1));
  const A l3 = at<A>(a.lam[3], c) + diffT<A, ORDER>(a.G[3], c, sx, 0);
  const A l4 = at<A>(a.lam[4], c) + diffT<A, ORDER>(a.G[4], c, sy, 0);
  const A l5 = at<A>(a.lam[5], c) + diffT<A, ORDER>(a.G[5], c, sz, 0);
  at<A>(a.lam[0], c) = l0;
  at<A>(a.lam[1], c) = l1;
  at<A>(a.lam[2], c) = l2;
  at<A>(a.lam[3], c) = l3;
  at<A>(a.lam[4], c) = l4;
  at<A>(a.lam[5], c) = l5;
}

template <typename S, int ORDER, class AT>
__device__ __forceinline__ void adj_velocity_local(const AT &a, int q, u32 t)
{
  using A = typename AdjT<S>::A;
  constexpr int h = ORDER / 2;
  CellIdx ci;
  if (t >= a.nupd || !cell_of(a, h, t, ci)) return;
  const u32 c = ci.c;
  constexpr int kU1[3] = {K_U1X, K_U1Y, K_U1Z}, kU2[3] = {K_U2X, K_U2Y, K_U2Z};
  constexpr int kW1[3] = {K_W1X, K_W1Y, K_W1Z}, kW2[3] = {K_W2X, K_W2Y, K_W2Z};
  constexpr int kW3[3] = {K_W3X, K_W3Y, K_W3Y};  // qz takes W3y (kernel.cu:220)
  // loads: adjoint state, coefficients, forward q of states q and q+1, derivative-table
  // entries, gradient sums
  A kv[14];  // K[0..13]: U1 xyz, U2 xyz, W1 xyz, W2 xyz, W3x, W3y
  kblock<S, 14>(a, a.kv, c, 0, kv);
  A lv[3], lq[3], U1[3], U2[3], W1[3], W2[3], W3[3];
  double q0[3], q1[3], gU1[3], gU2[3], gW1[3], gW2[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    lv[m] = at<A>(a.lam[m], c);
    lq[m] = at<A>(a.lam[3 + m], c);
    U1[m] = kv[kU1[m]], U2[m] = kv[kU2[m]];
    W1[m] = kv[kW1[m]], W2[m] = kv[kW2[m]], W3[m] = kv[kW3[m]];
    q0[m] = (double)at<typename AdjT<S>::T>(fw_field(a, q, 3 + m), c);  // q carries no residual
    q1[m] = (double)at<typename AdjT<S>::T>(fw_field(a, q + 1, 3 + m), c);
  }
  PmlIn<A> in[12];
#pragma unroll
  for (int d = 0; d < 12; ++d) pml_load<S, ORDER>(a, VDER[d], ci, q, q, in[d]);
  double gv[14];  // sums of K[0..13]
  gload<A, 14>(a, a.gpv, c, 0, gv);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    gU1[m] = gv[kU1[m]], gU2[m] = gv[kU2[m]];
    gW1[m] = gv[kW1[m]], gW2[m] = gv[kW2[m]];
  }
  double gW3x = gv[K_W3X], gW3y = gv[K_W3Y];

  // per component: lambda_qe = lambda_q' - U2 lambda_v'; lambda_S = U1 lambda_v' - W2 lambda_qe;
  // lambda_Ds = -W3 lambda_qe (qz takes W3y, kernel.cu:220)
  A lS[3], lDs[3], lqe[3], lqn[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    lqe[m] = lq[m] - U2[m] * lv[m];
    lS[m] = U1[m] * lv[m] - W2[m] * lqe[m];
    lDs[m] = -W3[m] * lqe[m];
    lqn[m] = U2[m] * lv[m] + W1[m] * lqe[m];  // lambda_q at step n (lambda_v is unchanged)
  }
  // derivative order x1 x2 x3 s1 y1 y2 y3 s2 z1 z2 z3 s3
  const A lD[12] = {lS[0], lS[0], lS[0], lDs[0], lS[1], lS[1], lS[1], lDs[1],
                    lS[2], lS[2], lS[2], lDs[2]};
  A D[12], G[12];
#pragma unroll
  for (int d = 0; d < 12; ++d) pml_apply(in[d], lD[d], D[d], G[d]);
  // FP64 gradient sums with the forward terms of kernel.cu:218-225
  const double Ssum[3] = {(double)D[0] + D[1] + D[2], (double)D[5] + D[4] + D[6],
                          (double)D[9] + D[10] + D[8]};
  const double Ds[3] = {(double)D[3], (double)D[7], (double)D[11]};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    gU1[m] += (double)lv[m] * Ssum[m];
    gU2[m] += -(double)lv[m] * (q1[m] - q0[m]);
    gW1[m] += (double)lqe[m] * q0[m];
    gW2[m] += -(double)lqe[m] * Ssum[m];
  }
  gW3x += -(double)lqe[0] * Ds[0];
  gW3y += -(double)lqe[1] * Ds[1];  // qy's term, then qz's, into the one W3y sum
  gW3y += -(double)lqe[2] * Ds[2];
  // stores
#pragma unroll
  for (int m = 0; m < 3; ++m) at<A>(a.lam[3 + m], c) = lqn[m];
#pragma unroll
  for (int d = 0; d < 12; ++d) {
    at<A>(a.G[d], c) = G[d];
    pml_store(a, VDER[d], in[d]);
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    gv[kU1[m]] = gU1[m], gv[kU2[m]] = gU2[m];
    gv[kW1[m]] = gW1[m], gv[kW2[m]] = gW2[m];
  }
  gv[K_W3X] = gW3x;
  gv[K_W3Y] = gW3y;
  gstore<A, 14>(a, a.gpv, c, 0, gv);
}

// lambda(tau, p) += transposed differences of the velocity-side derivative adjoints
template <typename S, int ORDER, class AT>
__device__ __forceinline__ void adj_velocity_gather(const AT &a, int, u32 t)
{
  using A = typename AdjT<S>::A;
  constexpr int h = ORDER / 2;
  CellIdx ci;
  if (t >= a.nupd || !cell_of(a, h, t, ci)) return;
  const u32 c = ci.c, sx = 1, sy = (u32)a.sy, sz = (u32)a.sz;
  const A l6 = at<A>(a.lam[6], c) + diffT<A, ORDER>(a.G[0], c, sx, 1);
  const A l9 = at<A>(a.lam[9], c) + (diffT<A, ORDER>(a.G[1], c, sy, 0) + diffT<A, ORDER>(a.G[5], c, sx, 0));
  const A l10 = at<A>(a.lam[10], c) + (diffT<A, ORDER>(a.G[2], c, sz, 0) + diffT<A, ORDER>(a.G[9], c, sx, 0));
  const A l7 = at<A>(a.lam[7], c) + diffT<A, ORDER>(a.G[4], c, sy, 1);
  const A l11 = at<A>(a.lam[11], c) + (diffT<A, ORDER>(a.G[6], c, sz, 0) + diffT<A, ORDER>(a.G[10], c, sy, 0));
  const A l8 = at<A>(a.lam[8], c) + diffT<A, ORDER>(a.G[8], c, sz, 1);
  const A l12 = at<A>(a.lam[12], c) + (diffT<A, ORDER>(a.G[3], c, sx, 1) + diffT<A, ORDER>(a.G[7], c, sy, 1) +
                                       diffT<A, ORDER>(a.G[11], c, sz, 1));
  at<A>(a.lam[6], c) = l6;
  at<A>(a.lam[9], c) = l9;
  at<A>(a.lam[10], c) = l10;
  at<A>(a.lam[7], c) = l7;
  at<A>(a.lam[11], c) = l11;
  at<A>(a.lam[8], c) = l8;
  at<A>(a.lam[12], c) = l12;
}

// the float partial sums of a segment into the FP64 gradient sums, cleared for the next one
template <typename S, int ORDER, class AT>
__device__ __forceinline__ void adj_flush(const AT &a, u32 t)
{
  constexpr int h = ORDER / 2;
  CellIdx ci;
  if (t >= a.nupd || !cell_of(a, h, t, ci)) return;
  const u32 c = ci.c;
  const uint64_t n = (uint64_t)a.nx * a.ny * a.nz;
  uint32_t dv[14], dt[8];
  packed::load<float, 14, false>(a.gpv, n, c, dv);
  packed::load<float, 8, false>(a.gpt, n, c, dt);
  double tot[NKC];
#pragma unroll
  for (int k = 0; k < NKC; ++k) tot[k] = a.grad[k][c];
#pragma unroll
  for (int k = 0; k < NKC; ++k)
    a.grad[k][c] = tot[k] + (double)(k < 14 ? packed::value<float>(dv, k) : packed::value<float>(dt, k - 14));
  const uint32_t zero[14] = {};
  packed::store<float, 14>(a.gpv, n, c, zero);
  packed::store<float, 8>(a.gpt, n, c, zero);
}

template <typename S, int ORDER, int WHICH, class AT>
__device__ __forceinline__ void adj_pass(const AT &a, int q, u32 t)
{
  if constexpr (WHICH == 0) adj_stress_local<S, ORDER>(a, q, t);
  else if constexpr (WHICH == 1) adj_stress_gather<S, ORDER>(a, q, t);
  else if constexpr (WHICH == 2) adj_velocity_local<S, ORDER>(a, q, t);
  else if constexpr (WHICH == 3) adj_velocity_gather<S, ORDER>(a, q, t);
  else adj_flush<S, ORDER>(a, t);
}

template <typename S, int ORDER, int WHICH>
__global__ void __launch_bounds__(256) k_adj(const AdjArgs a, int q)
{
  adj_pass<S, ORDER, WHICH>(a, q, blockIdx.x * blockDim.x + threadIdx.x);
}

// batched: argument tables read through the constant address space (uniform per block)
typedef __attribute__((address_space(4))) const AdjArgs CAdjArgs;
typedef __attribute__((address_space(4))) const int CInt;
typedef __attribute__((address_space(4))) const InjectArgs CInjectArgs;
typedef __attribute__((address_space(4))) const CopyEntry CCopyEntry;

__device__ __forceinline__ int shot_of_block(const AdjBatch &b, int blk)
{
  CInt *bs = (CInt *)b.block_start;
  int lo = 0, hi = b.nshots - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bs[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename S, int ORDER, int WHICH>
__global__ void __launch_bounds__(256) k_adj_batch(const AdjBatch b, int q)
{
  const int blk = blockIdx.x;
  const int s = shot_of_block(b, blk);
  const CAdjArgs &a = ((CAdjArgs *)b.shots)[s];
  const u32 t = (u32)(blk - ((CInt *)b.block_start)[s]) * 256u + threadIdx.x;
  adj_pass<S, ORDER, WHICH>(a, q, t);
}

// lambda(recorded field at receiver) += scale x adjoint source sample of step it
template <typename A, class IT>
__device__ __forceinline__ void inject_one(const IT &a, int64_t r, int ci, int it)
{
  if (r >= a.nrec || ci >= a.ncomp) return;
  const float v = a.src[((int64_t)ci * a.nt + it) * a.nrec + r] * a.scale[ci];
  atomicAdd(&((A *)a.lam[a.comp[ci]])[a.rec[r]], (A)v);
}
template <typename A>
__global__ void k_adj_inject(const InjectArgs a, int it)
{
  inject_one<A>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, (int)blockIdx.y, it);
}
template <typename A>
__global__ void k_adj_inject_batch(const InjectArgs *tab, int it)
{
  const CInjectArgs &a = ((CInjectArgs *)tab)[blockIdx.z];
  inject_one<A>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, (int)blockIdx.y, it);
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

// ---- chain rule (the device form of the host chain in biot_params.cpp) ----
__device__ __forceinline__ double harmonic4(double a, double b, double c, double d)
{
  if (a == 0 || b == 0 || c == 0 || d == 0) return 0.0;
  return 1.0 / (0.25 * (1 / a + 1 / b + 1 / c + 1 / d));
}

__global__ void __launch_bounds__(256) k_chain(const ChainArgs c)
{
  const int64_t wsy = c.wx, wsz = (int64_t)c.wx * c.wy;
  const int64_t nw = wsz * c.wz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nw) return;
  const int i = (int)(t % c.wx);
  const int j = (int)((t / c.wx) % c.wy);
  const int k = (int)(t / wsz);
  if (i < 1 || i >= c.wx - 1 || j < 1 || j >= c.wy - 1 || k < 1 || k >= c.wz - 1) return;
  // p: window cell (parameters, coefficient gradients); g: model cell (output)
  const int64_t p = t;
  const int64_t g = (int64_t)(c.x0 + i) + (int64_t)(c.y0 + j) * c.msy + (int64_t)(c.z0 + k) * c.msz;
  auto gK = [&](int kk) { return c.grad[kk][t] * c.kfac[kk]; };
  const double *mu = c.par[1], *alp = c.par[2], *M = c.par[3];
  const double *rho = c.par[4], *rhof = c.par[5], *C1 = c.par[6], *tphi = c.par[7];
  const double k1 = c.Cv * c.dt / c.Cp;   // velocity coefficient scale
  const double kk = c.dt * c.Cp / c.Cv;   // stress coefficient scale
  // stress side: K_P1 = C kk/dx, K_P2 = M kk/dx, K_S1 = lambda_c .., K_S2 = H .., K_S3 = C ..
  const double al = alp[p], Mp = M[p];
  const double gC = (gK(K_P1) + gK(K_S3)) * kk * c.Hinv;
  const double glc = gK(K_S1) * kk * c.Hinv;
  const double gH = gK(K_S2) * kk * c.Hinv;
  const double gMd = gK(K_P2) * kk * c.Hinv;
  atomicAdd(&c.out[0][g], glc + gH);
  double gmu_c = 2 * gH;
  atomicAdd(&c.out[2][g], gC * Mp + (glc + gH) * 2 * al * Mp);
  atomicAdd(&c.out[3][g], gMd + gC * al + (glc + gH) * al * al);
  // shear: 4-cell harmonic means, d mu_bar / d mu_i = mu_bar^2 / (4 mu_i^2)
  const int64_t wq[3][4] = {{0, 1, wsy, wsy + 1}, {0, 1, wsz, wsz + 1}, {0, wsy, wsz, wsz + wsy}};
  const int64_t mq[3][4] = {{0, 1, c.msy, c.msy + 1}, {0, 1, c.msz, c.msz + 1},
                            {0, c.msy, c.msz, c.msz + c.msy}};
  const int kidx[3] = {K_SXY, K_SXZ, K_SYZ};
  for (int m = 0; m < 3; ++m) {
    const double m0 = mu[p + wq[m][0]], m1 = mu[p + wq[m][1]], m2 = mu[p + wq[m][2]], m3 = mu[p + wq[m][3]];
    const double hm = harmonic4(m0, m1, m2, m3);
    if (hm == 0) continue;
    const double gg = gK(kidx[m]) * kk * c.Hinv;
    for (int e = 0; e < 4; ++e) {
      const double me = mu[p + wq[m][e]];
      const double v = gg * hm * hm / (4 * me * me);
      if (e == 0) gmu_c += v;
      else atomicAdd(&c.out[1][g + mq[m][e]], v);
    }
  }
  atomicAdd(&c.out[1][g], gmu_c);
  // velocity side, per staggering direction
  const int64_t wstep[3] = {1, wsy, wsz}, mstep[3] = {1, c.msy, c.msz};
  double A[3], B[3], c1[3], D[3];
  bool c1zero = false, azero = false;
  for (int m = 0; m < 3; ++m) {
    const int64_t p2 = p + wstep[m];
    A[m] = 0.5 * (rho[p] + rho[p2]);
    B[m] = 0.5 * (rhof[p] + rhof[p2]);
    c1[m] = 0.5 * (C1[p] + C1[p2]);
    const double c2 = 0.5 * (tphi[p] * rhof[p] + tphi[p2] * rhof[p2]);
    D[m] = (c2 - B[m] * B[m] / A[m]) + 0.5 * c.dt * c1[m];
    c1zero |= c1[m] == 0;
    azero |= A[m] == 0;
  }
  double grho_c = 0, grf_c = 0;
  for (int m = 0; m < 3; ++m) {
    double gA = 0, gB = 0, gc2 = 0;
    if (!c1zero) {
      const double gW1 = gK(K_W1X + m), gW2 = gK(K_W2X + m);
      // W3y multiplies the pressure term of both qy and qz, so its gradient holds both;
      // W3z is never read
      const double gW3 = m < 2 ? gK(K_W3X + m) : 0.0;
      const double D2 = D[m] * D[m];
      const double dW1_dmp = c.dt * c1[m] / D2;
      const double dW3_dmp = -k1 * c.Hinv / D2;
      const double dW2_dmp = -k1 * c.Hinv * (B[m] / A[m]) / D2;
      const double gmp = gW1 * dW1_dmp + gW2 * dW2_dmp + gW3 * dW3_dmp;
      gA += gmp * B[m] * B[m] / (A[m] * A[m]) - gW2 * k1 * c.Hinv * B[m] / (A[m] * A[m] * D[m]);
      gB += gmp * (-2 * B[m] / A[m]) + gW2 * k1 * c.Hinv / (A[m] * D[m]);
      gc2 += gmp;
    }
    if (!azero) {
      const double gU1 = gK(K_U1X + m), gU2 = gK(K_U2X + m);
      gA += -gU1 * k1 * c.Hinv / (A[m] * A[m]) - gU2 * B[m] / (A[m] * A[m]);
      gB += gU2 / A[m];
    }
    grho_c += 0.5 * gA;
    atomicAdd(&c.out[4][g + mstep[m]], 0.5 * gA);
    grf_c += 0.5 * gB + 0.5 * gc2 * tphi[p];
    atomicAdd(&c.out[5][g + mstep[m]], 0.5 * gB + 0.5 * gc2 * tphi[p + wstep[m]]);
  }
  atomicAdd(&c.out[4][g], grho_c);
  atomicAdd(&c.out[5][g], grf_c);
}

// ---- batched state copies ----
__global__ void __launch_bounds__(256) k_copy_batch(const CopyEntry *tab)
{
  const CCopyEntry &e = ((CCopyEntry *)tab)[blockIdx.y];
  const uint64_t words = e.bytes >> 2;
  const uint32_t *src = (const uint32_t *)e.src;
  uint32_t *dst = (uint32_t *)e.dst;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words;
       w += (uint64_t)gridDim.x * blockDim.x)
    dst[w] = src[w];
  const uint32_t rem = (uint32_t)(e.bytes & 3);
  if (blockIdx.x == 0 && threadIdx.x < rem)
    ((uint8_t *)e.dst)[(words << 2) + threadIdx.x] = ((const uint8_t *)e.src)[(words << 2) + threadIdx.x];
}

template <typename S, int ORDER>
hipError_t adj_launch(int which, const AdjArgs &a, int q, hipStream_t s)
{
  const unsigned blocks = (unsigned)((a.nupd + 255) / 256);
  switch (which) {
    case 0: hipLaunchKernelGGL((k_adj<S, ORDER, 0>), dim3(blocks), dim3(256), 0, s, a, q); break;
    case 1: hipLaunchKernelGGL((k_adj<S, ORDER, 1>), dim3(blocks), dim3(256), 0, s, a, q); break;
    case 2: hipLaunchKernelGGL((k_adj<S, ORDER, 2>), dim3(blocks), dim3(256), 0, s, a, q); break;
    case 3: hipLaunchKernelGGL((k_adj<S, ORDER, 3>), dim3(blocks), dim3(256), 0, s, a, q); break;
    case 4: hipLaunchKernelGGL((k_adj<S, ORDER, 4>), dim3(blocks), dim3(256), 0, s, a, q); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <typename S, int ORDER>
hipError_t adj_launch_batch(int which, const AdjBatch &b, int q, hipStream_t s)
{
  const unsigned blocks = (unsigned)b.total_blocks;
  switch (which) {
    case 0: hipLaunchKernelGGL((k_adj_batch<S, ORDER, 0>), dim3(blocks), dim3(256), 0, s, b, q); break;
    case 1: hipLaunchKernelGGL((k_adj_batch<S, ORDER, 1>), dim3(blocks), dim3(256), 0, s, b, q); break;
    case 2: hipLaunchKernelGGL((k_adj_batch<S, ORDER, 2>), dim3(blocks), dim3(256), 0, s, b, q); break;
    case 3: hipLaunchKernelGGL((k_adj_batch<S, ORDER, 3>), dim3(blocks), dim3(256), 0, s, b, q); break;
    case 4: hipLaunchKernelGGL((k_adj_batch<S, ORDER, 4>), dim3(blocks), dim3(256), 0, s, b, q); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

struct SingleRun {
  const AdjArgs &a;
  int q;
  hipStream_t s;
  template <typename S, int O>
  hipError_t run(int which) { return adj_launch<S, O>(which, a, q, s); }
};
struct BatchRun {
  const AdjBatch &b;
  int q;
  hipStream_t s;
  template <typename S, int O>
  hipError_t run(int which) { return adj_launch_batch<S, O>(which, b, q, s); }
};

// passes first..last-1 (0-3: the adjoint step, 4: the flush), in order, for the storage
// type of prec
template <class F>
hipError_t passes(int prec, int order, F f, int first, int last)
{
  if (order != 2 && order != 4) return hipErrorInvalidValue;
  for (int which = first; which < last; ++which) {
    hipError_t e;
    if (prec == PREC_FP32) e = order == 2 ? f.template run<float, 2>(which) : f.template run<float, 4>(which);
    else if (prec == PREC_FP64) e = order == 2 ? f.template run<double, 2>(which) : f.template run<double, 4>(which);
    else if (prec == PREC_FP16) e = order == 2 ? f.template run<_Float16, 2>(which) : f.template run<_Float16, 4>(which);
    else if (prec == PREC_FP16T || prec == PREC_FP16C)
      e = order == 2 ? f.template run<H16R, 2>(which) : f.template run<H16R, 4>(which);
    else if (prec == PREC_FP16M)
      e = order == 2 ? f.template run<H16P, 2>(which) : f.template run<H16P, 4>(which);
    else return hipErrorInvalidValue;
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

hipError_t launch_adjoint_step(int prec, int order, const AdjArgs &a, int q, hipStream_t s)
{
  if (a.nupd == 0) return hipSuccess;
  return passes(prec, order, SingleRun{a, q, s}, 0, 4);
}

hipError_t launch_adjoint_flush(int prec, int order, const AdjArgs &a, hipStream_t s)
{
  if (a.nupd == 0 || prec == PREC_FP64) return hipSuccess;
  return passes(prec, order, SingleRun{a, 0, s}, 4, 5);
}

hipError_t launch_adjoint_flush_batch(int prec, int order, const AdjBatch &b, hipStream_t s)
{
  if (b.total_blocks <= 0 || prec == PREC_FP64) return hipSuccess;
  return passes(prec, order, BatchRun{b, 0, s}, 4, 5);
}

hipError_t launch_adjoint_step_batch(int prec, int order, const AdjBatch &b, int q, hipStream_t s)
{
  if (b.total_blocks <= 0) return hipSuccess;
  return passes(prec, order, BatchRun{b, q, s}, 0, 4);
}

hipError_t launch_adjoint_inject(int prec, const InjectArgs &a, int it, hipStream_t s)
{
  if (a.nrec <= 0 || a.ncomp <= 0) return hipSuccess;
  dim3 grid((unsigned)((a.nrec + 255) / 256), (unsigned)a.ncomp);
  if (prec == PREC_FP64) hipLaunchKernelGGL(k_adj_inject<double>, grid, dim3(256), 0, s, a, it);
  else if (prec == PREC_FP32 || prec == PREC_FP16 || prec == PREC_FP16T || prec == PREC_FP16C ||
           prec == PREC_FP16M)
    hipLaunchKernelGGL(k_adj_inject<float>, grid, dim3(256), 0, s, a, it);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_adjoint_inject_batch(int prec, const InjectArgs *tab, int nshots, int64_t max_nrec,
                                       int max_ncomp, int it, hipStream_t s)
{
  if (nshots <= 0 || max_nrec <= 0 || max_ncomp <= 0) return hipSuccess;
  dim3 grid((unsigned)((max_nrec + 255) / 256), (unsigned)max_ncomp, (unsigned)nshots);
  if (prec == PREC_FP64) hipLaunchKernelGGL(k_adj_inject_batch<double>, grid, dim3(256), 0, s, tab, it);
  else if (prec == PREC_FP32 || prec == PREC_FP16 || prec == PREC_FP16T || prec == PREC_FP16C ||
           prec == PREC_FP16M)
    hipLaunchKernelGGL(k_adj_inject_batch<float>, grid, dim3(256), 0, s, tab, it);
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

hipError_t launch_chain(const ChainArgs &c, hipStream_t s)
{
  const int64_t nw = (int64_t)c.wx * c.wy * c.wz;
  if (nw <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_chain, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_copy_batch(const CopyEntry *tab, int n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_batch, dim3(1024, (unsigned)n), dim3(256), 0, s, tab);
  return hipGetLastError();
}

}  // namespace biot
