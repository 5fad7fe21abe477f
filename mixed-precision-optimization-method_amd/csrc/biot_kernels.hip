// HIP kernels (gfx950) of the 13-field staggered-grid Biot time step.
//
// Velocity half step (reference kernel.cu:145-229) and stress half step (kernel.cu:405-493)
// are each one z-marching kernel: a 2-D thread block covers an x-y tile with x fastest
// (64 consecutive cells per wavefront -> 256 B coalesced rows in FP32), every thread walks
// `zchunk` planes of its column and keeps the z-differenced fields in a register window
// that rotates by one plane per iteration.  The C-PML recursion is applied only in the
// absorbing slabs, on the folded carry variable described in biot_device.hpp.
//
// Storage is chosen by a policy (float / binary16 / binary16 + int8 residual / double);
// arithmetic is float, double for the double policy.
//
// Addressing: every array is reached from its uniform (kernel-argument) base pointer with
// a 32-bit byte offset, so a load is SGPR base + one 32-bit VGPR offset instead of a
// 64-bit VGPR address per load; the host guarantees every array spans < 4 GiB.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "biot_device.hpp"
#include "biot_packed.hpp"

namespace biot {
namespace {

using u32 = uint32_t;

// cmask: the fields that carry a residual (packed: in FP16M's packed record); nt_v / nt_t:
// the velocity / stress kernel's own-value loads and field stores are nontemporal (same-box
// A/B: FP32 velocity kernel -3.5 % time, FP16 velocity kernel -1 % step time; the residual
// policies 2-3 % slower with them, FP16's stress kernel neutral); medium: the coefficients
// are formed from the stored moduli (FP16M) instead of read from the packed arrays; outer16:
// the residual-carrying fields are read at the 4th-order outer taps as binary16
struct P32 {
  using S = float;
  using SC = float;
  using A = float;
  static constexpr bool comp = false, scaled = false;
  static constexpr uint32_t cmask = 0u;
  static constexpr bool nt_v = true, nt_t = true;
  static constexpr bool packed = false, medium = false;
  static constexpr bool outer16 = false;
};
struct P16 {
  using S = _Float16;
  using SC = _Float16;
  using A = float;
  static constexpr bool comp = false, scaled = true;
  static constexpr uint32_t cmask = 0u;
  static constexpr bool nt_v = true, nt_t = false;
  static constexpr bool packed = false, medium = false;
  static constexpr bool outer16 = false;
};
struct P16C {
  using S = _Float16;
  using SC = float;
  using A = float;
  static constexpr bool comp = true, scaled = true;
  static constexpr uint32_t cmask = COMP_FIELDS_C;
  static constexpr bool nt_v = false, nt_t = false;
  static constexpr bool packed = false, medium = false;
  static constexpr bool outer16 = false;
};
struct P16T {
  using S = _Float16;
  using SC = float;
  using A = float;
  static constexpr bool comp = true, scaled = true;
  static constexpr uint32_t cmask = COMP_FIELDS_T;
  static constexpr bool nt_v = false, nt_t = false;
  static constexpr bool packed = false, medium = false;
  static constexpr bool outer16 = true;
};
// v, q binary16; tau, p binary16 + 4-bit residual (the binary16 part in the field array, the
// residual in the cell's record); coefficients from the stored moduli
struct P16M {
  using S = _Float16;
  using SC = float;
  using A = float;
  static constexpr bool comp = true, scaled = true;
  static constexpr uint32_t cmask = COMP_FIELDS_T;
  static constexpr bool nt_v = false, nt_t = false;
  static constexpr bool packed = true, medium = true;
  static constexpr bool outer16 = true;
};
struct P64 {
  using S = double;
  using SC = double;
  using A = double;
  static constexpr bool comp = false, scaled = false;
  static constexpr uint32_t cmask = 0u;
  static constexpr bool nt_v = false, nt_t = false;
  static constexpr bool packed = false, medium = false;
  static constexpr bool outer16 = false;
};

// element idx of an array, byte offset formed in 32 bits
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

// binary16 residual scale: ulp(h)/256 = 2^(e-33) and its inverse 2^(33-e), e the biased
// exponent of h (subnormals use e = 1), built directly as float bit patterns
__device__ __forceinline__ int h_exp(_Float16 h)
{
  const u32 u = __builtin_bit_cast(uint16_t, h);
  return max((int)((u >> 10) & 31u), 1);
}
__device__ __forceinline__ float res_unit(int e) { return __builtin_bit_cast(float, (u32)(e + 94) << 23); }
__device__ __forceinline__ float res_inv(int e) { return __builtin_bit_cast(float, (u32)(160 - e) << 23); }

// FP16M's tau and p: the binary16 h nearest to the value (on the field's binary16 storage
// scale) and a 4-bit residual r in -7..7, value = h + r ulp(h)/16 (a 15-bit significand; units
// of the subnormal ulp below binary16's normal range).  h sits in the field's own array
// (2 bytes per cell, like a binary16 field); r in the cell's 4-byte residual record shared by
// the seven fields (nibble pk_slot(f)).  A stencil point of several of those fields loads the
// record once, the stress step writes it with one 4-byte store, and every access is aligned.
// A stencil tap that needs no more than binary16 precision reads h alone.  The record array
// is 256-byte aligned, so a field's residual pointer carries the field's nibble in its low 3
// bits (the runtime loaders below).
typedef uint32_t dw2_t __attribute__((ext_vector_type(2)));
// h and r of x: r as its nibble (two's complement)
__device__ __forceinline__ void pk_split(float x, uint16_t &hb, u32 &rb)
{
  const _Float16 h = (_Float16)x;
  hb = __builtin_bit_cast(uint16_t, h);
  // 2^(29 - e) = 16 / ulp(h), e the biased exponent of h (at least 1)
  const float inv = __builtin_bit_cast(float, (u32)(156 - h_exp(h)) << 23);
  const float r = rintf((x - (float)h) * inv);
  rb = (u32)(int)fminf(fmaxf(r, -7.0f), 7.0f) & 0xFu;
}
// the binary16 part alone, as a float
__device__ __forceinline__ float pk_hi(const void *hi, u32 idx)
{
  return (float)at<_Float16>(hi, idx);
}
// value of field F at cell idx: hi the field's array, rec the residual record array
template <int F>
__device__ __forceinline__ float pk_at(const void *hi, const void *rec, u32 idx)
{
  return pk_join(pk_hi(hi, idx), pk_nib(at<u32>(rec, idx), pk_slot(F)));
}
// the same value from a record already loaded
template <int F>
__device__ __forceinline__ float pk_rec(const void *hi, u32 r, u32 idx)
{
  return pk_join(pk_hi(hi, idx), pk_nib(r, pk_slot(F)));
}
// hf (the binary16 part alone, as a float) completed with field F's residual of record r
template <int F>
__device__ __forceinline__ float pk_fill(float hf, u32 r)
{
  return pk_join(hf, pk_nib(r, pk_slot(F)));
}
// runtime field: lo is the record array's base with the field's nibble in its low 3 bits
__device__ __forceinline__ float pk_load(const void *hi, const int8_t *lo, u32 idx)
{
  const uintptr_t t = reinterpret_cast<uintptr_t>(lo);
  const int s = (int)(t & 7u);
  return pk_join(pk_hi(hi, idx), pk_nib(at<u32>(reinterpret_cast<const void *>(t - s), idx), s));
}
// One field's value: its binary16 part, and its nibble set with two atomics that touch only
// that nibble's bits (lanes storing other fields of the same cell at once keep theirs).
__device__ __forceinline__ void pk_store(void *hi, int8_t *lo, u32 idx, float x)
{
  uint16_t hb;
  u32 rb;
  pk_split(x, hb, rb);
  at<uint16_t>(hi, idx) = hb;
  const uintptr_t t = reinterpret_cast<uintptr_t>(lo);
  const int s = (int)(t & 7u);
  u32 *w = &at<u32>(reinterpret_cast<void *>(t - s), idx);
  atomicAnd(w, ~(0xFu << (4 * s)));
  atomicOr(w, rb << (4 * s));
}
// the residual record of seven nibbles, b[s] for nibble s
__device__ __forceinline__ u32 pk_record(const u32 (&b)[7])
{
  return b[0] | (b[1] << 4) | (b[2] << 8) | (b[3] << 12) | (b[4] << 16) | (b[5] << 20) | (b[6] << 24);
}

// value = stored part (+ residual for the compensated fields of the FP16C / FP16T policies;
// FP16M's tau and p with the residual of their packed record)
template <class P>
__device__ __forceinline__ typename P::A load_raw(const void *f, const int8_t *lo, bool comp_f,
                                                  u32 idx)
{
  using S = typename P::S;
  if constexpr (P::packed) {
    if (comp_f) return pk_load(f, lo, idx);
  }
  const S h = at<S>(f, idx);
  typename P::A v = (typename P::A)h;
  if constexpr (P::comp && !P::packed) {
    if (comp_f) v += (float)at<int8_t>(lo, idx) * res_unit(h_exp(h));
  }
  return v;
}
// the value of a cell that only the updating thread reads (its own old value)
template <class P, bool NT>
__device__ __forceinline__ typename P::A load_own(const void *f, const int8_t *lo, bool comp_f, u32 idx)
{
  using S = typename P::S;
  if constexpr (P::packed) {
    if (comp_f) return pk_load(f, lo, idx);
  }
  S h;
  if constexpr (NT) h = __builtin_nontemporal_load(&at<S>(f, idx));
  else h = at<S>(f, idx);
  typename P::A v = (typename P::A)h;
  if constexpr (P::comp && !P::packed) {
    if (comp_f) v += (float)at<int8_t>(lo, idx) * res_unit(h_exp(h));
  }
  return v;
}
template <class P, bool NT = false>
__device__ __forceinline__ void store_raw(void *f, int8_t *lo, bool comp_f, u32 idx,
                                          typename P::A x)
{
  using S = typename P::S;
  if constexpr (P::packed) {
    if (comp_f) {
      pk_store(f, lo, idx, x);
      return;
    }
  }
  const S h = (S)x;
  if constexpr (NT) __builtin_nontemporal_store(h, &at<S>(f, idx));
  else at<S>(f, idx) = h;
  if constexpr (P::comp && !P::packed) {
    if (comp_f) {
      const float r = rintf((x - (float)h) * res_inv(h_exp(h)));
      at<int8_t>(lo, idx) = (int8_t)fminf(fmaxf(r, -127.0f), 127.0f);
    }
  }
}
template <class P, class AT>
__device__ __forceinline__ typename P::A load_rt(const AT &a, int f, u32 idx)
{
  return load_raw<P>(a.f[f], a.lo[f], (P::cmask >> f) & 1u, idx);
}
// compile-time field index: the residual test folds away
// (FP16M's tau and p read the whole residual record, from the base that field PK_FIRST's
// pointer holds, so the fields of one stencil point share one record load)
template <class P, int F, class AT>
__device__ __forceinline__ typename P::A L(const AT &a, u32 idx)
{
  if constexpr (P::packed && ((PK_FIELDS >> F) & 1u)) return pk_at<F>(a.f[F], a.lo[PK_FIRST], idx);
  else return load_rt<P>(a, F, idx);
}
// FP16M and FP16T read their tau and p at a 4th-order stencil's outer taps (the pair with
// weight 1/24 against the inner pair's 9/8) as the binary16 part, without the residual
// (FP16C reads every tap whole)
template <class P, int F>
constexpr bool outer16 = P::outer16 && ((P::cmask >> F) & 1u);
// the value of field F at an outer tap: the binary16 part of FP16M's and FP16T's tau and p,
// the full value of every other field
template <class P, int F, class AT>
__device__ __forceinline__ typename P::A LF(const AT &a, u32 idx)
{
  if constexpr (outer16<P, F>) return pk_hi(a.f[F], idx);
  else return L<P, F>(a, idx);
}
template <class P, int F, bool NT, class AT>
__device__ __forceinline__ typename P::A LO(const AT &a, u32 idx)
{
  if constexpr (P::packed && ((PK_FIELDS >> F) & 1u)) return pk_at<F>(a.f[F], a.lo[PK_FIRST], idx);
  else return load_own<P, NT>(a.f[F], a.lo[F], (P::cmask >> F) & 1u, idx);
}
template <class P, int F, bool NT, class AT>
__device__ __forceinline__ void ST(const AT &a, u32 idx, typename P::A x)
{
  store_raw<P, NT>(a.f[F], a.lo[F], (P::cmask >> F) & 1u, idx, x);
}

// staggered difference across the centre of a 2h-point window w[0..2h-1]
template <int ORDER, typename A>
__device__ __forceinline__ A dwin(const A *w)
{
  if constexpr (ORDER == 2) {
    return w[1] - w[0];
  } else {
    return A(9.0 / 8.0) * (w[2] - w[1]) + A(-1.0 / 24.0) * (w[3] - w[0]);
  }
}

// forward difference of field F at c + s/2 (value at c supplied)
template <class P, int F, int ORDER, class AT>
__device__ __forceinline__ typename P::A dfw(const AT &a, u32 c, u32 s, typename P::A fc)
{
  using A = typename P::A;
  if constexpr (ORDER == 2) {
    return L<P, F>(a, c + s) - fc;
  } else {
    if constexpr (outer16<P, F>) {
      // FP16M, FP16T: both outer taps (c + 2s, c - s) as binary16, each an fma on the binary16 value
      const float o2 = (float)at<_Float16>(a.f[F], c + 2 * s), o1 = (float)at<_Float16>(a.f[F], c - s);
      return fmaf(A(-1.0 / 24.0), o2, fmaf(A(1.0 / 24.0), o1, A(9.0 / 8.0) * (L<P, F>(a, c + s) - fc)));
    } else {
      return A(9.0 / 8.0) * (L<P, F>(a, c + s) - fc) +
             A(-1.0 / 24.0) * (L<P, F>(a, c + 2 * s) - L<P, F>(a, c - s));
    }
  }
}
// backward difference of field F at c - s/2 (value at c supplied)
template <class P, int F, int ORDER, class AT>
__device__ __forceinline__ typename P::A dbw(const AT &a, u32 c, u32 s, typename P::A fc)
{
  using A = typename P::A;
  if constexpr (ORDER == 2) {
    return fc - L<P, F>(a, c - s);
  } else {
    if constexpr (outer16<P, F>) {
      // FP16M, FP16T: both outer taps (c + s, c - 2s) as binary16, each an fma on the binary16 value
      const float o1 = (float)at<_Float16>(a.f[F], c + s), o2 = (float)at<_Float16>(a.f[F], c - 2 * s);
      return fmaf(A(-1.0 / 24.0), o1, fmaf(A(1.0 / 24.0), o2, A(9.0 / 8.0) * (fc - L<P, F>(a, c - s))));
    } else {
      return A(9.0 / 8.0) * (fc - L<P, F>(a, c - s)) +
             A(-1.0 / 24.0) * (L<P, F>(a, c + s) - L<P, F>(a, c - 2 * s));
    }
  }
}

// The N coefficients of cell c from a packed block (biot_packed.hpp layout; nontemporal:
// read once per half step), each times its array's scale for 16-bit storage.
template <class P, int N, class AT>
__device__ __forceinline__ void coef_block(const AT &a, const void *base, u32 c, int k0,
                                           typename P::A (&out)[N])
{
  using S = typename P::S;
  using A = typename P::A;
  uint32_t d[packed::block_dwords<S, N>()];
  packed::load<S, N, true>(base, (uint64_t)a.nx * a.ny * a.nz, c, d);
#pragma unroll
  for (int e = 0; e < N; ++e) {
    if constexpr (P::scaled) out[e] = (A)packed::value<S>(d, e) * a.ks[k0 + e];
    else out[e] = (A)packed::value<S>(d, e);
  }
}

// FP16M medium record of cell c: 4 binary16 values in 8 bytes (mv, mt)
__device__ __forceinline__ float h2f(uint32_t d, int half)
{
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(d >> (16 * half)));
}
__device__ __forceinline__ dw2_t med_rec(const void *base, u32 c) { return at<dw2_t>(base, c); }
__device__ __forceinline__ float med_v(const dw2_t &r, int q) { return h2f(q < 2 ? r.x : r.y, q & 1); }

// The velocity half step's 14 coefficients (packed order K_U1X..K_W3Y) of a cell from the
// FP16M moduli records of the cell (m0) and of its forward neighbours along x, y, z (mn): per
// axis the staggered averages of rho, rho_f, C1, C2 (main.cpp:339-353), then the coefficient
// formulas of main.cpp:447-481 with their zero guards (every W when a C1 average is zero,
// every U when a rho average is).
template <class AT>
__device__ __forceinline__ void medium_velocity(const AT &a, const dw2_t &m0, const dw2_t (&mn)[3],
                                                float (&kv)[14])
{
  bool zw = false, zu = false;
  float W3[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float rho = a.mc[MC_HRHO] * (med_v(m0, 0) + med_v(mn[d], 0));
    const float rhof = a.mc[MC_HRHOF] * (med_v(m0, 1) + med_v(mn[d], 1));
    const float c1 = a.mc[MC_HC1] * (med_v(m0, 2) + med_v(mn[d], 2));
    const float c2 = a.mc[MC_HC2] * (med_v(m0, 3) + med_v(mn[d], 3));
    zw = zw || c1 == 0.0f;
    zu = zu || rho == 0.0f;
    const float rr = __builtin_amdgcn_rcpf(rho);
    const float ratio = rhof * rr;
    const float m = c2 - rhof * ratio, am = a.mc[MC_AH] * c1;
    const float rD = __builtin_amdgcn_rcpf(m + am);
    kv[K_U1X + d] = a.mc[MC_CU1] * rr;
    kv[K_U2X + d] = a.mc[MC_KU2] * ratio;
    kv[K_W1X + d] = (m - am) * rD;
    kv[K_W2X + d] = a.mc[MC_CW2] * ratio * rD;
    W3[d] = a.mc[MC_CW3] * rD;
  }
  kv[K_W3X] = W3[0];
  kv[K_W3Y] = W3[1];
#pragma unroll
  for (int k = K_W1X; k <= K_W3Y; ++k) kv[k] = zw ? 0.0f : kv[k];
#pragma unroll
  for (int k = K_U1X; k <= K_U2Z; ++k) kv[k] = zu ? 0.0f : kv[k];
}

// FP16M: the reciprocals of the stored mu (scaled binary16) of the 2 x 2 cells (i, j),
// (i+1, j), (i, j+1), (i+1, j+1) of cell c's plane; a zero mu gives an infinite reciprocal
template <class AT>
__device__ __forceinline__ void mu_quad(const AT &a, u32 c, u32 sy, float (&r)[4])
{
  r[0] = __builtin_amdgcn_rcpf((float)at<_Float16>(a.mu, c));
  r[1] = __builtin_amdgcn_rcpf((float)at<_Float16>(a.mu, c + 1u));
  r[2] = __builtin_amdgcn_rcpf((float)at<_Float16>(a.mu, c + sy));
  r[3] = __builtin_amdgcn_rcpf((float)at<_Float16>(a.mu, c + 1u + sy));
}

// The stress half step's 8 coefficients (K_P1..K_SYZ) of cell c: P1, P2, S1, S2, S3 from its
// FP16M record (C, M, lambda_c, H; main.cpp:483-490), the shear ones from the 4-cell harmonic
// means of mu of main.cpp:326-337 (xy over the cell's plane, xz and yz over the cell's and the
// next plane; the constant MC_KSH holds the 4 of 4 / sum(1/mu)).  A zero mu in a quad makes
// its sum infinite and the mean zero, the reference's guard.  rq holds the quad of the cell's
// plane and is advanced to the next plane's.
template <class AT>
__device__ __forceinline__ void medium_stress(const AT &a, u32 c, u32 sy, u32 sz, float (&rq)[4],
                                              float (&kt)[8])
{
  const dw2_t d = at<dw2_t>(a.mt, c);
  kt[0] = a.mc[MC_KP1] * h2f(d.x, 0);
  kt[1] = a.mc[MC_KP2] * h2f(d.x, 1);
  kt[2] = a.mc[MC_KS1] * h2f(d.y, 0);
  kt[3] = a.mc[MC_KS2] * h2f(d.y, 1);
  kt[4] = a.mc[MC_KS3] * h2f(d.x, 0);
  float rn[4];
  mu_quad(a, c + sz, sy, rn);
  kt[5] = a.mc[MC_KSH] * __builtin_amdgcn_rcpf((rq[0] + rq[1]) + (rq[2] + rq[3]));
  kt[6] = a.mc[MC_KSH] * __builtin_amdgcn_rcpf((rq[0] + rq[1]) + (rn[0] + rn[1]));
  kt[7] = a.mc[MC_KSH] * __builtin_amdgcn_rcpf((rq[0] + rq[2]) + (rn[0] + rn[2]));
#pragma unroll
  for (int q = 0; q < 4; ++q) rq[q] = rn[q];
}

// One folded C-PML update on a carry held in a register: psi = carry + a D;
// carry <- e (psi + a D); returns D + psi.
template <class P>
__device__ __forceinline__ typename P::A pml(typename P::SC &cv, typename P::A D, typename P::A e,
                                             typename P::A aa)
{
  using A = typename P::A;
  using SC = typename P::SC;
  const A psi = (A)cv + aa * D;
  cv = (SC)(e * (psi + aa * D));
  return D + psi;
}
// the record of four carries of slab cell o in group g (0: velocity half step, 1: stress half
// step): one vector load and one vector store (ordinary: nontemporal accesses measured slower
// here, the records are reused from the memory-side cache between steps)
template <class SC>
__device__ __forceinline__ void carry_load(SC (&cv)[4], const void *base, u32 g, u32 o, u32 sl)
{
  typedef SC V4 __attribute__((ext_vector_type(4)));
  const V4 v = *reinterpret_cast<const V4 *>(reinterpret_cast<const char *>(base) +
                                             (g * sl + o) * (u32)(4 * sizeof(SC)));
  cv[0] = v.x, cv[1] = v.y, cv[2] = v.z, cv[3] = v.w;
}
template <class SC>
__device__ __forceinline__ void carry_store(const SC (&cv)[4], void *base, u32 g, u32 o, u32 sl)
{
  typedef SC V4 __attribute__((ext_vector_type(4)));
  V4 v;
  v.x = cv[0], v.y = cv[1], v.z = cv[2], v.w = cv[3];
  *reinterpret_cast<V4 *>(reinterpret_cast<char *>(base) + (g * sl + o) * (u32)(4 * sizeof(SC))) = v;
}

// slab position of index i on an axis with widths L (low) and R (high), or -1
__device__ __forceinline__ int slab_pos(int i, int n, int Lw, int Rw)
{
  if (i < Lw) return i;
  if (i >= n - Rw) return Lw + i - (n - Rw);
  return -1;
}

// XCD-aware tile assignment.  The dispatcher hands workgroup b of a 1-D grid to XCD
// b % 8 in order; each XCD here gets one contiguous range of tiles, x fastest then y then
// z, so the y- and z-adjacent tiles whose halos overlap are processed by the same XCD,
// close in time, and the halo re-reads hit that XCD's L2.
__device__ __forceinline__ bool tile_of(const StepArgs &a, int &tx, int &ty, int &tz)
{
  const int b = blockIdx.x;
  const int per = (a.ntiles + 7) >> 3;
  const int t = (b & 7) * per + (b >> 3);
  if (t >= a.ntiles) return false;
  tx = t % a.gx;
  const int r = t / a.gx;
  ty = r % a.gy;
  tz = r / a.gy;
  return true;
}

template <typename A>
__device__ __forceinline__ A prof(const void *p, int i)
{
  return at<A>(p, (u32)i);
}

enum { FVX = 0, FVY, FVZ, FQX, FQY, FQZ, FTXX, FTYY, FTZZ, FTXY, FTXZ, FTYZ, FP };

// one x-y tile of the update box, planes kb..ke-1 of z chunk tz
// NT: the launch allows the policy's nontemporal accesses (the single-grid kernels; the
// batched kernels, whose shot windows are reused from the caches, do not)
template <class P, int ORDER, bool NT, class AT>
__device__ __forceinline__ void velocity_tile(const AT &a, int tx, int ty, int tz)
{
  using A = typename P::A;
  constexpr int h = ORDER / 2;
  constexpr int W = 2 * h;
  const int i = tx * blockDim.x + threadIdx.x;
  const int j = ty * blockDim.y + threadIdx.y;
  if (i < h || i >= a.nx - h || j < h || j >= a.ny - h) return;
  const int kb = h + tz * a.zchunk;
  const int ke = min(kb + a.zchunk, a.nz - h);
  if (kb >= ke) return;
  const u32 sy = (u32)a.sy, sz = (u32)a.sz;
  const u32 col = (u32)i + (u32)j * sy;

  // x / y slab positions and profile values are constant along the column
  const int px = slab_pos(i, a.nx, a.L[0], a.R[0]);
  const int py = slab_pos(j, a.ny, a.L[1], a.R[1]);
  A exi = 0, axi = 0, exh = 0, axh = 0, eyi = 0, ayi = 0, eyh = 0, ayh = 0;
  if (px >= 0) {
    exi = prof<A>(a.prof[0][0], i);
    axi = prof<A>(a.prof[0][1], i);
    exh = prof<A>(a.prof[0][2], i);
    axh = prof<A>(a.prof[0][3], i);
  }
  if (py >= 0) {
    eyi = prof<A>(a.prof[1][0], j);
    ayi = prof<A>(a.prof[1][1], j);
    eyh = prof<A>(a.prof[1][2], j);
    ayh = prof<A>(a.prof[1][3], j);
  }
  const u32 wxs = (u32)(a.L[0] + a.R[0]), wys = (u32)(a.L[1] + a.R[1]);
  const u32 ny = (u32)a.ny, nx = (u32)a.nx;
  const u32 sl0 = (u32)a.slab[0], sl1 = (u32)a.slab[1], sl2 = (u32)a.slab[2];

  // z windows: backward-differenced txz, tyz hold planes k-h..k+h-1,
  // forward-differenced tzz, p hold planes k-h+1..k+h.
  // FP16M at order 4 (zrec): the forward windows' newest plane is an outer tap and holds the
  // binary16 part alone until the next plane completes it, the first one included
  constexpr bool zrec = P::packed && ORDER == 4;
  A wxz[W], wyz[W], wzz[W], wp[W];
#pragma unroll
  for (int m = 0; m < W - 1; ++m) {
    wxz[m] = L<P, FTXZ>(a, col + (u32)(kb - h + m) * sz);
    wyz[m] = L<P, FTYZ>(a, col + (u32)(kb - h + m) * sz);
    if (zrec && m == W - 2) {
      wzz[m] = LF<P, FTZZ>(a, col + (u32)(kb - h + 1 + m) * sz);
      wp[m] = LF<P, FP>(a, col + (u32)(kb - h + 1 + m) * sz);
    } else {
      wzz[m] = L<P, FTZZ>(a, col + (u32)(kb - h + 1 + m) * sz);
      wp[m] = L<P, FP>(a, col + (u32)(kb - h + 1 + m) * sz);
    }
  }

  // FP16M: the cell's moduli record, rotated along z (the forward-z neighbour's record of one
  // plane is the cell's own record of the next)
  dw2_t mown = {0u, 0u};
  if constexpr (P::medium) mown = med_rec(a.mv, col + (u32)kb * sz);
  // FP16M at order 4: the cell's residual record, rotated along z like the moduli record
  u32 rown = 0u;
  if constexpr (zrec) rown = at<u32>(a.lo[PK_FIRST], col + (u32)kb * sz);

  // Every load of a plane is written before the plane's first store.  The compiler cannot
  // prove the arrays distinct, so it never moves a load above a store; in this order all of
  // a plane's loads (fields, coefficients, C-PML carries) are in flight together and a
  // plane costs one memory round trip instead of one per store.
  using SC = typename P::SC;
  for (int k = kb; k < ke; ++k) {
    const u32 c = col + (u32)k * sz;
    // twelve raw staggered differences (kernel.cu:172-185 without the 1/dx factor)
    A x1, x2, x3, s1, y1, y2, y3, s2, z1, z2, z3, s3;
    {
      A txx_c, tyy_c, txy_c;
      if constexpr (zrec) {
        // One residual record a plane, plane k+1's: it completes the backward windows' new
        // plane and the forward windows' plane k+1 (read as its binary16 part a plane
        // earlier), and is the next plane's own record.  The forward windows' new plane k+2
        // is an outer tap: its binary16 part.
        const u32 r1 = at<u32>(a.lo[PK_FIRST], c + sz);
        wxz[W - 1] = pk_rec<FTXZ>(a.f[FTXZ], r1, c + sz);
        wyz[W - 1] = pk_rec<FTYZ>(a.f[FTYZ], r1, c + sz);
        wzz[W - 2] = pk_fill<FTZZ>(wzz[W - 2], r1);
        wp[W - 2] = pk_fill<FP>(wp[W - 2], r1);
        wzz[W - 1] = pk_hi(a.f[FTZZ], c + 2u * sz);
        wp[W - 1] = pk_hi(a.f[FP], c + 2u * sz);
        txx_c = pk_rec<FTXX>(a.f[FTXX], rown, c);
        tyy_c = pk_rec<FTYY>(a.f[FTYY], rown, c);
        txy_c = pk_rec<FTXY>(a.f[FTXY], rown, c);
        rown = r1;
      } else {
        wxz[W - 1] = L<P, FTXZ>(a, c + (u32)(h - 1) * sz);
        wyz[W - 1] = L<P, FTYZ>(a, c + (u32)(h - 1) * sz);
        wzz[W - 1] = L<P, FTZZ>(a, c + (u32)h * sz);
        wp[W - 1] = L<P, FP>(a, c + (u32)h * sz);
        txx_c = L<P, FTXX>(a, c), tyy_c = L<P, FTYY>(a, c), txy_c = L<P, FTXY>(a, c);
      }
      const A txz_c = wxz[h], tyz_c = wyz[h], p_c = wp[h - 1];
      x1 = dfw<P, FTXX, ORDER>(a, c, 1u, txx_c);
      x2 = dbw<P, FTXY, ORDER>(a, c, sy, txy_c);
      x3 = dwin<ORDER>(wxz);
      s1 = dfw<P, FP, ORDER>(a, c, 1u, p_c);
      y1 = dfw<P, FTYY, ORDER>(a, c, sy, tyy_c);
      y2 = dbw<P, FTXY, ORDER>(a, c, 1u, txy_c);
      y3 = dwin<ORDER>(wyz);
      s2 = dfw<P, FP, ORDER>(a, c, sy, p_c);
      z1 = dwin<ORDER>(wzz);
      z2 = dbw<P, FTXZ, ORDER>(a, c, 1u, txz_c);
      z3 = dbw<P, FTYZ, ORDER>(a, c, sy, tyz_c);
      s3 = dwin<ORDER>(wp);
    }

    const A qx0 = LO<P, FQX, (P::nt_v && NT)>(a, c), qy0 = LO<P, FQY, (P::nt_v && NT)>(a, c), qz0 = LO<P, FQZ, (P::nt_v && NT)>(a, c);
    const A vx0 = LO<P, FVX, (P::nt_v && NT)>(a, c), vy0 = LO<P, FVY, (P::nt_v && NT)>(a, c), vz0 = LO<P, FVZ, (P::nt_v && NT)>(a, c);
    A kv[14];  // K[0..13]: U1 xyz, U2 xyz, W1 xyz, W2 xyz, W3x, W3y
    dw2_t mn[3];
    if constexpr (P::medium) {
      mn[0] = med_rec(a.mv, c + 1u);
      mn[1] = med_rec(a.mv, c + sy);
      mn[2] = med_rec(a.mv, c + sz);
      medium_velocity(a, mown, mn, kv);
    } else {
      coef_block<P, 14>(a, a.kv, c, 0, kv);
    }
    const A w1x = kv[K_W1X], w2x = kv[K_W2X], w3x = kv[K_W3X];
    const A w1y = kv[K_W1Y], w2y = kv[K_W2Y], w3y = kv[K_W3Y];
    const A w1z = kv[K_W1Z], w2z = kv[K_W2Z];
    const A u1x = kv[K_U1X], u2x = kv[K_U2X];
    const A u1y = kv[K_U1Y], u2y = kv[K_U2Y];
    const A u1z = kv[K_U1Z], u2z = kv[K_U2Z];

    // C-PML carries of the slabs this cell lies in (kernel.cu:188-216): x memory of x1, s1
    // (half node), y2, z2 (integer); y memory of x2, y1, s2 (half), z3 (integer); z memory of
    // x3, y3 (integer), z1, s3 (half)
    const int pz = slab_pos(k, a.nz, a.L[2], a.R[2]);
    SC cx[4], cy[4], cz[4];
    u32 ox = 0, oy = 0, oz = 0;
    A ezi = 0, azi = 0, ezh = 0, azh = 0;
    if (px >= 0) {
      ox = (u32)px + wxs * ((u32)j + ny * (u32)k);
      carry_load(cx, a.carry[0], 0u, ox, sl0);
    }
    if (py >= 0) {
      oy = (u32)i + nx * ((u32)py + wys * (u32)k);
      carry_load(cy, a.carry[1], 0u, oy, sl1);
    }
    if (pz >= 0) {
      ezi = prof<A>(a.prof[2][0], k), azi = prof<A>(a.prof[2][1], k);
      ezh = prof<A>(a.prof[2][2], k), azh = prof<A>(a.prof[2][3], k);
      oz = (u32)i + nx * ((u32)j + ny * (u32)pz);
      carry_load(cz, a.carry[2], 0u, oz, sl2);
    }
    if (px >= 0) {
      x1 = pml<P>(cx[0], x1, exh, axh);
      s1 = pml<P>(cx[1], s1, exh, axh);
      y2 = pml<P>(cx[2], y2, exi, axi);
      z2 = pml<P>(cx[3], z2, exi, axi);
      carry_store(cx, a.carry[0], 0u, ox, sl0);
    }
    if (py >= 0) {
      x2 = pml<P>(cy[0], x2, eyi, ayi);
      y1 = pml<P>(cy[1], y1, eyh, ayh);
      s2 = pml<P>(cy[2], s2, eyh, ayh);
      z3 = pml<P>(cy[3], z3, eyi, ayi);
      carry_store(cy, a.carry[1], 0u, oy, sl1);
    }
    if (pz >= 0) {
      x3 = pml<P>(cz[0], x3, ezi, azi);
      y3 = pml<P>(cz[1], y3, ezi, azi);
      z1 = pml<P>(cz[2], z1, ezh, azh);
      s3 = pml<P>(cz[3], s3, ezh, azh);
      carry_store(cz, a.carry[2], 0u, oz, sl2);
    }

    // relative fluid velocity (kernel.cu:218-220; qz takes W3y as the reference does) and
    // solid velocity with the stored new q (kernel.cu:223-225)
    const A sx = x1 + x2 + x3, sy_ = y1 + y2 + y3, sz_ = z1 + z2 + z3;
    const A qx1 = w1x * qx0 - w2x * sx - w3x * s1;
    const A qy1 = w1y * qy0 - w2y * sy_ - w3y * s2;
    const A qz1 = w1z * qz0 - w2z * sz_ - w3y * s3;
    ST<P, FQX, (P::nt_v && NT)>(a, c, qx1);
    ST<P, FQY, (P::nt_v && NT)>(a, c, qy1);
    ST<P, FQZ, (P::nt_v && NT)>(a, c, qz1);
    // the solid update sees q as stored (rounded), as the reference re-reads it; q is
    // never a compensated field, so the stored value is the plain storage rounding
    using S = typename P::S;
    const A qx1s = (A)(S)qx1, qy1s = (A)(S)qy1, qz1s = (A)(S)qz1;
    ST<P, FVX, (P::nt_v && NT)>(a, c, vx0 + u1x * sx - u2x * (qx1s - qx0));
    ST<P, FVY, (P::nt_v && NT)>(a, c, vy0 + u1y * sy_ - u2y * (qy1s - qy0));
    ST<P, FVZ, (P::nt_v && NT)>(a, c, vz0 + u1z * sz_ - u2z * (qz1s - qz0));

#pragma unroll
    for (int m = 0; m < W - 1; ++m) {
      wxz[m] = wxz[m + 1];
      wyz[m] = wyz[m + 1];
      wzz[m] = wzz[m + 1];
      wp[m] = wp[m + 1];
    }
    if constexpr (P::medium) {
      mown = mn[2];
    }
  }
}

// one x-y tile of the update box, planes kb..ke-1 of z chunk tz
template <class P, int ORDER, bool NT, class AT>
__device__ __forceinline__ void stress_tile(const AT &a, int tx, int ty, int tz)
{
  using A = typename P::A;
  constexpr int h = ORDER / 2;
  constexpr int W = 2 * h;
  const int i = tx * blockDim.x + threadIdx.x;
  const int j = ty * blockDim.y + threadIdx.y;
  if (i < h || i >= a.nx - h || j < h || j >= a.ny - h) return;
  const int kb = h + tz * a.zchunk;
  const int ke = min(kb + a.zchunk, a.nz - h);
  if (kb >= ke) return;
  const u32 sy = (u32)a.sy, sz = (u32)a.sz;
  const u32 col = (u32)i + (u32)j * sy;

  const int px = slab_pos(i, a.nx, a.L[0], a.R[0]);
  const int py = slab_pos(j, a.ny, a.L[1], a.R[1]);
  A exi = 0, axi = 0, exh = 0, axh = 0, eyi = 0, ayi = 0, eyh = 0, ayh = 0;
  if (px >= 0) {
    exi = prof<A>(a.prof[0][0], i);
    axi = prof<A>(a.prof[0][1], i);
    exh = prof<A>(a.prof[0][2], i);
    axh = prof<A>(a.prof[0][3], i);
  }
  if (py >= 0) {
    eyi = prof<A>(a.prof[1][0], j);
    ayi = prof<A>(a.prof[1][1], j);
    eyh = prof<A>(a.prof[1][2], j);
    ayh = prof<A>(a.prof[1][3], j);
  }
  const u32 wxs = (u32)(a.L[0] + a.R[0]), wys = (u32)(a.L[1] + a.R[1]);
  const u32 ny = (u32)a.ny, nx = (u32)a.nx;
  const u32 sl0 = (u32)a.slab[0], sl1 = (u32)a.slab[1], sl2 = (u32)a.slab[2];

  // z windows: backward-differenced vz, qz hold planes k-h..k+h-1,
  // forward-differenced vx, vy hold planes k-h+1..k+h.
  A wvz[W], wqz[W], wvx[W], wvy[W];
#pragma unroll
  for (int m = 0; m < W - 1; ++m) {
    wvz[m] = L<P, FVZ>(a, col + (u32)(kb - h + m) * sz);
    wqz[m] = L<P, FQZ>(a, col + (u32)(kb - h + m) * sz);
    wvx[m] = L<P, FVX>(a, col + (u32)(kb - h + 1 + m) * sz);
    wvy[m] = L<P, FVY>(a, col + (u32)(kb - h + 1 + m) * sz);
  }

  // FP16M: the mu quad of the first plane, advanced along z by medium_stress
  float rq[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if constexpr (P::medium) mu_quad(a, col + (u32)kb * sz, sy, rq);

  // all of a plane's loads before its first store (see velocity_tile)
  using SC = typename P::SC;
  for (int k = kb; k < ke; ++k) {
    const u32 c = col + (u32)k * sz;
    wvz[W - 1] = L<P, FVZ>(a, c + (u32)(h - 1) * sz);
    wqz[W - 1] = L<P, FQZ>(a, c + (u32)(h - 1) * sz);
    wvx[W - 1] = L<P, FVX>(a, c + (u32)h * sz);
    wvy[W - 1] = L<P, FVY>(a, c + (u32)h * sz);
    const A vz_c = wvz[h], vx_c = wvx[h - 1], vy_c = wvy[h - 1];
    const A qx_c = L<P, FQX>(a, c), qy_c = L<P, FQY>(a, c);

    // twelve raw differences (kernel.cu:430-445 without 1/dx)
    A uxx = dbw<P, FVX, ORDER>(a, c, 1u, vx_c);
    A uyy = dbw<P, FVY, ORDER>(a, c, sy, vy_c);
    A uzz = dwin<ORDER>(wvz);
    A wx = dbw<P, FQX, ORDER>(a, c, 1u, qx_c);
    A wy = dbw<P, FQY, ORDER>(a, c, sy, qy_c);
    A wz = dwin<ORDER>(wqz);
    A uxy = dfw<P, FVX, ORDER>(a, c, sy, vx_c);
    A uyx = dfw<P, FVY, ORDER>(a, c, 1u, vy_c);
    A uxz = dwin<ORDER>(wvx);
    A uzx = dfw<P, FVZ, ORDER>(a, c, 1u, vz_c);
    A uyz = dwin<ORDER>(wvy);
    A uzy = dfw<P, FVZ, ORDER>(a, c, sy, vz_c);

    const A p0 = LO<P, FP, (P::nt_t && NT)>(a, c), txx0 = LO<P, FTXX, (P::nt_t && NT)>(a, c), tyy0 = LO<P, FTYY, (P::nt_t && NT)>(a, c);
    const A tzz0 = LO<P, FTZZ, (P::nt_t && NT)>(a, c), txy0 = LO<P, FTXY, (P::nt_t && NT)>(a, c), tyz0 = LO<P, FTYZ, (P::nt_t && NT)>(a, c);
    const A txz0 = LO<P, FTXZ, (P::nt_t && NT)>(a, c);
    A kt[8];  // K[14..21]: P1, P2, S1, S2, S3, Sxy, Sxz, Syz
    if constexpr (P::medium) medium_stress(a, c, sy, sz, rq, kt);
    else coef_block<P, 8>(a, a.kt, c, K_P1, kt);
    const A s1 = kt[K_S1 - K_P1], s2 = kt[K_S2 - K_P1], s3 = kt[K_S3 - K_P1];
    const A kp1 = kt[0], kp2 = kt[K_P2 - K_P1];
    const A kxy = kt[K_SXY - K_P1], kyz = kt[K_SYZ - K_P1], kxz = kt[K_SXZ - K_P1];

    // C-PML carries 4..7 of each axis (kernel.cu:449-482)
    const int pz = slab_pos(k, a.nz, a.L[2], a.R[2]);
    SC cx[4], cy[4], cz[4];
    u32 ox = 0, oy = 0, oz = 0;
    A ezi = 0, azi = 0, ezh = 0, azh = 0;
    if (px >= 0) {
      ox = (u32)px + wxs * ((u32)j + ny * (u32)k);
      carry_load(cx, a.carry[0], 1u, ox, sl0);
    }
    if (py >= 0) {
      oy = (u32)i + nx * ((u32)py + wys * (u32)k);
      carry_load(cy, a.carry[1], 1u, oy, sl1);
    }
    if (pz >= 0) {
      ezi = prof<A>(a.prof[2][0], k), azi = prof<A>(a.prof[2][1], k);
      ezh = prof<A>(a.prof[2][2], k), azh = prof<A>(a.prof[2][3], k);
      oz = (u32)i + nx * ((u32)j + ny * (u32)pz);
      carry_load(cz, a.carry[2], 1u, oz, sl2);
    }
    if (px >= 0) {
      uxx = pml<P>(cx[0], uxx, exi, axi);
      wx = pml<P>(cx[1], wx, exi, axi);
      uyx = pml<P>(cx[2], uyx, exh, axh);
      uzx = pml<P>(cx[3], uzx, exh, axh);
      carry_store(cx, a.carry[0], 1u, ox, sl0);
    }
    if (py >= 0) {
      uyy = pml<P>(cy[0], uyy, eyi, ayi);
      wy = pml<P>(cy[1], wy, eyi, ayi);
      uxy = pml<P>(cy[2], uxy, eyh, ayh);
      uzy = pml<P>(cy[3], uzy, eyh, ayh);
      carry_store(cy, a.carry[1], 1u, oy, sl1);
    }
    if (pz >= 0) {
      uzz = pml<P>(cz[0], uzz, ezi, azi);
      wz = pml<P>(cz[1], wz, ezi, azi);
      uxz = pml<P>(cz[2], uxz, ezh, azh);
      uyz = pml<P>(cz[3], uyz, ezh, azh);
      carry_store(cz, a.carry[2], 1u, oz, sl2);
    }

    // pore pressure and stresses (kernel.cu:484-490)
    const A divw = wx + wy + wz;
    const A pn = p0 - kp1 * (uxx + uyy + uzz) - kp2 * divw;
    const A txxn = txx0 + s1 * (uyy + uzz) + s2 * uxx + s3 * divw;
    const A tyyn = tyy0 + s1 * (uxx + uzz) + s2 * uyy + s3 * divw;
    const A tzzn = tzz0 + s1 * (uxx + uyy) + s2 * uzz + s3 * divw;
    const A txyn = txy0 + kxy * (uxy + uyx);
    const A tyzn = tyz0 + kyz * (uyz + uzy);
    const A txzn = txz0 + kxz * (uxz + uzx);
    if constexpr (P::packed) {
      // binary16 parts to the field arrays, the seven residuals in one record store
      uint16_t hb[7];
      u32 rb[7];
      pk_split(txxn, hb[pk_slot(FTXX)], rb[pk_slot(FTXX)]);
      pk_split(tyyn, hb[pk_slot(FTYY)], rb[pk_slot(FTYY)]);
      pk_split(tzzn, hb[pk_slot(FTZZ)], rb[pk_slot(FTZZ)]);
      pk_split(txyn, hb[pk_slot(FTXY)], rb[pk_slot(FTXY)]);
      pk_split(txzn, hb[pk_slot(FTXZ)], rb[pk_slot(FTXZ)]);
      pk_split(tyzn, hb[pk_slot(FTYZ)], rb[pk_slot(FTYZ)]);
      pk_split(pn, hb[pk_slot(FP)], rb[pk_slot(FP)]);
      at<uint16_t>(a.f[FTXX], c) = hb[pk_slot(FTXX)];
      at<uint16_t>(a.f[FTYY], c) = hb[pk_slot(FTYY)];
      at<uint16_t>(a.f[FTZZ], c) = hb[pk_slot(FTZZ)];
      at<uint16_t>(a.f[FTXY], c) = hb[pk_slot(FTXY)];
      at<uint16_t>(a.f[FTXZ], c) = hb[pk_slot(FTXZ)];
      at<uint16_t>(a.f[FTYZ], c) = hb[pk_slot(FTYZ)];
      at<uint16_t>(a.f[FP], c) = hb[pk_slot(FP)];
      at<u32>(a.lo[PK_FIRST], c) = pk_record(rb);
    } else {
      ST<P, FP, (P::nt_t && NT)>(a, c, pn);
      ST<P, FTXX, (P::nt_t && NT)>(a, c, txxn);
      ST<P, FTYY, (P::nt_t && NT)>(a, c, tyyn);
      ST<P, FTZZ, (P::nt_t && NT)>(a, c, tzzn);
      ST<P, FTXY, (P::nt_t && NT)>(a, c, txyn);
      ST<P, FTYZ, (P::nt_t && NT)>(a, c, tyzn);
      ST<P, FTXZ, (P::nt_t && NT)>(a, c, txzn);
    }

#pragma unroll
    for (int m = 0; m < W - 1; ++m) {
      wvz[m] = wvz[m + 1];
      wqz[m] = wqz[m + 1];
      wvx[m] = wvx[m + 1];
      wvy[m] = wvy[m + 1];
    }
  }
}

template <class P, int ORDER>
__global__ void __launch_bounds__(256) k_velocity(const StepArgs a)
{
  int tx, ty, tz;
  if (!tile_of(a, tx, ty, tz)) return;
  velocity_tile<P, ORDER, true>(a, tx, ty, tz);
}
template <class P, int ORDER>
__global__ void __launch_bounds__(256) k_stress(const StepArgs a)
{
  int tx, ty, tz;
  if (!tile_of(a, tx, ty, tz)) return;
  stress_tile<P, ORDER, true>(a, tx, ty, tz);
}

// Batched forms: one persistent grid walks the tiles of every shot.  The shots' tile
// lists are concatenated (tile_start[s] = first global tile of shot s); each XCD takes one
// contiguous range of that list, so a shot's neighbouring tiles share an L2, and a block
// strides through its XCD's range until it is exhausted (every wave exits).
// The argument tables are read-only for a launch and indexed uniformly per block: read
// them through the constant address space, so the compiler keeps them in scalar registers
// (as it does kernel arguments) instead of a VGPR pair per pointer.
typedef __attribute__((address_space(4))) const StepArgs CStepArgs;
typedef __attribute__((address_space(4))) const int CInt;
__device__ __forceinline__ int shot_of(const BatchArgs &b, int t)
{
  CInt *ts = ((CInt *)b.tile_start);
  int lo = 0, hi = b.nshots - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ts[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
template <class P, int ORDER, bool VEL>
__global__ void __launch_bounds__(256) k_step_batch(const BatchArgs b)
{
  // workgroup b goes to XCD b % 8; XCD group x owns the contiguous tile range [begin, end) of
  // the concatenated tile lists
  const int x = blockIdx.x & 7;
  const int per = (b.total_tiles + 7) >> 3;
  const int begin = x * per, end = min((x + 1) * per, b.total_tiles);
  auto tile = [&](int t) {
    const int s = shot_of(b, t);
    const CStepArgs &a = ((CStepArgs *)b.shots)[s];
    const int lt = t - ((CInt *)b.tile_start)[s];
    const int tx = lt % a.gx, r = lt / a.gx, ty = r % a.gy, tz = r / a.gy;
    if (VEL) velocity_tile<P, ORDER, false>(a, tx, ty, tz);
    else stress_tile<P, ORDER, false>(a, tx, ty, tz);
  };
  // Both mappings run one loop, so the tile body is inlined once (two inlined copies took the
  // residual policies' kernels to 248-256 VGPRs).
  // Fixed mapping (no queue): block l of the group takes tiles l, l + per_x, ...
  // Work queue: a block takes the group's next tile from the group's counter (one atomic a
  // tile, the next one claimed while the current one runs), so blocks that drew cheap tiles
  // take more.  Every block ends on one claim past the range; the last block of the group to
  // make it sets the group's two counters back to zero for the next launch (no other block
  // of the group touches them after its own last claim).  The branches on the queue are
  // uniform over the launch, so the barriers inside them are reached by every thread.
  __shared__ int s_next;
  const bool queue = b.queue != nullptr;
  const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
  int *ctr = queue ? b.queue + x : nullptr, *done = queue ? b.queue + 8 + x : nullptr;
  int t = begin + (blockIdx.x >> 3);
  if (queue) {
    if (lead) s_next = atomicAdd(ctr, 1);
    __syncthreads();
    t = begin + s_next;
  }
  while (t < end) {
    if (queue) {
      __syncthreads();  // every thread has read s_next
      if (lead) s_next = atomicAdd(ctr, 1);
    }
    tile(t);
    if (queue) {
      __syncthreads();  // the next claim is in s_next
      t = begin + s_next;
    } else {
      t += gridDim.x >> 3;
    }
  }
  if (queue && lead && atomicAdd(done, 1) == (int)(gridDim.x >> 3) - 1) {
    atomicExch(ctr, 0);
    atomicExch(done, 0);
  }
}

// Source injection (kernel.cu:130-142) for one cell: one lane per field.
template <class P>
__device__ __forceinline__ void source_cell(const SourceArgs &sa, int it)
{
  using A = typename P::A;
  const int f = threadIdx.x;
  if (f >= NFIELD || !(sa.mask & (1u << f))) return;
  const u32 c = (u32)sa.cell;
  const bool cf = (P::cmask >> f) & 1u;
  const A amp = (A)sa.wavelet[it] * (A)sa.inv_scale[f];
  store_raw<P>(sa.f[f], sa.lo[f], cf, c, load_raw<P>(sa.f[f], sa.lo[f], cf, c) + amp);
}
template <class P>
__global__ void k_source(const SourceArgs sa)
{
  source_cell<P>(sa, sa.it);
}

// Receiver sampling: traces[comp][it][rec] = field value * storage scale.
template <class P>
__device__ __forceinline__ void record_one(const RecordArgs &ra, int64_t r, int ci, int it)
{
  if (r >= ra.nrec || ci >= ra.ncomp) return;
  const int f = ra.comp[ci];
  // the product of two floats is exact in double, so (float)v is the rounded float product
  const double v = (double)load_raw<P>(ra.f[f], ra.lo[f], (P::cmask >> f) & 1u, (u32)ra.rec[r]) *
                   (double)ra.scale[ci];
  const int64_t o = ((int64_t)ci * ra.nt_max + it) * ra.nrec + r;
  const float vf = (float)v;
  ra.traces[o] = vf;
  if (ra.traces64) ra.traces64[o] = v;
  // the watchdog of the reference's abort on a bad state (kernel.cu:42-52): the first step
  // at which a receiver sees NaN or Inf
  if (ra.nonfinite && !__builtin_isfinite(vf)) atomicMin(ra.nonfinite, it);
}
template <class P>
__global__ void k_record(const RecordArgs ra)
{
  record_one<P>(ra, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, (int)blockIdx.y, ra.it);
}

// Batched source / record: one block (z) per shot; the step index comes from a device
// counter that k_advance moves on, so a captured sequence of steps can be replayed.
template <class P>
__global__ void k_source_batch(const SourceArgs *tab, const int *itp)
{
  const SourceArgs &sa = tab[blockIdx.x];
  const int it = *itp;
  if (!sa.wavelet || it >= sa.nt_w) return;
  source_cell<P>(sa, it);
}
template <class P>
__global__ void k_record_batch(const RecordArgs *tab, const int *itp)
{
  const RecordArgs &ra = tab[blockIdx.z];
  const int it = *itp;
  if (!ra.traces || it >= ra.nt_max) return;
  record_one<P>(ra, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, (int)blockIdx.y, it);
}
__global__ void k_advance(int *itp)
{
  if (threadIdx.x == 0) atomicAdd(itp, 1);
}
__global__ void k_set_counter(int *itp, int v)
{
  if (threadIdx.x == 0) atomicExch(itp, v);
}

// In-place multiply of an FP16M tau or p field (lo: its tagged residual pointer).
__global__ void k_rescale_pk(void *p, int8_t *lo, int64_t n, double f)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    pk_store(p, lo, (u32)i, pk_load(p, lo, (u32)i) * (float)f);
}

// In-place multiply of a stored array (re-expresses values when a storage scale changes).
template <typename S>
__global__ void k_rescale(S *p, int8_t *lo, int64_t n, double f)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if constexpr (std::is_same<S, double>::value) {
      p[i] = p[i] * f;
    } else if constexpr (std::is_same<S, _Float16>::value) {
      float x = (float)p[i];
      if (lo) x += (float)lo[i] * res_unit(h_exp(p[i]));
      x *= (float)f;
      const _Float16 hh = (_Float16)x;
      p[i] = hh;
      if (lo) lo[i] = (int8_t)fminf(fmaxf(rintf((x - (float)hh) * res_inv(h_exp(hh))), -127.0f), 127.0f);
    } else {
      p[i] = (S)((float)p[i] * (float)f);
    }
  }
}

// xz slice of a field in physical units; the product of the stored value and the scale is
// formed in double and rounded once, as the host read-back does
template <class P>
__global__ void k_slice_xz(const void *f, const int8_t *lo, bool comp, double scale, int j, int nx, int ny,
                           float *out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
  if (i >= nx) return;
  const u32 c = (u32)i + (u32)nx * ((u32)j + (u32)ny * (u32)k);
  out[(int64_t)k * nx + i] = (float)((double)load_raw<P>(f, lo, comp, c) * scale);
}

// packing of the coefficient arrays into the velocity block (K[0..13]) and the stress block
// (K[14..21]), each in the group-major 16-byte layout coef_block reads
template <typename S>
__device__ __forceinline__ void pack_block(const S *soa, S *blk, int k0, int N, int64_t n, int64_t c)
{
  constexpr int G = 16 / (int)sizeof(S);
  const int NG = N / G, T = N % G;
  for (int e = 0; e < N; ++e) {
    const int g = e / G;
    const S v = soa[(int64_t)(k0 + e) * n + c];
    if (g < NG) blk[(int64_t)g * n * G + c * G + (e % G)] = v;
    else blk[(int64_t)NG * n * G + c * T + (e - NG * G)] = v;
  }
}
template <typename S>
__global__ void k_pack_coef(const S *soa, S *kv, S *kt, int64_t n)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) {
    pack_block(soa, kv, 0, 14, n, c);
    pack_block(soa, kt, 14, 8, n, c);
  }
}

template <typename S>
__global__ void k_rescale4(S *p, int64_t nrec, double f0, double f1, double f2, double f3)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 4 * nrec; i += stride) {
    const int e = (int)(i & 3);
    const double f = e == 0 ? f0 : (e == 1 ? f1 : (e == 2 ? f2 : f3));
    if constexpr (std::is_same<S, double>::value) p[i] = p[i] * f;
    else p[i] = (S)((float)p[i] * (float)f);
  }
}

template <class P, int ORDER, bool VEL>
hipError_t launch_step(const StepArgs &in, LaunchShape sh, hipStream_t s)
{
  StepArgs a = in;
  const int h = ORDER / 2;
  const int nzi = a.nz - 2 * h;
  if (nzi <= 0 || a.nx <= 2 * h || a.ny <= 2 * h) return hipSuccess;
  if (sh.zchunk <= 0) return hipErrorInvalidValue;
  a.zchunk = sh.zchunk;
  a.gx = (a.nx + sh.bx - 1) / sh.bx;
  a.gy = (a.ny + sh.by - 1) / sh.by;
  a.gz = (nzi + a.zchunk - 1) / a.zchunk;
  a.ntiles = a.gx * a.gy * a.gz;
  dim3 block(sh.bx, sh.by, 1);
  dim3 grid((unsigned)(((a.ntiles + 7) / 8) * 8));
  if (VEL)
    hipLaunchKernelGGL((k_velocity<P, ORDER>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_stress<P, ORDER>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <bool VEL, int ORDER>
hipError_t dispatch_order(int prec, const StepArgs &a, LaunchShape sh, hipStream_t s)
{
  switch (prec) {
    case PREC_FP32: return launch_step<P32, ORDER, VEL>(a, sh, s);
    case PREC_FP16: return launch_step<P16, ORDER, VEL>(a, sh, s);
    case PREC_FP16C: return launch_step<P16C, ORDER, VEL>(a, sh, s);
    case PREC_FP16T: return launch_step<P16T, ORDER, VEL>(a, sh, s);
    case PREC_FP16M: return launch_step<P16M, ORDER, VEL>(a, sh, s);
    case PREC_FP64: return launch_step<P64, ORDER, VEL>(a, sh, s);
  }
  return hipErrorInvalidValue;
}

template <bool VEL>
hipError_t dispatch(int prec, int order, const StepArgs &a, LaunchShape sh, hipStream_t s)
{
  if (sh.bx * sh.by > 256 || sh.bx <= 0 || sh.by <= 0) return hipErrorInvalidValue;
  if (order == 2) return dispatch_order<VEL, 2>(prec, a, sh, s);
  if (order == 4) return dispatch_order<VEL, 4>(prec, a, sh, s);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_step_batch(int prec, int order, bool vel, const BatchArgs &b, LaunchShape sh,
                             int blocks, hipStream_t s)
{
  if (sh.bx * sh.by > 256 || sh.bx <= 0 || sh.by <= 0 || blocks <= 0 || blocks % 8) return hipErrorInvalidValue;
  if (b.total_tiles <= 0) return hipSuccess;
  dim3 grid((unsigned)blocks), block(sh.bx, sh.by, 1);
#define BIOT_BATCH_CASE(PP, OO)                                                                 \
  if (vel) hipLaunchKernelGGL((k_step_batch<PP, OO, true>), grid, block, 0, s, b);              \
  else hipLaunchKernelGGL((k_step_batch<PP, OO, false>), grid, block, 0, s, b);
  if (order != 2 && order != 4) return hipErrorInvalidValue;
  switch (prec) {
    case PREC_FP32:
      if (order == 2) { BIOT_BATCH_CASE(P32, 2) } else { BIOT_BATCH_CASE(P32, 4) }
      break;
    case PREC_FP16:
      if (order == 2) { BIOT_BATCH_CASE(P16, 2) } else { BIOT_BATCH_CASE(P16, 4) }
      break;
    case PREC_FP16C:
      if (order == 2) { BIOT_BATCH_CASE(P16C, 2) } else { BIOT_BATCH_CASE(P16C, 4) }
      break;
    case PREC_FP16T:
      if (order == 2) { BIOT_BATCH_CASE(P16T, 2) } else { BIOT_BATCH_CASE(P16T, 4) }
      break;
    case PREC_FP16M:
      if (order == 2) { BIOT_BATCH_CASE(P16M, 2) } else { BIOT_BATCH_CASE(P16M, 4) }
      break;
    case PREC_FP64:
      if (order == 2) { BIOT_BATCH_CASE(P64, 2) } else { BIOT_BATCH_CASE(P64, 4) }
      break;
    default: return hipErrorInvalidValue;
  }
#undef BIOT_BATCH_CASE
  return hipGetLastError();
}

hipError_t launch_source_batch(int prec, const SourceArgs *tab, int nshots, const int *it, hipStream_t s)
{
  if (nshots <= 0) return hipSuccess;
  switch (prec) {
    case PREC_FP32: hipLaunchKernelGGL(k_source_batch<P32>, dim3(nshots), dim3(64), 0, s, tab, it); break;
    case PREC_FP16: hipLaunchKernelGGL(k_source_batch<P16>, dim3(nshots), dim3(64), 0, s, tab, it); break;
    case PREC_FP16C: hipLaunchKernelGGL(k_source_batch<P16C>, dim3(nshots), dim3(64), 0, s, tab, it); break;
    case PREC_FP16T: hipLaunchKernelGGL(k_source_batch<P16T>, dim3(nshots), dim3(64), 0, s, tab, it); break;
    case PREC_FP16M: hipLaunchKernelGGL(k_source_batch<P16M>, dim3(nshots), dim3(64), 0, s, tab, it); break;
    case PREC_FP64: hipLaunchKernelGGL(k_source_batch<P64>, dim3(nshots), dim3(64), 0, s, tab, it); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_record_batch(int prec, const RecordArgs *tab, int nshots, int64_t max_nrec,
                               int max_ncomp, const int *it, hipStream_t s)
{
  if (nshots <= 0 || max_nrec <= 0 || max_ncomp <= 0) return hipSuccess;
  dim3 grid((unsigned)((max_nrec + 255) / 256), (unsigned)max_ncomp, (unsigned)nshots), block(256);
  switch (prec) {
    case PREC_FP32: hipLaunchKernelGGL(k_record_batch<P32>, grid, block, 0, s, tab, it); break;
    case PREC_FP16: hipLaunchKernelGGL(k_record_batch<P16>, grid, block, 0, s, tab, it); break;
    case PREC_FP16C: hipLaunchKernelGGL(k_record_batch<P16C>, grid, block, 0, s, tab, it); break;
    case PREC_FP16T: hipLaunchKernelGGL(k_record_batch<P16T>, grid, block, 0, s, tab, it); break;
    case PREC_FP16M: hipLaunchKernelGGL(k_record_batch<P16M>, grid, block, 0, s, tab, it); break;
    case PREC_FP64: hipLaunchKernelGGL(k_record_batch<P64>, grid, block, 0, s, tab, it); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_advance(int *it, hipStream_t s)
{
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, it);
  return hipGetLastError();
}

hipError_t launch_set_counter(int *it, int v, hipStream_t s)
{
  hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(64), 0, s, it, v);
  return hipGetLastError();
}

hipError_t launch_velocity(int prec, int order, const StepArgs &args, LaunchShape sh, hipStream_t s)
{
  return dispatch<true>(prec, order, args, sh, s);
}
hipError_t launch_stress(int prec, int order, const StepArgs &args, LaunchShape sh, hipStream_t s)
{
  return dispatch<false>(prec, order, args, sh, s);
}

hipError_t launch_source(int prec, const SourceArgs &a, hipStream_t s)
{
  switch (prec) {
    case PREC_FP32: hipLaunchKernelGGL(k_source<P32>, dim3(1), dim3(64), 0, s, a); break;
    case PREC_FP16: hipLaunchKernelGGL(k_source<P16>, dim3(1), dim3(64), 0, s, a); break;
    case PREC_FP16C: hipLaunchKernelGGL(k_source<P16C>, dim3(1), dim3(64), 0, s, a); break;
    case PREC_FP16T: hipLaunchKernelGGL(k_source<P16T>, dim3(1), dim3(64), 0, s, a); break;
    case PREC_FP16M: hipLaunchKernelGGL(k_source<P16M>, dim3(1), dim3(64), 0, s, a); break;
    case PREC_FP64: hipLaunchKernelGGL(k_source<P64>, dim3(1), dim3(64), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rescale(int prec, void *p, int8_t *lo, int64_t n, double f, hipStream_t s)
{
  if (n <= 0 || !p) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  switch (prec) {
    case PREC_FP32:
      hipLaunchKernelGGL(k_rescale<float>, dim3(blocks), dim3(256), 0, s, (float *)p, lo, n, f);
      break;
    case PREC_FP16:
    case PREC_FP16C:
    case PREC_FP16T:
      hipLaunchKernelGGL(k_rescale<_Float16>, dim3(blocks), dim3(256), 0, s, (_Float16 *)p, lo, n, f);
      break;
    case PREC_FP16M:  // the binary16 arrays (v, q); launch_rescale_pk takes tau and p
      hipLaunchKernelGGL(k_rescale<_Float16>, dim3(blocks), dim3(256), 0, s, (_Float16 *)p, lo, n, f);
      break;
    case PREC_FP64:
      hipLaunchKernelGGL(k_rescale<double>, dim3(blocks), dim3(256), 0, s, (double *)p, lo, n, f);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rescale_pk(void *p, int8_t *lo, int64_t n, double f, hipStream_t s)
{
  if (n <= 0 || !p || !lo) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_rescale_pk, dim3(blocks), dim3(256), 0, s, p, lo, n, f);
  return hipGetLastError();
}

hipError_t launch_slice_xz(int prec, const void *f, const int8_t *lo, bool comp, double scale, int j, int nx,
                           int ny, int nz, float *out, hipStream_t s)
{
  if (nx <= 0 || nz <= 0) return hipSuccess;
  dim3 block(256), grid((unsigned)((nx + 255) / 256), (unsigned)nz);
  switch (prec) {
    case PREC_FP32: hipLaunchKernelGGL(k_slice_xz<P32>, grid, block, 0, s, f, lo, comp, scale, j, nx, ny, out); break;
    case PREC_FP16: hipLaunchKernelGGL(k_slice_xz<P16>, grid, block, 0, s, f, lo, comp, scale, j, nx, ny, out); break;
    case PREC_FP16C: hipLaunchKernelGGL(k_slice_xz<P16C>, grid, block, 0, s, f, lo, comp, scale, j, nx, ny, out); break;
    case PREC_FP16T: hipLaunchKernelGGL(k_slice_xz<P16T>, grid, block, 0, s, f, lo, comp, scale, j, nx, ny, out); break;
    case PREC_FP16M: hipLaunchKernelGGL(k_slice_xz<P16M>, grid, block, 0, s, f, lo, comp, scale, j, nx, ny, out); break;
    case PREC_FP64: hipLaunchKernelGGL(k_slice_xz<P64>, grid, block, 0, s, f, lo, comp, scale, j, nx, ny, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pack_coef(int prec, const void *soa, void *kv, void *kt, int64_t n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  switch (prec) {
    case PREC_FP32:
      hipLaunchKernelGGL(k_pack_coef<float>, dim3(blocks), dim3(256), 0, s, (const float *)soa, (float *)kv,
                         (float *)kt, n);
      break;
    case PREC_FP16:
    case PREC_FP16C:
    case PREC_FP16T:
      hipLaunchKernelGGL(k_pack_coef<_Float16>, dim3(blocks), dim3(256), 0, s, (const _Float16 *)soa,
                         (_Float16 *)kv, (_Float16 *)kt, n);
      break;
    case PREC_FP64:
      hipLaunchKernelGGL(k_pack_coef<double>, dim3(blocks), dim3(256), 0, s, (const double *)soa, (double *)kv,
                         (double *)kt, n);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rescale4(int prec, void *p, int64_t nrec, const double f[4], hipStream_t s)
{
  if (nrec <= 0 || !p) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<int64_t>((4 * nrec + 255) / 256, 65536);
  switch (prec) {
    case PREC_FP32:
    case PREC_FP16C:
    case PREC_FP16T:
      hipLaunchKernelGGL(k_rescale4<float>, dim3(blocks), dim3(256), 0, s, (float *)p, nrec, f[0], f[1], f[2], f[3]);
      break;
    case PREC_FP16:
      hipLaunchKernelGGL(k_rescale4<_Float16>, dim3(blocks), dim3(256), 0, s, (_Float16 *)p, nrec, f[0], f[1], f[2],
                         f[3]);
      break;
    case PREC_FP64:
      hipLaunchKernelGGL(k_rescale4<double>, dim3(blocks), dim3(256), 0, s, (double *)p, nrec, f[0], f[1], f[2], f[3]);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_record(int prec, const RecordArgs &a, hipStream_t s)
{
  if (a.nrec <= 0 || a.ncomp <= 0) return hipSuccess;
  dim3 block(256), grid((unsigned)((a.nrec + 255) / 256), (unsigned)a.ncomp);
  switch (prec) {
    case PREC_FP32: hipLaunchKernelGGL(k_record<P32>, grid, block, 0, s, a); break;
    case PREC_FP16: hipLaunchKernelGGL(k_record<P16>, grid, block, 0, s, a); break;
    case PREC_FP16C: hipLaunchKernelGGL(k_record<P16C>, grid, block, 0, s, a); break;
    case PREC_FP16T: hipLaunchKernelGGL(k_record<P16T>, grid, block, 0, s, a); break;
    case PREC_FP16M: hipLaunchKernelGGL(k_record<P16M>, grid, block, 0, s, a); break;
    case PREC_FP64: hipLaunchKernelGGL(k_record<P64>, grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace biot
