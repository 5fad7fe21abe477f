// x-vectorised forms of the velocity and stress half steps (reference kernel.cu:145-229,
// 405-493): every lane updates XV consecutive cells of an x row.
//
// Why: with 16-bit storage the one-cell-per-lane kernels issue one 2-byte load per lane
// per access, so a wave's load moves 128 B and the kernels are bound by load issue and
// address processing rather than by HBM bandwidth.  Here every access of a lane is one
// dword-multiple load of XV values (2 or 4 cells), and each x-differenced field is read
// once per plane as a segment of cells i0-2 .. i0+XV+1 that serves all XV cells (no
// separate neighbour loads in x).  y neighbours are row loads of XV values; z neighbours
// sit in a register window as in the scalar kernels.  The arithmetic per cell is the
// scalar kernels' (same operations in the same order).
//
// Alignment: a segment starts at element c-2 and every other access at c, with c = i0 +
// row start, i0 a multiple of XV; the host only selects these kernels when nx is a multiple
// of 4, so every access is 4-byte aligned (FP16C residual segments start at c-4, 4-byte
// aligned for XV = 4, the only width the FP16C policy uses).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "biot_device.hpp"

namespace biot {
namespace {

using u32 = uint32_t;

typedef uint32_t dw2_t __attribute__((ext_vector_type(2), aligned(4)));
typedef uint32_t dw3_t __attribute__((ext_vector_type(3), aligned(4)));
typedef uint32_t dw4_t __attribute__((ext_vector_type(4), aligned(4)));

struct P32 {
  using S = float;
  using SC = float;
  using A = float;
  static constexpr bool comp = false, scaled = false;
};
struct P16 {
  using S = _Float16;
  using SC = _Float16;
  using A = float;
  static constexpr bool comp = false, scaled = true;
};
struct P16C {
  using S = _Float16;
  using SC = float;
  using A = float;
  static constexpr bool comp = true, scaled = true;
};

// ND dwords from base + byte offset (4-byte aligned), as dwordx4/x3/x2/x1 loads
template <int ND>
__device__ __forceinline__ void ld_dw(const void *base, u32 off, uint32_t (&d)[ND])
{
  const char *p = reinterpret_cast<const char *>(base) + off;
#pragma unroll
  for (int o = 0; o < ND;) {
    const int r = ND - o;
    if (r >= 4) {
      const dw4_t v = *reinterpret_cast<const dw4_t *>(p + 4 * o);
      d[o] = v.x, d[o + 1] = v.y, d[o + 2] = v.z, d[o + 3] = v.w;
      o += 4;
    } else if (r == 3) {
      const dw3_t v = *reinterpret_cast<const dw3_t *>(p + 4 * o);
      d[o] = v.x, d[o + 1] = v.y, d[o + 2] = v.z;
      o += 3;
    } else if (r == 2) {
      const dw2_t v = *reinterpret_cast<const dw2_t *>(p + 4 * o);
      d[o] = v.x, d[o + 1] = v.y;
      o += 2;
    } else {
      d[o] = *reinterpret_cast<const uint32_t *>(p + 4 * o);
      o += 1;
    }
  }
}
template <int ND>
__device__ __forceinline__ void st_dw(void *base, u32 off, const uint32_t (&d)[ND])
{
  char *p = reinterpret_cast<char *>(base) + off;
#pragma unroll
  for (int o = 0; o < ND;) {
    const int r = ND - o;
    if (r >= 4) {
      *reinterpret_cast<dw4_t *>(p + 4 * o) = dw4_t{d[o], d[o + 1], d[o + 2], d[o + 3]};
      o += 4;
    } else if (r == 3) {
      *reinterpret_cast<dw3_t *>(p + 4 * o) = dw3_t{d[o], d[o + 1], d[o + 2]};
      o += 3;
    } else if (r == 2) {
      *reinterpret_cast<dw2_t *>(p + 4 * o) = dw2_t{d[o], d[o + 1]};
      o += 2;
    } else {
      *reinterpret_cast<uint32_t *>(p + 4 * o) = d[o];
      o += 1;
    }
  }
}

// binary16 residual unit ulp(h)/256 = 2^(e-33) and its inverse (e the biased exponent of
// h, subnormals e = 1), as in the scalar kernels
__device__ __forceinline__ int h_exp(_Float16 h)
{
  const u32 u = __builtin_bit_cast(uint16_t, h);
  return max((int)((u >> 10) & 31u), 1);
}
__device__ __forceinline__ float res_unit(int e) { return __builtin_bit_cast(float, (u32)(e + 94) << 23); }
__device__ __forceinline__ float res_inv(int e) { return __builtin_bit_cast(float, (u32)(160 - e) << 23); }

__device__ __forceinline__ _Float16 half_of(uint32_t d, int hi)
{
  return __builtin_bit_cast(_Float16, (uint16_t)(hi ? (d >> 16) : (d & 0xffffu)));
}
__device__ __forceinline__ float byte_of(uint32_t d, int b)
{
  return (float)(int)(int8_t)(uint8_t)(d >> (8 * b));
}

// N values of field F starting at element idx (idx*sizeof(S) 4-byte aligned).  For FP16C
// compensated fields the residual bytes come from the 4-byte aligned window starting at
// element idx - RO (RO = 0 or 2).
template <class P, int F, int N, int RO>
__device__ __forceinline__ void ldv(const StepArgs &a, u32 idx, typename P::A (&v)[N])
{
  using S = typename P::S;
  constexpr int ND = N * (int)sizeof(S) / 4;
  static_assert(N * sizeof(S) % 4 == 0, "dword multiple");
  uint32_t d[ND];
  ld_dw<ND>(a.f[F], idx * (u32)sizeof(S), d);
  _Float16 hs[N];
  if constexpr (std::is_same<S, float>::value) {
#pragma unroll
    for (int e = 0; e < N; ++e) v[e] = __builtin_bit_cast(float, d[e]);
  } else {
#pragma unroll
    for (int e = 0; e < N; ++e) {
      hs[e] = half_of(d[e >> 1], e & 1);
      v[e] = (float)hs[e];
    }
  }
  if constexpr (P::comp && ((COMP_FIELDS_C >> F) & 1u)) {
    constexpr int NB = ((N + RO + 3) / 4) * 4;  // residual bytes loaded
    uint32_t r[NB / 4];
    ld_dw<NB / 4>(a.lo[F], idx - (u32)RO, r);
#pragma unroll
    for (int e = 0; e < N; ++e) {
      const int b = e + RO;
      v[e] += byte_of(r[b >> 2], b & 3) * res_unit(h_exp(hs[e]));
    }
  }
}

// XV values of field F stored at idx; all cells valid (full) or the masked ones only
template <class P, int F, int XV>
__device__ __forceinline__ void stv(const StepArgs &a, u32 idx, const typename P::A (&x)[XV],
                                    const bool (&on)[XV], bool full)
{
  using S = typename P::S;
  if constexpr (std::is_same<S, float>::value) {
    if (full) {
      uint32_t d[XV];
#pragma unroll
      for (int e = 0; e < XV; ++e) d[e] = __builtin_bit_cast(uint32_t, x[e]);
      st_dw<XV>(a.f[F], idx * 4u, d);
    } else {
#pragma unroll
      for (int e = 0; e < XV; ++e)
        if (on[e]) *reinterpret_cast<float *>(reinterpret_cast<char *>(a.f[F]) + (idx + e) * 4u) = x[e];
    }
  } else {
    _Float16 hs[XV];
#pragma unroll
    for (int e = 0; e < XV; ++e) hs[e] = (_Float16)x[e];
    constexpr bool cf = P::comp && ((COMP_FIELDS_C >> F) & 1u);
    int8_t rs[XV];
    if constexpr (cf) {
#pragma unroll
      for (int e = 0; e < XV; ++e) {
        const float r = rintf((x[e] - (float)hs[e]) * res_inv(h_exp(hs[e])));
        rs[e] = (int8_t)fminf(fmaxf(r, -127.0f), 127.0f);
      }
    }
    if (full) {
      uint32_t d[XV / 2];
#pragma unroll
      for (int e = 0; e < XV / 2; ++e)
        d[e] = (uint32_t)__builtin_bit_cast(uint16_t, hs[2 * e]) |
               ((uint32_t)__builtin_bit_cast(uint16_t, hs[2 * e + 1]) << 16);
      st_dw<XV / 2>(a.f[F], idx * 2u, d);
      if constexpr (cf) {
        static_assert(!cf || XV == 4, "FP16C stores four residual bytes at once");
        uint32_t r = 0;
#pragma unroll
        for (int e = 0; e < XV; ++e) r |= (uint32_t)(uint8_t)rs[e] << (8 * e);
        *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(a.lo[F]) + idx) = r;
      }
    } else {
#pragma unroll
      for (int e = 0; e < XV; ++e) {
        if (!on[e]) continue;
        *reinterpret_cast<_Float16 *>(reinterpret_cast<char *>(a.f[F]) + (idx + e) * 2u) = hs[e];
        if constexpr (cf) *(reinterpret_cast<int8_t *>(a.lo[F]) + idx + e) = rs[e];
      }
    }
  }
}

// XV coefficient values of array K at c (scaled by the array's scale for 16-bit storage)
template <class P, int XV>
__device__ __forceinline__ void coefv(const StepArgs &a, int K, u32 c, typename P::A (&v)[XV])
{
  using S = typename P::S;
  constexpr int ND = XV * (int)sizeof(S) / 4;
  uint32_t d[ND];
  ld_dw<ND>(a.k[K], c * (u32)sizeof(S), d);
  if constexpr (std::is_same<S, float>::value) {
#pragma unroll
    for (int e = 0; e < XV; ++e) v[e] = __builtin_bit_cast(float, d[e]);
  } else {
    const float s = a.ks[K];
#pragma unroll
    for (int e = 0; e < XV; ++e) v[e] = (float)half_of(d[e >> 1], e & 1) * s;
  }
}

// one folded C-PML update (psi = carry + a D; carry <- e (psi + a D); returns D + psi)
template <class P>
__device__ __forceinline__ typename P::A pml(void *base, u32 idx, typename P::A D, typename P::A e,
                                             typename P::A aa)
{
  using A = typename P::A;
  using SC = typename P::SC;
  SC &cp = *reinterpret_cast<SC *>(reinterpret_cast<char *>(base) + idx * (u32)sizeof(SC));
  const A psi = (A)cp + aa * D;
  cp = (SC)(e * (psi + aa * D));
  return D + psi;
}

__device__ __forceinline__ int slab_pos(int i, int n, int Lw, int Rw)
{
  if (i < Lw) return i;
  if (i >= n - Rw) return Lw + i - (n - Rw);
  return -1;
}

// XCD-aware tile assignment (the scalar kernels' mapping)
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
__device__ __forceinline__ A profv(const void *p, int i)
{
  return *reinterpret_cast<const A *>(reinterpret_cast<const char *>(p) + (u32)i * (u32)sizeof(A));
}

// staggered differences of cell e from a segment s[m] = value at i0 - 2 + m
template <int ORDER, typename A, int N>
__device__ __forceinline__ A sfw(const A (&s)[N], int e)
{
  if constexpr (ORDER == 2) return s[e + 3] - s[e + 2];
  return A(9.0 / 8.0) * (s[e + 3] - s[e + 2]) + A(-1.0 / 24.0) * (s[e + 4] - s[e + 1]);
}
template <int ORDER, typename A, int N>
__device__ __forceinline__ A sbw(const A (&s)[N], int e)
{
  if constexpr (ORDER == 2) return s[e + 2] - s[e + 1];
  return A(9.0 / 8.0) * (s[e + 2] - s[e + 1]) + A(-1.0 / 24.0) * (s[e + 3] - s[e]);
}
// from rows: m1, p1, p2 / m2 are the rows at -1, +1, +2 / -2 (order 4 only uses the outer)
template <int ORDER, typename A>
__device__ __forceinline__ A rfw(A c, A m1, A p1, A p2)
{
  if constexpr (ORDER == 2) return p1 - c;
  return A(9.0 / 8.0) * (p1 - c) + A(-1.0 / 24.0) * (p2 - m1);
}
template <int ORDER, typename A>
__device__ __forceinline__ A rbw(A c, A m1, A m2, A p1)
{
  if constexpr (ORDER == 2) return c - m1;
  return A(9.0 / 8.0) * (c - m1) + A(-1.0 / 24.0) * (p1 - m2);
}
template <int ORDER, typename A>
__device__ __forceinline__ A dwin(const A *w)
{
  if constexpr (ORDER == 2) return w[1] - w[0];
  return A(9.0 / 8.0) * (w[2] - w[1]) + A(-1.0 / 24.0) * (w[3] - w[0]);
}

enum { FVX = 0, FVY, FVZ, FQX, FQY, FQZ, FTXX, FTYY, FTZZ, FTXY, FTXZ, FTYZ, FP };

// rows j-2, j-1, j+1, j+2 of field F (XV values each) around plane cell c; only the rows a
// forward (or backward) difference of the order needs are loaded
template <class P, int F, int XV, int ORDER, bool FWD>
__device__ __forceinline__ void rows(const StepArgs &a, u32 c, u32 sy, typename P::A (&m2)[XV],
                                     typename P::A (&m1)[XV], typename P::A (&p1)[XV],
                                     typename P::A (&p2)[XV])
{
  if (FWD) {
    ldv<P, F, XV, 0>(a, c + sy, p1);
    if constexpr (ORDER == 4) {
      ldv<P, F, XV, 0>(a, c + 2 * sy, p2);
      ldv<P, F, XV, 0>(a, c - sy, m1);
    }
  } else {
    ldv<P, F, XV, 0>(a, c - sy, m1);
    if constexpr (ORDER == 4) {
      ldv<P, F, XV, 0>(a, c - 2 * sy, m2);
      ldv<P, F, XV, 0>(a, c + sy, p1);
    }
  }
}

template <class P, int ORDER, int XV>
__global__ void __launch_bounds__(256) k_velocity_xv(const StepArgs a)
{
  using A = typename P::A;
  constexpr int h = ORDER / 2;
  constexpr int W = 2 * h;
  constexpr int NS = XV + 4;            // segment: cells i0-2 .. i0+XV+1
  constexpr int RO = P::comp ? 2 : 0;   // residual window offset of a segment
  int tx, ty, tz;
  if (!tile_of(a, tx, ty, tz)) return;
  const int i0 = (tx * (int)blockDim.x + (int)threadIdx.x) * XV;
  const int j = ty * (int)blockDim.y + (int)threadIdx.y;
  if (i0 + XV <= h || i0 >= a.nx - h || j < h || j >= a.ny - h) return;
  const int kb = h + tz * a.zchunk;
  const int ke = min(kb + a.zchunk, a.nz - h);
  if (kb >= ke) return;
  bool on[XV];
  bool full = true;
#pragma unroll
  for (int e = 0; e < XV; ++e) {
    on[e] = i0 + e >= h && i0 + e < a.nx - h;
    full = full && on[e];
  }
  const u32 sy = (u32)a.sy, sz = (u32)a.sz;
  const u32 col = (u32)i0 + (u32)j * sy;

  int px[XV];
  A exi[XV], axi[XV], exh[XV], axh[XV];
#pragma unroll
  for (int e = 0; e < XV; ++e) {
    px[e] = on[e] ? slab_pos(i0 + e, a.nx, a.L[0], a.R[0]) : -1;
    exi[e] = axi[e] = exh[e] = axh[e] = 0;
    if (px[e] >= 0) {
      exi[e] = profv<A>(a.prof[0][0], i0 + e);
      axi[e] = profv<A>(a.prof[0][1], i0 + e);
      exh[e] = profv<A>(a.prof[0][2], i0 + e);
      axh[e] = profv<A>(a.prof[0][3], i0 + e);
    }
  }
  const int py = slab_pos(j, a.ny, a.L[1], a.R[1]);
  A eyi = 0, ayi = 0, eyh = 0, ayh = 0;
  if (py >= 0) {
    eyi = profv<A>(a.prof[1][0], j);
    ayi = profv<A>(a.prof[1][1], j);
    eyh = profv<A>(a.prof[1][2], j);
    ayh = profv<A>(a.prof[1][3], j);
  }
  const u32 wxs = (u32)(a.L[0] + a.R[0]), wys = (u32)(a.L[1] + a.R[1]);
  const u32 ny = (u32)a.ny, nx = (u32)a.nx;
  const u32 sl0 = (u32)a.slab[0], sl1 = (u32)a.slab[1], sl2 = (u32)a.slab[2];

  // z windows (XV cells each): backward-differenced txz, tyz hold planes k-h..k+h-1,
  // forward-differenced tzz, p hold planes k-h+1..k+h
  A wxz[W][XV], wyz[W][XV], wzz[W][XV], wp[W][XV];
#pragma unroll
  for (int m = 0; m < W - 1; ++m) {
    ldv<P, FTXZ, XV, 0>(a, col + (u32)(kb - h + m) * sz, wxz[m]);
    ldv<P, FTYZ, XV, 0>(a, col + (u32)(kb - h + m) * sz, wyz[m]);
    ldv<P, FTZZ, XV, 0>(a, col + (u32)(kb - h + 1 + m) * sz, wzz[m]);
    ldv<P, FP, XV, 0>(a, col + (u32)(kb - h + 1 + m) * sz, wp[m]);
  }

  for (int k = kb; k < ke; ++k) {
    const u32 c = col + (u32)k * sz;
    ldv<P, FTXZ, XV, 0>(a, c + (u32)(h - 1) * sz, wxz[W - 1]);
    ldv<P, FTYZ, XV, 0>(a, c + (u32)(h - 1) * sz, wyz[W - 1]);
    ldv<P, FTZZ, XV, 0>(a, c + (u32)h * sz, wzz[W - 1]);
    ldv<P, FP, XV, 0>(a, c + (u32)h * sz, wp[W - 1]);
    // x segments of the plane
    A sxx[NS], sxy[NS], sxz[NS], sp[NS];
    ldv<P, FTXX, NS, RO>(a, c - 2, sxx);
    ldv<P, FTXY, NS, RO>(a, c - 2, sxy);
    ldv<P, FTXZ, NS, RO>(a, c - 2, sxz);
    ldv<P, FP, NS, RO>(a, c - 2, sp);
    // y rows
    A tyy_c[XV], yy_m2[XV], yy_m1[XV], yy_p1[XV], yy_p2[XV];
    A xy_m2[XV], xy_m1[XV], xy_p1[XV], xy_p2[XV];
    A yz_m2[XV], yz_m1[XV], yz_p1[XV], yz_p2[XV];
    A p_m2[XV], p_m1[XV], p_p1[XV], p_p2[XV];
    ldv<P, FTYY, XV, 0>(a, c, tyy_c);
    rows<P, FTYY, XV, ORDER, true>(a, c, sy, yy_m2, yy_m1, yy_p1, yy_p2);
    rows<P, FTXY, XV, ORDER, false>(a, c, sy, xy_m2, xy_m1, xy_p1, xy_p2);
    rows<P, FTYZ, XV, ORDER, false>(a, c, sy, yz_m2, yz_m1, yz_p1, yz_p2);
    rows<P, FP, XV, ORDER, true>(a, c, sy, p_m2, p_m1, p_p1, p_p2);
    A qx0[XV], qy0[XV], qz0[XV], vx0[XV], vy0[XV], vz0[XV];
    ldv<P, FQX, XV, 0>(a, c, qx0);
    ldv<P, FQY, XV, 0>(a, c, qy0);
    ldv<P, FQZ, XV, 0>(a, c, qz0);
    ldv<P, FVX, XV, 0>(a, c, vx0);
    ldv<P, FVY, XV, 0>(a, c, vy0);
    ldv<P, FVZ, XV, 0>(a, c, vz0);
    A W1x[XV], W1y[XV], W1z[XV], W2x[XV], W2y[XV], W2z[XV], W3x[XV], W3y[XV];
    A U1x[XV], U1y[XV], U1z[XV], U2x[XV], U2y[XV], U2z[XV];
    coefv<P, XV>(a, K_W1X, c, W1x);
    coefv<P, XV>(a, K_W1Y, c, W1y);
    coefv<P, XV>(a, K_W1Z, c, W1z);
    coefv<P, XV>(a, K_W2X, c, W2x);
    coefv<P, XV>(a, K_W2Y, c, W2y);
    coefv<P, XV>(a, K_W2Z, c, W2z);
    coefv<P, XV>(a, K_W3X, c, W3x);
    coefv<P, XV>(a, K_W3Y, c, W3y);
    coefv<P, XV>(a, K_U1X, c, U1x);
    coefv<P, XV>(a, K_U1Y, c, U1y);
    coefv<P, XV>(a, K_U1Z, c, U1z);
    coefv<P, XV>(a, K_U2X, c, U2x);
    coefv<P, XV>(a, K_U2Y, c, U2y);
    coefv<P, XV>(a, K_U2Z, c, U2z);

    const int pz = slab_pos(k, a.nz, a.L[2], a.R[2]);
    A ezi = 0, azi = 0, ezh = 0, azh = 0;
    if (pz >= 0) {
      ezi = profv<A>(a.prof[2][0], k);
      azi = profv<A>(a.prof[2][1], k);
      ezh = profv<A>(a.prof[2][2], k);
      azh = profv<A>(a.prof[2][3], k);
    }
    A qx1[XV], qy1[XV], qz1[XV], vx1[XV], vy1[XV], vz1[XV];
#pragma unroll
    for (int e = 0; e < XV; ++e) {
      // twelve raw staggered differences (kernel.cu:172-185 without the 1/dx factor)
      A x1 = sfw<ORDER>(sxx, e);
      A x2 = rbw<ORDER>(sxy[e + 2], xy_m1[e], xy_m2[e], xy_p1[e]);
      A x3;
      {
        A w[W];
#pragma unroll
        for (int m = 0; m < W; ++m) w[m] = wxz[m][e];
        x3 = dwin<ORDER>(w);
      }
      A s1 = sfw<ORDER>(sp, e);
      A y1 = rfw<ORDER>(tyy_c[e], yy_m1[e], yy_p1[e], yy_p2[e]);
      A y2 = sbw<ORDER>(sxy, e);
      A y3, z1, s3;
      {
        A w[W], u[W], v[W];
#pragma unroll
        for (int m = 0; m < W; ++m) {
          w[m] = wyz[m][e];
          u[m] = wzz[m][e];
          v[m] = wp[m][e];
        }
        y3 = dwin<ORDER>(w);
        z1 = dwin<ORDER>(u);
        s3 = dwin<ORDER>(v);
      }
      A s2 = rfw<ORDER>(sp[e + 2], p_m1[e], p_p1[e], p_p2[e]);
      A z2 = sbw<ORDER>(sxz, e);
      A z3 = rbw<ORDER>(wyz[h][e], yz_m1[e], yz_m2[e], yz_p1[e]);

      // C-PML memory (kernel.cu:188-216), same carries and order as the scalar kernel
      if (px[e] >= 0) {
        const u32 o = 4u * ((u32)px[e] + wxs * ((u32)j + ny * (u32)k));  // record of carries 0..3
        x1 = pml<P>(a.carry[0], o, x1, exh[e], axh[e]);
        s1 = pml<P>(a.carry[0], o + 1u, s1, exh[e], axh[e]);
        y2 = pml<P>(a.carry[0], o + 2u, y2, exi[e], axi[e]);
        z2 = pml<P>(a.carry[0], o + 3u, z2, exi[e], axi[e]);
      }
      if (py >= 0 && on[e]) {
        const u32 o = 4u * ((u32)(i0 + e) + nx * ((u32)py + wys * (u32)k));
        x2 = pml<P>(a.carry[1], o, x2, eyi, ayi);
        y1 = pml<P>(a.carry[1], o + 1u, y1, eyh, ayh);
        s2 = pml<P>(a.carry[1], o + 2u, s2, eyh, ayh);
        z3 = pml<P>(a.carry[1], o + 3u, z3, eyi, ayi);
      }
      if (pz >= 0 && on[e]) {
        const u32 o = 4u * ((u32)(i0 + e) + nx * ((u32)j + ny * (u32)pz));
        x3 = pml<P>(a.carry[2], o, x3, ezi, azi);
        y3 = pml<P>(a.carry[2], o + 1u, y3, ezi, azi);
        z1 = pml<P>(a.carry[2], o + 2u, z1, ezh, azh);
        s3 = pml<P>(a.carry[2], o + 3u, s3, ezh, azh);
      }
      // relative fluid velocity (kernel.cu:218-220; qz takes W3y) and solid velocity
      // with the stored new q (kernel.cu:223-225)
      const A sx = x1 + x2 + x3, sy_ = y1 + y2 + y3, sz_ = z1 + z2 + z3;
      qx1[e] = W1x[e] * qx0[e] - W2x[e] * sx - W3x[e] * s1;
      qy1[e] = W1y[e] * qy0[e] - W2y[e] * sy_ - W3y[e] * s2;
      qz1[e] = W1z[e] * qz0[e] - W2z[e] * sz_ - W3y[e] * s3;
      using S = typename P::S;
      const A qx1s = (A)(S)qx1[e], qy1s = (A)(S)qy1[e], qz1s = (A)(S)qz1[e];
      vx1[e] = vx0[e] + U1x[e] * sx - U2x[e] * (qx1s - qx0[e]);
      vy1[e] = vy0[e] + U1y[e] * sy_ - U2y[e] * (qy1s - qy0[e]);
      vz1[e] = vz0[e] + U1z[e] * sz_ - U2z[e] * (qz1s - qz0[e]);
    }
    stv<P, FQX, XV>(a, c, qx1, on, full);
    stv<P, FQY, XV>(a, c, qy1, on, full);
    stv<P, FQZ, XV>(a, c, qz1, on, full);
    stv<P, FVX, XV>(a, c, vx1, on, full);
    stv<P, FVY, XV>(a, c, vy1, on, full);
    stv<P, FVZ, XV>(a, c, vz1, on, full);

#pragma unroll
    for (int m = 0; m < W - 1; ++m)
#pragma unroll
      for (int e = 0; e < XV; ++e) {
        wxz[m][e] = wxz[m + 1][e];
        wyz[m][e] = wyz[m + 1][e];
        wzz[m][e] = wzz[m + 1][e];
        wp[m][e] = wp[m + 1][e];
      }
  }
}

template <class P, int ORDER, int XV>
__global__ void __launch_bounds__(256) k_stress_xv(const StepArgs a)
{
  using A = typename P::A;
  constexpr int h = ORDER / 2;
  constexpr int W = 2 * h;
  constexpr int NS = XV + 4;
  constexpr int RO = P::comp ? 2 : 0;
  int tx, ty, tz;
  if (!tile_of(a, tx, ty, tz)) return;
  const int i0 = (tx * (int)blockDim.x + (int)threadIdx.x) * XV;
  const int j = ty * (int)blockDim.y + (int)threadIdx.y;
  if (i0 + XV <= h || i0 >= a.nx - h || j < h || j >= a.ny - h) return;
  const int kb = h + tz * a.zchunk;
  const int ke = min(kb + a.zchunk, a.nz - h);
  if (kb >= ke) return;
  bool on[XV];
  bool full = true;
#pragma unroll
  for (int e = 0; e < XV; ++e) {
    on[e] = i0 + e >= h && i0 + e < a.nx - h;
    full = full && on[e];
  }
  const u32 sy = (u32)a.sy, sz = (u32)a.sz;
  const u32 col = (u32)i0 + (u32)j * sy;

  int px[XV];
  A exi[XV], axi[XV], exh[XV], axh[XV];
#pragma unroll
  for (int e = 0; e < XV; ++e) {
    px[e] = on[e] ? slab_pos(i0 + e, a.nx, a.L[0], a.R[0]) : -1;
    exi[e] = axi[e] = exh[e] = axh[e] = 0;
    if (px[e] >= 0) {
      exi[e] = profv<A>(a.prof[0][0], i0 + e);
      axi[e] = profv<A>(a.prof[0][1], i0 + e);
      exh[e] = profv<A>(a.prof[0][2], i0 + e);
      axh[e] = profv<A>(a.prof[0][3], i0 + e);
    }
  }
  const int py = slab_pos(j, a.ny, a.L[1], a.R[1]);
  A eyi = 0, ayi = 0, eyh = 0, ayh = 0;
  if (py >= 0) {
    eyi = profv<A>(a.prof[1][0], j);
    ayi = profv<A>(a.prof[1][1], j);
    eyh = profv<A>(a.prof[1][2], j);
    ayh = profv<A>(a.prof[1][3], j);
  }
  const u32 wxs = (u32)(a.L[0] + a.R[0]), wys = (u32)(a.L[1] + a.R[1]);
  const u32 ny = (u32)a.ny, nx = (u32)a.nx;
  const u32 sl0 = (u32)a.slab[0], sl1 = (u32)a.slab[1], sl2 = (u32)a.slab[2];

  // z windows: backward-differenced vz, qz hold planes k-h..k+h-1, forward-differenced
  // vx, vy hold planes k-h+1..k+h
  A wvz[W][XV], wqz[W][XV], wvx[W][XV], wvy[W][XV];
#pragma unroll
  for (int m = 0; m < W - 1; ++m) {
    ldv<P, FVZ, XV, 0>(a, col + (u32)(kb - h + m) * sz, wvz[m]);
    ldv<P, FQZ, XV, 0>(a, col + (u32)(kb - h + m) * sz, wqz[m]);
    ldv<P, FVX, XV, 0>(a, col + (u32)(kb - h + 1 + m) * sz, wvx[m]);
    ldv<P, FVY, XV, 0>(a, col + (u32)(kb - h + 1 + m) * sz, wvy[m]);
  }

  for (int k = kb; k < ke; ++k) {
    const u32 c = col + (u32)k * sz;
    ldv<P, FVZ, XV, 0>(a, c + (u32)(h - 1) * sz, wvz[W - 1]);
    ldv<P, FQZ, XV, 0>(a, c + (u32)(h - 1) * sz, wqz[W - 1]);
    ldv<P, FVX, XV, 0>(a, c + (u32)h * sz, wvx[W - 1]);
    ldv<P, FVY, XV, 0>(a, c + (u32)h * sz, wvy[W - 1]);
    // x segments: vx (uxx bwd), qx (wx bwd), vy (uyx fwd), vz (uzx fwd)
    A svx[NS], sqx[NS], svy[NS], svz[NS];
    ldv<P, FVX, NS, RO>(a, c - 2, svx);
    ldv<P, FQX, NS, RO>(a, c - 2, sqx);
    ldv<P, FVY, NS, RO>(a, c - 2, svy);
    ldv<P, FVZ, NS, RO>(a, c - 2, svz);
    // y rows: vy (uyy bwd), qy (wy bwd), vx (uxy fwd), vz (uzy fwd)
    A qy_c[XV];
    ldv<P, FQY, XV, 0>(a, c, qy_c);
    A vy_m2[XV], vy_m1[XV], vy_p1[XV], vy_p2[XV];
    A qy_m2[XV], qy_m1[XV], qy_p1[XV], qy_p2[XV];
    A vx_m2[XV], vx_m1[XV], vx_p1[XV], vx_p2[XV];
    A vz_m2[XV], vz_m1[XV], vz_p1[XV], vz_p2[XV];
    rows<P, FVY, XV, ORDER, false>(a, c, sy, vy_m2, vy_m1, vy_p1, vy_p2);
    rows<P, FQY, XV, ORDER, false>(a, c, sy, qy_m2, qy_m1, qy_p1, qy_p2);
    rows<P, FVX, XV, ORDER, true>(a, c, sy, vx_m2, vx_m1, vx_p1, vx_p2);
    rows<P, FVZ, XV, ORDER, true>(a, c, sy, vz_m2, vz_m1, vz_p1, vz_p2);
    A tp[XV], txx[XV], tyy[XV], tzz[XV], txy[XV], tyz[XV], txz[XV];
    ldv<P, FP, XV, 0>(a, c, tp);
    ldv<P, FTXX, XV, 0>(a, c, txx);
    ldv<P, FTYY, XV, 0>(a, c, tyy);
    ldv<P, FTZZ, XV, 0>(a, c, tzz);
    ldv<P, FTXY, XV, 0>(a, c, txy);
    ldv<P, FTYZ, XV, 0>(a, c, tyz);
    ldv<P, FTXZ, XV, 0>(a, c, txz);
    A P1[XV], P2[XV], S1[XV], S2[XV], S3[XV], Sxy[XV], Syz[XV], Sxz[XV];
    coefv<P, XV>(a, K_P1, c, P1);
    coefv<P, XV>(a, K_P2, c, P2);
    coefv<P, XV>(a, K_S1, c, S1);
    coefv<P, XV>(a, K_S2, c, S2);
    coefv<P, XV>(a, K_S3, c, S3);
    coefv<P, XV>(a, K_SXY, c, Sxy);
    coefv<P, XV>(a, K_SYZ, c, Syz);
    coefv<P, XV>(a, K_SXZ, c, Sxz);

    const int pz = slab_pos(k, a.nz, a.L[2], a.R[2]);
    A ezi = 0, azi = 0, ezh = 0, azh = 0;
    if (pz >= 0) {
      ezi = profv<A>(a.prof[2][0], k);
      azi = profv<A>(a.prof[2][1], k);
      ezh = profv<A>(a.prof[2][2], k);
      azh = profv<A>(a.prof[2][3], k);
    }
#pragma unroll
    for (int e = 0; e < XV; ++e) {
      // twelve raw differences (kernel.cu:430-445 without 1/dx)
      A uxx = sbw<ORDER>(svx, e);
      A uyy = rbw<ORDER>(svy[e + 2], vy_m1[e], vy_m2[e], vy_p1[e]);
      A uzz, wz, uxz, uyz;
      {
        A w0[W], w1[W], w2[W], w3[W];
#pragma unroll
        for (int m = 0; m < W; ++m) {
          w0[m] = wvz[m][e];
          w1[m] = wqz[m][e];
          w2[m] = wvx[m][e];
          w3[m] = wvy[m][e];
        }
        uzz = dwin<ORDER>(w0);
        wz = dwin<ORDER>(w1);
        uxz = dwin<ORDER>(w2);
        uyz = dwin<ORDER>(w3);
      }
      A wx = sbw<ORDER>(sqx, e);
      A wy = rbw<ORDER>(qy_c[e], qy_m1[e], qy_m2[e], qy_p1[e]);
      A uxy = rfw<ORDER>(svx[e + 2], vx_m1[e], vx_p1[e], vx_p2[e]);
      A uyx = sfw<ORDER>(svy, e);
      A uzx = sfw<ORDER>(svz, e);
      A uzy = rfw<ORDER>(svz[e + 2], vz_m1[e], vz_p1[e], vz_p2[e]);

      // C-PML memory (kernel.cu:449-482): carries 4..7 of each axis
      if (px[e] >= 0) {
        const u32 o = 4 * sl0 + 4u * ((u32)px[e] + wxs * ((u32)j + ny * (u32)k));  // carries 4..7
        uxx = pml<P>(a.carry[0], o, uxx, exi[e], axi[e]);
        wx = pml<P>(a.carry[0], o + 1u, wx, exi[e], axi[e]);
        uyx = pml<P>(a.carry[0], o + 2u, uyx, exh[e], axh[e]);
        uzx = pml<P>(a.carry[0], o + 3u, uzx, exh[e], axh[e]);
      }
      if (py >= 0 && on[e]) {
        const u32 o = 4 * sl1 + 4u * ((u32)(i0 + e) + nx * ((u32)py + wys * (u32)k));
        uyy = pml<P>(a.carry[1], o, uyy, eyi, ayi);
        wy = pml<P>(a.carry[1], o + 1u, wy, eyi, ayi);
        uxy = pml<P>(a.carry[1], o + 2u, uxy, eyh, ayh);
        uzy = pml<P>(a.carry[1], o + 3u, uzy, eyh, ayh);
      }
      if (pz >= 0 && on[e]) {
        const u32 o = 4 * sl2 + 4u * ((u32)(i0 + e) + nx * ((u32)j + ny * (u32)pz));
        uzz = pml<P>(a.carry[2], o, uzz, ezi, azi);
        wz = pml<P>(a.carry[2], o + 1u, wz, ezi, azi);
        uxz = pml<P>(a.carry[2], o + 2u, uxz, ezh, azh);
        uyz = pml<P>(a.carry[2], o + 3u, uyz, ezh, azh);
      }
      // pore pressure and stresses (kernel.cu:484-490)
      const A divw = wx + wy + wz;
      tp[e] = tp[e] - P1[e] * (uxx + uyy + uzz) - P2[e] * divw;
      txx[e] = txx[e] + S1[e] * (uyy + uzz) + S2[e] * uxx + S3[e] * divw;
      tyy[e] = tyy[e] + S1[e] * (uxx + uzz) + S2[e] * uyy + S3[e] * divw;
      tzz[e] = tzz[e] + S1[e] * (uxx + uyy) + S2[e] * uzz + S3[e] * divw;
      txy[e] = txy[e] + Sxy[e] * (uxy + uyx);
      tyz[e] = tyz[e] + Syz[e] * (uyz + uzy);
      txz[e] = txz[e] + Sxz[e] * (uxz + uzx);
    }
    stv<P, FP, XV>(a, c, tp, on, full);
    stv<P, FTXX, XV>(a, c, txx, on, full);
    stv<P, FTYY, XV>(a, c, tyy, on, full);
    stv<P, FTZZ, XV>(a, c, tzz, on, full);
    stv<P, FTXY, XV>(a, c, txy, on, full);
    stv<P, FTYZ, XV>(a, c, tyz, on, full);
    stv<P, FTXZ, XV>(a, c, txz, on, full);

#pragma unroll
    for (int m = 0; m < W - 1; ++m)
#pragma unroll
      for (int e = 0; e < XV; ++e) {
        wvz[m][e] = wvz[m + 1][e];
        wqz[m][e] = wqz[m + 1][e];
        wvx[m][e] = wvx[m + 1][e];
        wvy[m][e] = wvy[m + 1][e];
      }
  }
}

template <class P, int ORDER, int XV, bool VEL>
hipError_t launch_xv(const StepArgs &in, LaunchShape sh, hipStream_t s)
{
  StepArgs a = in;
  const int h = ORDER / 2;
  const int nzi = a.nz - 2 * h;
  if (nzi <= 0 || a.nx <= 2 * h || a.ny <= 2 * h) return hipSuccess;
  if (sh.zchunk <= 0 || a.nx % 4 != 0) return hipErrorInvalidValue;
  a.zchunk = sh.zchunk;
  a.gx = (a.nx + sh.bx * XV - 1) / (sh.bx * XV);
  a.gy = (a.ny + sh.by - 1) / sh.by;
  a.gz = (nzi + a.zchunk - 1) / a.zchunk;
  a.ntiles = a.gx * a.gy * a.gz;
  dim3 block(sh.bx, sh.by, 1);
  dim3 grid((unsigned)(((a.ntiles + 7) / 8) * 8));
  if (VEL)
    hipLaunchKernelGGL((k_velocity_xv<P, ORDER, XV>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_stress_xv<P, ORDER, XV>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <bool VEL, int ORDER>
hipError_t dispatch_xv(int prec, const StepArgs &a, LaunchShape sh, hipStream_t s)
{
  switch (prec) {
    case PREC_FP32:
      if (sh.xv == 2) return launch_xv<P32, ORDER, 2, VEL>(a, sh, s);
      if (sh.xv == 4) return launch_xv<P32, ORDER, 4, VEL>(a, sh, s);
      break;
    case PREC_FP16:
      if (sh.xv == 2) return launch_xv<P16, ORDER, 2, VEL>(a, sh, s);
      if (sh.xv == 4) return launch_xv<P16, ORDER, 4, VEL>(a, sh, s);
      break;
    case PREC_FP16C:
      if (sh.xv == 4) return launch_xv<P16C, ORDER, 4, VEL>(a, sh, s);
      break;
  }
  return hipErrorInvalidValue;
}

}  // namespace

bool xv_supported(int prec, int xv, int nx)
{
  if (nx % 4 != 0) return false;
  if (prec == PREC_FP32 || prec == PREC_FP16) return xv == 2 || xv == 4;
  if (prec == PREC_FP16C) return xv == 4;
  return false;
}

hipError_t launch_velocity_xv(int prec, int order, const StepArgs &a, LaunchShape sh, hipStream_t s)
{
  if (sh.bx * sh.by > 256 || sh.bx <= 0 || sh.by <= 0) return hipErrorInvalidValue;
  if (order == 2) return dispatch_xv<true, 2>(prec, a, sh, s);
  if (order == 4) return dispatch_xv<true, 4>(prec, a, sh, s);
  return hipErrorInvalidValue;
}
hipError_t launch_stress_xv(int prec, int order, const StepArgs &a, LaunchShape sh, hipStream_t s)
{
  if (sh.bx * sh.by > 256 || sh.bx <= 0 || sh.by <= 0) return hipErrorInvalidValue;
  if (order == 2) return dispatch_xv<false, 2>(prec, a, sh, s);
  if (order == 4) return dispatch_xv<false, 4>(prec, a, sh, s);
  return hipErrorInvalidValue;
}

}  // namespace biot
