// HIP kernels (gfx950) of the 13-field staggered-grid Biot time step.
//
// Velocity half step (reference kernel.cu:145-229) and stress half step (kernel.cu:405-493)
// are each one z-marching kernel: a 2-D thread block covers an x-y tile with x fastest
// (64 consecutive cells per wavefront -> 256 B coalesced rows in FP32), every thread walks
// `zchunk` planes of its column and keeps the z-differenced fields in a register window
// that rotates by one plane per iteration, so each of those planes is fetched once.
// Storage is float, _Float16 or double; arithmetic is float (double for FP64 storage).
// The C-PML recursion is applied only in the absorbing slabs, on the folded carry
// variable described in biot_device.hpp.
//
// Addressing: every array is reached from its uniform (kernel-argument) base pointer with
// a 32-bit byte offset, so a load is SGPR base + one 32-bit VGPR offset instead of a
// 64-bit VGPR address per load; the host guarantees every array spans < 4 GiB.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "biot_device.hpp"

namespace biot {
namespace {

using u32 = uint32_t;

template <typename S>
struct Arith {
  using A = float;
};
template <>
struct Arith<double> {
  using A = double;
};

// element idx of an array, byte offset formed in 32 bits
template <typename T>
__device__ __forceinline__ T &at(T *base, u32 idx)
{
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (u32)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T &at(const T *base, u32 idx)
{
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + idx * (u32)sizeof(T));
}

template <typename S>
__device__ __forceinline__ typename Arith<S>::A ld(S x)
{
  return (typename Arith<S>::A)x;
}
template <typename S, typename A>
__device__ __forceinline__ S st(A a)
{
  return (S)a;
}
template <typename S>
__device__ __forceinline__ typename Arith<S>::A ldi(const S *f, u32 idx)
{
  return ld(at(f, idx));
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

// forward difference at c + s/2 (value at c supplied)
template <int ORDER, typename S, typename A>
__device__ __forceinline__ A dfw(const S *f, u32 c, u32 s, A fc)
{
  if constexpr (ORDER == 2) {
    return ldi(f, c + s) - fc;
  } else {
    return A(9.0 / 8.0) * (ldi(f, c + s) - fc) + A(-1.0 / 24.0) * (ldi(f, c + 2 * s) - ldi(f, c - s));
  }
}
// backward difference at c - s/2 (value at c supplied)
template <int ORDER, typename S, typename A>
__device__ __forceinline__ A dbw(const S *f, u32 c, u32 s, A fc)
{
  if constexpr (ORDER == 2) {
    return fc - ldi(f, c - s);
  } else {
    return A(9.0 / 8.0) * (fc - ldi(f, c - s)) + A(-1.0 / 24.0) * (ldi(f, c + s) - ldi(f, c - 2 * s));
  }
}

template <typename S, typename A>
__device__ __forceinline__ A coef(const StepArgs<S> &a, int idx, u32 c)
{
  if constexpr (std::is_same<S, _Float16>::value) {
    return ldi(a.k[idx], c) * a.ks[idx];
  } else {
    return ldi(a.k[idx], c);
  }
}

// One folded C-PML update: psi = carry + a D; carry <- e (psi + a D); returns D + psi.
template <typename S, typename A>
__device__ __forceinline__ A pml(S *base, u32 idx, A D, A e, A aa)
{
  S &cp = at(base, idx);
  const A psi = ld(cp) + aa * D;
  cp = st<S>(e * (psi + aa * D));
  return D + psi;
}

// slab position of index i on an axis with widths L (low) and R (high), or -1
__device__ __forceinline__ int slab_pos(int i, int n, int L, int R)
{
  if (i < L) return i;
  if (i >= n - R) return L + i - (n - R);
  return -1;
}

// XCD-aware tile assignment.  The dispatcher hands workgroup b of a 1-D grid to XCD
// b % 8 in order; each XCD here gets one contiguous range of tiles, x fastest then y then
// z, so the y- and z-adjacent tiles whose halos overlap are processed by the same XCD,
// close in time, and the halo re-reads hit that XCD's L2.
template <typename S>
__device__ __forceinline__ bool tile_of(const StepArgs<S> &a, int &tx, int &ty, int &tz)
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
  return at(static_cast<const A *>(p), (u32)i);
}

template <typename S, int ORDER>
__global__ void __launch_bounds__(256) k_velocity(const StepArgs<S> a)
{
  using A = typename Arith<S>::A;
  constexpr int h = ORDER / 2;
  constexpr int W = 2 * h;
  int tx, ty, tz;
  if (!tile_of(a, tx, ty, tz)) return;
  const int i = tx * blockDim.x + threadIdx.x;
  const int j = ty * blockDim.y + threadIdx.y;
  if (i < h || i >= a.nx - h || j < h || j >= a.ny - h) return;
  const int kb = h + tz * a.zchunk;
  const int ke = min(kb + a.zchunk, a.nz - h);
  if (kb >= ke) return;
  const u32 sy = (u32)a.sy, sz = (u32)a.sz;
  const u32 col = (u32)i + (u32)j * sy;

  S *vx = a.f[0], *vy = a.f[1], *vz = a.f[2], *qx = a.f[3], *qy = a.f[4], *qz = a.f[5];
  const S *txx = a.f[6], *tyy = a.f[7], *tzz = a.f[8], *txy = a.f[9], *txz = a.f[10],
          *tyz = a.f[11], *p = a.f[12];

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
  A wxz[W], wyz[W], wzz[W], wp[W];
#pragma unroll
  for (int m = 0; m < W - 1; ++m) {
    wxz[m] = ldi(txz, col + (u32)(kb - h + m) * sz);
    wyz[m] = ldi(tyz, col + (u32)(kb - h + m) * sz);
    wzz[m] = ldi(tzz, col + (u32)(kb - h + 1 + m) * sz);
    wp[m] = ldi(p, col + (u32)(kb - h + 1 + m) * sz);
  }

  for (int k = kb; k < ke; ++k) {
    const u32 c = col + (u32)k * sz;
    wxz[W - 1] = ldi(txz, c + (u32)(h - 1) * sz);
    wyz[W - 1] = ldi(tyz, c + (u32)(h - 1) * sz);
    wzz[W - 1] = ldi(tzz, c + (u32)h * sz);
    wp[W - 1] = ldi(p, c + (u32)h * sz);
    const A txz_c = wxz[h], tyz_c = wyz[h], p_c = wp[h - 1];
    const A txx_c = ldi(txx, c), tyy_c = ldi(tyy, c), txy_c = ldi(txy, c);

    // twelve raw staggered differences (kernel.cu:172-185 without the 1/dx factor)
    A x1 = dfw<ORDER>(txx, c, 1u, txx_c);
    A x2 = dbw<ORDER>(txy, c, sy, txy_c);
    A x3 = dwin<ORDER>(wxz);
    A s1 = dfw<ORDER>(p, c, 1u, p_c);
    A y1 = dfw<ORDER>(tyy, c, sy, tyy_c);
    A y2 = dbw<ORDER>(txy, c, 1u, txy_c);
    A y3 = dwin<ORDER>(wyz);
    A s2 = dfw<ORDER>(p, c, sy, p_c);
    A z1 = dwin<ORDER>(wzz);
    A z2 = dbw<ORDER>(txz, c, 1u, txz_c);
    A z3 = dbw<ORDER>(tyz, c, sy, tyz_c);
    A s3 = dwin<ORDER>(wp);

    // C-PML memory (kernel.cu:188-216): x memory of x1, s1 (half node), y2, z2 (integer)
    if (px >= 0) {
      const u32 o = (u32)px + wxs * ((u32)j + ny * (u32)k);
      x1 = pml(a.carry[0], o, x1, exh, axh);
      s1 = pml(a.carry[0], o + sl0, s1, exh, axh);
      y2 = pml(a.carry[0], o + 2 * sl0, y2, exi, axi);
      z2 = pml(a.carry[0], o + 3 * sl0, z2, exi, axi);
    }
    // y memory of x2, z3 (integer), y1, s2 (half)
    if (py >= 0) {
      const u32 o = (u32)i + nx * ((u32)py + wys * (u32)k);
      x2 = pml(a.carry[1], o, x2, eyi, ayi);
      y1 = pml(a.carry[1], o + sl1, y1, eyh, ayh);
      s2 = pml(a.carry[1], o + 2 * sl1, s2, eyh, ayh);
      z3 = pml(a.carry[1], o + 3 * sl1, z3, eyi, ayi);
    }
    // z memory of x3, y3 (integer), z1, s3 (half)
    const int pz = slab_pos(k, a.nz, a.L[2], a.R[2]);
    if (pz >= 0) {
      const A ezi = prof<A>(a.prof[2][0], k), azi = prof<A>(a.prof[2][1], k);
      const A ezh = prof<A>(a.prof[2][2], k), azh = prof<A>(a.prof[2][3], k);
      const u32 o = (u32)i + nx * ((u32)j + ny * (u32)pz);
      x3 = pml(a.carry[2], o, x3, ezi, azi);
      y3 = pml(a.carry[2], o + sl2, y3, ezi, azi);
      z1 = pml(a.carry[2], o + 2 * sl2, z1, ezh, azh);
      s3 = pml(a.carry[2], o + 3 * sl2, s3, ezh, azh);
    }

    // relative fluid velocity (kernel.cu:218-220; qz takes W3y as the reference does) and
    // solid velocity with the stored new q (kernel.cu:223-225)
    const A sx = x1 + x2 + x3, sy_ = y1 + y2 + y3, sz_ = z1 + z2 + z3;
    const A qx0 = ldi(qx, c), qy0 = ldi(qy, c), qz0 = ldi(qz, c);
    const S qx1 = st<S>(coef<S, A>(a, K_W1X, c) * qx0 - coef<S, A>(a, K_W2X, c) * sx -
                        coef<S, A>(a, K_W3X, c) * s1);
    const S qy1 = st<S>(coef<S, A>(a, K_W1Y, c) * qy0 - coef<S, A>(a, K_W2Y, c) * sy_ -
                        coef<S, A>(a, K_W3Y, c) * s2);
    const S qz1 = st<S>(coef<S, A>(a, K_W1Z, c) * qz0 - coef<S, A>(a, K_W2Z, c) * sz_ -
                        coef<S, A>(a, K_W3Y, c) * s3);
    at(qx, c) = qx1;
    at(qy, c) = qy1;
    at(qz, c) = qz1;
    at(vx, c) = st<S>(ldi(vx, c) + coef<S, A>(a, K_U1X, c) * sx -
                      coef<S, A>(a, K_U2X, c) * (ld(qx1) - qx0));
    at(vy, c) = st<S>(ldi(vy, c) + coef<S, A>(a, K_U1Y, c) * sy_ -
                      coef<S, A>(a, K_U2Y, c) * (ld(qy1) - qy0));
    at(vz, c) = st<S>(ldi(vz, c) + coef<S, A>(a, K_U1Z, c) * sz_ -
                      coef<S, A>(a, K_U2Z, c) * (ld(qz1) - qz0));

#pragma unroll
    for (int m = 0; m < W - 1; ++m) {
      wxz[m] = wxz[m + 1];
      wyz[m] = wyz[m + 1];
      wzz[m] = wzz[m + 1];
      wp[m] = wp[m + 1];
    }
  }
}

template <typename S, int ORDER>
__global__ void __launch_bounds__(256) k_stress(const StepArgs<S> a)
{
  using A = typename Arith<S>::A;
  constexpr int h = ORDER / 2;
  constexpr int W = 2 * h;
  int tx, ty, tz;
  if (!tile_of(a, tx, ty, tz)) return;
  const int i = tx * blockDim.x + threadIdx.x;
  const int j = ty * blockDim.y + threadIdx.y;
  if (i < h || i >= a.nx - h || j < h || j >= a.ny - h) return;
  const int kb = h + tz * a.zchunk;
  const int ke = min(kb + a.zchunk, a.nz - h);
  if (kb >= ke) return;
  const u32 sy = (u32)a.sy, sz = (u32)a.sz;
  const u32 col = (u32)i + (u32)j * sy;

  const S *vx = a.f[0], *vy = a.f[1], *vz = a.f[2], *qx = a.f[3], *qy = a.f[4], *qz = a.f[5];
  S *txx = a.f[6], *tyy = a.f[7], *tzz = a.f[8], *txy = a.f[9], *txz = a.f[10], *tyz = a.f[11],
    *p = a.f[12];

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
    wvz[m] = ldi(vz, col + (u32)(kb - h + m) * sz);
    wqz[m] = ldi(qz, col + (u32)(kb - h + m) * sz);
    wvx[m] = ldi(vx, col + (u32)(kb - h + 1 + m) * sz);
    wvy[m] = ldi(vy, col + (u32)(kb - h + 1 + m) * sz);
  }

  for (int k = kb; k < ke; ++k) {
    const u32 c = col + (u32)k * sz;
    wvz[W - 1] = ldi(vz, c + (u32)(h - 1) * sz);
    wqz[W - 1] = ldi(qz, c + (u32)(h - 1) * sz);
    wvx[W - 1] = ldi(vx, c + (u32)h * sz);
    wvy[W - 1] = ldi(vy, c + (u32)h * sz);
    const A vz_c = wvz[h], vx_c = wvx[h - 1], vy_c = wvy[h - 1];
    const A qx_c = ldi(qx, c), qy_c = ldi(qy, c);

    // twelve raw differences (kernel.cu:430-445 without 1/dx)
    A uxx = dbw<ORDER>(vx, c, 1u, vx_c);
    A uyy = dbw<ORDER>(vy, c, sy, vy_c);
    A uzz = dwin<ORDER>(wvz);
    A wx = dbw<ORDER>(qx, c, 1u, qx_c);
    A wy = dbw<ORDER>(qy, c, sy, qy_c);
    A wz = dwin<ORDER>(wqz);
    A uxy = dfw<ORDER>(vx, c, sy, vx_c);
    A uyx = dfw<ORDER>(vy, c, 1u, vy_c);
    A uxz = dwin<ORDER>(wvx);
    A uzx = dfw<ORDER>(vz, c, 1u, vz_c);
    A uyz = dwin<ORDER>(wvy);
    A uzy = dfw<ORDER>(vz, c, sy, vz_c);

    // C-PML memory (kernel.cu:449-482): carries 4..7 of each axis
    if (px >= 0) {
      const u32 o = 4 * sl0 + (u32)px + wxs * ((u32)j + ny * (u32)k);
      uxx = pml(a.carry[0], o, uxx, exi, axi);
      wx = pml(a.carry[0], o + sl0, wx, exi, axi);
      uyx = pml(a.carry[0], o + 2 * sl0, uyx, exh, axh);
      uzx = pml(a.carry[0], o + 3 * sl0, uzx, exh, axh);
    }
    if (py >= 0) {
      const u32 o = 4 * sl1 + (u32)i + nx * ((u32)py + wys * (u32)k);
      uyy = pml(a.carry[1], o, uyy, eyi, ayi);
      wy = pml(a.carry[1], o + sl1, wy, eyi, ayi);
      uxy = pml(a.carry[1], o + 2 * sl1, uxy, eyh, ayh);
      uzy = pml(a.carry[1], o + 3 * sl1, uzy, eyh, ayh);
    }
    const int pz = slab_pos(k, a.nz, a.L[2], a.R[2]);
    if (pz >= 0) {
      const A ezi = prof<A>(a.prof[2][0], k), azi = prof<A>(a.prof[2][1], k);
      const A ezh = prof<A>(a.prof[2][2], k), azh = prof<A>(a.prof[2][3], k);
      const u32 o = 4 * sl2 + (u32)i + nx * ((u32)j + ny * (u32)pz);
      uzz = pml(a.carry[2], o, uzz, ezi, azi);
      wz = pml(a.carry[2], o + sl2, wz, ezi, azi);
      uxz = pml(a.carry[2], o + 2 * sl2, uxz, ezh, azh);
      uyz = pml(a.carry[2], o + 3 * sl2, uyz, ezh, azh);
    }

    // pore pressure and stresses (kernel.cu:484-490)
    const A divw = wx + wy + wz;
    const A s1 = coef<S, A>(a, K_S1, c), s2 = coef<S, A>(a, K_S2, c), s3 = coef<S, A>(a, K_S3, c);
    at(p, c) = st<S>(ldi(p, c) - coef<S, A>(a, K_P1, c) * (uxx + uyy + uzz) -
                     coef<S, A>(a, K_P2, c) * divw);
    at(txx, c) = st<S>(ldi(txx, c) + s1 * (uyy + uzz) + s2 * uxx + s3 * divw);
    at(tyy, c) = st<S>(ldi(tyy, c) + s1 * (uxx + uzz) + s2 * uyy + s3 * divw);
    at(tzz, c) = st<S>(ldi(tzz, c) + s1 * (uxx + uyy) + s2 * uzz + s3 * divw);
    at(txy, c) = st<S>(ldi(txy, c) + coef<S, A>(a, K_SXY, c) * (uxy + uyx));
    at(tyz, c) = st<S>(ldi(tyz, c) + coef<S, A>(a, K_SYZ, c) * (uyz + uzy));
    at(txz, c) = st<S>(ldi(txz, c) + coef<S, A>(a, K_SXZ, c) * (uxz + uzx));

#pragma unroll
    for (int m = 0; m < W - 1; ++m) {
      wvz[m] = wvz[m + 1];
      wqz[m] = wqz[m + 1];
      wvx[m] = wvx[m + 1];
      wvy[m] = wvy[m + 1];
    }
  }
}

// Source injection (kernel.cu:130-142) for one cell: one lane per field.
template <typename S>
__global__ void k_source(const SourceArgs a)
{
  using A = typename Arith<S>::A;
  const int f = threadIdx.x;
  if (f >= NFIELD || !(a.mask & (1u << f))) return;
  S *x = (S *)a.f[f];
  const A amp = (A)a.wavelet[a.it] * (A)a.inv_scale[f];
  x[a.cell] = st<S>(ld(x[a.cell]) + amp);
}

// Receiver sampling: traces[comp][it][rec] = field value * storage scale.
template <typename S>
__global__ void k_record(const RecordArgs a)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ci = blockIdx.y;
  if (r >= a.nrec || ci >= a.ncomp) return;
  const S *x = (const S *)a.f[a.comp[ci]];
  const float v = (float)x[a.rec[r]] * a.scale[ci];
  a.traces[((int64_t)ci * a.nt_max + a.it) * a.nrec + r] = v;
}

// In-place multiply of a stored array (re-expresses values when a storage scale changes).
template <typename S>
__global__ void k_rescale(S *p, int64_t n, double f)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if constexpr (std::is_same<S, double>::value)
      p[i] = p[i] * f;
    else
      p[i] = (S)((float)p[i] * (float)f);
  }
}

template <typename S, int ORDER, bool VEL>
hipError_t launch_step(const void *args, LaunchShape sh, hipStream_t s)
{
  StepArgs<S> a = *static_cast<const StepArgs<S> *>(args);
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
    hipLaunchKernelGGL((k_velocity<S, ORDER>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_stress<S, ORDER>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <bool VEL>
hipError_t dispatch(int prec, int order, const void *args, LaunchShape sh, hipStream_t s)
{
  if (sh.bx * sh.by > 256 || sh.bx <= 0 || sh.by <= 0) return hipErrorInvalidValue;
  if (order == 2) {
    if (prec == 0) return launch_step<float, 2, VEL>(args, sh, s);
    if (prec == 1) return launch_step<_Float16, 2, VEL>(args, sh, s);
    if (prec == 2) return launch_step<double, 2, VEL>(args, sh, s);
  } else if (order == 4) {
    if (prec == 0) return launch_step<float, 4, VEL>(args, sh, s);
    if (prec == 1) return launch_step<_Float16, 4, VEL>(args, sh, s);
    if (prec == 2) return launch_step<double, 4, VEL>(args, sh, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_velocity(int prec, int order, const void *args, LaunchShape sh, hipStream_t s)
{
  return dispatch<true>(prec, order, args, sh, s);
}
hipError_t launch_stress(int prec, int order, const void *args, LaunchShape sh, hipStream_t s)
{
  return dispatch<false>(prec, order, args, sh, s);
}

hipError_t launch_source(int prec, const SourceArgs &a, hipStream_t s)
{
  if (prec == 0) hipLaunchKernelGGL(k_source<float>, dim3(1), dim3(64), 0, s, a);
  else if (prec == 1) hipLaunchKernelGGL(k_source<_Float16>, dim3(1), dim3(64), 0, s, a);
  else if (prec == 2) hipLaunchKernelGGL(k_source<double>, dim3(1), dim3(64), 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_rescale(int prec, void *p, int64_t n, double f, hipStream_t s)
{
  if (n <= 0 || !p) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  if (prec == 0) hipLaunchKernelGGL(k_rescale<float>, dim3(blocks), dim3(256), 0, s, (float *)p, n, f);
  else if (prec == 1) hipLaunchKernelGGL(k_rescale<_Float16>, dim3(blocks), dim3(256), 0, s, (_Float16 *)p, n, f);
  else if (prec == 2) hipLaunchKernelGGL(k_rescale<double>, dim3(blocks), dim3(256), 0, s, (double *)p, n, f);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_record(int prec, const RecordArgs &a, hipStream_t s)
{
  if (a.nrec <= 0 || a.ncomp <= 0) return hipSuccess;
  dim3 block(256), grid((unsigned)((a.nrec + 255) / 256), (unsigned)a.ncomp);
  if (prec == 0) hipLaunchKernelGGL(k_record<float>, grid, block, 0, s, a);
  else if (prec == 1) hipLaunchKernelGGL(k_record<_Float16>, grid, block, 0, s, a);
  else if (prec == 2) hipLaunchKernelGGL(k_record<double>, grid, block, 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace biot
