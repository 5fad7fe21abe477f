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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "biot_device.hpp"

namespace biot {
namespace {

template <typename S>
struct Arith {
  using A = float;
};
template <>
struct Arith<double> {
  using A = double;
};

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
__device__ __forceinline__ A dfw(const S *__restrict__ f, int64_t c, int64_t s, A fc)
{
  if constexpr (ORDER == 2) {
    return ld(f[c + s]) - fc;
  } else {
    return A(9.0 / 8.0) * (ld(f[c + s]) - fc) + A(-1.0 / 24.0) * (ld(f[c + 2 * s]) - ld(f[c - s]));
  }
}
// backward difference at c - s/2 (value at c supplied)
template <int ORDER, typename S, typename A>
__device__ __forceinline__ A dbw(const S *__restrict__ f, int64_t c, int64_t s, A fc)
{
  if constexpr (ORDER == 2) {
    return fc - ld(f[c - s]);
  } else {
    return A(9.0 / 8.0) * (fc - ld(f[c - s])) + A(-1.0 / 24.0) * (ld(f[c + s]) - ld(f[c - 2 * s]));
  }
}

template <typename S, typename A>
__device__ __forceinline__ A coef(const StepArgs<S> &a, int idx, int64_t c)
{
  if constexpr (std::is_same<S, _Float16>::value) {
    return ld(a.k[idx][c]) * a.ks[idx];
  } else {
    return ld(a.k[idx][c]);
  }
}

// One folded C-PML update: psi = carry + a D; carry <- e (psi + a D); returns D + psi.
template <typename S, typename A>
__device__ __forceinline__ A pml(S *__restrict__ cp, A D, A e, A aa)
{
  const A psi = ld(*cp) + aa * D;
  *cp = st<S>(e * (psi + aa * D));
  return D + psi;
}

// slab position of index i on an axis with widths L (low) and R (high), or -1
__device__ __forceinline__ int slab_pos(int i, int n, int L, int R)
{
  if (i < L) return i;
  if (i >= n - R) return L + i - (n - R);
  return -1;
}

template <typename S, int ORDER>
__global__ void __launch_bounds__(256) k_velocity(const StepArgs<S> a)
{
  using A = typename Arith<S>::A;
  constexpr int h = ORDER / 2;
  constexpr int W = 2 * h;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i < h || i >= a.nx - h || j < h || j >= a.ny - h) return;
  const int kb = h + blockIdx.z * a.zchunk;
  const int ke = min(kb + a.zchunk, a.nz - h);
  if (kb >= ke) return;
  const int64_t sy = a.sy, sz = a.sz;
  const int64_t col = i + j * sy;

  S *__restrict__ vx = a.f[0];
  S *__restrict__ vy = a.f[1];
  S *__restrict__ vz = a.f[2];
  S *__restrict__ qx = a.f[3];
  S *__restrict__ qy = a.f[4];
  S *__restrict__ qz = a.f[5];
  const S *__restrict__ txx = a.f[6];
  const S *__restrict__ tyy = a.f[7];
  const S *__restrict__ tzz = a.f[8];
  const S *__restrict__ txy = a.f[9];
  const S *__restrict__ txz = a.f[10];
  const S *__restrict__ tyz = a.f[11];
  const S *__restrict__ p = a.f[12];

  // x / y slab positions and profile values are constant along the column
  const int px = slab_pos(i, a.nx, a.L[0], a.R[0]);
  const int py = slab_pos(j, a.ny, a.L[1], a.R[1]);
  A exi = 0, axi = 0, exh = 0, axh = 0, eyi = 0, ayi = 0, eyh = 0, ayh = 0;
  if (px >= 0) {
    const A *pr = (const A *)a.prof[0][0];
    exi = pr[i];
    axi = ((const A *)a.prof[0][1])[i];
    exh = ((const A *)a.prof[0][2])[i];
    axh = ((const A *)a.prof[0][3])[i];
  }
  if (py >= 0) {
    eyi = ((const A *)a.prof[1][0])[j];
    ayi = ((const A *)a.prof[1][1])[j];
    eyh = ((const A *)a.prof[1][2])[j];
    ayh = ((const A *)a.prof[1][3])[j];
  }
  const int wxs = a.L[0] + a.R[0], wys = a.L[1] + a.R[1];

  // z windows: backward-differenced txz, tyz hold planes k-h..k+h-1,
  // forward-differenced tzz, p hold planes k-h+1..k+h.
  A wxz[W], wyz[W], wzz[W], wp[W];
#pragma unroll
  for (int m = 0; m < W - 1; ++m) {
    wxz[m] = ld(txz[col + (int64_t)(kb - h + m) * sz]);
    wyz[m] = ld(tyz[col + (int64_t)(kb - h + m) * sz]);
    wzz[m] = ld(tzz[col + (int64_t)(kb - h + 1 + m) * sz]);
    wp[m] = ld(p[col + (int64_t)(kb - h + 1 + m) * sz]);
  }

  for (int k = kb; k < ke; ++k) {
    const int64_t c = col + (int64_t)k * sz;
    wxz[W - 1] = ld(txz[c + (int64_t)(h - 1) * sz]);
    wyz[W - 1] = ld(tyz[c + (int64_t)(h - 1) * sz]);
    wzz[W - 1] = ld(tzz[c + (int64_t)h * sz]);
    wp[W - 1] = ld(p[c + (int64_t)h * sz]);
    const A txz_c = wxz[h], tyz_c = wyz[h], p_c = wp[h - 1];
    const A txx_c = ld(txx[c]), tyy_c = ld(tyy[c]), txy_c = ld(txy[c]);

    // twelve raw staggered differences (kernel.cu:172-185 without the 1/dx factor)
    A x1 = dfw<ORDER>(txx, c, 1, txx_c);
    A x2 = dbw<ORDER>(txy, c, sy, txy_c);
    A x3 = dwin<ORDER>(wxz);
    A s1 = dfw<ORDER>(p, c, 1, p_c);
    A y1 = dfw<ORDER>(tyy, c, sy, tyy_c);
    A y2 = dbw<ORDER>(txy, c, 1, txy_c);
    A y3 = dwin<ORDER>(wyz);
    A s2 = dfw<ORDER>(p, c, sy, p_c);
    A z1 = dwin<ORDER>(wzz);
    A z2 = dbw<ORDER>(txz, c, 1, txz_c);
    A z3 = dbw<ORDER>(tyz, c, sy, tyz_c);
    A s3 = dwin<ORDER>(wp);

    // C-PML memory (kernel.cu:188-216): x memory of x1, s1 (half node), y2, z2 (integer)
    if (px >= 0) {
      S *cp = a.carry[0] + px + (int64_t)wxs * (j + (int64_t)a.ny * k);
      const int64_t st_ = a.slab[0];
      x1 = pml(cp, x1, exh, axh);
      s1 = pml(cp + st_, s1, exh, axh);
      y2 = pml(cp + 2 * st_, y2, exi, axi);
      z2 = pml(cp + 3 * st_, z2, exi, axi);
    }
    // y memory of x2, z3 (integer), y1, s2 (half)
    if (py >= 0) {
      S *cp = a.carry[1] + i + (int64_t)a.nx * (py + (int64_t)wys * k);
      const int64_t st_ = a.slab[1];
      x2 = pml(cp, x2, eyi, ayi);
      y1 = pml(cp + st_, y1, eyh, ayh);
      s2 = pml(cp + 2 * st_, s2, eyh, ayh);
      z3 = pml(cp + 3 * st_, z3, eyi, ayi);
    }
    // z memory of x3, y3 (integer), z1, s3 (half)
    const int pz = slab_pos(k, a.nz, a.L[2], a.R[2]);
    if (pz >= 0) {
      const A ezi = ((const A *)a.prof[2][0])[k], azi = ((const A *)a.prof[2][1])[k];
      const A ezh = ((const A *)a.prof[2][2])[k], azh = ((const A *)a.prof[2][3])[k];
      S *cp = a.carry[2] + i + (int64_t)a.nx * (j + (int64_t)a.ny * pz);
      const int64_t st_ = a.slab[2];
      x3 = pml(cp, x3, ezi, azi);
      y3 = pml(cp + st_, y3, ezi, azi);
      z1 = pml(cp + 2 * st_, z1, ezh, azh);
      s3 = pml(cp + 3 * st_, s3, ezh, azh);
    }

    // relative fluid velocity (kernel.cu:218-220; qz takes W3y as the reference does) and
    // solid velocity with the stored new q (kernel.cu:223-225)
    const A sx = x1 + x2 + x3, sy_ = y1 + y2 + y3, sz_ = z1 + z2 + z3;
    const A qx0 = ld(qx[c]), qy0 = ld(qy[c]), qz0 = ld(qz[c]);
    const S qx1 = st<S>(coef<S, A>(a, K_W1X, c) * qx0 - coef<S, A>(a, K_W2X, c) * sx -
                        coef<S, A>(a, K_W3X, c) * s1);
    const S qy1 = st<S>(coef<S, A>(a, K_W1Y, c) * qy0 - coef<S, A>(a, K_W2Y, c) * sy_ -
                        coef<S, A>(a, K_W3Y, c) * s2);
    const S qz1 = st<S>(coef<S, A>(a, K_W1Z, c) * qz0 - coef<S, A>(a, K_W2Z, c) * sz_ -
                        coef<S, A>(a, K_W3Y, c) * s3);
    qx[c] = qx1;
    qy[c] = qy1;
    qz[c] = qz1;
    vx[c] = st<S>(ld(vx[c]) + coef<S, A>(a, K_U1X, c) * sx - coef<S, A>(a, K_U2X, c) * (ld(qx1) - qx0));
    vy[c] = st<S>(ld(vy[c]) + coef<S, A>(a, K_U1Y, c) * sy_ - coef<S, A>(a, K_U2Y, c) * (ld(qy1) - qy0));
    vz[c] = st<S>(ld(vz[c]) + coef<S, A>(a, K_U1Z, c) * sz_ - coef<S, A>(a, K_U2Z, c) * (ld(qz1) - qz0));

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
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i < h || i >= a.nx - h || j < h || j >= a.ny - h) return;
  const int kb = h + blockIdx.z * a.zchunk;
  const int ke = min(kb + a.zchunk, a.nz - h);
  if (kb >= ke) return;
  const int64_t sy = a.sy, sz = a.sz;
  const int64_t col = i + j * sy;

  const S *__restrict__ vx = a.f[0];
  const S *__restrict__ vy = a.f[1];
  const S *__restrict__ vz = a.f[2];
  const S *__restrict__ qx = a.f[3];
  const S *__restrict__ qy = a.f[4];
  const S *__restrict__ qz = a.f[5];
  S *__restrict__ txx = a.f[6];
  S *__restrict__ tyy = a.f[7];
  S *__restrict__ tzz = a.f[8];
  S *__restrict__ txy = a.f[9];
  S *__restrict__ txz = a.f[10];
  S *__restrict__ tyz = a.f[11];
  S *__restrict__ p = a.f[12];

  const int px = slab_pos(i, a.nx, a.L[0], a.R[0]);
  const int py = slab_pos(j, a.ny, a.L[1], a.R[1]);
  A exi = 0, axi = 0, exh = 0, axh = 0, eyi = 0, ayi = 0, eyh = 0, ayh = 0;
  if (px >= 0) {
    exi = ((const A *)a.prof[0][0])[i];
    axi = ((const A *)a.prof[0][1])[i];
    exh = ((const A *)a.prof[0][2])[i];
    axh = ((const A *)a.prof[0][3])[i];
  }
  if (py >= 0) {
    eyi = ((const A *)a.prof[1][0])[j];
    ayi = ((const A *)a.prof[1][1])[j];
    eyh = ((const A *)a.prof[1][2])[j];
    ayh = ((const A *)a.prof[1][3])[j];
  }
  const int wxs = a.L[0] + a.R[0], wys = a.L[1] + a.R[1];

  // z windows: backward-differenced vz, qz hold planes k-h..k+h-1,
  // forward-differenced vx, vy hold planes k-h+1..k+h.
  A wvz[W], wqz[W], wvx[W], wvy[W];
#pragma unroll
  for (int m = 0; m < W - 1; ++m) {
    wvz[m] = ld(vz[col + (int64_t)(kb - h + m) * sz]);
    wqz[m] = ld(qz[col + (int64_t)(kb - h + m) * sz]);
    wvx[m] = ld(vx[col + (int64_t)(kb - h + 1 + m) * sz]);
    wvy[m] = ld(vy[col + (int64_t)(kb - h + 1 + m) * sz]);
  }

  for (int k = kb; k < ke; ++k) {
    const int64_t c = col + (int64_t)k * sz;
    wvz[W - 1] = ld(vz[c + (int64_t)(h - 1) * sz]);
    wqz[W - 1] = ld(qz[c + (int64_t)(h - 1) * sz]);
    wvx[W - 1] = ld(vx[c + (int64_t)h * sz]);
    wvy[W - 1] = ld(vy[c + (int64_t)h * sz]);
    const A vz_c = wvz[h], vx_c = wvx[h - 1], vy_c = wvy[h - 1];
    const A qx_c = ld(qx[c]), qy_c = ld(qy[c]);

    // twelve raw differences (kernel.cu:430-445 without 1/dx)
    A uxx = dbw<ORDER>(vx, c, 1, vx_c);
    A uyy = dbw<ORDER>(vy, c, sy, vy_c);
    A uzz = dwin<ORDER>(wvz);
    A wx = dbw<ORDER>(qx, c, 1, qx_c);
    A wy = dbw<ORDER>(qy, c, sy, qy_c);
    A wz = dwin<ORDER>(wqz);
    A uxy = dfw<ORDER>(vx, c, sy, vx_c);
    A uyx = dfw<ORDER>(vy, c, 1, vy_c);
    A uxz = dwin<ORDER>(wvx);
    A uzx = dfw<ORDER>(vz, c, 1, vz_c);
    A uyz = dwin<ORDER>(wvy);
    A uzy = dfw<ORDER>(vz, c, sy, vz_c);

    // C-PML memory (kernel.cu:449-482): carries 4..7 of each axis
    if (px >= 0) {
      const int64_t st_ = a.slab[0];
      S *cp = a.carry[0] + 4 * st_ + px + (int64_t)wxs * (j + (int64_t)a.ny * k);
      uxx = pml(cp, uxx, exi, axi);
      wx = pml(cp + st_, wx, exi, axi);
      uyx = pml(cp + 2 * st_, uyx, exh, axh);
      uzx = pml(cp + 3 * st_, uzx, exh, axh);
    }
    if (py >= 0) {
      const int64_t st_ = a.slab[1];
      S *cp = a.carry[1] + 4 * st_ + i + (int64_t)a.nx * (py + (int64_t)wys * k);
      uyy = pml(cp, uyy, eyi, ayi);
      wy = pml(cp + st_, wy, eyi, ayi);
      uxy = pml(cp + 2 * st_, uxy, eyh, ayh);
      uzy = pml(cp + 3 * st_, uzy, eyh, ayh);
    }
    const int pz = slab_pos(k, a.nz, a.L[2], a.R[2]);
    if (pz >= 0) {
      const A ezi = ((const A *)a.prof[2][0])[k], azi = ((const A *)a.prof[2][1])[k];
      const A ezh = ((const A *)a.prof[2][2])[k], azh = ((const A *)a.prof[2][3])[k];
      const int64_t st_ = a.slab[2];
      S *cp = a.carry[2] + 4 * st_ + i + (int64_t)a.nx * (j + (int64_t)a.ny * pz);
      uzz = pml(cp, uzz, ezi, azi);
      wz = pml(cp + st_, wz, ezi, azi);
      uxz = pml(cp + 2 * st_, uxz, ezh, azh);
      uyz = pml(cp + 3 * st_, uyz, ezh, azh);
    }

    // pore pressure and stresses (kernel.cu:484-490)
    const A divw = wx + wy + wz;
    const A s1 = coef<S, A>(a, K_S1, c), s2 = coef<S, A>(a, K_S2, c), s3 = coef<S, A>(a, K_S3, c);
    p[c] = st<S>(ld(p[c]) - coef<S, A>(a, K_P1, c) * (uxx + uyy + uzz) - coef<S, A>(a, K_P2, c) * divw);
    txx[c] = st<S>(ld(txx[c]) + s1 * (uyy + uzz) + s2 * uxx + s3 * divw);
    tyy[c] = st<S>(ld(tyy[c]) + s1 * (uxx + uzz) + s2 * uyy + s3 * divw);
    tzz[c] = st<S>(ld(tzz[c]) + s1 * (uxx + uyy) + s2 * uzz + s3 * divw);
    txy[c] = st<S>(ld(txy[c]) + coef<S, A>(a, K_SXY, c) * (uxy + uyx));
    tyz[c] = st<S>(ld(tyz[c]) + coef<S, A>(a, K_SYZ, c) * (uyz + uzy));
    txz[c] = st<S>(ld(txz[c]) + coef<S, A>(a, K_SXZ, c) * (uxz + uzx));

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
  const StepArgs<S> &a = *static_cast<const StepArgs<S> *>(args);
  const int h = ORDER / 2;
  const int nzi = a.nz - 2 * h;
  if (nzi <= 0 || a.nx <= 2 * h || a.ny <= 2 * h) return hipSuccess;
  dim3 block(sh.bx, sh.by, 1);
  dim3 grid((a.nx + sh.bx - 1) / sh.bx, (a.ny + sh.by - 1) / sh.by,
            (nzi + a.zchunk - 1) / a.zchunk);
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
