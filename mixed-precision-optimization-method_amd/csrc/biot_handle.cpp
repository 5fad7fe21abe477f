// Device memory management and coefficient/profile uploads of a biot_handle.
#include "biot_handle.hpp"
#include "biot_setup.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

using namespace biot;

// One time-stepping state (wavefields + C-PML carries) in one device buffer: the fields
// block at offset 0, the residual block of the 16-bit policies that have one, then
// the three carry blocks, each piece starting on a 256-byte boundary.
static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }
size_t biot_handle::state_lo_off() const { return round256((size_t)NFIELD * n * esz); }
size_t biot_handle::state_off(int piece) const
{
  size_t off = state_lo_off() + round256(lo_bytes());
  for (int ax = 0; ax < piece - 1; ++ax) off += round256((size_t)NCARRY * slab[ax] * csz);
  return piece == 0 ? 0 : off;
}
size_t biot_handle::state_bytes() const { return state_off(4); }

void biot_handle::save_state(void *dst)
{
  char *p = (char *)dst;
  (void)hipMemcpyAsync(p, d_fields, (size_t)NFIELD * n * esz, hipMemcpyDeviceToDevice, stream);
  if (d_lo) (void)hipMemcpyAsync(p + state_lo_off(), d_lo, lo_bytes(), hipMemcpyDeviceToDevice, stream);
  for (int ax = 0; ax < 3; ++ax) {
    const size_t b = (size_t)NCARRY * slab[ax] * csz;
    if (b) (void)hipMemcpyAsync(p + state_off(1 + ax), d_carry[ax], b, hipMemcpyDeviceToDevice, stream);
  }
}

void biot_handle::load_state(const void *src)
{
  const char *p = (const char *)src;
  (void)hipMemcpyAsync(d_fields, p, (size_t)NFIELD * n * esz, hipMemcpyDeviceToDevice, stream);
  if (d_lo) (void)hipMemcpyAsync(d_lo, p + state_lo_off(), lo_bytes(), hipMemcpyDeviceToDevice, stream);
  for (int ax = 0; ax < 3; ++ax) {
    const size_t b = (size_t)NCARRY * slab[ax] * csz;
    if (b) (void)hipMemcpyAsync(d_carry[ax], p + state_off(1 + ax), b, hipMemcpyDeviceToDevice, stream);
  }
}

void biot_handle::source_args(SourceArgs &sa) const
{
  std::memset(&sa, 0, sizeof sa);
  for (int f = 0; f < NFIELD; ++f) {
    sa.f[f] = field_ptr(f);
    sa.lo[f] = lo_ptr(f);
    sa.inv_scale[f] = 1.0f / fscale[f];
  }
  sa.cell = src_cell;
  sa.mask = src_mask;
  sa.wavelet = (const float *)d_wavelet;
  sa.nt_w = (int)wavelet.size();
}

void biot_handle::record_args(RecordArgs &ra) const
{
  std::memset(&ra, 0, sizeof ra);
  for (int f = 0; f < NFIELD; ++f) {
    ra.f[f] = field_ptr(f);
    ra.lo[f] = lo_ptr(f);
  }
  ra.rec = (const int64_t *)d_rec;
  ra.nrec = nrec;
  ra.ncomp = ncomp;
  for (int c = 0; c < ncomp; ++c) {
    ra.comp[c] = comp[c];
    ra.scale[c] = fscale[comp[c]];
  }
  ra.traces = (float *)d_traces;
  ra.traces64 = d_traces64;
  ra.nonfinite = d_flag;
  ra.nt_max = nt_max;
}

int biot_handle::clear_nonfinite()
{
  if (hipMemsetAsync(d_flag, 0x7F, sizeof(int), stream) != hipSuccess) {
    err = "clearing the non-finite flag failed";
    return BIOT_ERR_HIP;
  }
  return BIOT_OK;
}

int biot_handle::check_nonfinite()
{
  int first = NONFINITE_NONE;
  if (hipStreamSynchronize(stream) != hipSuccess ||
      hipMemcpy(&first, d_flag, sizeof first, hipMemcpyDeviceToHost) != hipSuccess) {
    err = "reading the non-finite flag failed";
    return BIOT_ERR_HIP;
  }
  if (first == NONFINITE_NONE) return BIOT_OK;
  char buf[160];
  snprintf(buf, sizeof buf, "non-finite value on the receivers at step %d (the steps ran; traces and fields hold it)",
           first);
  err = buf;
  return BIOT_ERR_NONFINITE;
}

size_t biot_handle::lam_bytes() const
{
  const size_t asz = prec == BIOT_FP64 ? 8 : 4;
  size_t b = (size_t)NFIELD * n * asz + (size_t)12 * n * asz;
  for (int ax = 0; ax < 3; ++ax) b += (size_t)NCARRY * slab[ax] * asz;
  return b;
}

int biot_handle::alloc_adjoint(int every)
{
  const int64_t m = (int64_t)ncomp * nt_max * nrec;
  // the adjoint kernels address each multiplier, derivative and carry-multiplier array with
  // 32-bit byte offsets from its base, as the forward kernels do their arrays
  const int64_t asz = prec == BIOT_FP64 ? 8 : 4, lim = (int64_t)1 << 32;
  bool fits = n * asz < lim;
  for (int ax = 0; ax < 3; ++ax) fits = fits && NCARRY * slab[ax] * asz < lim;
  if (!fits) {
    err = "grid too large for the adjoint (an adjoint array would exceed 4 GiB)";
    return BIOT_ERR_ARG;
  }
  auto get = [&](void **p, size_t bytes) -> bool {
    if (*p) return true;
    if (hipMalloc(p, bytes) != hipSuccess) {
      *p = nullptr;
      return false;
    }
    return true;
  };
  if ((every > 0 && !get(&d_slots, state_bytes() * (size_t)(every + 1))) || !get(&d_lam, lam_bytes()) ||
      !get((void **)&d_grad, (size_t)NKC * n * sizeof(double)) ||
      (prec != BIOT_FP64 && !get((void **)&d_gpart, (size_t)NKC * n * sizeof(float))) ||
      !get((void **)&d_adjsrc, (size_t)std::max<int64_t>(m, 1) * sizeof(float))) {
    err = "adjoint buffers do not fit in device memory";
    return BIOT_ERR_NOMEM;
  }
  return BIOT_OK;
}

void biot_handle::adj_args(AdjArgs &A) const
{
  const size_t asz = prec == BIOT_FP64 ? 8 : 4;
  const int hh = g.order / 2;
  std::memset(&A, 0, sizeof A);
  A.slots = (const char *)d_slots;
  A.sb = state_bytes();
  A.fbytes = (uint64_t)n * esz;
  // FP16T / FP16C residual arrays, FP16M's residual records (tau and p)
  const bool res = prec == BIOT_FP16T || prec == BIOT_FP16C || prec == BIOT_FP16M;
  A.loff = res ? state_lo_off() : 0;
  A.lomask = res ? comp_fields(prec) : 0u;
  for (int ax = 0; ax < 3; ++ax) A.coff[ax] = state_off(1 + ax);
  char *lp = (char *)d_lam;
  for (int f = 0; f < NFIELD; ++f) A.lam[f] = lp + (size_t)f * n * asz;
  lp += (size_t)NFIELD * n * asz;
  for (int ax = 0; ax < 3; ++ax) {
    A.lcarry[ax] = lp;
    lp += (size_t)NCARRY * slab[ax] * asz;
  }
  for (int d = 0; d < 12; ++d) A.G[d] = lp + (size_t)d * n * asz;
  for (int k = 0; k < NKC; ++k) {
    A.grad[k] = d_grad + (size_t)k * n;
    A.gacc[k] = prec == BIOT_FP64 ? (void *)A.grad[k] : (void *)(d_gpart + (size_t)k * n);
    // (FP16M keeps no per-array coefficients; the adjoint reads the packed copy only)
    A.k[k] = prec == BIOT_FP16M ? nullptr : (const char *)d_coef + (size_t)k * n * esz;
    A.ks[k] = ks[k];
  }
  A.gpv = prec == BIOT_FP64 ? nullptr : (void *)d_gpart;
  A.gpt = prec == BIOT_FP64 ? nullptr : (void *)(d_gpart + (size_t)14 * n);
  A.kv = d_kpack;
  A.kt = (const char *)d_kpack + (size_t)14 * n * esz;
  for (int ax = 0; ax < 3; ++ax) {
    for (int q = 0; q < 4; ++q) A.prof[ax][q] = d_prof[ax][q];
    A.L[ax] = L[ax];
    A.R[ax] = R[ax];
    A.slab[ax] = slab[ax];
  }
  A.nx = g.nx;
  A.ny = g.ny;
  A.nz = g.nz;
  A.sy = g.nx;
  A.sz = (int64_t)g.nx * g.ny;
  A.nupd = (uint32_t)((int64_t)(g.nx - 2 * hh) * (g.ny - 2 * hh) * (g.nz - 2 * hh));
}

void biot_handle::inject_args(InjectArgs &I) const
{
  const size_t asz = prec == BIOT_FP64 ? 8 : 4;
  std::memset(&I, 0, sizeof I);
  for (int f = 0; f < NFIELD; ++f) I.lam[f] = (char *)d_lam + (size_t)f * n * asz;
  I.rec = (const int64_t *)d_rec;
  I.nrec = nrec;
  I.ncomp = ncomp;
  for (int c = 0; c < ncomp; ++c) {
    I.comp[c] = comp[c];
    I.scale[c] = fscale[comp[c]];  // traces hold scale x stored value
  }
  I.src = d_adjsrc;
  I.nt = nt_max;
}

void biot_handle::free_adjoint()
{
  if (stream) (void)hipStreamSynchronize(stream);
  for (void *p : ckpt)
    if (p) (void)hipFree(p);
  ckpt.clear();
  ckpt_every = ckpt_nt = 0;
  if (prec == BIOT_FP16M && d_kpack) {  // the adjoint's coefficient copy of the last parameters
    (void)hipFree(d_kpack);
    d_kpack = nullptr;
  }
  for (void **p : {&d_slots, &d_lam, (void **)&d_grad, (void **)&d_gpart, (void **)&d_adjsrc})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
}

void biot_handle::forget_model()
{
  free_adjoint();
  for (auto &v : mod) std::vector<float>().swap(v);
  for (auto &v : par) std::vector<double>().swap(v);
  for (auto &v : K_last) std::vector<double>().swap(v);
}

biot_handle::~biot_handle()
{
  free_adjoint();
  if (stream) (void)hipStreamSynchronize(stream);
  if (d_fields) (void)hipFree(d_fields);
  if (d_lo) (void)hipFree(d_lo);
  if (d_coef) (void)hipFree(d_coef);
  if (d_kpack) (void)hipFree(d_kpack);
  if (d_flag) (void)hipFree(d_flag);
  for (auto *p : d_carry)
    if (p) (void)hipFree(p);
  for (auto &ax : d_prof)
    for (auto *p : ax)
      if (p) (void)hipFree(p);
  if (d_wavelet) (void)hipFree(d_wavelet);
  if (d_rec) (void)hipFree(d_rec);
  if (d_traces) (void)hipFree(d_traces);
  if (d_traces64) (void)hipFree(d_traces64);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

size_t biot_handle::fields_bytes() const { return (size_t)NFIELD * n * esz; }

void *biot_handle::field_ptr(int f) const { return (char *)d_fields + (size_t)f * n * esz; }

size_t biot_handle::lo_bytes() const
{
  if (!comp_fields(prec)) return 0;
  return prec == BIOT_FP16M ? (size_t)PK_RECORD * n : (size_t)NFIELD * n;
}

int8_t *biot_handle::lo_ptr(int f) const
{
  if (!d_lo || !((comp_fields(prec) >> f) & 1u)) return nullptr;
  if (packed(f)) return d_lo + pk_slot(f);  // the record base (256-byte aligned) tagged with the nibble
  return d_lo + (size_t)f * n;
}

void biot_handle::fill_args(StepArgs &a) const
{
  std::memset(&a, 0, sizeof a);
  for (int f = 0; f < NFIELD; ++f) {
    a.f[f] = field_ptr(f);
    a.lo[f] = lo_ptr(f);
  }
  for (int k = 0; k < NKC; ++k) {
    a.k[k] = (const char *)d_coef + (size_t)k * n * esz;
    a.ks[k] = ks[k];
  }
  a.kv = d_kpack;
  a.kt = d_kpack ? (const char *)d_kpack + (size_t)14 * n * esz : nullptr;
  if (prec == BIOT_FP16M) {
    for (int k = 0; k < NKC; ++k) a.k[k] = nullptr;
    a.mv = d_coef;
    a.mt = (const char *)d_coef + mt_offset();
    a.mu = (const char *)d_coef + mu_offset();
    for (int q = 0; q < NMC; ++q) a.mc[q] = mc[q];
  }
  for (int ax = 0; ax < 3; ++ax) {
    a.carry[ax] = d_carry[ax];
    a.slab[ax] = slab[ax];
    a.L[ax] = L[ax];
    a.R[ax] = R[ax];
    for (int q = 0; q < 4; ++q) a.prof[ax][q] = d_prof[ax][q];
  }
  a.nx = g.nx;
  a.ny = g.ny;
  a.nz = g.nz;
  a.sy = g.nx;
  a.sz = (int64_t)g.nx * g.ny;
}

LaunchShape default_shape(int prec, int order)
{
  (void)prec;
  (void)order;
  return LaunchShape{64, 4, 16};
}

int biot_handle::allocate()
{
  auto alloc = [&](void **p, size_t bytes) -> bool {
    if (bytes == 0) return true;
    if (hipMalloc(p, bytes) != hipSuccess) {
      *p = nullptr;
      return false;
    }
    device_bytes += (int64_t)bytes;
    return hipMemsetAsync(*p, 0, bytes, stream) == hipSuccess;
  };
  // FP16M keeps its medium (8 + 8 bytes of records and 2 of mu per cell) in place of the
  // coefficient arrays
  const size_t coef_bytes = prec == BIOT_FP16M ? mu_offset() + (size_t)2 * n : (size_t)NKC * n * esz;
  const size_t pack_bytes = prec == BIOT_FP16M ? 0 : (size_t)NKC * n * esz;
  bool ok = alloc(&d_fields, fields_bytes()) && alloc(&d_coef, coef_bytes) &&
            alloc(&d_kpack, pack_bytes) && alloc((void **)&d_flag, 256) && clear_nonfinite() == BIOT_OK;
  if (ok) ok = alloc((void **)&d_lo, lo_bytes());
  for (int ax = 0; ax < 3 && ok; ++ax) ok = alloc(&d_carry[ax], (size_t)NCARRY * slab[ax] * csz);
  const int dims[3] = {g.nx, g.ny, g.nz};
  const size_t asz = prec == BIOT_FP64 ? 8 : 4;
  for (int ax = 0; ax < 3 && ok; ++ax)
    for (int q = 0; q < 4 && ok; ++q) ok = alloc(&d_prof[ax][q], (size_t)dims[ax] * asz);
  if (!ok) {
    err = "device allocation failed";
    return BIOT_ERR_NOMEM;
  }
  if (hipStreamSynchronize(stream) != hipSuccess) {
    err = "hipStreamSynchronize failed after allocation";
    return BIOT_ERR_HIP;
  }
  return BIOT_OK;
}

// e and a = -dt*d/2 as the kernels use them.  The reference stores e and d multiplied by
// Cpml = 0.01f (main.cpp:1512-1523) and the kernels multiply back by 100.0 in double
// (kernel.cu:188): the effective values are reproduced here.
int biot_handle::upload_profiles()
{
  const int dims[3] = {g.nx, g.ny, g.nz};
  const float Cpml = 0.01f;
  for (int ax = 0; ax < 3; ++ax) {
    const int n_ = dims[ax];
    std::vector<double> ei(n_), ai(n_), eh(n_), ah(n_);
    for (int i = 0; i < n_; ++i) {
      const float d_s = prof[ax][0][i] * Cpml, d2_s = prof[ax][1][i] * Cpml;
      const float e_s = prof[ax][2][i] * Cpml, e2_s = prof[ax][3][i] * Cpml;
      ei[i] = (double)e_s * 100.0;
      eh[i] = (double)e2_s * 100.0;
      if (prec == BIOT_FP64) {  // no float rounding of -DT*d on the double path
        ai[i] = -(double)g.dt * (double)d_s * 100.0 * 0.5;
        ah[i] = -(double)g.dt * (double)d2_s * 100.0 * 0.5;
      } else {                  // float product first, as the reference's expression
        ai[i] = (double)(float)(-g.dt * d_s) * 100.0 * 0.5;
        ah[i] = (double)(float)(-g.dt * d2_s) * 100.0 * 0.5;
      }
    }
    const std::vector<double> *src[4] = {&ei, &ai, &eh, &ah};
    for (int q = 0; q < 4; ++q) {
      if (prec == BIOT_FP64) {
        if (hipMemcpy(d_prof[ax][q], src[q]->data(), n_ * 8, hipMemcpyHostToDevice) != hipSuccess)
          return BIOT_ERR_HIP;
      } else {
        std::vector<float> f(n_);
        for (int i = 0; i < n_; ++i) f[i] = (float)(*src[q])[i];
        if (hipMemcpy(d_prof[ax][q], f.data(), n_ * 4, hipMemcpyHostToDevice) != hipSuccess)
          return BIOT_ERR_HIP;
      }
    }
  }
  return BIOT_OK;
}

namespace {
struct KernCtx {
  const std::vector<float> *ref;
  double Hinv;
  std::vector<double> *K;  // NKC arrays
};
// reference arrays -> kernel coefficients: 1/dx folded into every coefficient that
// multiplies a difference, the kernels' 0.01 (d_Cpp*, d_Csp*, kernel.cu:26-31) folded into
// the pressure/stress ones.
void kern_body(int64_t lo, int64_t hi, void *p)
{
  auto *cx = static_cast<KernCtx *>(p);
  const std::vector<float> *R = cx->ref;
  std::vector<double> *K = cx->K;
  const double H = cx->Hinv;
  for (int64_t c = lo; c < hi; ++c) {
    for (int d = 0; d < 3; ++d) {
      K[K_U1X + d][c] = (double)R[BIOT_U1X + d][c] * H;
      K[K_U2X + d][c] = (double)R[BIOT_U2X + d][c];
      K[K_W1X + d][c] = (double)R[BIOT_W1X + d][c];
      K[K_W2X + d][c] = (double)R[BIOT_W2X + d][c] * H;
    }
    K[K_W3X][c] = (double)R[BIOT_W3X][c] * H;
    K[K_W3Y][c] = (double)R[BIOT_W3Y][c] * H;
    K[K_P1][c] = 0.01 * R[BIOT_PP1][c] * H;
    K[K_P2][c] = 0.01 * R[BIOT_PP2][c] * H;
    K[K_S1][c] = 0.01 * R[BIOT_SP1][c] * H;
    K[K_S2][c] = 0.01 * R[BIOT_SP2][c] * H;
    K[K_S3][c] = 0.01 * R[BIOT_SP3][c] * H;
    K[K_SXY][c] = 0.01 * R[BIOT_SPXY][c] * H;
    K[K_SXZ][c] = 0.01 * R[BIOT_SPXZ][c] * H;
    K[K_SYZ][c] = 0.01 * R[BIOT_SPYZ][c] * H;
  }
}
double amax(const std::vector<double> &v)
{
  double m = 0;
  for (double x : v) m = std::max(m, std::fabs(x));
  return m;
}
}  // namespace

// host form of the reference -> kernel coefficient conversion (before storage scaling)
void host_kernel_coefs(double Hinv, const std::vector<float> (&ref)[BIOT_NCOEF], std::vector<double> (&K)[NKC])
{
  for (auto &v : K) v.assign(ref[0].size(), 0.0);
  KernCtx cx{ref, Hinv, K};
  parallel_for((int64_t)ref[0].size(), kern_body, &cx);
}
namespace {
}  // namespace

// Builds the 22 kernel coefficient arrays from the reference arrays and uploads them: the
// reference arrays go to the device once and the conversion (1/dx, 0.01, storage scales,
// storage type) runs there (upload_from_ref).
int biot_handle::upload_coefficients()
{
  if (coef_ref[0].empty()) {
    if (!K_last[0].empty()) return upload_kernel_coefs(K_last);  // parameter-built set
    err = "no coefficients to upload";
    return BIOT_ERR_STATE;
  }
  float *d_ref = nullptr;
  if (hipMalloc((void **)&d_ref, (size_t)BIOT_NCOEF * n * sizeof(float)) != hipSuccess) {
    err = "reference coefficient buffer does not fit";
    return BIOT_ERR_NOMEM;
  }
  int rc = BIOT_OK;
  for (int q = 0; q < BIOT_NCOEF && !rc; ++q)
    if (hipMemcpy(d_ref + (size_t)q * n, coef_ref[q].data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice) !=
        hipSuccess)
      rc = BIOT_ERR_HIP;
  if (rc) err = "reference coefficient upload failed";
  else rc = upload_from_ref(d_ref);
  (void)hipFree(d_ref);
  return rc;
}

// 16-bit storage scales.  Each field group is stored divided by a scale (stresses/pressure
// by the peak source amplitude A; velocities by A*sqrt(max U1 / max S2), which makes the
// velocity and stress update coefficients both equal to the CFL number c*dt/dx; the
// relative fluid velocity by its forced response A * max((|W2|+|W3|)/(1-W1)) * 0.2, the
// 0.2 being k*dx of a wave sampled at ~30 points per wavelength); every coefficient array
// is multiplied by the ratio of the scales it connects (kfac) and stored divided by its
// largest magnitude (ks) so it lies in [-1, 1].  amaxK: per-array max |K| before scaling.
void biot_handle::set_scales(const double amaxK[NKC], double resp)
{
  for (int f = 0; f < NFIELD; ++f) fscale[f] = 1.0f;
  for (int k = 0; k < NKC; ++k) ks[k] = 1.0f, kfac[k] = 1.0;
  if (!half()) return;
  double A = 0;
  for (float w : wavelet) A = std::max(A, (double)std::fabs(w));
  if (!(A > 0)) A = 1.0;
  // 16x headroom: the stored source amplitude is 16, so the far field stays clear of the
  // binary16 subnormal range; peaks at the source cell stay below ~1/16 of the largest
  // binary16 (measured < 256 A on the layered and perturbed media)
  A = std::ldexp(A, -4);
  const double u1 = std::max({amaxK[K_U1X], amaxK[K_U1Y], amaxK[K_U1Z]});
  const double s2 = amaxK[K_S2];
  const double sig_t = A, sig_p = A;
  const double sig_v = (u1 > 0 && s2 > 0) ? A * std::sqrt(u1 / s2) : A;
  const double sig_q = resp > 0 ? A * resp * 0.2 : sig_v;
  const double sv = sig_v, sq = sig_q, stau = sig_t, sp = sig_p;
  for (int d = 0; d < 3; ++d) {
    kfac[K_U1X + d] = stau / sv;
    kfac[K_U2X + d] = sq / sv;
    kfac[K_W2X + d] = stau / sq;
  }
  kfac[K_W3X] = kfac[K_W3Y] = sp / sq;
  kfac[K_P1] = sv / sp;
  kfac[K_P2] = sq / sp;
  kfac[K_S1] = kfac[K_S2] = kfac[K_SXY] = kfac[K_SXZ] = kfac[K_SYZ] = sv / stau;
  kfac[K_S3] = sq / stau;
  for (int f = 0; f < 3; ++f) fscale[f] = (float)sv;
  for (int f = 3; f < 6; ++f) fscale[f] = (float)sq;
  for (int f = 6; f < 12; ++f) fscale[f] = (float)stau;
  fscale[12] = (float)sp;
  for (int k = 0; k < NKC; ++k) {
    const double m = amaxK[k] * kfac[k];  // the largest scaled magnitude (rounding is monotonic)
    if (m > 0) ks[k] = (float)m;
  }
}

// Stored wavefields and C-PML carries are expressed in the storage scales: when a scale
// changes (e.g. a new source amplitude), re-express what the handle already holds.
// Carries 0..3 of each axis are differences of stresses/pressure, 4..7 of solid
// velocities except carry 5 (relative fluid velocity).
int biot_handle::rescale_state(const float old_scale[NFIELD])
{
  auto rescale = [&](int p_, void *p, int8_t *lo, int64_t cnt, double f) -> bool {
    return f == 1.0 || launch_rescale(p_, p, lo, cnt, f, stream) == hipSuccess;
  };
  bool ok = true;
  for (int f = 0; f < NFIELD && ok; ++f) {
    const double r = (double)old_scale[f] / (double)fscale[f];
    if (packed(f)) ok = r == 1.0 || launch_rescale_pk(field_ptr(f), lo_ptr(f), n, r, stream) == hipSuccess;
    else ok = rescale(prec, field_ptr(f), lo_ptr(f), n, r);
  }
  const double rt = (double)old_scale[6] / fscale[6], rv = (double)old_scale[0] / fscale[0];
  const double rq = (double)old_scale[3] / fscale[3];
  for (int ax = 0; ax < 3 && ok; ++ax) {
    if (slab[ax] == 0) continue;
    // records of carries 0..3 (all in stress units), then 4..7 (velocity units, 5 in q units)
    const int cp = prec == BIOT_FP64 ? BIOT_FP64 : prec == BIOT_FP16 ? BIOT_FP16 : BIOT_FP32;
    ok = rescale(cp, d_carry[ax], nullptr, 4 * slab[ax], rt);
    const double f4[4] = {rv, rq, rv, rv};
    if (ok && (rv != 1.0 || rq != 1.0))
      ok = launch_rescale4(cp, (char *)d_carry[ax] + (size_t)4 * slab[ax] * csz, slab[ax], f4, stream) == hipSuccess;
  }
  if (!ok || hipStreamSynchronize(stream) != hipSuccess) {
    err = "rescaling the stored wavefields failed";
    return BIOT_ERR_HIP;
  }
  return BIOT_OK;
}

// Device-side conversion of the 23 reference arrays (d_ref, contiguous) into the kernel
// arrays: statistics for the scales, then K = R (x 1/dx, x 0.01) x ratio / ks per cell.
int biot_handle::upload_from_ref(const float *d_ref)
{
  void *stats = nullptr;
  if (hipMalloc(&stats, 256) != hipSuccess) {
    err = "statistics buffer";
    return BIOT_ERR_NOMEM;
  }
  RefArgs ra;
  std::memset(&ra, 0, sizeof ra);
  for (int q = 0; q < BIOT_NCOEF; ++q) ra.ref[q] = d_ref + (size_t)q * n;
  ra.n = n;
  ra.H = (double)(1.0f / g.dx);  // 1/dx as the float the reference's kernels hold (kernel.cu:158)
  ra.stats_u32 = (uint32_t *)stats;
  ra.stats_u64 = (unsigned long long *)((char *)stats + 128);
  uint32_t st32[BIOT_NCOEF];
  unsigned long long st64 = 0;
  int rc = BIOT_OK;
  if (hipMemsetAsync(stats, 0, 256, stream) != hipSuccess || launch_ref_stats(ra, stream) != hipSuccess ||
      hipMemcpyAsync(st32, stats, sizeof st32, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipMemcpyAsync(&st64, (char *)stats + 128, sizeof st64, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    rc = BIOT_ERR_HIP;
  (void)hipFree(stats);
  if (rc) {
    err = "coefficient statistics failed";
    return rc;
  }
  // per-array max |K| from max |R| (every conversion is a monotonic rounding)
  auto amaxR = [&](int q) { return (double)__builtin_bit_cast(float, st32[q]); };
  double amaxK[NKC];
  const double H = ra.H;
  for (int d = 0; d < 3; ++d) {
    amaxK[K_U1X + d] = amaxR(BIOT_U1X + d) * H;
    amaxK[K_U2X + d] = amaxR(BIOT_U2X + d);
    amaxK[K_W1X + d] = amaxR(BIOT_W1X + d);
    amaxK[K_W2X + d] = amaxR(BIOT_W2X + d) * H;
  }
  amaxK[K_W3X] = amaxR(BIOT_W3X) * H;
  amaxK[K_W3Y] = amaxR(BIOT_W3Y) * H;
  const int sref[8] = {BIOT_PP1, BIOT_PP2, BIOT_SP1, BIOT_SP2, BIOT_SP3, BIOT_SPXY, BIOT_SPXZ, BIOT_SPYZ};
  for (int q = 0; q < 8; ++q) amaxK[K_P1 + q] = 0.01 * amaxR(sref[q]) * H;
  const double resp = __builtin_bit_cast(double, st64);
  float old_scale[NFIELD];
  std::memcpy(old_scale, fscale, sizeof old_scale);
  set_scales(amaxK, resp);
  if ((rc = rescale_state(old_scale))) return rc;
  ra.kout = d_coef;
  ra.scaled = half() ? 1 : 0;
  for (int k = 0; k < NKC; ++k) {
    ra.ratio[k] = kfac[k];
    ra.inv_ks[k] = 1.0 / (double)ks[k];
  }
  if (prec == BIOT_FP16M) {
    if ((rc = upload_medium())) return rc;
    have_coef = true;
    return BIOT_OK;
  }
  if (launch_ref_to_kernel(prec, ra, stream) != hipSuccess || hipStreamSynchronize(stream) != hipSuccess) {
    err = "coefficient conversion failed";
    return BIOT_ERR_HIP;
  }
  if ((rc = pack_coefficients())) return rc;
  have_coef = true;
  return BIOT_OK;
}

namespace {
struct AdjCoefCtx {
  const biot_handle *h;
  _Float16 *out;  // NKC arrays of n
};
void adj_coef_body(int64_t lo, int64_t hi, void *p)
{
  auto *cx = static_cast<AdjCoefCtx *>(p);
  const biot_handle *h = cx->h;
  for (int k = 0; k < NKC; ++k) {
    const double f = h->kfac[k] / (double)h->ks[k];
    const double *K = h->K_last[k].data();
    _Float16 *o = cx->out + (size_t)k * h->n;
    for (int64_t c = lo; c < hi; ++c) o[c] = (_Float16)(float)(K[c] * f);
  }
}
}  // namespace

// FP16M keeps moduli, not coefficient arrays: the adjoint reads the coefficients of the last
// parameter set as FP16T stores them (K kfac / ks rounded to binary16, packed per cell).  The
// forward forms them from the binary16 moduli in float, so the two differ by the rounding of
// either.  Built when a checkpointed forward run starts, freed with the adjoint buffers.
int biot_handle::build_adjoint_coefs()
{
  if (prec != BIOT_FP16M || d_kpack) return BIOT_OK;
  if (K_last[0].empty()) {
    err = "FP16M gradients need biot_set_params (its coefficients come from the parameters)";
    return BIOT_ERR_STATE;
  }
  const size_t bytes = (size_t)NKC * n * sizeof(_Float16);
  void *soa = nullptr;
  if (hipMalloc(&soa, bytes) != hipSuccess) {
    err = "the adjoint's coefficient arrays do not fit in device memory";
    return BIOT_ERR_NOMEM;
  }
  if (hipMalloc(&d_kpack, bytes) != hipSuccess) {
    d_kpack = nullptr;
    (void)hipFree(soa);
    err = "the adjoint's coefficient arrays do not fit in device memory";
    return BIOT_ERR_NOMEM;
  }
  // the 22 arrays converted on the host threads, then one upload
  std::vector<_Float16> buf((size_t)NKC * n);
  AdjCoefCtx cx{this, buf.data()};
  parallel_for(n, adj_coef_body, &cx);
  int rc = hipMemcpy(soa, buf.data(), bytes, hipMemcpyHostToDevice) == hipSuccess ? BIOT_OK : BIOT_ERR_HIP;
  if (!rc && (launch_pack_coef(PREC_FP16T, soa, d_kpack, (char *)d_kpack + (size_t)14 * n * sizeof(_Float16),
                               n, stream) != hipSuccess ||
              hipStreamSynchronize(stream) != hipSuccess))
    rc = BIOT_ERR_HIP;
  (void)hipFree(soa);
  if (rc) {
    (void)hipFree(d_kpack);
    d_kpack = nullptr;
    err = "building the adjoint's coefficient copy failed";
  }
  return rc;
}

int biot_handle::pack_coefficients()
{
  const void *kv = d_kpack;
  void *kt = (char *)d_kpack + (size_t)14 * n * esz;
  if (launch_pack_coef(prec, d_coef, (void *)kv, kt, n, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess) {
    err = "coefficient packing failed";
    return BIOT_ERR_HIP;
  }
  return BIOT_OK;
}

// Host-side form for coefficient sets built in double from inversion parameters.
int biot_handle::upload_kernel_coefs(const std::vector<double> (&K_in)[NKC])
{
  double amaxK[NKC];
  for (int k = 0; k < NKC; ++k) amaxK[k] = amax(K_in[k]);
  double resp = 0;
  for (int d = 0; d < 3 && half(); ++d) {
    const auto &w1 = K_in[K_W1X + d], &w2 = K_in[K_W2X + d], &w3 = K_in[d == 2 ? K_W3Y : K_W3X + d];
    for (int64_t c = 0; c < n; ++c) {
      const double den = 1.0 - w1[c];
      if (den > 0) resp = std::max(resp, (std::fabs(w2[c]) + std::fabs(w3[c])) / den);
    }
  }
  float old_scale[NFIELD];
  std::memcpy(old_scale, fscale, sizeof old_scale);
  set_scales(amaxK, resp);
  int rc = rescale_state(old_scale);
  if (rc) return rc;
  if (prec == BIOT_FP16M) {
    if ((rc = upload_medium())) return rc;
    have_coef = true;
    return BIOT_OK;
  }
  std::vector<double> K[NKC];
  for (int k = 0; k < NKC; ++k) {
    K[k] = K_in[k];
    if (half()) {
      const double inv = 1.0 / (double)ks[k];
      for (auto &x : K[k]) {
        x *= kfac[k];
        x *= inv;  // the kernels multiply ks back
      }
    }
  }

  std::vector<unsigned char> buf((size_t)n * esz);
  for (int k = 0; k < NKC; ++k) {
    const std::vector<double> &src = K[k];
    if (prec == BIOT_FP32) {
      float *d = (float *)buf.data();
      for (int64_t c = 0; c < n; ++c) d[c] = (float)src[c];
    } else if (half()) {
      _Float16 *d = (_Float16 *)buf.data();
      for (int64_t c = 0; c < n; ++c) d[c] = (_Float16)(float)src[c];
    } else {
      std::memcpy(buf.data(), src.data(), (size_t)n * 8);
    }
    if (hipMemcpy((char *)d_coef + (size_t)k * n * esz, buf.data(), (size_t)n * esz,
                  hipMemcpyHostToDevice) != hipSuccess) {
      err = "coefficient upload failed";
      return BIOT_ERR_HIP;
    }
  }
  if ((rc = pack_coefficients())) return rc;
  have_coef = true;
  return BIOT_OK;
}

// FP16M: the medium records from the 9 cell moduli.  Velocity record per cell: rho, rho_f, C1,
// C2; stress record: C, M, lambda_c, H; and the cell's mu, from which the stress kernel forms
// the three 4-cell harmonic means of main.cpp:326-337 (zero when a mu of the quad is zero).
// Each stored quantity is scaled by its largest magnitude and rounded to binary16; the
// kernels' constants fold the scales back together with the reference's coefficient factors
// (Cv, Cp, DT, 1/dx as the kernels hold it, the 0.01 x 100 of the stress update) and the
// field-storage ratios kfac.
namespace {
enum { MD_M = 0, MD_HH, MD_H2U, MD_C, MD_C1, MD_C2, MD_RHO, MD_RHOF, MD_MU };
struct MedCtx {
  const std::vector<float> *mod;
  double sc[9];        // scales: rho rho_f C1 C2 | C M H2u HH | mu
  uint16_t *mv, *mt, *mu;
  double mx[64][9];    // per-thread maxima (first pass)
  std::atomic<int> slot{0};
};
inline uint16_t h16(double x)
{
  const _Float16 h = (_Float16)(float)x;
  return __builtin_bit_cast(uint16_t, h);
}
const int MED_SRC[8] = {MD_RHO, MD_RHOF, MD_C1, MD_C2, MD_C, MD_M, MD_H2U, MD_HH};
void med_max_body(int64_t lo, int64_t hi, void *p)
{
  auto *cx = static_cast<MedCtx *>(p);
  double m[9] = {};
  for (int64_t c = lo; c < hi; ++c) {
    for (int q = 0; q < 8; ++q) m[q] = std::max(m[q], (double)std::fabs(cx->mod[MED_SRC[q]][c]));
    m[8] = std::max(m[8], (double)std::fabs(cx->mod[MD_MU][c]));
  }
  const int t = cx->slot.fetch_add(1);
  for (int q = 0; q < 9; ++q) cx->mx[t][q] = m[q];
}
void med_pack_body(int64_t lo, int64_t hi, void *p)
{
  auto *cx = static_cast<MedCtx *>(p);
  const std::vector<float> *M = cx->mod;
  for (int64_t c = lo; c < hi; ++c) {
    for (int q = 0; q < 4; ++q) {
      cx->mv[4 * c + q] = h16(M[MED_SRC[q]][c] / cx->sc[q]);
      cx->mt[4 * c + q] = h16(M[MED_SRC[4 + q]][c] / cx->sc[4 + q]);
    }
    cx->mu[c] = h16(M[MD_MU][c] / cx->sc[8]);
  }
}
}  // namespace

int biot_handle::upload_medium()
{
  for (const auto &v : mod)
    if ((int64_t)v.size() != n) {
      err = "FP16M needs the medium moduli (biot_set_medium, biot_set_medium_device or biot_set_params)";
      return BIOT_ERR_STATE;
    }
  auto cx = std::make_unique<MedCtx>();
  cx->mod = mod;
  std::memset(cx->mx, 0, sizeof cx->mx);
  parallel_for(n, med_max_body, cx.get());
  double *sc = cx->sc;
  for (int q = 0; q < 9; ++q) {
    sc[q] = 0.0;
    for (int t = 0; t < 64; ++t) sc[q] = std::max(sc[q], cx->mx[t][q]);
    if (!(sc[q] > 0) || !std::isfinite(sc[q])) sc[q] = 1.0;
  }
  std::vector<uint16_t> mv((size_t)4 * n), mt((size_t)4 * n), mu((size_t)n);
  cx->mv = mv.data();
  cx->mt = mt.data();
  cx->mu = mu.data();
  parallel_for(n, med_pack_body, cx.get());
  if (hipMemcpy(d_coef, mv.data(), mv.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy((char *)d_coef + mt_offset(), mt.data(), mt.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy((char *)d_coef + mu_offset(), mu.data(), mu.size() * 2, hipMemcpyHostToDevice) != hipSuccess) {
    err = "medium upload failed";
    return BIOT_ERR_HIP;
  }
  const double Hinv = (double)(1.0f / g.dx), DT = (double)g.dt, Cv = (double)g.Cv, Cp = (double)g.Cp;
  const double vel = Cv * DT / Cp * Hinv, str = Cp * DT / Cv * Hinv;
  mc[MC_HRHO] = (float)(0.5 * sc[0]);
  mc[MC_HRHOF] = (float)(0.5 * sc[1]);
  mc[MC_HC1] = (float)(0.5 * sc[2]);
  mc[MC_HC2] = (float)(0.5 * sc[3]);
  mc[MC_KU2] = (float)kfac[K_U2X];
  mc[MC_CU1] = (float)(vel * kfac[K_U1X]);
  mc[MC_CW2] = (float)(vel * kfac[K_W2X]);
  mc[MC_CW3] = (float)(vel * kfac[K_W3X]);
  mc[MC_AH] = (float)(0.5 * DT);
  mc[MC_KP1] = (float)(sc[4] * str * kfac[K_P1]);
  mc[MC_KP2] = (float)(sc[5] * str * kfac[K_P2]);
  mc[MC_KS1] = (float)(sc[6] * str * kfac[K_S1]);
  mc[MC_KS2] = (float)(sc[7] * str * kfac[K_S2]);
  mc[MC_KS3] = (float)(sc[4] * str * kfac[K_S3]);
  mc[MC_KSH] = (float)(4.0 * sc[8] * str * kfac[K_SXY]);  // the kernel forms 1 / sum(1/mu) of a quad
  return BIOT_OK;
}
