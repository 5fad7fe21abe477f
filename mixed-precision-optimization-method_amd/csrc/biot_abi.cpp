// C-ABI of the Biot solver (include/biot.h): handle lifetime, device memory, set-up
// uploads and the time loop.  Replaces the reference's callCUDAKernel (kernel.cu:1728)
// and the caller-side allocation/copy of blockInfo (main.cpp:1376-1743).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/biot.h"
#include "biot_device.hpp"
#include "biot_handle.hpp"
#include "biot_setup.hpp"

using namespace biot;

namespace {

int fail(biot_handle *h, int code, const char *fmt, ...)
{
  if (h) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    h->err = buf;
  }
  return code;
}

#define HIPCHK(h, call)                                                                       \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail((h), BIOT_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                  __FILE__, __LINE__);                                                        \
  } while (0)

// binary16 residual unit ulp(h)/256 = 2^(e-33), e the biased exponent (>= 1) of h; FP16M's
// 4-bit residuals count ulp(h)/16 = 2^(e-29)
double half_res_unit(_Float16 h, int bits = 8)
{
  uint16_t u;
  std::memcpy(&u, &h, 2);
  const int e = std::max((u >> 10) & 31, 1);
  return std::ldexp(1.0, e - 25 - bits);
}

// host storage-type conversion; lo: the residuals of a compensated field (stride 1), or for an
// FP16M field (packed) the cells' residual records (4 bytes each) and the field's nibble slot,
// else null
void to_storage(int prec, const double *src, int64_t n, double scale_inv, void *dst, int8_t *lo,
                bool packed = false, int slot = 0)
{
  if (packed) {  // binary16 + 4-bit residual, patched into nibble `slot` of each cell's record
    _Float16 *d = (_Float16 *)dst;
    uint32_t *rec = (uint32_t *)lo;
    for (int64_t c = 0; c < n; ++c) {
      const float x = (float)(src[c] * scale_inv);
      d[c] = (_Float16)x;
      const double r = std::nearbyint(((double)x - (double)(float)d[c]) / half_res_unit(d[c], 4));
      const uint32_t nib = (uint32_t)(int)std::min(7.0, std::max(-7.0, r)) & 0xFu;
      rec[c] = (rec[c] & ~(0xFu << (4 * slot))) | (nib << (4 * slot));
    }
    return;
  }
  if (prec == BIOT_FP32) {
    float *d = (float *)dst;
    for (int64_t c = 0; c < n; ++c) d[c] = (float)(src[c] * scale_inv);
  } else if (prec == BIOT_FP16 || prec == BIOT_FP16C || prec == BIOT_FP16T || prec == BIOT_FP16M) {
    _Float16 *d = (_Float16 *)dst;
    for (int64_t c = 0; c < n; ++c) {
      const float x = (float)(src[c] * scale_inv);
      d[c] = (_Float16)x;
      if (lo) {
        const double r = std::nearbyint(((double)x - (double)(float)d[c]) / half_res_unit(d[c]));
        lo[c] = (int8_t)std::min(127.0, std::max(-127.0, r));
      }
    }
  } else {
    double *d = (double *)dst;
    for (int64_t c = 0; c < n; ++c) d[c] = src[c] * scale_inv;
  }
}
void from_storage(int prec, const void *src, const int8_t *lo, int64_t n, double scale, double *dst,
                  bool packed = false, int slot = 0)
{
  if (packed) {
    const _Float16 *s = (const _Float16 *)src;
    const uint32_t *rec = (const uint32_t *)lo;
    for (int64_t c = 0; c < n; ++c) {
      const int r = (int)(rec[c] << (28 - 4 * slot)) >> 28;
      dst[c] = ((double)(float)s[c] + (double)r * half_res_unit(s[c], 4)) * scale;
    }
    return;
  }
  if (prec == BIOT_FP32) {
    const float *s = (const float *)src;
    for (int64_t c = 0; c < n; ++c) dst[c] = (double)s[c] * scale;
  } else if (prec == BIOT_FP16 || prec == BIOT_FP16C || prec == BIOT_FP16T || prec == BIOT_FP16M) {
    const _Float16 *s = (const _Float16 *)src;
    for (int64_t c = 0; c < n; ++c) {
      double v = (double)(float)s[c];
      if (lo) v += (double)lo[c] * half_res_unit(s[c]);
      dst[c] = v * scale;
    }
  } else {
    const double *s = (const double *)src;
    for (int64_t c = 0; c < n; ++c) dst[c] = s[c] * scale;
  }
}

// bytes spanned by the residuals of n cells of field f from lo_copy_ptr(f) (whole records for
// FP16M)
size_t lo_span(const biot_handle *h, int f, int64_t n) { return (size_t)n * h->lo_stride(f); }

}  // namespace

// ------------------------------------------------------------------------------------
// lifetime
// ------------------------------------------------------------------------------------
extern "C" int biot_abi_version(void) { return BIOT_ABI_VERSION; }

extern "C" const char *biot_status_string(int s)
{
  switch (s) {
    case BIOT_OK: return "ok";
    case BIOT_ERR_ARG: return "invalid argument";
    case BIOT_ERR_HIP: return "HIP runtime error";
    case BIOT_ERR_STATE: return "call out of order";
    case BIOT_ERR_NONFINITE: return "non-finite value detected";
    case BIOT_ERR_RANGE: return "value outside the storage range";
    case BIOT_ERR_NOMEM: return "out of memory";
    default: return "unknown status";
  }
}

// message of the calling thread's last failed biot_create (biot_last_error(NULL))
static thread_local std::string t_create_err;

extern "C" const char *biot_last_error(const biot_handle *h)
{
  return h ? h->err.c_str() : t_create_err.c_str();
}

static int validate_grid(const biot_grid *g, std::string &why)
{
  char buf[256];
  if (g->order != 2 && g->order != 4) {
    snprintf(buf, sizeof buf, "order must be 2 or 4, got %d", g->order);
    why = buf;
    return BIOT_ERR_ARG;
  }
  const int h = g->order / 2;
  if (g->nx < 2 * h + 1 || g->ny < 2 * h + 1 || g->nz < 2 * h + 1) {
    snprintf(buf, sizeof buf, "grid %dx%dx%d too small for order %d", g->nx, g->ny, g->nz,
             g->order);
    why = buf;
    return BIOT_ERR_ARG;
  }
  if ((int64_t)g->nx * g->ny * g->nz > (int64_t)1 << 34) {
    why = "grid too large";
    return BIOT_ERR_ARG;
  }
  if (g->precision < 0 || g->precision > 5) {
    why = "precision must be BIOT_FP32, BIOT_FP16, BIOT_FP64, BIOT_FP16C, BIOT_FP16T or BIOT_FP16M";
    return BIOT_ERR_ARG;
  }
  if (!(g->dx > 0) || !(g->dt > 0) || !(g->Cv > 0) || !(g->Cp > 0)) {
    why = "dx, dt, Cv and Cp must be positive";
    return BIOT_ERR_ARG;
  }
  if (g->npml < 0 || 2 * g->npml + 2 > std::min({g->nx, g->ny, g->nz})) {
    why = "npml out of range for the grid";
    return BIOT_ERR_ARG;
  }
  if (g->npml > 0 && (!(g->pml_v0 > 0) || !(g->pml_R > 0) || !(g->pml_R < 1))) {
    why = "pml_v0 must be > 0 and 0 < pml_R < 1";
    return BIOT_ERR_ARG;
  }
  return BIOT_OK;
}

extern "C" int biot_create(const biot_grid *g, biot_handle **out)
{
  t_create_err.clear();
  auto refuse = [](biot_handle *h, int rc, const char *why) {
    delete h;
    t_create_err = why;
    return rc;
  };
  if (!g || !out) return refuse(nullptr, BIOT_ERR_ARG, "grid and output pointers are required");
  *out = nullptr;
  std::string why;
  int rc = validate_grid(g, why);
  if (rc) return refuse(nullptr, rc, why.c_str());
  biot_handle *h = new biot_handle();
  h->g = *g;
  h->prec = g->precision;
  h->esz = field_bytes(g->precision);
  h->csz = carry_bytes(g->precision);
  h->n = (int64_t)g->nx * g->ny * g->nz;
  for (int k = 0; k < NKC; ++k) h->ks[k] = 1.0f;
  for (int f = 0; f < NFIELD; ++f) h->fscale[f] = 1.0f;
  // PML profiles (main.cpp:537-566) and slab widths
  const int dims[3] = {g->nx, g->ny, g->nz};
  for (int ax = 0; ax < 3; ++ax) {
    const int n = dims[ax];
    auto &P = h->prof[ax];
    for (auto &v : P) v.assign(n, 0.0f);
    if (g->npml > 0)
      build_pml_profile(n, g->dx, g->dt, g->npml, g->pml_v0, g->pml_R, P[0].data(), P[1].data(),
                        P[2].data(), P[3].data());
    int L = 0, R = 0;
    for (int i = 0; i < n; ++i) {
      const bool nz_ = P[0][i] != 0.0f || P[1][i] != 0.0f;
      if (!nz_) continue;
      if (i < n / 2) L = std::max(L, i + 1);
      else R = std::max(R, n - i);
    }
    h->L[ax] = L;
    h->R[ax] = R;
  }
  h->slab[0] = (int64_t)(h->L[0] + h->R[0]) * g->ny * g->nz;
  h->slab[1] = (int64_t)g->nx * (h->L[1] + h->R[1]) * g->nz;
  h->slab[2] = (int64_t)g->nx * g->ny * (h->L[2] + h->R[2]);
  // the kernels address every array with 32-bit byte offsets from its base
  const int64_t lim = (int64_t)1 << 32;
  if (h->n * (int64_t)h->esz >= lim)
    return refuse(h, BIOT_ERR_ARG, "grid too large: one field array would exceed 4 GiB");
  if (h->prec == BIOT_FP16M && h->n * 8 >= lim)
    return refuse(h, BIOT_ERR_ARG, "grid too large: the FP16M medium records (8 bytes a cell) would exceed 4 GiB");
  for (int ax = 0; ax < 3; ++ax)
    if (NCARRY * h->slab[ax] * (int64_t)h->csz >= lim)
      return refuse(h, BIOT_ERR_ARG, "absorbing slabs too large: one axis's C-PML memory would exceed 4 GiB");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return refuse(h, BIOT_ERR_HIP, "no HIP device");
  if (g->device < 0 || g->device >= ndev) return refuse(h, BIOT_ERR_ARG, "device index out of range");
  if (hipSetDevice(g->device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess ||
      hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
    return refuse(h, BIOT_ERR_HIP, "device, stream or event set-up failed");
  h->shape = default_shape(h->prec, g->order);
  rc = h->allocate();
  if (rc) {
    const std::string e = h->err;
    return refuse(h, rc, e.c_str());
  }
  rc = h->upload_profiles();
  if (rc) {
    const std::string e = h->err.empty() ? std::string("PML profile upload failed") : h->err;
    return refuse(h, rc, e.c_str());
  }
  *out = h;
  return BIOT_OK;
}

extern "C" int biot_destroy(biot_handle *h)
{
  if (!h) return BIOT_OK;
  delete h;
  return BIOT_OK;
}

// ------------------------------------------------------------------------------------
// medium / coefficients
// ------------------------------------------------------------------------------------
static int check_medium(biot_handle *h, const biot_medium *m)
{
  if (!m || !m->Ks || !m->Kf || !m->Kb || !m->mu || !m->phi || !m->rhos || !m->rhof || !m->eta ||
      !m->perm)
    return fail(h, BIOT_ERR_ARG, "biot_set_medium: every per-cell array must be given");
  for (int64_t c = 0; c < h->n; ++c) {
    const bool fluid = m->mu[c] == 0.0 && m->Kb[c] == 0.0;
    if (!(m->rhos[c] > 0) || !(m->rhof[c] > 0) || !(m->Kf[c] > 0))
      return fail(h, BIOT_ERR_ARG, "cell %lld: rho_s, rho_f and Kf must be > 0", (long long)c);
    if (!fluid && (!(m->phi[c] > 0) || !(m->phi[c] <= 1) || !(m->Ks[c] > 0) || !(m->perm[c] > 0) ||
                   !(m->eta[c] >= 0) || !(m->mu[c] >= 0) || !(m->Kb[c] >= 0)))
      return fail(h, BIOT_ERR_ARG, "cell %lld: invalid porous-medium parameters", (long long)c);
  }
  return BIOT_OK;
}

extern "C" int biot_set_medium(biot_handle *h, const biot_medium *m)
{
  if (!h) return BIOT_ERR_ARG;
  int rc0 = check_medium(h, m);
  if (rc0) return rc0;
  h->forget_model();
  Moduli md;
  Averages av;
  build_moduli(h->n, *m, md);
  build_averages(h->g.nx, h->g.ny, h->g.nz, md, av);
  build_coefficients(h->g.nx, h->g.ny, h->g.nz, h->g.dt, h->g.Cv, h->g.Cp, md, av, h->coef_ref);
  if (h->prec == BIOT_FP16M) {
    std::vector<float> *src[9] = {&md.M, &md.HH, &md.H2u, &md.C, &md.C1, &md.C2, &md.rho, &md.rhof, &md.mu};
    for (int q = 0; q < 9; ++q) h->mod[q].swap(*src[q]);
  }
  return h->upload_coefficients();
}

// The same set-up on the device: medium arrays uploaded once, moduli, averages and the 23
// coefficient arrays built by kernels (bit-identical to the host builder), converted to the
// kernel arrays there; the reference arrays are copied back for biot_get_coefficients.
extern "C" int biot_set_medium_device(biot_handle *h, const biot_medium *m)
{
  if (!h) return BIOT_ERR_ARG;
  int rc = check_medium(h, m);
  if (rc) return rc;
  h->forget_model();
  const int64_t n = h->n;
  double *d_med = nullptr;
  float *d_mod = nullptr, *d_ref = nullptr;
  if (hipMalloc((void **)&d_med, (size_t)9 * n * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&d_mod, (size_t)9 * n * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&d_ref, (size_t)BIOT_NCOEF * n * sizeof(float)) != hipSuccess) {
    for (void *p : {(void *)d_med, (void *)d_mod, (void *)d_ref})
      if (p) (void)hipFree(p);
    return fail(h, BIOT_ERR_NOMEM, "set-up buffers do not fit");
  }
  const double *src[9] = {m->Ks, m->Kf, m->Kb, m->mu, m->phi, m->rhos, m->rhof, m->eta, m->perm};
  for (int q = 0; q < 9 && !rc; ++q)
    if (hipMemcpy(d_med + (size_t)q * n, src[q], (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      rc = BIOT_ERR_HIP;
  MediumArgs ma;
  std::memset(&ma, 0, sizeof ma);
  const double **mp[9] = {&ma.Ks, &ma.Kf, &ma.Kb, &ma.mu, &ma.phi, &ma.rhos, &ma.rhof, &ma.eta, &ma.perm};
  for (int q = 0; q < 9; ++q) *mp[q] = d_med + (size_t)q * n;
  ma.tao = m->tao;
  ma.porousm = m->porousm;
  for (int q = 0; q < 9; ++q) ma.out[q] = d_mod + (size_t)q * n;
  CoefRefArgs ca;
  std::memset(&ca, 0, sizeof ca);
  for (int q = 0; q < 9; ++q) ca.mod[q] = d_mod + (size_t)q * n;
  for (int q = 0; q < BIOT_NCOEF; ++q) ca.coef[q] = d_ref + (size_t)q * n;
  ca.nx = h->g.nx;
  ca.ny = h->g.ny;
  ca.nz = h->g.nz;
  ca.DT = h->g.dt;
  ca.Cv = h->g.Cv;
  ca.Cp = h->g.Cp;
  if (!rc && (launch_moduli(ma, n, h->stream) != hipSuccess || launch_coef_ref(ca, h->stream) != hipSuccess ||
              hipStreamSynchronize(h->stream) != hipSuccess))
    rc = BIOT_ERR_HIP;
  for (int q = 0; q < BIOT_NCOEF && !rc; ++q) {
    h->coef_ref[q].resize((size_t)n);
    if (hipMemcpy(h->coef_ref[q].data(), d_ref + (size_t)q * n, (size_t)n * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess)
      rc = BIOT_ERR_HIP;
  }
  // FP16M builds its medium records from the moduli (M HH H2u C C1 C2 rho rhof mu)
  for (int q = 0; q < 9 && !rc && h->prec == BIOT_FP16M; ++q) {
    h->mod[q].resize((size_t)n);
    if (hipMemcpy(h->mod[q].data(), d_mod + (size_t)q * n, (size_t)n * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess)
      rc = BIOT_ERR_HIP;
  }
  if (rc) {
    for (void *p : {(void *)d_med, (void *)d_mod, (void *)d_ref}) (void)hipFree(p);
    return fail(h, rc, "device set-up failed");
  }
  (void)hipFree(d_med);
  (void)hipFree(d_mod);
  rc = h->upload_from_ref(d_ref);
  (void)hipFree(d_ref);
  return rc;
}

extern "C" int biot_kernel_coefficients(float dx, int64_t n, const float *const coef[BIOT_NCOEF],
                                        double *const K[22])
{
  if (!(dx > 0) || n <= 0 || !coef || !K) return BIOT_ERR_ARG;
  std::vector<float> ref[BIOT_NCOEF];
  for (int q = 0; q < BIOT_NCOEF; ++q) {
    if (!coef[q] && q != BIOT_W3Z) return BIOT_ERR_ARG;
    if (coef[q]) ref[q].assign(coef[q], coef[q] + n);
    else ref[q].assign((size_t)n, 0.0f);
  }
  for (int k = 0; k < NKC; ++k)
    if (!K[k]) return BIOT_ERR_ARG;
  std::vector<double> Kv[NKC];
  host_kernel_coefs((double)(1.0f / dx), ref, Kv);
  for (int k = 0; k < NKC; ++k) std::memcpy(K[k], Kv[k].data(), (size_t)n * sizeof(double));
  return BIOT_OK;
}

extern "C" int biot_read_kernel_coefficients(biot_handle *h, int32_t k, double *out)
{
  if (!h || !out) return BIOT_ERR_ARG;
  if (k < 0 || k >= NKC) return fail(h, BIOT_ERR_ARG, "kernel coefficient %d out of range", k);
  if (!h->have_coef) return fail(h, BIOT_ERR_STATE, "no coefficients set");
  if (h->prec == BIOT_FP16M)
    return fail(h, BIOT_ERR_STATE, "FP16M handles store the medium moduli, not the coefficient arrays");
  std::vector<unsigned char> buf((size_t)h->n * h->esz);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(buf.data(), (const char *)h->d_coef + (size_t)k * h->n * h->esz, buf.size(),
                      hipMemcpyDeviceToHost));
  for (int64_t c = 0; c < h->n; ++c) {
    if (h->prec == BIOT_FP64) out[c] = ((const double *)buf.data())[c];
    else if (h->prec == BIOT_FP32) out[c] = ((const float *)buf.data())[c];
    else out[c] = (double)(float)((const _Float16 *)buf.data())[c] * (double)h->ks[k];
  }
  return BIOT_OK;
}

extern "C" int biot_build_coefficients(int32_t nx, int32_t ny, int32_t nz, float dt, float Cv,
                                       float Cp, const biot_medium *m, float *const coef[BIOT_NCOEF])
{
  if (!m || !coef || nx < 3 || ny < 3 || nz < 3) return BIOT_ERR_ARG;
  for (int a = 0; a < BIOT_NCOEF; ++a)
    if (!coef[a]) return BIOT_ERR_ARG;
  const int64_t n = (int64_t)nx * ny * nz;
  Moduli md;
  Averages av;
  std::vector<float> K[BIOT_NCOEF];
  build_moduli(n, *m, md);
  build_averages(nx, ny, nz, md, av);
  build_coefficients(nx, ny, nz, dt, Cv, Cp, md, av, K);
  for (int a = 0; a < BIOT_NCOEF; ++a) std::memcpy(coef[a], K[a].data(), (size_t)n * sizeof(float));
  return BIOT_OK;
}

extern "C" int biot_pml_profile_1d(int32_t n, float dx, float dt, int32_t npml, float v0, float R,
                                   float *d, float *d2, float *e, float *e2)
{
  if (n <= 0 || npml <= 0 || !(dx > 0) || !d || !d2 || !e || !e2) return BIOT_ERR_ARG;
  build_pml_profile(n, dx, dt, npml, v0, R, d, d2, e, e2);
  return BIOT_OK;
}

extern "C" int biot_set_coefficients(biot_handle *h, const float *const coef[BIOT_NCOEF])
{
  if (!h) return BIOT_ERR_ARG;
  if (!coef) return fail(h, BIOT_ERR_ARG, "coef is null");
  for (int a = 0; a < BIOT_NCOEF; ++a)
    if (!coef[a] && a != BIOT_W3Z) return fail(h, BIOT_ERR_ARG, "coefficient array %d is null", a);
  if (h->prec == BIOT_FP16M)
    return fail(h, BIOT_ERR_STATE,
                "FP16M needs the medium moduli (biot_set_medium, biot_set_medium_device or biot_set_params)");
  h->forget_model();
  for (int a = 0; a < BIOT_NCOEF; ++a) {
    if (coef[a])
      h->coef_ref[a].assign(coef[a], coef[a] + h->n);
    else
      h->coef_ref[a].assign((size_t)h->n, 0.0f);
  }
  return h->upload_coefficients();
}

extern "C" int biot_get_coefficients(biot_handle *h, float *const coef[BIOT_NCOEF])
{
  if (!h || !coef) return BIOT_ERR_ARG;
  if (!h->have_coef) return fail(h, BIOT_ERR_STATE, "no coefficients set");
  if (h->coef_ref[0].empty())
    return fail(h, BIOT_ERR_STATE, "coefficients built by biot_set_params have no reference-representation copy");
  for (int a = 0; a < BIOT_NCOEF; ++a)
    if (coef[a]) std::memcpy(coef[a], h->coef_ref[a].data(), (size_t)h->n * sizeof(float));
  return BIOT_OK;
}

extern "C" int biot_get_pml_profile(biot_handle *h, int32_t axis, float *d, float *d2, float *e,
                                    float *e2)
{
  if (!h) return BIOT_ERR_ARG;
  if (axis < 0 || axis > 2) return fail(h, BIOT_ERR_ARG, "axis must be 0, 1 or 2");
  const auto &P = h->prof[axis];
  const size_t n = P[0].size();
  if (d) std::memcpy(d, P[0].data(), n * 4);
  if (d2) std::memcpy(d2, P[1].data(), n * 4);
  if (e) std::memcpy(e, P[2].data(), n * 4);
  if (e2) std::memcpy(e2, P[3].data(), n * 4);
  return BIOT_OK;
}

// ------------------------------------------------------------------------------------
// source / receivers
// ------------------------------------------------------------------------------------
extern "C" int biot_ricker(float f0, float dt, int32_t nt, float *out)
{
  if (!out || nt < 0 || !(f0 > 0)) return BIOT_ERR_ARG;
  for (int it = 0; it < nt; ++it) out[it] = ricker(it, dt, f0);
  return BIOT_OK;
}

extern "C" int biot_set_source(biot_handle *h, int64_t cell, uint32_t mask, const float *w,
                               int32_t nt)
{
  if (!h) return BIOT_ERR_ARG;
  h->free_adjoint();  // checkpoints of another source are stale
  if (cell < 0 || cell >= h->n) return fail(h, BIOT_ERR_ARG, "source cell %lld outside grid", (long long)cell);
  if (mask >> NFIELD) return fail(h, BIOT_ERR_ARG, "source field mask has bits above field 12");
  if (nt < 0 || (nt > 0 && !w)) return fail(h, BIOT_ERR_ARG, "wavelet null or nt < 0");
  h->src_cell = cell;
  h->src_mask = mask;
  h->wavelet.assign(w, w + nt);
  if (h->d_wavelet) (void)hipFree(h->d_wavelet);
  h->d_wavelet = nullptr;
  if (nt > 0) {
    HIPCHK(h, hipMalloc(&h->d_wavelet, (size_t)nt * sizeof(float)));
    HIPCHK(h, hipMemcpy(h->d_wavelet, w, (size_t)nt * sizeof(float), hipMemcpyHostToDevice));
  }
  if (h->half() && h->have_coef) return h->upload_coefficients();  // storage scales follow |w|
  return BIOT_OK;
}

extern "C" int biot_set_receivers(biot_handle *h, const int64_t *cells, int64_t nrec,
                                  uint32_t mask, int32_t nt_max)
{
  if (!h) return BIOT_ERR_ARG;
  h->free_adjoint();  // the adjoint source layout follows the receivers
  if (nrec < 0 || (nrec > 0 && !cells)) return fail(h, BIOT_ERR_ARG, "receiver list invalid");
  if (mask >> NFIELD) return fail(h, BIOT_ERR_ARG, "receiver field mask has bits above field 12");
  if (nt_max < 0) return fail(h, BIOT_ERR_ARG, "nt_max < 0");
  for (int64_t r = 0; r < nrec; ++r)
    if (cells[r] < 0 || cells[r] >= h->n)
      return fail(h, BIOT_ERR_ARG, "receiver %lld at cell %lld outside grid", (long long)r,
                  (long long)cells[r]);
  if (h->d_rec) (void)hipFree(h->d_rec);
  if (h->d_traces) (void)hipFree(h->d_traces);
  if (h->d_traces64) (void)hipFree(h->d_traces64);
  h->d_rec = nullptr;
  h->d_traces = nullptr;
  h->d_traces64 = nullptr;
  h->ncomp = 0;
  for (int f = 0; f < NFIELD; ++f)
    if (mask & (1u << f)) h->comp[h->ncomp++] = f;
  h->nrec = nrec;
  h->rec_mask = mask;
  h->nt_max = nt_max;
  const size_t tr = (size_t)h->ncomp * nt_max * nrec;
  if (nrec > 0) {
    HIPCHK(h, hipMalloc(&h->d_rec, (size_t)nrec * sizeof(int64_t)));
    HIPCHK(h, hipMemcpy(h->d_rec, cells, (size_t)nrec * sizeof(int64_t), hipMemcpyHostToDevice));
  }
  if (tr > 0) {
    HIPCHK(h, hipMalloc(&h->d_traces, tr * sizeof(float)));
    HIPCHK(h, hipMemsetAsync(h->d_traces, 0, tr * sizeof(float), h->stream));
    if (h->prec == BIOT_FP64) {
      HIPCHK(h, hipMalloc((void **)&h->d_traces64, tr * sizeof(double)));
      HIPCHK(h, hipMemsetAsync(h->d_traces64, 0, tr * sizeof(double), h->stream));
    }
    if (int rc = h->clear_nonfinite()) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return BIOT_OK;
}

// ------------------------------------------------------------------------------------
// time loop
// ------------------------------------------------------------------------------------
extern "C" int biot_reset(biot_handle *h)
{
  if (!h) return BIOT_ERR_ARG;
  HIPCHK(h, hipMemsetAsync(h->d_fields, 0, h->fields_bytes(), h->stream));
  if (h->d_lo) HIPCHK(h, hipMemsetAsync(h->d_lo, 0, h->lo_bytes(), h->stream));
  for (int ax = 0; ax < 3; ++ax)
    if (h->slab[ax] > 0)
      HIPCHK(h, hipMemsetAsync(h->d_carry[ax], 0, (size_t)NCARRY * h->slab[ax] * h->csz, h->stream));
  if (h->d_traces)
    HIPCHK(h, hipMemsetAsync(h->d_traces, 0, (size_t)h->ncomp * h->nt_max * h->nrec * sizeof(float),
                             h->stream));
  if (h->d_traces64)
    HIPCHK(h, hipMemsetAsync(h->d_traces64, 0,
                             (size_t)h->ncomp * h->nt_max * h->nrec * sizeof(double), h->stream));
  if (int rc = h->clear_nonfinite()) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->it = 0;
  return BIOT_OK;
}

int biot_handle::enqueue_steps(int nsteps, hipEvent_t *marks, char *states)
{
  const size_t sb = states ? state_bytes() : 0;
  SourceArgs sa;
  source_args(sa);
  RecordArgs ra;
  record_args(ra);

  StepArgs args;
  fill_args(args);

  for (int s = 0; s < nsteps; ++s) {
    if (d_wavelet && src_mask && it < (int)wavelet.size()) {
      sa.it = it;
      hipError_t e = launch_source(prec, sa, stream);
      if (e != hipSuccess) return fail(this, BIOT_ERR_HIP, "source launch: %s", hipGetErrorString(e));
    }
    if (states) save_state(states + (size_t)s * sb);
    if (marks) (void)hipEventRecord(marks[3 * s], stream);
    hipError_t e = launch_velocity(prec, g.order, args, shape, stream);
    if (e != hipSuccess) return fail(this, BIOT_ERR_HIP, "velocity launch: %s", hipGetErrorString(e));
    if (marks) (void)hipEventRecord(marks[3 * s + 1], stream);
    e = launch_stress(prec, g.order, args, shape, stream);
    if (e != hipSuccess) return fail(this, BIOT_ERR_HIP, "stress launch: %s", hipGetErrorString(e));
    if (marks) (void)hipEventRecord(marks[3 * s + 2], stream);
    if (d_traces && it < nt_max) {
      ra.it = it;
      e = launch_record(prec, ra, stream);
      if (e != hipSuccess) return fail(this, BIOT_ERR_HIP, "record launch: %s", hipGetErrorString(e));
    }
    ++it;
  }
  return BIOT_OK;
}

extern "C" int biot_run(biot_handle *h, int32_t nsteps)
{
  if (!h) return BIOT_ERR_ARG;
  if (nsteps < 0) return fail(h, BIOT_ERR_ARG, "nsteps < 0");
  if (!h->have_coef) return fail(h, BIOT_ERR_STATE, "biot_run before biot_set_medium/coefficients");
  HIPCHK(h, hipSetDevice(h->g.device));
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  int rc = h->enqueue_steps(nsteps, nullptr);
  if (rc) {
    (void)hipStreamSynchronize(h->stream);
    return rc;
  }
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  return h->check_nonfinite();
}

extern "C" int biot_profile_steps(biot_handle *h, int32_t nsteps, double *ms_v, double *ms_t)
{
  if (!h || !ms_v || !ms_t) return BIOT_ERR_ARG;
  if (nsteps <= 0 || nsteps > 100000) return fail(h, BIOT_ERR_ARG, "nsteps must be in 1..100000");
  if (!h->have_coef) return fail(h, BIOT_ERR_STATE, "biot_profile_steps before set_medium");
  HIPCHK(h, hipSetDevice(h->g.device));
  std::vector<hipEvent_t> ev((size_t)3 * nsteps, nullptr);
  int rc = BIOT_OK;
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) rc = BIOT_ERR_HIP;
  if (!rc) rc = h->enqueue_steps(nsteps, ev.data());
  double sv = 0, stt = 0;
  if (!rc && hipStreamSynchronize(h->stream) == hipSuccess) {
    for (int s = 0; s < nsteps; ++s) {
      float a = 0, b = 0;
      (void)hipEventElapsedTime(&a, ev[3 * s], ev[3 * s + 1]);
      (void)hipEventElapsedTime(&b, ev[3 * s + 1], ev[3 * s + 2]);
      sv += a;
      stt += b;
    }
  } else if (!rc) {
    rc = fail(h, BIOT_ERR_HIP, "stream synchronize failed while profiling");
  }
  (void)hipStreamSynchronize(h->stream);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc) return rc;
  *ms_v = sv / nsteps;
  *ms_t = stt / nsteps;
  return BIOT_OK;
}

// ------------------------------------------------------------------------------------
// launch shapes and the autotuner
// ------------------------------------------------------------------------------------
static bool shape_ok(int bx, int by, int zc)
{
  return bx >= 64 && bx % 64 == 0 && by >= 1 && bx * by <= 256 && zc >= 1 && zc <= 4096;
}
extern "C" int biot_set_launch_shape(biot_handle *h, int32_t bx, int32_t by, int32_t zchunk)
{
  if (!h) return BIOT_ERR_ARG;
  if (!shape_ok(bx, by, zchunk))
    return fail(h, BIOT_ERR_ARG, "launch shape %dx%d, zchunk %d not allowed", bx, by, zchunk);
  h->shape = LaunchShape{bx, by, zchunk};
  return BIOT_OK;
}

extern "C" int biot_get_launch_shape(biot_handle *h, int32_t *bx, int32_t *by, int32_t *zchunk)
{
  if (!h || !bx || !by || !zchunk) return BIOT_ERR_ARG;
  *bx = h->shape.bx;
  *by = h->shape.by;
  *zchunk = h->shape.zchunk;
  return BIOT_OK;
}

namespace {
std::string tune_cache_path()
{
  if (const char *p = std::getenv("BIOT_TUNE_CACHE")) return p;
  if (const char *home = std::getenv("HOME")) return std::string(home) + "/.cache/biot_tune.txt";
  return "biot_tune.txt";
}
std::string tune_key(const biot_handle *h)
{
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, h->g.device) == hipSuccess) arch = prop.gcnArchName;
  for (auto &ch : arch)
    if (ch == ' ') ch = '_';
  char buf[256];
  snprintf(buf, sizeof buf, "v3 %s p%d o%d %dx%dx%d", arch.c_str(), h->prec, h->g.order, h->g.nx,
           h->g.ny, h->g.nz);
  return buf;
}
bool tune_lookup(const std::string &key, LaunchShape &out)
{
  FILE *f = std::fopen(tune_cache_path().c_str(), "r");
  if (!f) return false;
  char line[512];
  bool found = false;
  while (std::fgets(line, sizeof line, f)) {
    std::string s(line);
    if (s.compare(0, key.size(), key) != 0 || s.size() <= key.size() || s[key.size()] != ' ')
      continue;
    int bx, by, zc;
    if (std::sscanf(s.c_str() + key.size(), " %d %d %d", &bx, &by, &zc) == 3 && shape_ok(bx, by, zc)) {
      out = LaunchShape{bx, by, zc};
      found = true;  // the last entry for a key wins
    }
  }
  std::fclose(f);
  return found;
}
void tune_store(const std::string &key, LaunchShape s, double ms)
{
  const std::string path = tune_cache_path();
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    std::string dir = path.substr(0, slash);
    // create the directory without a shell: walk the components
    for (size_t k = 1; k <= dir.size(); ++k)
      if (k == dir.size() || dir[k] == '/') (void)mkdir(dir.substr(0, k).c_str(), 0755);
  }
  if (FILE *f = std::fopen(path.c_str(), "a")) {
    std::fprintf(f, "%s %d %d %d %.6f\n", key.c_str(), s.bx, s.by, s.zchunk, ms);
    std::fclose(f);
  }
}
}  // namespace

extern "C" int biot_autotune(biot_handle *h, int32_t trial_steps)
{
  if (!h) return BIOT_ERR_ARG;
  if (trial_steps < 1 || trial_steps > 1000) return fail(h, BIOT_ERR_ARG, "trial_steps must be 1..1000");
  if (!h->have_coef) return fail(h, BIOT_ERR_STATE, "biot_autotune before set_medium");
  const std::string key = tune_key(h);
  LaunchShape best{};
  if (tune_lookup(key, best)) {
    h->shape = best;
    return biot_reset(h);
  }
  static const int shapes[][2] = {{64, 4}, {64, 2}, {64, 1}, {128, 2}, {128, 1}, {256, 1}};
  static const int chunks[] = {1, 2, 4, 8, 16, 32, 64};
  const int nzi = h->g.nz - h->g.order;
  double best_ms = 1e30;
  const LaunchShape keep = h->shape;
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  int rc = BIOT_OK;
  for (auto &sh : shapes) {
    for (int zc : chunks) {
      if (zc > 2 * nzi && zc != chunks[0]) continue;
      h->shape = LaunchShape{sh[0], sh[1], zc};
      rc = h->enqueue_steps(1, nullptr);  // warm-up
      if (!rc && hipEventRecord(e0, h->stream) != hipSuccess) rc = BIOT_ERR_HIP;
      if (!rc) rc = h->enqueue_steps(trial_steps, nullptr);
      if (!rc && hipEventRecord(e1, h->stream) != hipSuccess) rc = BIOT_ERR_HIP;
      if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = BIOT_ERR_HIP;
      float ms = 0;
      if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = BIOT_ERR_HIP;
      if (rc) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = h->shape;
      }
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    h->shape = keep;
    (void)hipStreamSynchronize(h->stream);
    return rc == BIOT_ERR_HIP ? fail(h, rc, "autotune timing failed") : rc;
  }
  h->shape = best;
  tune_store(key, best, best_ms / trial_steps);
  return biot_reset(h);
}

extern "C" int biot_step_count(biot_handle *h, int32_t *it)
{
  if (!h || !it) return BIOT_ERR_ARG;
  *it = h->it;
  return BIOT_OK;
}

extern "C" int biot_synchronize(biot_handle *h)
{
  if (!h) return BIOT_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return BIOT_OK;
}

extern "C" int biot_last_run_ms(biot_handle *h, double *ms)
{
  if (!h || !ms) return BIOT_ERR_ARG;
  *ms = h->last_ms;
  return BIOT_OK;
}

// ------------------------------------------------------------------------------------
// outputs
// ------------------------------------------------------------------------------------
extern "C" int biot_read_traces(biot_handle *h, float *out, int64_t count)
{
  if (!h || !out) return BIOT_ERR_ARG;
  const int64_t need = (int64_t)h->ncomp * h->nt_max * h->nrec;
  if (count < need)
    return fail(h, BIOT_ERR_ARG, "trace buffer holds %lld, need %lld", (long long)count,
                (long long)need);
  if (need == 0) return BIOT_OK;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(out, h->d_traces, (size_t)need * sizeof(float), hipMemcpyDeviceToHost));
  return BIOT_OK;
}

extern "C" int biot_read_traces_f64(biot_handle *h, double *out, int64_t count)
{
  if (!h || !out) return BIOT_ERR_ARG;
  if (h->prec != BIOT_FP64) return fail(h, BIOT_ERR_STATE, "double traces are kept by FP64 handles only");
  const int64_t need = (int64_t)h->ncomp * h->nt_max * h->nrec;
  if (count < need)
    return fail(h, BIOT_ERR_ARG, "trace buffer holds %lld, need %lld", (long long)count,
                (long long)need);
  if (need == 0) return BIOT_OK;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(out, h->d_traces64, (size_t)need * sizeof(double), hipMemcpyDeviceToHost));
  return BIOT_OK;
}

extern "C" int biot_read_field_f64(biot_handle *h, int32_t field, double *out)
{
  if (!h || !out) return BIOT_ERR_ARG;
  if (field < 0 || field >= NFIELD) return fail(h, BIOT_ERR_ARG, "field %d out of range", field);
  std::vector<unsigned char> tmp((size_t)h->n * h->esz);
  std::vector<int8_t> lo;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(tmp.data(), h->field_ptr(field), tmp.size(), hipMemcpyDeviceToHost));
  if (h->lo_ptr(field)) {
    lo.resize(lo_span(h, field, h->n));
    HIPCHK(h, hipMemcpy(lo.data(), h->lo_copy_ptr(field), lo.size(), hipMemcpyDeviceToHost));
  }
  from_storage(h->prec, tmp.data(), lo.empty() ? nullptr : lo.data(), h->n, h->fscale[field], out, h->packed(field),
               pk_slot(field));
  return BIOT_OK;
}

extern "C" int biot_read_field(biot_handle *h, int32_t field, float *out)
{
  if (!h || !out) return BIOT_ERR_ARG;
  std::vector<double> tmp((size_t)h->n);
  int rc = biot_read_field_f64(h, field, tmp.data());
  if (rc) return rc;
  for (int64_t c = 0; c < h->n; ++c) out[c] = (float)tmp[c];
  return BIOT_OK;
}

extern "C" int biot_write_field_f64(biot_handle *h, int32_t field, const double *in)
{
  if (!h || !in) return BIOT_ERR_ARG;
  if (field < 0 || field >= NFIELD) return fail(h, BIOT_ERR_ARG, "field %d out of range", field);
  std::vector<unsigned char> tmp((size_t)h->n * h->esz);
  std::vector<int8_t> lo(h->lo_ptr(field) ? lo_span(h, field, h->n) : 0);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  // FP16M's residuals share their records with the other six fields: read, patch, write back
  if (h->packed(field)) HIPCHK(h, hipMemcpy(lo.data(), h->lo_copy_ptr(field), lo.size(), hipMemcpyDeviceToHost));
  to_storage(h->prec, in, h->n, 1.0 / h->fscale[field], tmp.data(), lo.empty() ? nullptr : lo.data(),
             h->packed(field), pk_slot(field));
  HIPCHK(h, hipMemcpy(h->field_ptr(field), tmp.data(), tmp.size(), hipMemcpyHostToDevice));
  if (!lo.empty())
    HIPCHK(h, hipMemcpy(h->lo_copy_ptr(field), lo.data(), lo.size(), hipMemcpyHostToDevice));
  return BIOT_OK;
}

extern "C" int biot_snapshot_xz(biot_handle *h, int32_t field, int32_t j, float *out)
{
  if (!h || !out) return BIOT_ERR_ARG;
  if (field < 0 || field >= NFIELD) return fail(h, BIOT_ERR_ARG, "field %d out of range", field);
  if (j < 0 || j >= h->g.ny) return fail(h, BIOT_ERR_ARG, "y index %d outside grid", j);
  const int nx = h->g.nx, nz = h->g.nz;
  const size_t eb = h->esz;
  std::vector<unsigned char> row((size_t)nx * nz * eb);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  // one 2-D copy: nz rows of nx elements, source pitch nx*ny elements
  HIPCHK(h, hipMemcpy2D(row.data(), (size_t)nx * eb, (char *)h->field_ptr(field) + (size_t)j * nx * eb,
                        (size_t)nx * h->g.ny * eb, (size_t)nx * eb, (size_t)nz, hipMemcpyDeviceToHost));
  std::vector<int8_t> lo;
  const size_t ls = h->lo_stride(field);
  if (h->lo_ptr(field)) {
    // nz rows of nx cells' residuals (nx bytes, or nx 4-byte records for FP16M)
    lo.resize((size_t)nx * nz * ls);
    HIPCHK(h, hipMemcpy2D(lo.data(), (size_t)nx * ls, h->lo_copy_ptr(field) + (size_t)j * nx * ls,
                          (size_t)nx * h->g.ny * ls, lo_span(h, field, nx), (size_t)nz, hipMemcpyDeviceToHost));
  }
  std::vector<double> tmp((size_t)nx * nz);
  from_storage(h->prec, row.data(), lo.empty() ? nullptr : lo.data(), (int64_t)nx * nz,
               h->fscale[field], tmp.data(), h->packed(field), pk_slot(field));
  for (size_t c = 0; c < tmp.size(); ++c) out[c] = (float)tmp[c];
  return BIOT_OK;
}

extern "C" int biot_get_field_scales(biot_handle *h, float *scales13)
{
  if (!h || !scales13) return BIOT_ERR_ARG;
  for (int f = 0; f < NFIELD; ++f) scales13[f] = h->fscale[f];
  return BIOT_OK;
}

extern "C" int biot_device_bytes(biot_handle *h, int64_t *bytes)
{
  if (!h || !bytes) return BIOT_ERR_ARG;
  *bytes = h->device_bytes;
  return BIOT_OK;
}
