// C-ABI of the inversion parameterization, the checkpointed forward run, the FP64 misfit
// and the discrete adjoint (include/biot.h).  None of this exists in the reference, whose
// driver is forward-only (kernel.cu:1728-2697).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/biot.h"
#include "biot_adjoint.hpp"
#include "biot_handle.hpp"
#include "biot_params.hpp"

using namespace biot;

// with every step's state stored, the float partial gradient sums are added into the FP64
// sums every this many adjoint steps (the default checkpoint spacing: at most 50 float
// additions per partial sum)
constexpr int kStoredFlushEvery = 50;

namespace {

int afail(biot_handle *h, int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  h->err = buf;
  return code;
}

#define AHIP(h, call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return afail((h), BIOT_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
  } while (0)

ParamArrays param_view(const biot_handle *h)
{
  const auto &P = h->par;
  return ParamArrays{P[0].data(), P[1].data(), P[2].data(), P[3].data(),
                     P[4].data(), P[5].data(), P[6].data(), P[7].data()};
}

CoefScales coef_scales(const biot_handle *h)
{
  return CoefScales{(double)h->g.dt, (double)h->g.Cv, (double)h->g.Cp, (double)(1.0f / h->g.dx)};
}


}  // namespace

extern "C" int biot_set_params(biot_handle *h, const biot_params *p)
{
  if (!h) return BIOT_ERR_ARG;
  if (!p || !p->lambda || !p->mu || !p->alpha || !p->M || !p->rho || !p->rhof || !p->C1 || !p->tphi)
    return afail(h, BIOT_ERR_ARG, "biot_set_params: every per-cell array must be given");
  const double *src[8] = {p->lambda, p->mu, p->alpha, p->M, p->rho, p->rhof, p->C1, p->tphi};
  for (int64_t c = 0; c < h->n; ++c) {
    if (!(p->rho[c] > 0) || !(p->rhof[c] > 0) || !(p->M[c] > 0) || !(p->mu[c] >= 0) ||
        !(p->C1[c] >= 0) || !(p->tphi[c] > 0))
      return afail(h, BIOT_ERR_ARG, "cell %lld: invalid parameters", (long long)c);
  }
  h->free_adjoint();  // checkpoints of the previous model are stale
  for (int q = 0; q < 8; ++q) h->par[q].assign(src[q], src[q] + h->n);
  if (h->prec == BIOT_FP16M) {
    // the cell moduli of the parameters (outputfunc.cpp:210-217 in parameter form): C = alpha M,
    // lambda_c = lambda + alpha^2 M, H = lambda_c + 2 mu, C2 = tphi rho_f
    for (auto &v : h->mod) v.resize((size_t)h->n);
    for (int64_t c = 0; c < h->n; ++c) {
      const double al = p->alpha[c], M = p->M[c], lc = p->lambda[c] + al * al * M;
      h->mod[0][c] = (float)M;
      h->mod[1][c] = (float)(lc + 2.0 * p->mu[c]);
      h->mod[2][c] = (float)lc;
      h->mod[3][c] = (float)(al * M);
      h->mod[4][c] = (float)p->C1[c];
      h->mod[5][c] = (float)(p->tphi[c] * p->rhof[c]);
      h->mod[6][c] = (float)p->rho[c];
      h->mod[7][c] = (float)p->rhof[c];
      h->mod[8][c] = (float)p->mu[c];
    }
  }
  kernel_coefs_from_params(h->g.nx, h->g.ny, h->g.nz, coef_scales(h), param_view(h), h->K_last);
  for (auto &v : h->coef_ref) std::vector<float>().swap(v);  // the reference arrays no longer apply
  return h->upload_kernel_coefs(h->K_last);
}

extern "C" int biot_forward_checkpointed(biot_handle *h, int32_t nt, int32_t every)
{
  if (!h) return BIOT_ERR_ARG;
  if (nt <= 0 || every <= 0) return afail(h, BIOT_ERR_ARG, "nt and every must be positive");
  if (!h->have_coef) return afail(h, BIOT_ERR_STATE, "no medium set");
  if (h->nt_max < nt || h->ncomp == 0) return afail(h, BIOT_ERR_STATE, "receivers must record >= nt steps");
  for (uint32_t f = 0; f < NFIELD; ++f)
    if ((h->src_mask >> f) & 1u && !(f >= 6))
      return afail(h, BIOT_ERR_ARG, "the adjoint path supports sources on stresses and pressure only");
  h->free_adjoint();
  if (int rc = h->build_adjoint_coefs()) return rc;  // FP16M: the coefficient copy the adjoint reads
  const size_t sb = h->state_bytes();
  if (every == 1) {
    // every step's state, in one block the adjoint reads directly (no recomputation)
    void *p = nullptr;
    const size_t bytes = sb * (size_t)(nt + 1);
    if (hipMalloc(&p, bytes) != hipSuccess)
      return afail(h, BIOT_ERR_NOMEM, "the %d step states (%zu bytes) do not fit; use every > 1", nt + 1, bytes);
    h->ckpt.push_back(p);
    int rc = biot_reset(h);
    if (rc) return rc;
    AHIP(h, hipEventRecord(h->ev0, h->stream));
    rc = h->enqueue_steps(nt, nullptr, (char *)p);
    if (rc) return rc;
    h->save_state((char *)p + (size_t)nt * sb);
    AHIP(h, hipEventRecord(h->ev1, h->stream));
    AHIP(h, hipEventSynchronize(h->ev1));
    float ms = 0;
    AHIP(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    h->ckpt_every = 1;
    h->ckpt_nt = nt;
    return BIOT_OK;
  }
  const int nseg = (nt + every - 1) / every;
  for (int s = 0; s < nseg; ++s) {
    void *p = nullptr;
    if (hipMalloc(&p, sb) != hipSuccess) {
      h->free_adjoint();
      return afail(h, BIOT_ERR_NOMEM, "checkpoint %d of %d (%zu bytes) does not fit", s, nseg, sb);
    }
    h->ckpt.push_back(p);
  }
  int rc = biot_reset(h);
  if (rc) return rc;
  AHIP(h, hipEventRecord(h->ev0, h->stream));
  for (int s = 0; s < nseg; ++s) {
    h->save_state(h->ckpt[s]);
    rc = h->enqueue_steps(std::min(every, nt - s * every), nullptr);
    if (rc) return rc;
  }
  AHIP(h, hipEventRecord(h->ev1, h->stream));
  AHIP(h, hipEventSynchronize(h->ev1));
  float ms = 0;
  AHIP(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->ckpt_every = every;
  h->ckpt_nt = nt;
  return BIOT_OK;
}

extern "C" int biot_misfit_l2(biot_handle *h, const float *observed, float *residual, double *J)
{
  if (!h || !observed || !J) return BIOT_ERR_ARG;
  const int64_t m = (int64_t)h->ncomp * h->nt_max * h->nrec;
  if (m == 0) return afail(h, BIOT_ERR_STATE, "no receivers");
  float *d_obs = nullptr, *d_res = nullptr;
  double *d_J = nullptr;
  int rc = BIOT_OK;
  if (hipMalloc(&d_obs, m * sizeof(float)) != hipSuccess || hipMalloc(&d_res, m * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_J, sizeof(double)) != hipSuccess)
    rc = afail(h, BIOT_ERR_NOMEM, "misfit buffers");
  if (!rc && (hipMemcpyAsync(d_obs, observed, m * sizeof(float), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
              hipMemsetAsync(d_J, 0, sizeof(double), h->stream) != hipSuccess ||
              launch_misfit(h->d_traces64 ? (const void *)h->d_traces64 : h->d_traces, h->d_traces64 != nullptr,
                            d_obs, d_res, m, d_J, h->stream) != hipSuccess ||
              hipMemcpyAsync(J, d_J, sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
              (residual && hipMemcpyAsync(residual, d_res, m * sizeof(float), hipMemcpyDeviceToHost,
                                          h->stream) != hipSuccess) ||
              hipStreamSynchronize(h->stream) != hipSuccess))
    rc = afail(h, BIOT_ERR_HIP, "misfit reduction failed");
  for (void *p : {(void *)d_obs, (void *)d_res, (void *)d_J})
    if (p) (void)hipFree(p);
  return rc;
}

extern "C" int biot_adjoint(biot_handle *h, const float *adj_source, double *const grad_params[6],
                            double *const grad_coef[22])
{
  if (!h || !adj_source) return BIOT_ERR_ARG;
  if (h->ckpt_nt <= 0) return afail(h, BIOT_ERR_STATE, "biot_adjoint needs biot_forward_checkpointed first");
  if (grad_params && h->par[0].empty())
    return afail(h, BIOT_ERR_STATE, "parameter gradients need biot_set_params");
  if (grad_params)
    for (int q = 0; q < 6; ++q)
      if (!grad_params[q]) return afail(h, BIOT_ERR_ARG, "grad_params[%d] is null", q);
  const int nt = h->ckpt_nt, every = h->ckpt_every, order = h->g.order;
  const int64_t n = h->n;
  const int64_t m = (int64_t)h->ncomp * h->nt_max * h->nrec;
  const bool stored = every == 1;  // every step's state kept by biot_forward_checkpointed
  int rc = h->alloc_adjoint(stored ? 0 : every);
  if (rc) return rc;
  AHIP(h, hipMemsetAsync(h->d_lam, 0, h->lam_bytes(), h->stream));
  AHIP(h, hipMemsetAsync(h->d_grad, 0, (size_t)NKC * n * sizeof(double), h->stream));
  if (h->d_gpart) AHIP(h, hipMemsetAsync(h->d_gpart, 0, (size_t)NKC * n * sizeof(float), h->stream));
  AHIP(h, hipMemcpyAsync(h->d_adjsrc, adj_source, (size_t)m * sizeof(float), hipMemcpyHostToDevice, h->stream));

  AdjArgs A;
  h->adj_args(A);
  InjectArgs I;
  h->inject_args(I);
  SourceArgs sa;
  h->source_args(sa);
  StepArgs st;
  h->fill_args(st);
  const size_t sb = h->state_bytes();
  auto slot = [&](int q) { return (char *)h->d_slots + (size_t)q * sb; };

  // ---- backward sweep over the checkpoint segments
  AHIP(h, hipEventRecord(h->ev0, h->stream));
  if (stored) {
    // the stored states are the segment states: state it is slot it.  The float partial sums
    // go into the FP64 sums at the same steps as with 50-step segments (after it = 950, ...,
    // 0), so both modes give the same gradients bit for bit.
    A.slots = (const char *)h->ckpt[0];
    for (int it = nt - 1; it >= 0; --it) {
      AHIP(h, launch_adjoint_inject(h->prec, I, it, h->stream));
      AHIP(h, launch_adjoint_step(h->prec, order, A, it, h->stream));
      if (it % kStoredFlushEvery == 0) AHIP(h, launch_adjoint_flush(h->prec, order, A, h->stream));
    }
  }
  const int nseg = stored ? 0 : (int)h->ckpt.size();
  for (int s = nseg - 1; s >= 0; --s) {
    const int n0 = s * every, K = std::min(every, nt - n0);
    h->load_state(h->ckpt[s]);
    for (int q = 0; q < K; ++q) {
      const int it = n0 + q;
      if (h->d_wavelet && h->src_mask && it < (int)h->wavelet.size()) {
        sa.it = it;
        AHIP(h, launch_source(h->prec, sa, h->stream));
      }
      h->save_state(slot(q));  // state of step it after its source
      AHIP(h, launch_velocity(h->prec, order, st, h->shape, h->stream));
      AHIP(h, launch_stress(h->prec, order, st, h->shape, h->stream));
    }
    h->save_state(slot(K));
    for (int it = n0 + K - 1; it >= n0; --it) {
      AHIP(h, launch_adjoint_inject(h->prec, I, it, h->stream));
      AHIP(h, launch_adjoint_step(h->prec, order, A, it - n0, h->stream));
    }
    AHIP(h, launch_adjoint_flush(h->prec, order, A, h->stream));  // the segment's float partial sums
  }
  AHIP(h, hipEventRecord(h->ev1, h->stream));
  AHIP(h, hipEventSynchronize(h->ev1));
  float ms = 0;
  AHIP(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;

  // ---- gradients out: storage-unit coefficient gradients times the per-array factor
  std::vector<double> gK((size_t)NKC * n);
  AHIP(h, hipMemcpy(gK.data(), h->d_grad, gK.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < NKC; ++k)
    if (h->kfac[k] != 1.0)
      for (int64_t c = 0; c < n; ++c) gK[(size_t)k * n + c] *= h->kfac[k];
  if (grad_coef)
    for (int k = 0; k < NKC; ++k)
      if (grad_coef[k]) std::memcpy(grad_coef[k], gK.data() + (size_t)k * n, (size_t)n * sizeof(double));
  if (grad_params) {
    const double *gk[NKC];
    for (int k = 0; k < NKC; ++k) gk[k] = gK.data() + (size_t)k * n;
    for (int q = 0; q < 6; ++q) std::memset(grad_params[q], 0, (size_t)n * sizeof(double));
    chain_params(h->g.nx, h->g.ny, h->g.nz, coef_scales(h), param_view(h), gk, grad_params);
  }
  return BIOT_OK;
}

static_assert((int)BIOT_K_SYZ == (int)K_SYZ && (int)BIOT_NKCOEF == (int)NKC, "kernel coefficient order");

namespace {
bool host_grid_ok(const biot_grid *g)
{
  return g && g->nx >= 3 && g->ny >= 3 && g->nz >= 3 && g->dx > 0 && g->dt > 0 && g->Cv > 0 && g->Cp > 0;
}
bool params_ok(const biot_params *p)
{
  return p && p->lambda && p->mu && p->alpha && p->M && p->rho && p->rhof && p->C1 && p->tphi;
}
CoefScales grid_scales(const biot_grid *g)
{
  return CoefScales{(double)g->dt, (double)g->Cv, (double)g->Cp, (double)(1.0f / g->dx)};
}
ParamArrays param_arrays(const biot_params *p)
{
  return ParamArrays{p->lambda, p->mu, p->alpha, p->M, p->rho, p->rhof, p->C1, p->tphi};
}
}  // namespace

extern "C" int biot_params_coefficients(const biot_grid *g, const biot_params *p, double *const K[22])
{
  if (!host_grid_ok(g) || !params_ok(p) || !K) return BIOT_ERR_ARG;
  for (int k = 0; k < NKC; ++k)
    if (!K[k]) return BIOT_ERR_ARG;
  std::vector<double> Kv[NKC];
  kernel_coefs_from_params(g->nx, g->ny, g->nz, grid_scales(g), param_arrays(p), Kv);
  for (int k = 0; k < NKC; ++k) std::memcpy(K[k], Kv[k].data(), Kv[k].size() * sizeof(double));
  return BIOT_OK;
}

extern "C" int biot_params_chain(const biot_grid *g, const biot_params *p, const double *const grad_coef[22],
                                 double *const grad_params[6])
{
  if (!host_grid_ok(g) || !params_ok(p) || !grad_coef || !grad_params) return BIOT_ERR_ARG;
  for (int k = 0; k < NKC; ++k)
    if (!grad_coef[k]) return BIOT_ERR_ARG;
  const size_t n = (size_t)g->nx * g->ny * g->nz;
  for (int q = 0; q < 6; ++q) {
    if (!grad_params[q]) return BIOT_ERR_ARG;
    std::memset(grad_params[q], 0, n * sizeof(double));
  }
  chain_params(g->nx, g->ny, g->nz, grid_scales(g), param_arrays(p), grad_coef, grad_params);
  return BIOT_OK;
}
