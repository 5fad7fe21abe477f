/*
 * CPU oracle for the 13-field staggered-grid Biot time step (TEST INFRASTRUCTURE ONLY:
 * used by tests/, the smoke check and the benchmark's CPU-baseline leg as the checker;
 * never by the product).  See biot_oracle.h for the reference lines each part restates.
 *
 * Build: `make -C oracle` (gcc -O3 -fopenmp -ffp-contract=off: no FMA contraction, so the
 * float/double rounding of every restated expression is the one written).
 */
#include "biot_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------ */
/* IEEE-754 binary16, round to nearest even, subnormals kept, overflow to inf.           */
/* Written from the format definition: 1 sign, 5 exponent (bias 15), 10 fraction bits.   */
/* ------------------------------------------------------------------------------------ */
static uint16_t f16_from_parts(uint32_t sign, int exp2, uint64_t mant, int mbits)
{
  /* value = mant * 2^(exp2 - mbits), mant has its leading one at bit mbits (normalised),
   * or is zero.  Returns the RNE binary16 encoding. */
  if (mant == 0) return (uint16_t)(sign << 15);
  int e = exp2 + 15; /* biased exponent if normal */
  int shift;         /* bits of mant to drop to keep 10 fraction bits (+ implicit one) */
  uint32_t out_exp;
  if (e >= 1) {
    shift = mbits - 10;
    out_exp = (uint32_t)e;
  } else {
    shift = mbits - 10 + (1 - e); /* subnormal: extra right shift */
    out_exp = 0;
  }
  uint64_t q, rem, half;
  if (shift > 63) return (uint16_t)(sign << 15);
  if (shift <= 0) {
    q = mant << (-shift);
    rem = 0;
    half = 1;
  } else {
    q = mant >> shift;
    rem = mant & ((1ull << shift) - 1);
    half = 1ull << (shift - 1);
  }
  if (shift > 0 && (rem > half || (rem == half && (q & 1)))) q++;
  /* q now holds implicit one (bit 10) + fraction for normals, or the subnormal fraction */
  uint32_t bits;
  if (out_exp == 0) {
    /* subnormal; rounding may carry into the smallest normal, which the encoding absorbs */
    bits = (uint32_t)q;
  } else {
    if (q >= (1ull << 11)) { /* carry out of the mantissa */
      q >>= 1;
      out_exp++;
    }
    if (out_exp >= 31) return (uint16_t)((sign << 15) | 0x7c00u);
    bits = (out_exp << 10) | ((uint32_t)q & 0x3ffu);
  }
  return (uint16_t)((sign << 15) | bits);
}

uint16_t or_f64_to_f16(double d)
{
  uint64_t u;
  memcpy(&u, &d, 8);
  uint32_t sign = (uint32_t)(u >> 63);
  int ex = (int)((u >> 52) & 0x7ff);
  uint64_t frac = u & ((1ull << 52) - 1);
  if (ex == 0x7ff) return (uint16_t)((sign << 15) | 0x7c00u | (frac ? 0x200u : 0u));
  if (ex == 0) return (uint16_t)(sign << 15); /* double subnormals are far below f16 range */
  return f16_from_parts(sign, ex - 1023, frac | (1ull << 52), 52);
}

uint16_t or_f32_to_f16(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  uint32_t sign = u >> 31;
  int ex = (int)((u >> 23) & 0xff);
  uint64_t frac = u & 0x7fffffu;
  if (ex == 0xff) return (uint16_t)((sign << 15) | 0x7c00u | (frac ? 0x200u : 0u));
  if (ex == 0) return (uint16_t)(sign << 15); /* float subnormals are below f16 range */
  return f16_from_parts(sign, ex - 127, frac | (1ull << 23), 23);
}

float or_f16_to_f32(uint16_t h)
{
  uint32_t sign = (uint32_t)(h >> 15);
  int ex = (h >> 10) & 0x1f;
  uint32_t frac = h & 0x3ffu;
  float v;
  if (ex == 0x1f) {
    v = frac ? NAN : INFINITY;
  } else if (ex == 0) {
    v = ldexpf((float)frac, -24);
  } else {
    v = ldexpf((float)(frac | 0x400u), ex - 25);
  }
  return sign ? -v : v;
}

/* ------------------------------------------------------------------------------------ */
/* Set-up restatements                                                                   */
/* ------------------------------------------------------------------------------------ */
void or_biot_moduli(int64_t n, const double *Ks, const double *Kf, const double *Kb,
                    const double *mu_in, const double *phi, const double *rhos_in,
                    const double *rhof_in, const double *eta, const double *perm,
                    double tao, double porousm,
                    float *M, float *HH, float *H2u, float *C, float *C1, float *C2,
                    float *rho, float *rhof, float *mu)
{
  /* outputfunc.cpp:190-217: rhos/rhof/mu live in float arrays, the moduli, porosity,
   * permeability and viscosity are double scalars, results are stored as float. */
  for (int64_t c = 0; c < n; c++) {
    const float rhos_f = (float)rhos_in[c];
    const float rhof_f = (float)rhof_in[c];
    const double por = phi[c];
    rhof[c] = rhof_f;
    rho[c] = (float)((1 - por) * rhos_f + por * rhof_f);
    mu[c] = (float)mu_in[c];
    if (mu_in[c] == 0.0 && Kb[c] == 0.0) {
      /* fluid cell of the reference model (outputfunc.cpp:173-193): C=M=H=lambda_c=Kf,
       * no damping, C2 = rho_f */
      const float kf = (float)Kf[c];
      C[c] = kf; M[c] = kf; HH[c] = kf; H2u[c] = kf;
      C1[c] = 0.0f;
      C2[c] = rhof_f;
      rho[c] = rhof_f;
      continue;
    }
    const double Kb1 = Kb[c], Ks1 = Ks[c], Kf1 = Kf[c];
    const double a1 = 1 - Kb1 / Ks1;
    M[c] = (float)(Kf1 * Ks1 / (por * Ks1 + (a1 - por) * Kf1));
    HH[c] = (float)(a1 * a1 * M[c] + Kb1 + mu[c] * 4.0 / 3.0);
    H2u[c] = HH[c] - 2 * mu[c];
    C[c] = (float)(M[c] * a1);
    C1[c] = (float)(eta[c] / perm[c]);
    C2[c] = (float)((1 + 2 / porousm) * tao * rhof[c] / por);
  }
}

void or_stagger_average(int nx, int ny, int nz, const float *mu, const float *rho,
                        const float *rhof, const float *C1, const float *C2,
                        float *muxz, float *muxy, float *muyz,
                        float *rhox, float *rhoy, float *rhoz,
                        float *rhofx, float *rhofy, float *rhofz,
                        float *C1x, float *C1y, float *C1z,
                        float *C2x, float *C2y, float *C2z)
{
  const int64_t n = (int64_t)nx * ny * nz, sy = nx, sz = (int64_t)nx * ny;
  float *outs[] = {muxz, muxy, muyz, rhox, rhoy, rhoz, rhofx, rhofy, rhofz,
                   C1x, C1y, C1z, C2x, C2y, C2z};
  for (int a = 0; a < 15; a++) memset(outs[a], 0, (size_t)n * sizeof(float));
  for (int k = 1; k < nz - 1; k++)
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        const int64_t c = i + j * sy + k * sz;
        /* 4-cell harmonic means of mu on the three shear planes (main.cpp:326-337);
         * `1/mu` is int/float, the 0.25 promotes the sum to double. */
        const float m0 = mu[c], mx = mu[c + 1], my = mu[c + sy], mz = mu[c + sz];
        const float mxz = mu[c + sz + 1], myz = mu[c + sz + sy], mxy = mu[c + sy + 1];
        muxz[c] = (m0 == 0 || mx == 0 || mz == 0 || mxz == 0)
                      ? 0.0f
                      : (float)(1 / (0.25 * (1 / m0 + 1 / mx + 1 / mz + 1 / mxz)));
        muyz[c] = (m0 == 0 || my == 0 || mz == 0 || myz == 0)
                      ? 0.0f
                      : (float)(1 / (0.25 * (1 / m0 + 1 / my + 1 / mz + 1 / myz)));
        muxy[c] = (m0 == 0 || mx == 0 || my == 0 || mxy == 0)
                      ? 0.0f
                      : (float)(1 / (0.25 * (1 / m0 + 1 / mx + 1 / my + 1 / mxy)));
        /* arithmetic two-cell means (main.cpp:339-353) */
        rhofx[c] = (float)(0.5 * (rhof[c] + rhof[c + 1]));
        rhofy[c] = (float)(0.5 * (rhof[c] + rhof[c + sy]));
        rhofz[c] = (float)(0.5 * (rhof[c] + rhof[c + sz]));
        rhox[c] = (float)(0.5 * (rho[c] + rho[c + 1]));
        rhoy[c] = (float)(0.5 * (rho[c] + rho[c + sy]));
        rhoz[c] = (float)(0.5 * (rho[c] + rho[c + sz]));
        C1x[c] = (float)(0.5 * (C1[c] + C1[c + 1]));
        C1y[c] = (float)(0.5 * (C1[c] + C1[c + sy]));
        C1z[c] = (float)(0.5 * (C1[c] + C1[c + sz]));
        C2x[c] = (float)(0.5 * (C2[c] + C2[c + 1]));
        C2y[c] = (float)(0.5 * (C2[c] + C2[c + sy]));
        C2z[c] = (float)(0.5 * (C2[c] + C2[c + sz]));
      }
}

void or_coefficients(int nx, int ny, int nz, float DT, float Cv, float Cp,
                     const float *muxz, const float *muxy, const float *muyz,
                     const float *rhox, const float *rhoy, const float *rhoz,
                     const float *rhofx, const float *rhofy, const float *rhofz,
                     const float *C1x, const float *C1y, const float *C1z,
                     const float *C2x, const float *C2y, const float *C2z,
                     const float *M, const float *C, const float *HH, const float *H2u,
                     float *const coef[OR_NCOEF])
{
  /* Scale factors of main.cpp:248-262 (floats). */
  const float Cvwp1 = 1.0f, Cvwp2 = 1.0f, Cvwp3 = 1.0f, Cvup1 = 1.0f, Cvup2 = 1.0f;
  const float Cpp1 = 100, Cpp2 = 100, Csp1 = 100, Csp2 = 100, Csp3 = 100, Csp4 = 100;
  const int64_t n = (int64_t)nx * ny * nz;
  const float *rf[3] = {rhofx, rhofy, rhofz}, *rh[3] = {rhox, rhoy, rhoz};
  const float *c1[3] = {C1x, C1y, C1z}, *c2[3] = {C2x, C2y, C2z};
  for (int64_t c = 0; c < n; c++) {
    /* fluid-velocity coefficients (main.cpp:431-459), all zero if any C1 average is 0 */
    if (C1x[c] == 0 || C1y[c] == 0 || C1z[c] == 0) {
      for (int a = OR_W1X; a <= OR_W3Z; a++) coef[a][c] = 0;
    } else {
      for (int d = 0; d < 3; d++) {
        const float r = rf[d][c], q = rh[d][c], k1 = c1[d][c], k2 = c2[d][c];
        coef[OR_W1X + d][c] =
            (float)(Cvwp1 * (k2 - (r * r / q) - 0.5 * DT * k1) / (k2 - (r * r / q) + 0.5 * DT * k1));
        coef[OR_W2X + d][c] =
            (float)(Cv * DT * Cvwp2 * (r / q) / (Cp * (k2 - (r * r / q) + 0.5 * DT * k1)));
        coef[OR_W3X + d][c] = (float)(Cv * DT * Cvwp3 / (Cp * (k2 - r * r / q + 0.5 * DT * k1)));
      }
    }
    /* solid-velocity coefficients (main.cpp:462-481), zero if any rho average is 0 */
    if (rhox[c] == 0 || rhoy[c] == 0 || rhoz[c] == 0) {
      for (int a = OR_U1X; a <= OR_U2Z; a++) coef[a][c] = 0;
    } else {
      for (int d = 0; d < 3; d++) {
        coef[OR_U1X + d][c] = Cv * DT * Cvup1 / (rh[d][c] * Cp);
        coef[OR_U2X + d][c] = rf[d][c] * Cvup2 / (rh[d][c]);
      }
    }
    /* pressure and stress coefficients (main.cpp:483-494) */
    coef[OR_PP1][c] = C[c] * DT * Cp * Cpp1 / Cv;
    coef[OR_PP2][c] = M[c] * DT * Cp * Cpp2 / Cv;
    coef[OR_SP1][c] = Cp * DT * H2u[c] * Csp1 / Cv;
    coef[OR_SP2][c] = Cp * DT * HH[c] * Csp2 / Cv;
    coef[OR_SP3][c] = Cp * DT * C[c] * Csp3 / Cv;
    coef[OR_SPXY][c] = Cp * DT * muxy[c] * Csp4 / Cv;
    coef[OR_SPXZ][c] = Cp * DT * muxz[c] * Csp4 / Cv;
    coef[OR_SPYZ][c] = Cp * DT * muyz[c] * Csp4 / Cv;
  }
}

void or_pml_profile(int n, float H, float DT, int NP, float v0, float RC,
                    float *d, float *d2, float *e, float *e2)
{
  /* main.cpp:537-566 with the single-domain faces: left face [0, DP), right face
   * [(n-1)H - DP, nH).  Outside both, d = e = 0 (zero-initialised arrays). */
  const float DP = NP * H;
  const float xl = DP;
  const float xr = (n - 1) * H - DP;
  for (int i = 0; i < n; i++) {
    d[i] = d2[i] = e[i] = e2[i] = 0.0f;
    const float x = i * H;
    if (x >= 0 && x < xl) {
      const float d0 = (float)(3.0 * v0 * log(1.0 / RC) / (2.0 * DP));
      d[i] = (float)(d0 * pow(((xl - x) / DP), 2));
      d2[i] = (float)(d0 * pow(((xl - x - 0.5 * H) / DP), 2));
      /* float argument: under C++ <math.h> the reference's exp() resolves to the float
       * overload, i.e. expf */
      e[i] = expf(-(d[i]) * DT);
      e2[i] = expf(-(d2[i]) * DT);
    }
    if (x >= xr && x < n * H) {
      const float d0 = (float)(3.0 * v0 * log(1.0 / RC) / (2.0 * DP));
      d[i] = (float)(d0 * pow(((x - xr) / DP), 2));
      d2[i] = (float)(d0 * pow(((x + 0.5 * H - xr) / DP), 2));
      /* float argument: under C++ <math.h> the reference's exp() resolves to the float
       * overload, i.e. expf */
      e[i] = expf(-(d[i]) * DT);
      e2[i] = expf(-(d2[i]) * DT);
    }
  }
}

float or_ricker(int it, float DT, float F0)
{
  /* kernel.cu:1730 (T0) and 1897-1898: PI is a double literal. */
  const double PI_ = 3.1415926535;
  const float T0 = 1.2 / F0;
  const float tt = it * DT;
  return (float)(-(1 - 2 * PI_ * PI_ * F0 * F0 * (tt - T0) * (tt - T0)) *
                 exp(-PI_ * PI_ * F0 * F0 * (tt - T0) * (tt - T0)));
}

/* ------------------------------------------------------------------------------------ */
/* State                                                                                 */
/* ------------------------------------------------------------------------------------ */
struct or_state {
  int nx, ny, nz, order, mode;
  float DT, Hinv;
  size_t esz;
  void *f[OR_NFIELD];
  void *qprev[3];
  void *psi[24];
  void *dprev[24];
  void *coef[OR_NCOEF];
  void *pd[3], *pd2[3], *pe[3], *pe2[3];
};

#define CAT2_(a, b) a##b
#define CAT2(a, b) CAT2_(a, b)

/* mode F32 */
#define SFX f32
#define S float
#define A float
#define LD(x) ((float)(x))
#define STD(x) ((float)(x))
#define STA(x) ((float)(x))
#include "biot_oracle_core.inc"
#undef SFX
#undef S
#undef A
#undef LD
#undef STD
#undef STA

/* mode F16 (storage binary16, temporaries float) */
#define SFX f16
#define S uint16_t
#define A float
#define LD(x) or_f16_to_f32(x)
#define STD(x) or_f64_to_f16((double)(x))
#define STA(x) or_f32_to_f16((float)(x))
#include "biot_oracle_core.inc"
#undef SFX
#undef S
#undef A
#undef LD
#undef STD
#undef STA

/* mode F64 */
#define SFX f64
#define S double
#define A double
#define LD(x) ((double)(x))
#define STD(x) ((double)(x))
#define STA(x) ((double)(x))
#include "biot_oracle_core.inc"
#undef SFX
#undef S
#undef A
#undef LD
#undef STD
#undef STA

or_state *or_state_create(const or_config *cfg)
{
  if (cfg->order != 2 && cfg->order != 4) return NULL;
  if (cfg->mode < 0 || cfg->mode > 2) return NULL;
  or_state *st = (or_state *)calloc(1, sizeof(or_state));
  st->nx = cfg->nx; st->ny = cfg->ny; st->nz = cfg->nz;
  st->order = cfg->order; st->mode = cfg->mode;
  st->DT = cfg->DT; st->Hinv = cfg->Hinv;
  st->esz = cfg->mode == OR_MODE_F32 ? 4 : cfg->mode == OR_MODE_F16 ? 2 : 8;
  const size_t n = (size_t)cfg->nx * cfg->ny * cfg->nz;
  for (int a = 0; a < OR_NFIELD; a++) st->f[a] = calloc(n, st->esz);
  for (int a = 0; a < 3; a++) st->qprev[a] = calloc(n, st->esz);
  for (int a = 0; a < 24; a++) { st->psi[a] = calloc(n, st->esz); st->dprev[a] = calloc(n, st->esz); }
  for (int a = 0; a < OR_NCOEF; a++) st->coef[a] = calloc(n, st->esz);
  const int dims[3] = {cfg->nx, cfg->ny, cfg->nz};
  for (int a = 0; a < 3; a++) {
    st->pd[a] = calloc(dims[a], st->esz); st->pd2[a] = calloc(dims[a], st->esz);
    st->pe[a] = calloc(dims[a], st->esz); st->pe2[a] = calloc(dims[a], st->esz);
  }
  if (cfg->mode == OR_MODE_F32) pack_f32(st, cfg);
  else if (cfg->mode == OR_MODE_F16) pack_f16(st, cfg);
  else pack_f64(st, cfg);
  return st;
}

void or_state_destroy(or_state *st)
{
  if (!st) return;
  for (int a = 0; a < OR_NFIELD; a++) free(st->f[a]);
  for (int a = 0; a < 3; a++) free(st->qprev[a]);
  for (int a = 0; a < 24; a++) { free(st->psi[a]); free(st->dprev[a]); }
  for (int a = 0; a < OR_NCOEF; a++) free(st->coef[a]);
  for (int a = 0; a < 3; a++) { free(st->pd[a]); free(st->pd2[a]); free(st->pe[a]); free(st->pe2[a]); }
  free(st);
}

void or_state_zero(or_state *st)
{
  const size_t n = (size_t)st->nx * st->ny * st->nz;
  for (int a = 0; a < OR_NFIELD; a++) memset(st->f[a], 0, n * st->esz);
  for (int a = 0; a < 3; a++) memset(st->qprev[a], 0, n * st->esz);
  for (int a = 0; a < 24; a++) { memset(st->psi[a], 0, n * st->esz); memset(st->dprev[a], 0, n * st->esz); }
}

static double get_any(const or_state *st, const void *a, int64_t c)
{
  return st->mode == OR_MODE_F32 ? get_f32(a, c) : st->mode == OR_MODE_F16 ? get_f16(a, c) : get_f64(a, c);
}

void or_state_get_field(const or_state *st, int field, float *out)
{
  const int64_t n = (int64_t)st->nx * st->ny * st->nz;
  for (int64_t c = 0; c < n; c++) out[c] = (float)get_any(st, st->f[field], c);
}

void or_state_get_field_f64(const or_state *st, int field, double *out)
{
  const int64_t n = (int64_t)st->nx * st->ny * st->nz;
  for (int64_t c = 0; c < n; c++) out[c] = get_any(st, st->f[field], c);
}

void or_state_set_field_f64(or_state *st, int field, const double *in)
{
  /* A relative fluid velocity also sets its previous-step copy: between steps the
   * reference's vwx2..vwz2 always equal vwx..vwz (kernel.cu:226). */
  const int64_t n = (int64_t)st->nx * st->ny * st->nz;
  void *dst[2] = {st->f[field], (field >= 3 && field <= 5) ? st->qprev[field - 3] : NULL};
  for (int d = 0; d < 2; d++) {
    if (!dst[d]) continue;
    for (int64_t c = 0; c < n; c++) {
      if (st->mode == OR_MODE_F32) set_f32(dst[d], c, in[c]);
      else if (st->mode == OR_MODE_F16) set_f16(dst[d], c, in[c]);
      else set_f64(dst[d], c, in[c]);
    }
  }
}

void or_source(or_state *st, int64_t sn, unsigned mask, float amp)
{
  if (st->mode == OR_MODE_F32) source_f32(st, sn, mask, amp);
  else if (st->mode == OR_MODE_F16) source_f16(st, sn, mask, amp);
  else source_f64(st, sn, mask, amp);
}

void or_step_v(or_state *st)
{
  if (st->mode == OR_MODE_F32) step_v_f32(st);
  else if (st->mode == OR_MODE_F16) step_v_f16(st);
  else step_v_f64(st);
}

void or_step_t(or_state *st)
{
  if (st->mode == OR_MODE_F32) step_t_f32(st);
  else if (st->mode == OR_MODE_F16) step_t_f16(st);
  else step_t_f64(st);
}

static int run_generic(or_state *st, int it0, int nsteps, int nt_total, const float *wavelet,
                       int64_t src, unsigned src_mask, int nrec, const int64_t *rec,
                       unsigned rec_mask, float *tr32, double *tr64)
{
  int comps[OR_NFIELD], nc = 0;
  for (int f = 0; f < OR_NFIELD; f++)
    if (rec_mask & (1u << f)) comps[nc++] = f;
  for (int it = it0; it < it0 + nsteps; it++) {
    /* kernel.cu:2064-2071 then the V and T phases, then the per-step recording */
    if (src >= 0 && src_mask) or_source(st, src, src_mask, wavelet[it]);
    or_step_v(st);
    or_step_t(st);
    for (int ci = 0; ci < nc; ci++) {
      const void *a = st->f[comps[ci]];
      for (int r = 0; r < nrec; r++) {
        const double v = get_any(st, a, rec[r]);
        const size_t o = ((size_t)ci * nt_total + it) * nrec + r;
        if (tr32) tr32[o] = (float)v;
        if (tr64) tr64[o] = v;
      }
    }
  }
  return 0;
}

int or_run(or_state *st, int it0, int nsteps, int nt_total, const float *wavelet,
           int64_t src, unsigned src_mask, int nrec, const int64_t *rec, unsigned rec_mask,
           float *traces)
{
  return run_generic(st, it0, nsteps, nt_total, wavelet, src, src_mask, nrec, rec, rec_mask,
                     traces, NULL);
}

int or_run_f64(or_state *st, int it0, int nsteps, int nt_total, const float *wavelet,
               int64_t src, unsigned src_mask, int nrec, const int64_t *rec,
               unsigned rec_mask, double *traces)
{
  return run_generic(st, it0, nsteps, nt_total, wavelet, src, src_mask, nrec, rec, rec_mask,
                     NULL, traces);
}

int or_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
