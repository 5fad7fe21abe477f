// Host-side set-up of the Biot solver (see biot_setup.hpp).
#include "biot_setup.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

namespace biot {

void parallel_for(int64_t n, void (*body)(int64_t, int64_t, void *), void *ctx)
{
  unsigned hw = std::thread::hardware_concurrency();
  int64_t nth = std::max<int64_t>(1, std::min<int64_t>(hw ? hw : 1, 64));
  if (n < (1 << 16)) nth = 1;
  if (nth == 1) {
    body(0, n, ctx);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t chunk = (n + nth - 1) / nth;
  for (int64_t t = 0; t < nth; ++t) {
    const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    pool.emplace_back(body, lo, hi, ctx);
  }
  for (auto &th : pool) th.join();
}

// ------------------------------------------------------------------------------------
// Biot moduli: outputfunc.cpp:190-217.  The reference keeps rho_s, rho_f, mu and the
// results in float arrays and the moduli/porosity/permeability/viscosity in doubles.
// ------------------------------------------------------------------------------------
namespace {
struct ModCtx {
  const biot_medium *m;
  Moduli *o;
};
void moduli_body(int64_t lo, int64_t hi, void *p)
{
  auto *cx = static_cast<ModCtx *>(p);
  const biot_medium &m = *cx->m;
  Moduli &o = *cx->o;
  const double tortuosity = (1 + 2 / m.porousm) * m.tao;  // outputfunc.cpp:214 prefactor
  for (int64_t c = lo; c < hi; ++c) {
    const float rhos = (float)m.rhos[c];
    const float rhof = (float)m.rhof[c];
    const double phi = m.phi[c];
    o.rhof[c] = rhof;
    o.mu[c] = (float)m.mu[c];
    if (m.mu[c] == 0.0 && m.Kb[c] == 0.0) {
      // fluid cell (outputfunc.cpp:176-193): every modulus is Kf, no Darcy damping
      const float kf = (float)m.Kf[c];
      o.M[c] = o.HH[c] = o.H2u[c] = o.C[c] = kf;
      o.C1[c] = 0.0f;
      o.C2[c] = rhof;
      o.rho[c] = rhof;
      continue;
    }
    o.rho[c] = (float)((1 - phi) * rhos + phi * rhof);
    const double alpha = 1 - m.Kb[c] / m.Ks[c];
    const float M = (float)(m.Kf[c] * m.Ks[c] / (phi * m.Ks[c] + (alpha - phi) * m.Kf[c]));
    const float HH = (float)(alpha * alpha * M + m.Kb[c] + o.mu[c] * 4.0 / 3.0);
    o.M[c] = M;
    o.HH[c] = HH;
    o.H2u[c] = HH - 2 * o.mu[c];
    o.C[c] = (float)(M * alpha);
    o.C1[c] = (float)(m.eta[c] / m.perm[c]);
    // tortuosity prefactor (1 + 2/porousm)*tao is evaluated first, as in the reference
    o.C2[c] = (float)(tortuosity * rhof / phi);
  }
}
}  // namespace

void build_moduli(int64_t n, const biot_medium &m, Moduli &o)
{
  for (auto *v : {&o.M, &o.HH, &o.H2u, &o.C, &o.C1, &o.C2, &o.rho, &o.rhof, &o.mu})
    v->assign((size_t)n, 0.0f);
  ModCtx cx{&m, &o};
  parallel_for(n, moduli_body, &cx);
}

// ------------------------------------------------------------------------------------
// Staggered averages: main.cpp:322-358.  Harmonic 4-cell mean of mu on each shear plane
// (zero when any of the four is zero), arithmetic 2-cell means of rho, rho_f, C1, C2.
// ------------------------------------------------------------------------------------
namespace {
struct AvgCtx {
  int nx, ny, nz;
  const Moduli *md;
  Averages *av;
};
inline float harmonic4(float a, float b, float c, float d)
{
  if (a == 0 || b == 0 || c == 0 || d == 0) return 0.0f;
  return (float)(1 / (0.25 * (1 / a + 1 / b + 1 / c + 1 / d)));
}
inline float mean2(float a, float b) { return (float)(0.5 * (a + b)); }
void avg_body(int64_t lo, int64_t hi, void *p)
{
  auto *cx = static_cast<AvgCtx *>(p);
  const int nx = cx->nx, ny = cx->ny;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const Moduli &md = *cx->md;
  Averages &av = *cx->av;
  const int64_t step[3] = {1, sy, sz};
  for (int64_t k = lo; k < hi; ++k) {
    if (k < 1 || k >= cx->nz - 1) continue;
    for (int j = 1; j < ny - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) {
        const int64_t c = i + j * sy + k * sz;
        const float *mu = md.mu.data();
        av.muxz[c] = harmonic4(mu[c], mu[c + 1], mu[c + sz], mu[c + sz + 1]);
        av.muyz[c] = harmonic4(mu[c], mu[c + sy], mu[c + sz], mu[c + sz + sy]);
        av.muxy[c] = harmonic4(mu[c], mu[c + 1], mu[c + sy], mu[c + sy + 1]);
        for (int d = 0; d < 3; ++d) {
          const int64_t s = step[d];
          av.rhof[d][c] = mean2(md.rhof[c], md.rhof[c + s]);
          av.rho[d][c] = mean2(md.rho[c], md.rho[c + s]);
          av.C1[d][c] = mean2(md.C1[c], md.C1[c + s]);
          av.C2[d][c] = mean2(md.C2[c], md.C2[c + s]);
        }
      }
  }
}
}  // namespace

void build_averages(int nx, int ny, int nz, const Moduli &md, Averages &av)
{
  const size_t n = (size_t)nx * ny * nz;
  av.muxz.assign(n, 0.0f);
  av.muxy.assign(n, 0.0f);
  av.muyz.assign(n, 0.0f);
  for (int d = 0; d < 3; ++d) {
    av.rho[d].assign(n, 0.0f);
    av.rhof[d].assign(n, 0.0f);
    av.C1[d].assign(n, 0.0f);
    av.C2[d].assign(n, 0.0f);
  }
  AvgCtx cx{nx, ny, nz, &md, &av};
  parallel_for(nz, avg_body, &cx);
}

// ------------------------------------------------------------------------------------
// Update coefficients: main.cpp:431-494 with the scale factors of main.cpp:248-262.
// ------------------------------------------------------------------------------------
namespace {
struct CoefCtx {
  int nx, ny, nz;
  float DT, Cv, Cp;
  const Moduli *md;
  const Averages *av;
  std::vector<float> *coef;
};
void coef_body(int64_t lo, int64_t hi, void *p)
{
  auto *cx = static_cast<CoefCtx *>(p);
  const float DT = cx->DT, Cv = cx->Cv, Cp = cx->Cp;
  const float Cw1 = 1.0f, Cw2 = 1.0f, Cw3 = 1.0f, Cu1 = 1.0f, Cu2 = 1.0f;  // Cvwp*, Cvup*
  const float Cpp = 100, Csp = 100;                                        // Cpp*, Csp*
  const Averages &av = *cx->av;
  const Moduli &md = *cx->md;
  std::vector<float> *K = cx->coef;
  const int nx = cx->nx, ny = cx->ny, nz = cx->nz;
  const int64_t sz = (int64_t)nx * ny;
  for (int64_t c = lo; c < hi; ++c) {
    const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / sz);
    const bool interior = i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1;
    if (av.C1[0][c] == 0 || av.C1[1][c] == 0 || av.C1[2][c] == 0) {
      for (int a = BIOT_W1X; a <= BIOT_W3Z; ++a) K[a][c] = 0;
    } else {
      for (int d = 0; d < 3; ++d) {
        const float rf = av.rhof[d][c], r = av.rho[d][c], c1 = av.C1[d][c], c2 = av.C2[d][c];
        // (C2 - rho_f^2/rho) is float, the 0.5*DT*C1 term brings in double
        K[BIOT_W1X + d][c] =
            (float)(Cw1 * (c2 - (rf * rf / r) - 0.5 * DT * c1) / (c2 - (rf * rf / r) + 0.5 * DT * c1));
        K[BIOT_W2X + d][c] =
            (float)(Cv * DT * Cw2 * (rf / r) / (Cp * (c2 - (rf * rf / r) + 0.5 * DT * c1)));
        K[BIOT_W3X + d][c] = (float)(Cv * DT * Cw3 / (Cp * (c2 - rf * rf / r + 0.5 * DT * c1)));
      }
    }
    if (av.rho[0][c] == 0 || av.rho[1][c] == 0 || av.rho[2][c] == 0) {
      for (int a = BIOT_U1X; a <= BIOT_U2Z; ++a) K[a][c] = 0;
    } else {
      for (int d = 0; d < 3; ++d) {
        K[BIOT_U1X + d][c] = Cv * DT * Cu1 / (av.rho[d][c] * Cp);
        K[BIOT_U2X + d][c] = av.rhof[d][c] * Cu2 / (av.rho[d][c]);
      }
    }
    const float M = interior ? md.M[c] : 0.0f, C = interior ? md.C[c] : 0.0f;
    const float HH = interior ? md.HH[c] : 0.0f, H2u = interior ? md.H2u[c] : 0.0f;
    K[BIOT_PP1][c] = C * DT * Cp * Cpp / Cv;
    K[BIOT_PP2][c] = M * DT * Cp * Cpp / Cv;
    K[BIOT_SP1][c] = Cp * DT * H2u * Csp / Cv;
    K[BIOT_SP2][c] = Cp * DT * HH * Csp / Cv;
    K[BIOT_SP3][c] = Cp * DT * C * Csp / Cv;
    K[BIOT_SPXY][c] = Cp * DT * av.muxy[c] * Csp / Cv;
    K[BIOT_SPXZ][c] = Cp * DT * av.muxz[c] * Csp / Cv;
    K[BIOT_SPYZ][c] = Cp * DT * av.muyz[c] * Csp / Cv;
  }
}
}  // namespace

void build_coefficients(int nx, int ny, int nz, float DT, float Cv, float Cp, const Moduli &md,
                        const Averages &av, std::vector<float> (&coef)[BIOT_NCOEF])
{
  const int64_t n = (int64_t)nx * ny * nz;
  for (auto &v : coef) v.assign((size_t)n, 0.0f);
  CoefCtx cx{nx, ny, nz, DT, Cv, Cp, &md, &av, coef};
  parallel_for(n, coef_body, &cx);
}

// ------------------------------------------------------------------------------------
// C-PML profile of one axis (main.cpp:537-566; the 25 other per-rank cases repeat the
// same two face rules).  Quadratic d(x) = d0 ((distance into the layer)/L)^2 with
// d0 = 3 v0 ln(1/R) / (2 L), sampled at x = i*H and at x + H/2; e = exp(-d DT).
// ------------------------------------------------------------------------------------
void build_pml_profile(int n, float H, float DT, int NP, float v0, float RC, float *d, float *d2,
                       float *e, float *e2)
{
  const float L = NP * H;                 // DP
  const float low_edge = L;               // xoleft
  const float high_edge = (n - 1) * H - L;  // xoright
  const float d0 = (float)(3.0 * v0 * std::log(1.0 / RC) / (2.0 * L));
  for (int i = 0; i < n; ++i) {
    const float x = i * H;
    float di = 0.0f, dh = 0.0f, ei = 0.0f, eh = 0.0f;
    bool in = false;
    if (x >= 0 && x < low_edge) {
      di = (float)(d0 * std::pow((low_edge - x) / L, 2));
      dh = (float)(d0 * std::pow((low_edge - x - 0.5 * H) / L, 2));
      in = true;
    }
    if (x >= high_edge && x < n * H) {
      di = (float)(d0 * std::pow((x - high_edge) / L, 2));
      dh = (float)(d0 * std::pow((x + 0.5 * H - high_edge) / L, 2));
      in = true;
    }
    if (in) {
      ei = (float)std::exp(-(di)*DT);
      eh = (float)std::exp(-(dh)*DT);
    }
    d[i] = di;
    d2[i] = dh;
    e[i] = ei;
    e2[i] = eh;
  }
}

float ricker(int it, float DT, float F0)
{
  const double pi = 3.1415926535;  // the reference's PI literal
  const float T0 = 1.2 / F0;
  const float t = it * DT;
  const double a = pi * pi * F0 * F0 * (t - T0) * (t - T0);
  return (float)(-(1 - 2 * a) * std::exp(-a));
}

}  // namespace biot
