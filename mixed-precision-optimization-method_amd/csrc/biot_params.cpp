// Inversion parameterization of the Biot medium and its chain rule (see biot_params.hpp).
#include "biot_params.hpp"

#include <cmath>

namespace biot {

namespace {

struct Cell {
  int nx, ny, nz;
  int64_t sy, sz;
  bool interior(int64_t c) const
  {
    const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / sz);
    return i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1;
  }
};

inline double harmonic4(double a, double b, double c, double d)
{
  if (a == 0 || b == 0 || c == 0 || d == 0) return 0.0;
  return 1.0 / (0.25 * (1 / a + 1 / b + 1 / c + 1 / d));
}

// the V-side quantities of one staggering direction at one cell
struct VDir {
  double A, B, c1, c2, mp, D, N;
};
inline VDir vdir(const ParamArrays &p, int64_t c, int64_t s, double dt)
{
  VDir v;
  v.A = 0.5 * (p.rho[c] + p.rho[c + s]);
  v.B = 0.5 * (p.rhof[c] + p.rhof[c + s]);
  v.c1 = 0.5 * (p.C1[c] + p.C1[c + s]);
  v.c2 = 0.5 * (p.tphi[c] * p.rhof[c] + p.tphi[c + s] * p.rhof[c + s]);
  v.mp = v.c2 - v.B * v.B / v.A;
  v.D = v.mp + 0.5 * dt * v.c1;
  v.N = v.mp - 0.5 * dt * v.c1;
  return v;
}

}  // namespace

void kernel_coefs_from_params(int nx, int ny, int nz, const CoefScales &s, const ParamArrays &p,
                              std::vector<double> (&K)[NKC])
{
  const int64_t n = (int64_t)nx * ny * nz, sy = nx, sz = (int64_t)nx * ny;
  const Cell cl{nx, ny, nz, sy, sz};
  for (auto &v : K) v.assign((size_t)n, 0.0);
  const double k = s.Cv * s.dt / s.Cp;  // velocity coefficient scale
  const double kk = s.dt * s.Cp / s.Cv; // stress coefficient scale (x100 x0.01 cancel)
  const int64_t step[3] = {1, sy, sz};
  for (int64_t c = 0; c < n; ++c) {
    if (!cl.interior(c)) continue;  // averages and moduli are zero on the outer planes
    VDir d[3];
    bool c1zero = false, azero = false;
    for (int q = 0; q < 3; ++q) {
      d[q] = vdir(p, c, step[q], s.dt);
      c1zero |= d[q].c1 == 0;
      azero |= d[q].A == 0;
    }
    for (int q = 0; q < 3; ++q) {
      if (!c1zero) {
        K[K_W1X + q][c] = d[q].N / d[q].D;
        K[K_W2X + q][c] = k * (d[q].B / d[q].A) / d[q].D * s.Hinv;
        if (q < 2) K[K_W3X + q][c] = k / d[q].D * s.Hinv;  // W3z is never read
      }
      if (!azero) {
        K[K_U1X + q][c] = k / d[q].A * s.Hinv;
        K[K_U2X + q][c] = d[q].B / d[q].A;
      }
    }
    const double C = p.alpha[c] * p.M[c];
    const double lamc = p.lambda[c] + p.alpha[c] * p.alpha[c] * p.M[c];
    const double Hm = lamc + 2 * p.mu[c];
    K[K_P1][c] = C * kk * s.Hinv;
    K[K_P2][c] = p.M[c] * kk * s.Hinv;
    K[K_S1][c] = lamc * kk * s.Hinv;
    K[K_S2][c] = Hm * kk * s.Hinv;
    K[K_S3][c] = C * kk * s.Hinv;
    const double *mu = p.mu;
    K[K_SXY][c] = harmonic4(mu[c], mu[c + 1], mu[c + sy], mu[c + sy + 1]) * kk * s.Hinv;
    K[K_SXZ][c] = harmonic4(mu[c], mu[c + 1], mu[c + sz], mu[c + sz + 1]) * kk * s.Hinv;
    K[K_SYZ][c] = harmonic4(mu[c], mu[c + sy], mu[c + sz], mu[c + sz + sy]) * kk * s.Hinv;
  }
}

void chain_params(int nx, int ny, int nz, const CoefScales &s, const ParamArrays &p,
                  const double *const gK[NKC], double *const out[6])
{
  const int64_t n = (int64_t)nx * ny * nz, sy = nx, sz = (int64_t)nx * ny;
  const Cell cl{nx, ny, nz, sy, sz};
  double *gl = out[0], *gmu = out[1], *ga = out[2], *gM = out[3], *grho = out[4], *grf = out[5];
  const double k = s.Cv * s.dt / s.Cp;
  const double kk = s.dt * s.Cp / s.Cv;
  const int64_t step[3] = {1, sy, sz};
  for (int64_t c = 0; c < n; ++c) {
    if (!cl.interior(c)) continue;
    // ---- stress side: K_P1 = C kk/dx, K_P2 = M kk/dx, K_S1 = lambda_c .., K_S2 = H .., K_S3 = C ..
    const double al = p.alpha[c], M = p.M[c];
    const double gC = (gK[K_P1][c] + gK[K_S3][c]) * kk * s.Hinv;
    const double glc = gK[K_S1][c] * kk * s.Hinv;
    const double gH = gK[K_S2][c] * kk * s.Hinv;
    const double gMd = gK[K_P2][c] * kk * s.Hinv;
    gl[c] += glc + gH;
    gmu[c] += 2 * gH;
    ga[c] += gC * M + (glc + gH) * 2 * al * M;
    gM[c] += gMd + gC * al + (glc + gH) * al * al;
    // shear: 4-cell harmonic means, d mu_bar / d mu_i = mu_bar^2 / (4 mu_i^2)
    const int64_t quads[3][4] = {{c, c + 1, c + sy, c + sy + 1},
                                 {c, c + 1, c + sz, c + sz + 1},
                                 {c, c + sy, c + sz, c + sz + sy}};
    const int kidx[3] = {K_SXY, K_SXZ, K_SYZ};
    for (int q = 0; q < 3; ++q) {
      const int64_t *Q = quads[q];
      const double hm = harmonic4(p.mu[Q[0]], p.mu[Q[1]], p.mu[Q[2]], p.mu[Q[3]]);
      if (hm == 0) continue;
      const double g = gK[kidx[q]][c] * kk * s.Hinv;
      for (int m = 0; m < 4; ++m) gmu[Q[m]] += g * hm * hm / (4 * p.mu[Q[m]] * p.mu[Q[m]]);
    }
    // ---- velocity side, per staggering direction
    VDir d[3];
    bool c1zero = false, azero = false;
    for (int q = 0; q < 3; ++q) {
      d[q] = vdir(p, c, step[q], s.dt);
      c1zero |= d[q].c1 == 0;
      azero |= d[q].A == 0;
    }
    for (int q = 0; q < 3; ++q) {
      const VDir &v = d[q];
      double gA = 0, gB = 0, gc2 = 0;
      if (!c1zero) {
        const double gW1 = gK[K_W1X + q][c], gW2 = gK[K_W2X + q][c];
        // W3y multiplies the pressure term of both qy and qz (the reference's quirk), so its
        // gradient already holds both; W3z is never read and has none
        const double gW3 = q < 2 ? gK[K_W3X + q][c] : 0.0;
        const double D2 = v.D * v.D;
        const double dW1_dmp = s.dt * v.c1 / D2;
        const double dW3_dmp = -k * s.Hinv / D2;
        const double dW2_dmp = -k * s.Hinv * (v.B / v.A) / D2;
        const double gmp = gW1 * dW1_dmp + gW2 * dW2_dmp + gW3 * dW3_dmp;
        gA += gmp * v.B * v.B / (v.A * v.A) - gW2 * k * s.Hinv * v.B / (v.A * v.A * v.D);
        gB += gmp * (-2 * v.B / v.A) + gW2 * k * s.Hinv / (v.A * v.D);
        gc2 += gmp;
      }
      if (!azero) {
        const double gU1 = gK[K_U1X + q][c], gU2 = gK[K_U2X + q][c];
        gA += -gU1 * k * s.Hinv / (v.A * v.A) - gU2 * v.B / (v.A * v.A);
        gB += gU2 / v.A;
      }
      const int64_t c2 = c + step[q];
      grho[c] += 0.5 * gA;
      grho[c2] += 0.5 * gA;
      grf[c] += 0.5 * gB + 0.5 * gc2 * p.tphi[c];
      grf[c2] += 0.5 * gB + 0.5 * gc2 * p.tphi[c2];
    }
  }
}

}  // namespace biot
