/*
 * FP64 host adjoint of the Biot time step (TEST INFRASTRUCTURE ONLY: the checker of the
 * device adjoint; the reference is forward-only and has no counterpart).
 *
 * Forward step n (the device solver's order): source of step n, velocity half step, stress
 * half step, record.  Per interior cell, with the folded C-PML carry c of each derivative
 * (e, a the axis profile at the integer node for a backward difference, the half node for a
 * forward one):
 *   D' = (1 + a) D + c,   c <- e (c + 2 a D)
 *   V:  q' = W1 q - W2 S - W3 Ds,   v' = v + U1 S - U2 (q' - q)   (qz uses W3y)
 *   T:  p' = p - P1 div u' - P2 div w',  tau_ii' = tau_ii + S1 (..) + S2 u_ii' + S3 div w',
 *       tau_ij' = tau_ij + Sij (u_ij' + u_ji')
 * (kernel.cu:145-229 and 405-493 with the reference's psi/D_prev pair folded into one carry
 * and 1/dx and 0.01 folded into the coefficients K, as the device kernels use them).
 *
 * The adjoint runs the transposed step backward in time with checkpoints every `every`
 * steps and per-step states recomputed inside each segment.  The transposed staggered
 * differences are applied by scattering each derivative adjoint to the stencil points it
 * was formed from (the device gathers per cell instead), and dJ/dK is summed per coefficient
 * array in double.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "biot_oracle.h"

enum { FVX, FVY, FVZ, FQX, FQY, FQZ, FTXX, FTYY, FTZZ, FTXY, FTXZ, FTYZ, FP, NF };
enum {
  KU1X, KU1Y, KU1Z, KU2X, KU2Y, KU2Z, KW1X, KW1Y, KW1Z, KW2X, KW2Y, KW2Z, KW3X, KW3Y,
  KP1, KP2, KS1, KS2, KS3, KSXY, KSXZ, KSYZ, NK
};

typedef struct {
  int field, dir, fwd;
} hder;
/* velocity half step: x1 x2 x3 s1 y1 y2 y3 s2 z1 z2 z3 s3 (kernel.cu:172-185) */
static const hder VD[12] = {{FTXX, 0, 1}, {FTXY, 1, 0}, {FTXZ, 2, 0}, {FP, 0, 1},
                            {FTYY, 1, 1}, {FTXY, 0, 0}, {FTYZ, 2, 0}, {FP, 1, 1},
                            {FTZZ, 2, 1}, {FTXZ, 0, 0}, {FTYZ, 1, 0}, {FP, 2, 1}};
/* stress half step: uxx uyy uzz wx wy wz uxy uyx uxz uzx uyz uzy (kernel.cu:430-445) */
static const hder TD[12] = {{FVX, 0, 0}, {FVY, 1, 0}, {FVZ, 2, 0}, {FQX, 0, 0},
                            {FQY, 1, 0}, {FQZ, 2, 0}, {FVX, 1, 1}, {FVY, 0, 1},
                            {FVX, 2, 1}, {FVZ, 0, 1}, {FVY, 2, 1}, {FVZ, 1, 1}};

#define NSTATE (NF + 24) /* fields, 12 velocity-side carries, 12 stress-side carries */

typedef struct {
  const ha_problem *P;
  int64_t n, st[3];
  int h;
} hctx;

static double dif(const hctx *X, const double *f, int64_t c, int dir, int fwd)
{
  const int64_t s = X->st[dir];
  if (fwd) {
    if (X->P->order == 2) return f[c + s] - f[c];
    return 9.0 / 8.0 * (f[c + s] - f[c]) + (-1.0 / 24.0) * (f[c + 2 * s] - f[c - s]);
  }
  if (X->P->order == 2) return f[c] - f[c - s];
  return 9.0 / 8.0 * (f[c] - f[c - s]) + (-1.0 / 24.0) * (f[c + s] - f[c - 2 * s]);
}

/* transpose of dif: the adjoint G of one difference, scattered to its stencil points */
static void dif_t(const hctx *X, double *lf, int64_t c, int dir, int fwd, double G)
{
  const int64_t s = X->st[dir];
  if (fwd) {
    if (X->P->order == 2) {
      lf[c + s] += G;
      lf[c] -= G;
      return;
    }
    lf[c + s] += 9.0 / 8.0 * G;
    lf[c] -= 9.0 / 8.0 * G;
    lf[c + 2 * s] += (-1.0 / 24.0) * G;
    lf[c - s] -= (-1.0 / 24.0) * G;
    return;
  }
  if (X->P->order == 2) {
    lf[c] += G;
    lf[c - s] -= G;
    return;
  }
  lf[c] += 9.0 / 8.0 * G;
  lf[c - s] -= 9.0 / 8.0 * G;
  lf[c + s] += (-1.0 / 24.0) * G;
  lf[c - 2 * s] -= (-1.0 / 24.0) * G;
}

/* profile value e (q = 0) or a (q = 1) of derivative (dir, fwd) at cell (i, j, k) */
static double prof(const hctx *X, int dir, int fwd, int q, int i, int j, int k)
{
  const int pos = dir == 0 ? i : (dir == 1 ? j : k);
  return X->P->prof[dir][(fwd ? 2 : 0) + q][pos];
}

#define FOR_INTERIOR(X)                                                 \
  for (int k = (X)->h; k < (X)->P->nz - (X)->h; ++k)                    \
    for (int j = (X)->h; j < (X)->P->ny - (X)->h; ++j)                  \
      for (int i = (X)->h; i < (X)->P->nx - (X)->h; ++i)

static void fwd_v(const hctx *X, double **S)
{
  const double *const *K = X->P->K;
  FOR_INTERIOR(X)
  {
    const int64_t c = i + j * X->st[1] + k * X->st[2];
    double D[12];
    for (int m = 0; m < 12; ++m) {
      const hder d = VD[m];
      const double e = prof(X, d.dir, d.fwd, 0, i, j, k), a = prof(X, d.dir, d.fwd, 1, i, j, k);
      const double r = dif(X, S[d.field], c, d.dir, d.fwd);
      double *cm = &S[NF + m][c];
      const double psi = *cm + a * r;
      *cm = e * (psi + a * r);
      D[m] = r + psi;
    }
    const double Sv[3] = {D[0] + D[1] + D[2], D[4] + D[5] + D[6], D[8] + D[9] + D[10]};
    const double Ds[3] = {D[3], D[7], D[11]};
    const int kw3[3] = {KW3X, KW3Y, KW3Y};
    for (int d = 0; d < 3; ++d) {
      const double q0 = S[FQX + d][c];
      const double q1 = K[KW1X + d][c] * q0 - K[KW2X + d][c] * Sv[d] - K[kw3[d]][c] * Ds[d];
      S[FQX + d][c] = q1;
      S[FVX + d][c] += K[KU1X + d][c] * Sv[d] - K[KU2X + d][c] * (q1 - q0);
    }
  }
}

static void fwd_t(const hctx *X, double **S)
{
  const double *const *K = X->P->K;
  FOR_INTERIOR(X)
  {
    const int64_t c = i + j * X->st[1] + k * X->st[2];
    double E[12];
    for (int m = 0; m < 12; ++m) {
      const hder d = TD[m];
      const double e = prof(X, d.dir, d.fwd, 0, i, j, k), a = prof(X, d.dir, d.fwd, 1, i, j, k);
      const double r = dif(X, S[d.field], c, d.dir, d.fwd);
      double *cm = &S[NF + 12 + m][c];
      const double psi = *cm + a * r;
      *cm = e * (psi + a * r);
      E[m] = r + psi;
    }
    const double divu = E[0] + E[1] + E[2], divw = E[3] + E[4] + E[5];
    S[FP][c] += -K[KP1][c] * divu - K[KP2][c] * divw;
    S[FTXX][c] += K[KS1][c] * (E[1] + E[2]) + K[KS2][c] * E[0] + K[KS3][c] * divw;
    S[FTYY][c] += K[KS1][c] * (E[0] + E[2]) + K[KS2][c] * E[1] + K[KS3][c] * divw;
    S[FTZZ][c] += K[KS1][c] * (E[0] + E[1]) + K[KS2][c] * E[2] + K[KS3][c] * divw;
    S[FTXY][c] += K[KSXY][c] * (E[6] + E[7]);
    S[FTXZ][c] += K[KSXZ][c] * (E[8] + E[9]);
    S[FTYZ][c] += K[KSYZ][c] * (E[10] + E[11]);
  }
}

static void source(const hctx *X, double **S, int it)
{
  const ha_problem *P = X->P;
  if (!P->wavelet || it >= P->nt_w) return;
  for (int f = 0; f < NF; ++f)
    if (P->src_mask & (1u << f)) S[f][P->src_cell] += (double)P->wavelet[it];
}

static void record(const hctx *X, double **S, int it, int nt, double *traces)
{
  const ha_problem *P = X->P;
  if (!traces) return;
  for (int ci = 0; ci < P->ncomp; ++ci)
    for (int64_t r = 0; r < P->nrec; ++r)
      traces[((int64_t)ci * nt + it) * P->nrec + r] = S[P->comp[ci]][P->rec[r]];
}

/* transposed stress half step at the forward state F (fields after the velocity half step,
 * stress-side carries before the stress half step); L the adjoint state */
static void adj_t(const hctx *X, double **F, double **L, double *const gK[NK])
{
  const double *const *K = X->P->K;
  FOR_INTERIOR(X)
  {
    const int64_t c = i + j * X->st[1] + k * X->st[2];
    const double lp = L[FP][c], lxx = L[FTXX][c], lyy = L[FTYY][c], lzz = L[FTZZ][c];
    const double lxy = L[FTXY][c], lxz = L[FTXZ][c], lyz = L[FTYZ][c];
    const double S1 = K[KS1][c], S2 = K[KS2][c], S3 = K[KS3][c], P1 = K[KP1][c], P2 = K[KP2][c];
    double lE[12];
    lE[0] = S2 * lxx + S1 * (lyy + lzz) - P1 * lp;
    lE[1] = S2 * lyy + S1 * (lxx + lzz) - P1 * lp;
    lE[2] = S2 * lzz + S1 * (lxx + lyy) - P1 * lp;
    lE[3] = lE[4] = lE[5] = S3 * (lxx + lyy + lzz) - P2 * lp;
    lE[6] = lE[7] = K[KSXY][c] * lxy;
    lE[8] = lE[9] = K[KSXZ][c] * lxz;
    lE[10] = lE[11] = K[KSYZ][c] * lyz;
    double E[12];
    for (int m = 0; m < 12; ++m) {
      const hder d = TD[m];
      const double e = prof(X, d.dir, d.fwd, 0, i, j, k), a = prof(X, d.dir, d.fwd, 1, i, j, k);
      const double r = dif(X, F[d.field], c, d.dir, d.fwd);
      E[m] = (1.0 + a) * r + F[NF + 12 + m][c];
      double *lc = &L[NF + 12 + m][c];
      const double G = (1.0 + a) * lE[m] + 2.0 * a * e * *lc;
      *lc = lE[m] + e * *lc;
      dif_t(X, L[d.field], c, d.dir, d.fwd, G);
    }
    const double divu = E[0] + E[1] + E[2], divw = E[3] + E[4] + E[5];
    gK[KP1][c] += -lp * divu;
    gK[KP2][c] += -lp * divw;
    gK[KS1][c] += lxx * (E[1] + E[2]) + lyy * (E[0] + E[2]) + lzz * (E[0] + E[1]);
    gK[KS2][c] += lxx * E[0] + lyy * E[1] + lzz * E[2];
    gK[KS3][c] += (lxx + lyy + lzz) * divw;
    gK[KSXY][c] += lxy * (E[6] + E[7]);
    gK[KSXZ][c] += lxz * (E[8] + E[9]);
    gK[KSYZ][c] += lyz * (E[10] + E[11]);
  }
}

/* transposed velocity half step at the forward state F (before the velocity half step) */
static void adj_v(const hctx *X, double **F, double **L, double *const gK[NK])
{
  const double *const *K = X->P->K;
  const int kw3[3] = {KW3X, KW3Y, KW3Y};
  FOR_INTERIOR(X)
  {
    const int64_t c = i + j * X->st[1] + k * X->st[2];
    double D[12], a_[12], e_[12], r_[12];
    for (int m = 0; m < 12; ++m) {
      const hder d = VD[m];
      e_[m] = prof(X, d.dir, d.fwd, 0, i, j, k);
      a_[m] = prof(X, d.dir, d.fwd, 1, i, j, k);
      r_[m] = dif(X, F[d.field], c, d.dir, d.fwd);
      D[m] = (1.0 + a_[m]) * r_[m] + F[NF + m][c];
    }
    const double Sv[3] = {D[0] + D[1] + D[2], D[4] + D[5] + D[6], D[8] + D[9] + D[10]};
    const double Ds[3] = {D[3], D[7], D[11]};
    double lS[3], lDs[3];
    for (int d = 0; d < 3; ++d) {
      const double q0 = F[FQX + d][c];
      const double q1 = K[KW1X + d][c] * q0 - K[KW2X + d][c] * Sv[d] - K[kw3[d]][c] * Ds[d];
      const double lv = L[FVX + d][c], lq = L[FQX + d][c];
      const double lqe = lq - K[KU2X + d][c] * lv;
      L[FQX + d][c] = K[KU2X + d][c] * lv + K[KW1X + d][c] * lqe;
      lS[d] = K[KU1X + d][c] * lv - K[KW2X + d][c] * lqe;
      lDs[d] = -K[kw3[d]][c] * lqe;
      gK[KU1X + d][c] += lv * Sv[d];
      gK[KU2X + d][c] += -lv * (q1 - q0);
      gK[KW1X + d][c] += lqe * q0;
      gK[KW2X + d][c] += -lqe * Sv[d];
      gK[kw3[d]][c] += -lqe * Ds[d];
    }
    const double lD[12] = {lS[0], lS[0], lS[0], lDs[0], lS[1], lS[1], lS[1], lDs[1],
                           lS[2], lS[2], lS[2], lDs[2]};
    for (int m = 0; m < 12; ++m) {
      const hder d = VD[m];
      double *lc = &L[NF + m][c];
      const double G = (1.0 + a_[m]) * lD[m] + 2.0 * a_[m] * e_[m] * *lc;
      *lc = lD[m] + e_[m] * *lc;
      dif_t(X, L[d.field], c, d.dir, d.fwd, G);
    }
  }
}

static int ctx_init(hctx *X, const ha_problem *P)
{
  if (!P || (P->order != 2 && P->order != 4) || P->nx < 2 * P->order || P->ny < 2 * P->order ||
      P->nz < 2 * P->order)
    return -1;
  X->P = P;
  X->h = P->order / 2;
  X->n = (int64_t)P->nx * P->ny * P->nz;
  X->st[0] = 1;
  X->st[1] = P->nx;
  X->st[2] = (int64_t)P->nx * P->ny;
  return 0;
}

static double *state_alloc(const hctx *X, double **S)
{
  double *blk = (double *)calloc((size_t)NSTATE * X->n, sizeof(double));
  if (!blk) return NULL;
  for (int a = 0; a < NSTATE; ++a) S[a] = blk + (int64_t)a * X->n;
  return blk;
}

int ha_forward(const ha_problem *P, int nt, double *traces)
{
  hctx X;
  if (ctx_init(&X, P) || nt < 0) return -1;
  double *S[NSTATE];
  double *blk = state_alloc(&X, S);
  if (!blk) return -2;
  for (int it = 0; it < nt; ++it) {
    source(&X, S, it);
    fwd_v(&X, S);
    fwd_t(&X, S);
    record(&X, S, it, nt, traces);
  }
  free(blk);
  return 0;
}

int ha_adjoint(const ha_problem *P, int nt, int every, const double *adj, double *const gK[22],
               double *traces)
{
  hctx X;
  if (ctx_init(&X, P) || nt <= 0 || every <= 0 || !adj || !gK) return -1;
  for (int q = 0; q < NK; ++q)
    if (!gK[q]) return -1;
  const int nseg = (nt + every - 1) / every;
  const size_t sbytes = (size_t)NSTATE * X.n * sizeof(double);
  double *S[NSTATE], *F[NSTATE], *L[NSTATE];
  double *blk = state_alloc(&X, S), *fblk = state_alloc(&X, F), *lblk = state_alloc(&X, L);
  double *ck = (double *)malloc(sbytes * (size_t)nseg);
  double *seg = (double *)malloc(sbytes * (size_t)every);
  int rc = 0;
  if (!blk || !fblk || !lblk || !ck || !seg) {
    rc = -2;
    goto done;
  }
  for (int q = 0; q < NK; ++q) memset(gK[q], 0, (size_t)X.n * sizeof(double));
  /* forward with a checkpoint at the start of every segment */
  for (int it = 0; it < nt; ++it) {
    if (it % every == 0) memcpy(ck + (size_t)(it / every) * NSTATE * X.n, blk, sbytes);
    source(&X, S, it);
    fwd_v(&X, S);
    fwd_t(&X, S);
    record(&X, S, it, nt, traces);
  }
  /* backward: recompute each segment keeping the state after every step's source */
  for (int s = nseg - 1; s >= 0; --s) {
    const int n0 = s * every, n1 = n0 + every < nt ? n0 + every : nt;
    memcpy(blk, ck + (size_t)s * NSTATE * X.n, sbytes);
    for (int it = n0; it < n1; ++it) {
      source(&X, S, it);
      memcpy(seg + (size_t)(it - n0) * NSTATE * X.n, blk, sbytes);
      fwd_v(&X, S);
      fwd_t(&X, S);
    }
    for (int it = n1 - 1; it >= n0; --it) {
      for (int ci = 0; ci < P->ncomp; ++ci)
        for (int64_t r = 0; r < P->nrec; ++r)
          L[P->comp[ci]][P->rec[r]] += adj[((int64_t)ci * nt + it) * P->nrec + r];
      memcpy(fblk, seg + (size_t)(it - n0) * NSTATE * X.n, sbytes);
      fwd_v(&X, F); /* fields after the velocity half step, stress carries before */
      adj_t(&X, F, L, gK);
      memcpy(fblk, seg + (size_t)(it - n0) * NSTATE * X.n, sbytes);
      adj_v(&X, F, L, gK);
    }
  }
done:
  free(blk);
  free(fblk);
  free(lblk);
  free(ck);
  free(seg);
  return rc;
}
