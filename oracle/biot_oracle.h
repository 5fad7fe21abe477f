/*
 * CPU oracle for the 13-field staggered-grid Biot (poroelastic) time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: a plain-C restatement of the
 * reference's per-cell arithmetic and set-up, used by the parity tests, by the
 * smoke check and as the CPU baseline of the benchmark.  The product never links,
 * loads or calls it.
 *
 * Reference being restated (file:line of the upstream sources):
 *   Biot moduli ................ outputfunc.cpp:190-217
 *   staggered averages ......... main.cpp:322-358
 *   coefficient arrays ......... main.cpp:431-494   (scale factors main.cpp:248-262)
 *   C-PML profiles ............. main.cpp:537-566 (+ the 25 other per-rank cases, all the
 *                                same single-domain rule) and the x0.01 pack main.cpp:1512-1523
 *   Ricker wavelet ............. kernel.cu:1898, 2068
 *   source injection ........... kernel.cu:130-142
 *   velocity half step ......... kernel.cu:145-229 (FD_Vtop; FD_Vmid/FD_Vbottom identical)
 *   stress half step ........... kernel.cu:405-493 (FD_Ttop; FD_Tmid/FD_Tbottom identical)
 *   step ordering / recording .. kernel.cu:1891-2053, 2057-2243
 *
 * Three arithmetic modes:
 *   OR_MODE_F32  storage float, temporaries float, double-literal promotions exactly as
 *                the reference's expressions produce them (the FP32 drop-in semantics);
 *   OR_MODE_F16  storage IEEE binary16 (round-to-nearest-even, no flush), temporaries as
 *                OR_MODE_F32 (the reference as shipped, `typedef half data_t`, common.h:5);
 *   OR_MODE_F64  storage and arithmetic double (the FP64 truth for error measurement).
 *
 * Space order 2 is the reference's stencil; order 4 (coefficients 9/8, -1/24) has no
 * reference arithmetic and is pinned by closed-form tests only.
 *
 * Pins (tests/test_oracle_pins.py): moduli by Gassmann's equation and hand values, and bit for
 * bit against the reference's own outputfunc.cpp compiled from its sources (`make ref`,
 * oracle/_ref); averages by scipy's harmonic mean; PML profiles by their design reflection;
 * Ricker by its closed form; the fluid coupling (inertial term m - rho_f^2/rho, damping
 * eta/kappa, rho_f/rho) by the exact solution of the Biot relaxation ODE, momentum
 * conservation and the transient toward Darcy's flux along each axis; every difference of
 * both half steps by plane-wave eigenmodes along every axis and polarisation, and by the
 * C-PML closed form per derivative; the stress update by the constitutive law on linear
 * fields.  tests/test_oracle_mutants.py plants 22 plausible mistakes (signs, inverted
 * ratios, wrong strides or directions, wrong profile nodes) in copies of these sources and
 * checks that the pins reject every one.
 * Parity unpinned: the rounding of the original nvcc-built binary (FMA contraction choices,
 * cuda_fp16's double-to-half stores), which cannot be built or run here.
 *
 * Field numbering (shared with the product's C-ABI): 0 vx 1 vy 2 vz 3 qx 4 qy 5 qz
 * 6 txx 7 tyy 8 tzz 9 txy 10 txz 11 tyz 12 p.
 */
#ifndef BIOT_ORACLE_H
#define BIOT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_MODE_F32 = 0, OR_MODE_F16 = 1, OR_MODE_F64 = 2 };
enum { OR_NFIELD = 13, OR_NCOEF = 23 };

/* Coefficient array order (the 23 arrays of main.cpp:431-494). */
enum {
  OR_W1X, OR_W1Y, OR_W1Z, OR_W2X, OR_W2Y, OR_W2Z, OR_W3X, OR_W3Y, OR_W3Z,
  OR_U1X, OR_U1Y, OR_U1Z, OR_U2X, OR_U2Y, OR_U2Z,
  OR_PP1, OR_PP2, OR_SP1, OR_SP2, OR_SP3, OR_SPXY, OR_SPXZ, OR_SPYZ
};

/* --- IEEE binary16 helpers (round to nearest even, subnormals kept, overflow -> inf) --- */
uint16_t or_f32_to_f16(float f);
uint16_t or_f64_to_f16(double d);
float or_f16_to_f32(uint16_t h);

/* --- set-up --- */
/* outputfunc.cpp:205-217 restated for modulus inputs (no velocity-to-modulus step). */
void or_biot_moduli(int64_t n, const double *Ks, const double *Kf, const double *Kb,
                    const double *mu_in, const double *phi, const double *rhos_in,
                    const double *rhof_in, const double *eta, const double *perm,
                    double tao, double porousm,
                    float *M, float *HH, float *H2u, float *C, float *C1, float *C2,
                    float *rho, float *rhof, float *mu);

/* main.cpp:322-358: averages over cells 1..n-2 of every axis, zero elsewhere. */
void or_stagger_average(int nx, int ny, int nz, const float *mu, const float *rho,
                        const float *rhof, const float *C1, const float *C2,
                        float *muxz, float *muxy, float *muyz,
                        float *rhox, float *rhoy, float *rhoz,
                        float *rhofx, float *rhofy, float *rhofz,
                        float *C1x, float *C1y, float *C1z,
                        float *C2x, float *C2y, float *C2z);

/* main.cpp:431-494.  M, C, HH, H2u are the per-cell moduli masked to the interior the way
 * main.cpp:355-358 leaves them (zero on the outer planes).  coef[23] are outputs. */
void or_coefficients(int nx, int ny, int nz, float DT, float Cv, float Cp,
                     const float *muxz, const float *muxy, const float *muyz,
                     const float *rhox, const float *rhoy, const float *rhoz,
                     const float *rhofx, const float *rhofy, const float *rhofz,
                     const float *C1x, const float *C1y, const float *C1z,
                     const float *C2x, const float *C2y, const float *C2z,
                     const float *M, const float *C, const float *HH, const float *H2u,
                     float *const coef[OR_NCOEF]);

/* Single-domain C-PML profile of one axis (main.cpp:537-566 rule): d and e=exp(-d*DT)
 * at integer nodes (d, e) and half nodes (d2, e2), before the x0.01 pack. */
void or_pml_profile(int n, float H, float DT, int NP, float v0, float RC,
                    float *d, float *d2, float *e, float *e2);

/* kernel.cu:1898 Ricker, evaluated the way the reference does (double, stored float). */
float or_ricker(int it, float DT, float F0);

/* --- state and stepping --- */
typedef struct or_config {
  int nx, ny, nz;
  int order;               /* 2 or 4 */
  int mode;                /* OR_MODE_* */
  float DT;                /* time step (s) */
  float Hinv;              /* 1/dx, the kernels' `H` (kernel.cu:158) */
  const float *coef[OR_NCOEF];          /* full-volume, reference representation */
  /* 1-D PML profiles per axis, already multiplied by Cpml = 0.01f (main.cpp:1512-1523):
   * pd[a], pd2[a], pe[a], pe2[a] for a = 0 (x), 1 (y), 2 (z). */
  const float *pd[3], *pd2[3], *pe[3], *pe2[3];
} or_config;

typedef struct or_state or_state;

or_state *or_state_create(const or_config *cfg);
void or_state_destroy(or_state *st);
void or_state_zero(or_state *st);
/* Read one field into float (descaled: none; the reference representation). */
void or_state_get_field(const or_state *st, int field, float *out);
void or_state_get_field_f64(const or_state *st, int field, double *out);
void or_state_set_field_f64(or_state *st, int field, const double *in);

/* Add `amp` to every field whose bit is set in mask at linear cell `sn`
 * (kernel.cu:130-142: float add, stored back). */
void or_source(or_state *st, int64_t sn, unsigned mask, float amp);
void or_step_v(or_state *st);   /* kernel.cu:145-229 over cells h..n-1-h */
void or_step_t(or_state *st);   /* kernel.cu:405-493 over cells h..n-1-h */

/* Full loop: for it in [it0, it0+nsteps): source(wavelet[it]) -> V -> T -> record.
 * traces layout [comp][it][rec] with comp running over the set bits of rec_mask in
 * increasing field order; nt_total is the it-stride.  Returns 0. */
int or_run(or_state *st, int it0, int nsteps, int nt_total, const float *wavelet,
           int64_t src, unsigned src_mask, int nrec, const int64_t *rec, unsigned rec_mask,
           float *traces);
/* Same, traces in double (for OR_MODE_F64). */
int or_run_f64(or_state *st, int it0, int nsteps, int nt_total, const float *wavelet,
               int64_t src, unsigned src_mask, int nrec, const int64_t *rec,
               unsigned rec_mask, double *traces);

int or_num_threads(void);

/* ---- FP64 host adjoint (biot_host_adjoint.c; no counterpart in the reference) ---- */
typedef struct ha_problem {
  int nx, ny, nz, order;
  /* per axis (x, y, z): e and a = -dt d / 2 at integer nodes, then at half nodes */
  const double *prof[3][4];
  const double *K[22];            /* kernel coefficients (enum biot_kcoef order) */
  int64_t src_cell;
  unsigned src_mask;              /* fields the wavelet is added to */
  const float *wavelet;
  int nt_w;
  const int64_t *rec;             /* receiver cells */
  int64_t nrec;
  const int *comp;                /* recorded fields, trace component order */
  int ncomp;
} ha_problem;
/* nt steps from rest (source -> V -> T -> record), traces [ncomp][nt][nrec] */
int ha_forward(const ha_problem *p, int nt, double *traces);
/* dJ/dK of J with dJ/d(traces) = adj ([ncomp][nt][nrec]), checkpoints every `every` steps;
 * gK: 22 full-volume arrays (overwritten); traces (optional) receives the forward traces */
int ha_adjoint(const ha_problem *p, int nt, int every, const double *adj, double *const gK[22],
               double *traces);

#ifdef __cplusplus
}
#endif
#endif
