/*
 * biot.h - C-ABI of the MI355X (gfx950) solver for the 13-field staggered-grid Biot
 * (poroelastic) velocity-stress system with C-PML, source injection and receivers.
 *
 * This ABI replaces the reference's single entry point
 *     void callCUDAKernel(blockInfo *parameter, int nxt_pro, int nyt_pro, int nzt_pro,
 *                         float DT, float F0, int sn, int *coord, int myid, int up, ...,
 *                         int sxnew, int synew, int sznew, int NT)      (kernel.cu:1728)
 * and the 99 caller-filled device arrays of `blockInfo` (common.h:6-133, filled in
 * main.cpp:1376-1743) with an opaque handle that owns its device memory.  The MPI
 * arguments (coord, myid, up..right, MCW) have no counterpart: one GPU holds the grid.
 *
 * Conventions
 *   - Every call returns BIOT_OK (0) or a BIOT_ERR_* code; nothing exits or aborts
 *     (the reference's CUDA_CHECK -> exit(EXIT_FAILURE), kernel.cu:42-52, is replaced).
 *     biot_last_error(h) gives the message of the last failure on a handle;
 *     biot_last_error(NULL) that of the calling thread's last failed biot_create.
 *   - Arrays are host pointers, x fastest: cell (i, j, k) is i + j*nx + k*nx*ny
 *     (main.cpp:1489).  Values are in the reference representation: velocities scaled
 *     by Cv, stresses and pressure by Cp (main.cpp:248-249).
 *   - One handle is used by one host thread at a time; distinct handles are independent.
 *
 * Field numbering: 0 vx 1 vy 2 vz 3 qx 4 qy 5 qz 6 txx 7 tyy 8 tzz 9 txy 10 txz 11 tyz 12 p
 * (reference names vux vuy vuz vwx vwy vwz txx tyy tzz txy txz tyz ss).
 */
#ifndef BIOT_H
#define BIOT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BIOT_ABI_VERSION 2

enum biot_status {
  BIOT_OK = 0,
  BIOT_ERR_ARG = 1,        /* invalid argument (dims, order, indices, sizes) */
  BIOT_ERR_HIP = 2,        /* HIP runtime failure */
  BIOT_ERR_STATE = 3,      /* call out of order (e.g. run before set_medium) */
  BIOT_ERR_NONFINITE = 4,  /* NaN/Inf detected on the receivers: the steps ran, and the
                              message names the first step that recorded one */
  BIOT_ERR_RANGE = 5,      /* value outside the storage format's range (FP16 scaling) */
  BIOT_ERR_NOMEM = 6       /* device or host allocation failed */
};

enum biot_precision {
  BIOT_FP32 = 0,  /* float storage, float arithmetic: the drop-in path */
  BIOT_FP16 = 1,  /* IEEE binary16 wavefields, coefficients and C-PML memory with per-field
                     scaling, float arithmetic */
  BIOT_FP64 = 2,  /* double storage and arithmetic (accuracy reference on the GPU) */
  BIOT_FP16C = 3, /* as BIOT_FP16 with float C-PML memory, plus an 8-bit residual per value of
                     v, tau and p (value = h + r * ulp(h)/256): compensated half precision */
  BIOT_FP16T = 4, /* as BIOT_FP16C with the residual on tau and p only (velocities plain
                     binary16): faster, with a larger error than BIOT_FP16C */
  BIOT_FP16M = 5  /* mixed 16/20-bit storage: v, q binary16 (per-field scaling); tau, p binary16
                     plus a 4-bit residual in units of ulp/16, the seven residuals of a cell
                     packed in one 4-byte record (the velocity step reads the outer taps of its
                     4th-order differences as binary16); the medium as per-cell binary16
                     moduli (rho, rho_f, C1, C2; C, M, lambda_c, H; mu) from which the kernels
                     form the staggered averages, the harmonic means of mu and the
                     coefficients; float C-PML
                     memory and arithmetic.  Needs the moduli: biot_set_medium,
                     biot_set_medium_device or biot_set_params (not biot_set_coefficients); the
                     adjoint after biot_set_params. */
};

enum biot_field {
  BIOT_VX = 0, BIOT_VY, BIOT_VZ, BIOT_QX, BIOT_QY, BIOT_QZ,
  BIOT_TXX, BIOT_TYY, BIOT_TZZ, BIOT_TXY, BIOT_TXZ, BIOT_TYZ, BIOT_P,
  BIOT_NFIELD
};
#define BIOT_MASK(f) (1u << (f))
/* The reference source: Ricker added to txx, tyy, tzz (kernel.cu:138-140). */
#define BIOT_SRC_REFERENCE (BIOT_MASK(BIOT_TXX) | BIOT_MASK(BIOT_TYY) | BIOT_MASK(BIOT_TZZ))

/* The 23 coefficient arrays of main.cpp:431-494, in this order. */
enum biot_coef {
  BIOT_W1X = 0, BIOT_W1Y, BIOT_W1Z, BIOT_W2X, BIOT_W2Y, BIOT_W2Z, BIOT_W3X, BIOT_W3Y, BIOT_W3Z,
  BIOT_U1X, BIOT_U1Y, BIOT_U1Z, BIOT_U2X, BIOT_U2Y, BIOT_U2Z,
  BIOT_PP1, BIOT_PP2, BIOT_SP1, BIOT_SP2, BIOT_SP3, BIOT_SPXY, BIOT_SPXZ, BIOT_SPYZ,
  BIOT_NCOEF
};

typedef struct biot_grid {
  int32_t nx, ny, nz;   /* cells per axis including the PML (NX_ext.., main.cpp:73-75) */
  int32_t order;        /* space order: 2 (the reference stencil) or 4 */
  int32_t precision;    /* enum biot_precision */
  int32_t npml;         /* PML thickness in cells (NP, main.cpp:23) */
  float dx;             /* grid spacing in m (H, main.cpp:88) */
  float dt;             /* time step in s (DT, main.cpp:90) */
  float pml_v0;         /* PML reference velocity (v0 = 4000, main.cpp:555) */
  float pml_R;          /* PML reflection coefficient (RC = 1e-6, main.cpp:89) */
  float Cv, Cp;         /* representation scales (Cv = 1e8, Cp = 1, main.cpp:248-249) */
  int32_t device;       /* HIP device ordinal */
} biot_grid;

/* Per-cell material inputs, nx*ny*nz doubles each (x fastest).  The Biot moduli follow
 * outputfunc.cpp:205-217: alpha = 1 - Kb/Ks, M = Kf Ks / (phi Ks + (alpha - phi) Kf),
 * H = alpha^2 M + Kb + 4 mu / 3, lambda_c = H - 2 mu, C = alpha M, C1 = eta/perm,
 * C2 = (1 + 2/porousm) tao rho_f / phi, rho = (1 - phi) rho_s + phi rho_f.
 * A cell with mu == 0 and Kb == 0 is a fluid cell (outputfunc.cpp:176-193). */
typedef struct biot_medium {
  const double *Ks, *Kf, *Kb, *mu, *phi, *rhos, *rhof, *eta, *perm;
  double tao;       /* 3.0 in the reference (outputfunc.cpp:213) */
  double porousm;   /* 8.0 in the reference (outputfunc.cpp:164) */
} biot_medium;

typedef struct biot_handle biot_handle;

/* ---- lifetime ---- */
/* Sizes: the kernels address each array with 32-bit byte offsets from its base, so a grid
 * is refused (BIOT_ERR_ARG) when one field array (n cells x the storage width) or one axis's
 * C-PML memory (8 carries x the slab's cells x the carry width) would exceed 4 GiB; the
 * adjoint refuses a handle whose multiplier arrays would (n x 4 bytes, x 8 on FP64). */
int biot_create(const biot_grid *grid, biot_handle **out);
int biot_destroy(biot_handle *h);
const char *biot_last_error(const biot_handle *h);
const char *biot_status_string(int status);
int biot_abi_version(void);

/* ---- medium, coefficients, PML ---- */
/* Moduli -> staggered averages -> the 23 coefficient arrays (main.cpp:322-494), then
 * uploads them.  Replaces rank 0's create_model_all + averaging + per-rank coefficient
 * loops (main.cpp:303-494). */
int biot_set_medium(biot_handle *h, const biot_medium *m);
/* Caller-built coefficient arrays in the reference representation (the 23 coefficient
 * members of blockInfo, common.h:21-27). coef[BIOT_W3Z] may be NULL (never read). */
int biot_set_coefficients(biot_handle *h, const float *const coef[BIOT_NCOEF]);
/* Read back the coefficient arrays the handle holds (reference representation). */
int biot_get_coefficients(biot_handle *h, float *const coef[BIOT_NCOEF]);
/* The single-domain C-PML profile of one axis (main.cpp:537-566 rule): d and e=exp(-d dt)
 * at integer (d, e) and half (d2, e2) nodes, before the x0.01 pack of main.cpp:1512. */
int biot_get_pml_profile(biot_handle *h, int32_t axis, float *d, float *d2, float *e,
                         float *e2);

/* biot_set_medium with the moduli, averages and coefficients built on the device (the same
 * bits as the host builder); for large grids, where the host set-up dominates. */
int biot_set_medium_device(biot_handle *h, const biot_medium *m);
/* Host only: reference arrays -> the 22 kernel coefficient arrays before storage scaling
 * (1/dx as float, kernel.cu:158, and the 0.01 of kernel.cu:484-490 folded; n cells). */
int biot_kernel_coefficients(float dx, int64_t n, const float *const coef[BIOT_NCOEF],
                             double *const K[22]);
/* The stored kernel coefficient array k (enum biot_kcoef) as double, times its storage
 * scale for 16-bit storage. */
int biot_read_kernel_coefficients(biot_handle *h, int32_t k, double *out);

/* Host-only forms of the two set-up stages above (no device is touched): the 23
 * coefficient arrays of main.cpp:431-494 from per-cell moduli (nx*ny*nz floats each), and
 * the PML profile of one axis of length n. */
int biot_build_coefficients(int32_t nx, int32_t ny, int32_t nz, float dt, float Cv, float Cp,
                            const biot_medium *m, float *const coef[BIOT_NCOEF]);
int biot_pml_profile_1d(int32_t n, float dx, float dt, int32_t npml, float v0, float R, float *d,
                        float *d2, float *e, float *e2);

/* ---- source, receivers ---- */
/* Wavelet samples w[0..nt-1]; at step it, w[it] is added to each field of field_mask at
 * linear cell `cell` before the velocity half step (kernel.cu:130-142, 2064-2071). */
int biot_set_source(biot_handle *h, int64_t cell, uint32_t field_mask, const float *wavelet,
                    int32_t nt);
/* Reference Ricker (kernel.cu:1898): w[it] = -(1 - 2 pi^2 f^2 (t-T0)^2) exp(-pi^2 f^2 (t-T0)^2),
 * t = it*dt, T0 = 1.2/f, evaluated in double and rounded to float. */
int biot_ricker(float f0, float dt, int32_t nt, float *out);
/* Receivers sample every field of field_mask at each cell after every step
 * (kernel.cu:2197-2243).  Traces hold nt_max steps; layout [comp][it][rec], comp running
 * over the set bits of field_mask in increasing field order. */
int biot_set_receivers(biot_handle *h, const int64_t *cells, int64_t nrec, uint32_t field_mask,
                       int32_t nt_max);

/* ---- time loop ---- */
int biot_reset(biot_handle *h);             /* zero wavefields and PML memory, step = 0 */
/* Advance nsteps: source -> V -> T -> record.  Returns BIOT_ERR_NONFINITE when a receiver has
 * recorded NaN or Inf since the last biot_reset / biot_set_receivers (the watchdog that
 * replaces the reference's exit on a failed check, kernel.cu:42-52). */
int biot_run(biot_handle *h, int32_t nsteps);
int biot_step_count(biot_handle *h, int32_t *it);
int biot_synchronize(biot_handle *h);
/* Wall time (ms, HIP events on the handle's stream) of the last biot_run. */
int biot_last_run_ms(biot_handle *h, double *ms);
/* Advance nsteps like biot_run, with a HIP event pair around every launch of the two
 * stencil kernels on the handle's stream; returns their mean durations in ms. */
int biot_profile_steps(biot_handle *h, int32_t nsteps, double *ms_velocity, double *ms_stress);

/* ---- output ---- */
/* The reference's text format (wfile, outputfunc.cpp:35-50): nr rows of nc values, each value
 * printed "%e ", each row followed by a newline and an empty line.  data is row-major. */
int biot_write_text_2d(const char *path, const float *data, int32_t nr, int32_t nc);
/* The y = ny/2 slice of a (nz, ny, nx) volume in the same format (wfile1d / wfile3d,
 * outputfunc.cpp:97-151): nz rows of nx values. */
int biot_write_text_xz(const char *path, const float *vol, int32_t nz, int32_t ny, int32_t nx);
/* count float32 values, raw little-endian (the binary option beside the text files). */
int biot_write_binary(const char *path, const float *data, int64_t count);
/* nsteps from the current step, like biot_run, with the recorded traces copied into out
 * (trace layout, ncomp*nt_max*nrec floats) on a side stream chunk steps at a time while
 * the loop continues; returns when every copy has landed.  Replaces the reference's
 * per-step full-volume copy to the host (kernel.cu:2197-2211). */
int biot_run_stream_traces(biot_handle *h, int32_t nsteps, int32_t chunk, float *out);
/* The same loop with xz snapshots as well: after step it0 + snap_steps[s] (it0 the step at the
 * call; snap_steps ascending within [0, nsteps)) the y = j slice of `field` is written, in
 * physical units, into snaps + s*nz*nx (row k, column i) and copied to the host on the side
 * stream while the loop continues (the reference writes these slices at it = 1000, 2000, 3000,
 * kernel.cu:2247-2385).  traces may be NULL (no trace copies). */
int biot_run_stream(biot_handle *h, int32_t nsteps, int32_t chunk, float *traces, int32_t field, int32_t j,
                    const int32_t *snap_steps, int32_t nsnap, float *snaps);

/* ---- inversion parameterization and the adjoint (no counterpart in the reference) ---- */
/* Per-cell inversion parameters (nx*ny*nz doubles each).  lambda (drained Lame), mu, alpha,
 * M, rho, rho_f are the inverted quantities; C1 = eta/kappa and tphi = C2/rho_f
 * = (1 + 2/porousm) tao / phi stay fixed.  The update moduli are C = alpha M,
 * lambda_c = lambda + alpha^2 M, H = lambda_c + 2 mu (outputfunc.cpp:210-217) and the
 * coefficients follow main.cpp:322-494, evaluated in double. */
typedef struct biot_params {
  const double *lambda, *mu, *alpha, *M, *rho, *rhof;
  const double *C1, *tphi;
} biot_params;
int biot_set_params(biot_handle *h, const biot_params *p);
/* Resets, runs nt steps recording traces (receivers must hold >= nt steps) and keeps the
 * state every `every` steps on the device for the adjoint.  Every policy (FP16T and FP16C
 * states carry their residual arrays, FP16M states its residual records, and the adjoint
 * reads them).  An FP16M handle needs biot_set_params first (BIOT_ERR_STATE otherwise): it
 * keeps moduli, not coefficient arrays, so the adjoint reads the parameters' coefficients
 * as FP16T stores them, binary16 packed per cell, built here (44 bytes per cell, freed with
 * the checkpoints).
 * every = 1 keeps every step's state (nt + 1 states in one device block, BIOT_ERR_NOMEM if
 * it does not fit): biot_adjoint then reads them instead of recomputing each segment, and
 * gives the same gradients as every = 50 bit for bit. */
int biot_forward_checkpointed(biot_handle *h, int32_t nt, int32_t every);
/* J = 1/2 sum (d - observed)^2 over the recorded traces (the double traces on FP64
 * handles), reduced in FP64 on the device; residual (may be NULL) receives d - observed in
 * the trace layout. */
int biot_misfit_l2(biot_handle *h, const float *observed, float *residual, double *J);
/* No counterpart in the reference: biot_misfit_l2 without device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The observed and residual staging and the FP64 accumulator live in one static
 * 64 MiB device array per GPU, held until the process ends; a trace layout of more than
 * BIOT_MISFIT_NOALLOC_MAX_SAMPLES samples (ncomp * nt_max * nrec) returns BIOT_ERR_ARG.  Every
 * call on a GPU uses that GPU's one array: calls that overlap on one GPU race. */
#define BIOT_MISFIT_NOALLOC_MAX_SAMPLES 8388576
int biot_misfit_l2_noalloc(biot_handle *h, const float *observed, float *residual, double *J);
/* Discrete adjoint of the checkpointed run.  adj_source = dJ/d(trace samples), trace
 * layout.  Segments of `every` steps are recomputed from their checkpoints and swept
 * backward; coefficient gradients are summed in FP64 on the device.  grad_params (may be
 * NULL; needs biot_set_params) receives dJ/d(lambda, mu, alpha, M, rho, rho_f), 6 arrays
 * of nx*ny*nz; grad_coef (may be NULL) the 22 kernel-coefficient gradients.  The
 * wavefields are left in an unspecified state. */
int biot_adjoint(biot_handle *h, const float *adj_source, double *const grad_params[6],
                 double *const grad_coef[22]);
/* Order of the 22 kernel coefficient arrays (grad_coef, biot_params_coefficients): the
 * reference arrays without W3z, with 1/dx and the 0.01 of the stress update folded in. */
enum biot_kcoef {
  BIOT_K_U1X = 0, BIOT_K_U1Y, BIOT_K_U1Z, BIOT_K_U2X, BIOT_K_U2Y, BIOT_K_U2Z,
  BIOT_K_W1X, BIOT_K_W1Y, BIOT_K_W1Z, BIOT_K_W2X, BIOT_K_W2Y, BIOT_K_W2Z, BIOT_K_W3X, BIOT_K_W3Y,
  BIOT_K_P1, BIOT_K_P2, BIOT_K_S1, BIOT_K_S2, BIOT_K_S3, BIOT_K_SXY, BIOT_K_SXZ, BIOT_K_SYZ,
  BIOT_NKCOEF
};
/* Host only: the kernel coefficients of a parameter set (grid: nx ny nz dx dt Cv Cp), and
 * the chain rule taking their gradients to dJ/d(lambda, mu, alpha, M, rho, rho_f). */
int biot_params_coefficients(const biot_grid *g, const biot_params *p, double *const K[22]);
int biot_params_chain(const biot_grid *g, const biot_params *p, const double *const grad_coef[22],
                      double *const grad_params[6]);

/* ---- batched shots (no counterpart in the reference) ---- */
/* Many shots, each a window of one model grid with its own absorbing boundary, source and
 * receivers, advanced together: one launch per half step (a grid of one block per tile
 * unless a persistent grid is set) covers the tiles of every shot; runs of steps are
 * captured into a HIP graph and replayed.  The gradient sums every shot's parameter
 * gradient into one model-grid volume on the device.  Precisions FP32, FP16 (binary16
 * wavefields, float arithmetic, FP64 gradient sums), FP16T and FP16C (FP16 with the 8-bit
 * residual on tau and p, and on v too), FP16M, FP64. */
typedef struct biot_batch biot_batch;
typedef struct biot_batch_config {
  int32_t nx, ny, nz;        /* model grid the windows are cut from (gradient volume) */
  int32_t order, precision, npml;
  float dx, dt, pml_v0, pml_R, Cv, Cp;
  int32_t device;
} biot_batch_config;
typedef struct biot_shot {
  int32_t x0, y0, z0;        /* window origin in the model grid */
  int32_t nx, ny, nz;        /* window size (x0 + nx <= model nx, ...) */
  int64_t src_cell;          /* window-local linear cell */
  uint32_t src_mask;         /* stresses / pressure only for gradients */
  const int64_t *rec_cells;  /* window-local linear cells */
  int64_t nrec;
  uint32_t rec_mask;
} biot_shot;
int biot_batch_create(const biot_batch_config *c, biot_batch **out);
int biot_batch_destroy(biot_batch *b);
/* the message of the last failure on a batch; with NULL, that of the calling thread's last
 * failed biot_batch_create */
const char *biot_batch_last_error(const biot_batch *b);
/* window_params: the window's cells (its nx*ny*nz); receivers hold nt steps */
int biot_batch_add_shot(biot_batch *b, const biot_shot *s, const biot_params *window_params,
                        const float *wavelet, int32_t nt);
int biot_batch_shot_count(biot_batch *b, int32_t *n);
/* new model-grid parameters (nx*ny*nz each): every shot's window is cut out of them and its
 * coefficients rebuilt (an inversion's model update) */
int biot_batch_set_model(biot_batch *b, const biot_params *model_params);
/* the per-shot grid (for reading fields or traces through the single-grid calls) */
int biot_batch_shot_handle(biot_batch *b, int32_t shot, biot_handle **h);
/* tile shape and persistent grid size (blocks, a multiple of 8) of the batched launches */
int biot_batch_set_launch(biot_batch *b, int32_t bx, int32_t by, int32_t zchunk, int32_t blocks);
int biot_batch_get_launch(biot_batch *b, int32_t *bx, int32_t *by, int32_t *zchunk, int32_t *blocks);
/* times tile shapes and grid sizes over trial_steps steps and keeps the fastest; leaves
 * every shot reset */
int biot_batch_autotune(biot_batch *b, int32_t trial_steps);
/* every shot from rest for nt steps, recording traces; graph_steps > 0 captures that many
 * steps into a HIP graph and replays it (the remainder runs step by step) */
int biot_batch_forward(biot_batch *b, int32_t nt, int32_t graph_steps);
int biot_batch_read_traces(biot_batch *b, int32_t shot, float *out, int64_t count);
/* J = sum over shots of 1/2 |d - observed|^2 (FP64 on the device) and its gradient with
 * respect to (lambda, mu, alpha, M, rho, rho_f) on the model grid (6 arrays of nx*ny*nz),
 * through checkpoints every `every` steps and the adjoint of every shot at once */
int biot_batch_gradient(biot_batch *b, int32_t nt, int32_t every, const float *const observed[],
                        double *J, double *const grad[6]);
/* shots are processed in groups that fit the free device memory (all at once when they
 * fit); the number of groups the last biot_batch_gradient used */
int biot_batch_gradient_groups(biot_batch *b, int32_t *groups);
/* HIP-event time of the last biot_batch_forward or biot_batch_gradient */
int biot_batch_last_ms(biot_batch *b, double *ms);
/* frees the gradient's checkpoints and adjoint buffers */
int biot_batch_release(biot_batch *b);

/* ---- launch shapes ---- */
/* Tile shape of the two stencil kernels: bx*by threads per block (bx*by <= 256, bx a
 * multiple of 64), each thread marching zchunk planes of its column. */
int biot_set_launch_shape(biot_handle *h, int32_t bx, int32_t by, int32_t zchunk);
int biot_get_launch_shape(biot_handle *h, int32_t *bx, int32_t *by, int32_t *zchunk);
/* Times every candidate tile shape for trial_steps steps on this grid and keeps the
 * fastest; results are cached per (precision, order, grid, GPU arch) in the file named by
 * $BIOT_TUNE_CACHE (default $HOME/.cache/biot_tune.txt), and a cached entry is reused
 * without timing.  Leaves the handle reset (wavefields zero, step 0). */
int biot_autotune(biot_handle *h, int32_t trial_steps);

/* ---- outputs ---- */
int biot_read_traces(biot_handle *h, float *out, int64_t count);
/* FP64 handles also keep the traces in double (the receiver sampling of kernel.cu:2027-2031
 * without the float store); finite-difference checks of the adjoint read these. */
int biot_read_traces_f64(biot_handle *h, double *out, int64_t count);
int biot_read_field(biot_handle *h, int32_t field, float *out);
int biot_read_field_f64(biot_handle *h, int32_t field, double *out);
int biot_write_field_f64(biot_handle *h, int32_t field, const double *in);
/* xz slice at y index j of one field (the slice wfile3d writes, outputfunc.cpp:97-123):
 * out[k*nx + i]. */
int biot_snapshot_xz(biot_handle *h, int32_t field, int32_t j, float *out);

/* ---- introspection ---- */
/* FP16 storage scales: field value = stored * scale (reference representation). */
int biot_get_field_scales(biot_handle *h, float *scales13);
/* Device bytes held by the handle. */
int biot_device_bytes(biot_handle *h, int64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
