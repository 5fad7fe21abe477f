// Device arguments and launchers of the discrete adjoint (biot_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "biot_device.hpp"

namespace biot {

// Per-grid adjoint arguments.  The forward states of a recomputed segment sit in `slots`
// (state q at slots + q*sb: the fields block, field f at f*fbytes, then the carry blocks at
// coff[0..2]); the adjoint step of segment step q reads state q (after its source) and the
// v, q fields of state q+1.
struct AdjArgs {
  const char *slots;
  uint64_t sb, fbytes, coff[3];
  uint64_t loff;             // offset in a state of the FP16T / FP16C residual arrays (int8, field f at
                             // f*n) or of FP16M's residual records (4 bytes per cell), else 0
  uint32_t lomask;           // fields whose residual the adjoint reads (FP16T, FP16M tau, p; FP16C v, tau, p)
  void *lam[NFIELD];         // adjoint wavefields (arithmetic type), storage units
  void *lcarry[3];           // adjoint carries (arithmetic type), forward layout
  void *G[12];               // derivative adjoints of the current half step (arithmetic type)
  double *grad[NKC];         // dJ/dK in storage units (multiply by the handle's kfac), FP64
  // per-step accumulation of dJ/dK: float partial sums over one checkpoint segment (FP32 and
  // FP16 handles; launch_adjoint_flush adds them into grad and clears them), grad itself on
  // FP64 handles
  void *gacc[NKC];
  // FP32/FP16 handles: the same partial sums packed per cell (float; gpv: K[0..13] of the
  // velocity half step, gpt: K[14..21] of the stress half step, in the 16-byte group layout
  // of biot_packed.hpp); null on FP64 handles
  void *gpv, *gpt;
  const void *k[NKC];        // kernel coefficients (storage type)
  const void *kv, *kt;       // the coefficients packed per cell (storage type, same layout)
  float ks[NKC];             // per-array coefficient scale (16-bit storage), else 1
  const void *prof[3][4];    // e_int, a_int, e_half, a_half per axis (arithmetic type)
  int nx, ny, nz;
  int L[3], R[3];
  int64_t slab[3];
  int64_t sy, sz;
  uint32_t nupd;             // cells of the update box
};

struct InjectArgs {
  void *lam[NFIELD];
  const int64_t *rec;
  int64_t nrec;
  int comp[NFIELD];
  int ncomp;
  float scale[NFIELD];
  const float *src;          // adjoint source [comp][nt][rec]
  int nt;
};

// adjoint of one full step (stress then velocity half step, two passes each) for segment
// step q; prec is the storage precision (FP32, FP64 or FP16)
hipError_t launch_adjoint_step(int prec, int order, const AdjArgs &a, int q, hipStream_t s);
hipError_t launch_adjoint_inject(int prec, const InjectArgs &a, int it, hipStream_t s);
// grad += packed partial sums, partial sums = 0 over the update box (no-op on FP64 handles,
// which sum into grad)
hipError_t launch_adjoint_flush(int prec, int order, const AdjArgs &a, hipStream_t s);
// J += 1/2 sum (d - obs)^2 (FP64), res = d - obs (res may be null); d is double when
// d_double, else float
hipError_t launch_misfit(const void *d, bool d_double, const float *obs, float *res, int64_t n,
                         double *J, hipStream_t s);

// Batched forms: device tables of per-shot arguments.  block_start[s] is the first block
// of shot s in a grid of 256-thread blocks covering every shot's update box.
struct AdjBatch {
  const AdjArgs *shots;
  const int *block_start;    // nshots + 1
  int nshots, total_blocks;
};
hipError_t launch_adjoint_step_batch(int prec, int order, const AdjBatch &b, int q, hipStream_t s);
hipError_t launch_adjoint_flush_batch(int prec, int order, const AdjBatch &b, hipStream_t s);
hipError_t launch_adjoint_inject_batch(int prec, const InjectArgs *tab, int nshots, int64_t max_nrec,
                                       int max_ncomp, int it, hipStream_t s);

// Chain rule on the device: storage-unit coefficient gradients of one window (grad[k] times
// kfac[k] gives dJ/dK) -> dJ/d(lambda, mu, alpha, M, rho, rho_f), added (FP64 atomics) into
// the model-grid arrays out[6] at the window's position.  Parameters are the window's.
struct ChainArgs {
  const double *grad[NKC];
  double kfac[NKC];
  const double *par[8];      // lambda mu alpha M rho rhof C1 tphi (window cells)
  double *out[6];            // model grid
  int wx, wy, wz;            // window size
  int x0, y0, z0;            // window origin in the model grid
  int64_t msy, msz;          // model strides
  double dt, Cv, Cp, Hinv;
};
hipError_t launch_chain(const ChainArgs &c, hipStream_t s);

// Copies of whole states between live handles and slot buffers, many at once: entry e
// copies bytes[e] from src[e] to dst[e] (all 4-byte aligned).
struct CopyEntry {
  const void *src;
  void *dst;
  uint64_t bytes;
};
hipError_t launch_copy_batch(const CopyEntry *tab, int n, hipStream_t s);

}  // namespace biot
