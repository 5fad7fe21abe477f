// Device arguments and launchers of the discrete adjoint (biot_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "biot_device.hpp"

namespace biot {

struct AdjArgs {
  const void *fw[NFIELD];    // forward state of step n (after its source), storage type
  const void *fw1[NFIELD];   // forward state of step n+1 (v, q used)
  const void *fcarry[3];     // C-PML carries before step n (storage type)
  void *lam[NFIELD];         // adjoint wavefields (arithmetic type)
  void *lcarry[3];           // adjoint carries (arithmetic type), forward layout
  void *G[12];               // derivative adjoints of the current half step (arithmetic type)
  double *grad[NKC];         // dJ/dK, FP64
  const void *k[NKC];        // kernel coefficients (storage type)
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
  int it, nt;
};

// adjoint of one full step (stress then velocity half step, two passes each)
hipError_t launch_adjoint_step(int prec, int order, const AdjArgs &a, hipStream_t s);
hipError_t launch_adjoint_inject(int prec, const InjectArgs &a, hipStream_t s);
// J += 1/2 sum (d - obs)^2 (FP64), res = d - obs (res may be null); d is double when
// d_double, else float
hipError_t launch_misfit(const void *d, bool d_double, const float *obs, float *res, int64_t n,
                         double *J, hipStream_t s);

}  // namespace biot
