// Host-side set-up of the Biot solver: Biot moduli, staggered averages, the 23 update
// coefficients, C-PML profiles and the Ricker wavelet.  Each routine keeps the operand
// types of the reference statement it implements (float arrays, double scalars and
// double literals), so the values match the reference's set-up bit for bit.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/biot.h"

namespace biot {

// Work-sharing loop over [0, n) on the host's hardware threads.
void parallel_for(int64_t n, void (*body)(int64_t lo, int64_t hi, void *ctx), void *ctx);

struct Moduli {  // per-cell, float (outputfunc.cpp: every model array is float***)
  std::vector<float> M, HH, H2u, C, C1, C2, rho, rhof, mu;
};

struct Averages {  // main.cpp:322-358, zero on the outer planes
  std::vector<float> muxz, muxy, muyz, rho[3], rhof[3], C1[3], C2[3];
};

// outputfunc.cpp:205-217 for modulus inputs.
void build_moduli(int64_t n, const biot_medium &m, Moduli &out);
// main.cpp:322-358.
void build_averages(int nx, int ny, int nz, const Moduli &md, Averages &out);
// main.cpp:431-494.  Moduli M, C, HH, H2u enter only on interior cells (the reference
// copies them into its broadcast arrays inside the 1..n-2 loop, main.cpp:355-358).
void build_coefficients(int nx, int ny, int nz, float DT, float Cv, float Cp, const Moduli &md,
                        const Averages &av, std::vector<float> (&coef)[BIOT_NCOEF]);
// main.cpp:537-566 single-domain rule for one axis.
void build_pml_profile(int n, float H, float DT, int NP, float v0, float RC, float *d, float *d2,
                       float *e, float *e2);
// kernel.cu:1898.
float ricker(int it, float DT, float F0);

}  // namespace biot
