// Inversion parameterization of the Biot medium and its exact (double) chain rule.
//
// Per cell: drained Lame lambda, shear modulus mu, Biot coefficient alpha, Biot modulus M,
// bulk density rho, fluid density rho_f; fixed per cell: C1 = eta/kappa (Darcy damping) and
// tphi = C2 / rho_f = (1 + 2/porousm) tao / phi (tortuosity over porosity).  The moduli of
// the update (outputfunc.cpp:210-217) follow as
//     C = alpha M,  lambda_c = lambda + alpha^2 M,  H = lambda + 2 mu + alpha^2 M,
// and the staggered averages and coefficients as main.cpp:322-494, evaluated in double.
#pragma once
#include <cstdint>
#include <vector>

#include "biot_device.hpp"

namespace biot {

struct ParamArrays {
  const double *lambda, *mu, *alpha, *M, *rho, *rhof, *C1, *tphi;
};

struct CoefScales {
  double dt, Cv, Cp, Hinv;  // Hinv: 1/dx as the kernels hold it
};

// The 22 kernel coefficient arrays (kernel order K_*) of a parameter set, in double.
void kernel_coefs_from_params(int nx, int ny, int nz, const CoefScales &s, const ParamArrays &p,
                              std::vector<double> (&K)[NKC]);

// Chain rule: gradients with respect to the 22 kernel coefficient arrays -> gradients with
// respect to (lambda, mu, alpha, M, rho, rho_f), accumulated into out[6] (n doubles each).
void chain_params(int nx, int ny, int nz, const CoefScales &s, const ParamArrays &p,
                  const double *const gK[NKC], double *const out[6]);

}  // namespace biot
