// Device-side data layout of the Biot solver and the launchers of its HIP kernels.
//
// HBM layout (one handle, n = nx*ny*nz cells, x fastest, idx = i + j*nx + k*nx*ny):
//   fields  : 13 arrays of n elements of the storage type (vx vy vz qx qy qz txx tyy tzz txy
//             txz tyz p), one allocation.
//   coefs   : 22 arrays of n elements (the 23 reference arrays minus W3z, which the
//             reference never reads, kernel.cu:220), with 1/dx, the 0.01 kernel constants
//             and the storage scales folded in.
//   carries : C-PML memory, only inside the absorbing slabs.  Per axis, 8 arrays (4 for the
//             velocity half step, 4 for the stress half step) over the slab cells of that
//             axis: x-slab cells (i < Lx or i >= nx-Rx) are packed to ix in [0, Lx+Rx) and
//             stored at ix + (Lx+Rx)*(j + ny*k); y and z likewise.  One folded variable per
//             derivative replaces the reference's (psi, D_prev) pair:
//                 psi_n = c + a D_n,   D <- D + psi_n,   c <- e (psi_n + a D_n)
//             which is algebraically the recursion of kernel.cu:188-196.
//   profiles: per axis e and a = -dt d / 2 at integer and half nodes (1-D, arithmetic type).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace biot {

enum { NFIELD = 13, NKC = 22, NCARRY = 8 };
// kernel coefficient order
enum {
  K_U1X = 0, K_U1Y, K_U1Z, K_U2X, K_U2Y, K_U2Z, K_W1X, K_W1Y, K_W1Z, K_W2X, K_W2Y, K_W2Z,
  K_W3X, K_W3Y,
  K_P1, K_P2, K_S1, K_S2, K_S3, K_SXY, K_SXZ, K_SYZ
};

template <typename S>
struct StepArgs {
  S *f[NFIELD];
  const S *k[NKC];
  float ks[NKC];             // per-array scale (binary16 storage), 1 otherwise
  S *carry[3];               // per axis: NCARRY arrays of slab[ax] elements
  int64_t slab[3];           // elements per carry array per axis
  const void *prof[3][4];    // per axis: e_int, a_int, e_half, a_half (arithmetic type)
  int nx, ny, nz;
  int L[3], R[3];            // slab widths
  int64_t sy, sz;            // strides
  int zchunk;
  int gx, gy, gz, ntiles;    // tile grid (x-y tiles of the block shape, z chunks)
};

// Host-side launchers (biot_kernels.hip).  prec: 0 fp32, 1 fp16, 2 fp64.
// Storage pointers are passed type-erased as the matching StepArgs<S>.
struct LaunchShape {
  int bx, by, zchunk;
};

hipError_t launch_velocity(int prec, int order, const void *args, LaunchShape shape,
                           hipStream_t s);
hipError_t launch_stress(int prec, int order, const void *args, LaunchShape shape,
                         hipStream_t s);
struct SourceArgs {
  void *f[NFIELD];
  int64_t cell;
  uint32_t mask;
  const float *wavelet;     // device, one sample per step
  int it;
  float inv_scale[NFIELD];  // 1 / storage scale per field
};
struct RecordArgs {
  const void *f[NFIELD];
  const int64_t *rec;       // device receiver cells
  int64_t nrec;
  int comp[NFIELD];
  int ncomp;
  float scale[NFIELD];      // storage scale per recorded component
  float *traces;            // device [comp][nt_max][nrec]
  int it, nt_max;
};
hipError_t launch_source(int prec, const SourceArgs &a, hipStream_t s);
hipError_t launch_rescale(int prec, void *p, int64_t n, double f, hipStream_t s);
hipError_t launch_record(int prec, const RecordArgs &a, hipStream_t s);

}  // namespace biot
