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
//   residuals (FP16C, FP16T): one int8 array per compensated field (FP16C v, tau, p;
//             FP16T tau, p).
//
// Storage modes (enum biot_precision):
//   FP32  fields, coefficients, carries float; arithmetic float.
//   FP16  fields, coefficients and carries binary16 (scaled); arithmetic float.
//   FP16C as FP16, plus an int8 residual r per value of v, tau and p: the value is
//         h + r * ulp(h) / 256 (h the binary16 part), about a 19-bit significand, so the
//         rounding of the leapfrog-integrated fields does not random-walk.
//   FP16T as FP16C with the residual on tau and p only (the velocities plain binary16): the
//         stresses' rounding dominates the binary16 error, and the velocity residuals are
//         a third of FP16C's residual traffic.
//   FP16M v and q binary16 (scaled); tau and p binary16 (scaled) plus a 4-bit residual in
//         units of ulp/16, the binary16 part in the field's own array, the residuals of the
//         seven such fields in one 4-byte record per cell; the medium as per-cell binary16
//         moduli scaled to [-1, 1] (mv: rho, rho_f, C1, C2 in 8 bytes; mt: C, M, lambda_c, H
//         in 8 bytes; mu: 2 bytes), from which the kernels form the staggered averages, the
//         4-cell harmonic means of mu and the 22 coefficients in registers; float C-PML
//         carries.
//   FP64  everything double.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace biot {

enum { NFIELD = 13, NKC = 22, NCARRY = 8 };
enum { PREC_FP32 = 0, PREC_FP16 = 1, PREC_FP64 = 2, PREC_FP16C = 3, PREC_FP16T = 4, PREC_FP16M = 5 };
// fields carrying a residual: FP16C vx vy vz txx tyy tzz txy txz tyz p; FP16T txx .. tyz p
constexpr uint32_t COMP_FIELDS_C = 0x1FC7u, COMP_FIELDS_T = 0x1FC0u;
constexpr uint32_t comp_fields(int prec)
{
  return prec == PREC_FP16C ? COMP_FIELDS_C : (prec == PREC_FP16T || prec == PREC_FP16M) ? COMP_FIELDS_T : 0u;
}
// FP16M: the fields stored as binary16 + residual (txx .. tyz, p): the binary16 part in the
// field's own array, the 4-bit residual in the cell's 4-byte residual record, at nibble
// pk_slot(f) (the record's last nibble is unused)
constexpr uint32_t PK_FIELDS = COMP_FIELDS_T;
constexpr int PK_FIRST = 6, PK_RECORD = 4;
// nibble of field f in the record: the four fields the velocity step differences along x
// (txx, txy, txz, p) in the low half, tyy, tzz, tyz in the high half
constexpr int pk_slot(int f)
{
  return f == 6 ? 0 : f == 9 ? 1 : f == 10 ? 2 : f == 12 ? 3 : f == 7 ? 4 : f == 8 ? 5 : 6;
}
// the kernels take field PK_FIRST's residual pointer as the record base
static_assert(pk_slot(PK_FIRST) == 0, "field PK_FIRST must own nibble 0 of the record");
// ulp(h)/16 from h as a float (exponent below binary16's normal range: the subnormal ulp)
__device__ __forceinline__ float pk_unit(float hf)
{
  const uint32_t e = (__builtin_bit_cast(uint32_t, hf) >> 23) & 0xFFu;
  return __builtin_bit_cast(float, (max(e, 113u) - 14u) << 23);
}
// residual s (0..6) of a record, sign-extended
__device__ __forceinline__ int pk_nib(uint32_t r, int s)
{
  return (int)(r << (28 - 4 * s)) >> 28;
}
// h as a float plus residual n: the value the forward kernels and the adjoint read
__device__ __forceinline__ float pk_join(float hf, int n)
{
  return fmaf((float)n, pk_unit(hf), hf);
}
// kernel coefficient order
enum {
  K_U1X = 0, K_U1Y, K_U1Z, K_U2X, K_U2Y, K_U2Z, K_W1X, K_W1Y, K_W1Z, K_W2X, K_W2Y, K_W2Z,
  K_W3X, K_W3Y,
  K_P1, K_P2, K_S1, K_S2, K_S3, K_SXY, K_SXZ, K_SYZ
};

// FP16M: the constants that turn the stored moduli into kernel coefficients (in the storage
// units of the fields they connect, 1/dx folded in).  Averages along an axis are
// h * (stored(c) + stored(c + axis)); then U1 = cU1 / rho, U2 = kU2 rho_f / rho,
// m = C2 - rho_f^2 / rho, a = aH C1, W1 = (m - a) / (m + a), W2 = cW2 (rho_f / rho) / (m + a),
// W3 = cW3 / (m + a) (main.cpp:447-481); P1 = kP1 C, P2 = kP2 M, S1 = kS1 lambda_c, S2 = kS2 H,
// S3 = kS3 C, Sij = kSh mu_ij (the stored 4-cell harmonic mean, main.cpp:326-337, 483-494), all
// on the stored values.
enum {
  MC_HRHO = 0, MC_HRHOF, MC_HC1, MC_HC2, MC_KU2, MC_CU1, MC_CW2, MC_CW3, MC_AH,
  MC_KP1, MC_KP2, MC_KS1, MC_KS2, MC_KS3, MC_KSH, NMC
};

// Kernel arguments; pointers are type-erased, their element types follow the storage mode.
struct StepArgs {
  void *f[NFIELD];           // field storage
  int8_t *lo[NFIELD];        // FP16C / FP16T residuals (null for other modes / fields)
  const void *k[NKC];        // coefficient arrays (field storage type)
  // the same coefficients packed per cell (storage type): kv holds K[0..13] of the velocity
  // half step, kt K[14..21] of the stress half step, cell c at element c*14 / c*8 (addressed
  // in 64 bits: these arrays may span more than 4 GiB)
  const void *kv, *kt;
  float ks[NKC];             // per-array scale (binary16 storage), 1 otherwise
  // FP16M: the medium records (mv: 4 binary16 values per cell, mt: 4), the cell shear
  // moduli (mu: binary16) and constants
  const void *mv, *mt, *mu;
  float mc[NMC];
  // per axis: the C-PML carries (carry type) of the slab cells, in two groups of records — the
  // velocity half step's carries 0..3, then the stress half step's 4..7 — each group slab[ax]
  // records of 4 carries: carry m of slab cell o is element (m/4)*4*slab + 4*o + m%4
  void *carry[3];
  int64_t slab[3];           // elements per carry array per axis
  const void *prof[3][4];    // per axis: e_int, a_int, e_half, a_half (arithmetic type)
  int nx, ny, nz;
  int L[3], R[3];            // slab widths
  int64_t sy, sz;            // strides
  int zchunk;
  int gx, gy, gz, ntiles;    // tile grid (x-y tiles of the block shape, z chunks)
};

inline size_t field_bytes(int prec) { return prec == PREC_FP64 ? 8 : prec == PREC_FP32 ? 4 : 2; }
// C-PML carries: binary16 in the FP16 policy (all state 16-bit, as the reference stores every
// array in half), float in FP32, FP16C and FP16T, double in FP64
inline size_t carry_bytes(int prec) { return prec == PREC_FP64 ? 8 : prec == PREC_FP16 ? 2 : 4; }

// Host-side launchers (biot_kernels.hip).
struct LaunchShape {
  int bx, by, zchunk;
};

hipError_t launch_velocity(int prec, int order, const StepArgs &args, LaunchShape shape,
                           hipStream_t s);
hipError_t launch_stress(int prec, int order, const StepArgs &args, LaunchShape shape,
                         hipStream_t s);
struct SourceArgs {
  void *f[NFIELD];
  int8_t *lo[NFIELD];
  int64_t cell;
  uint32_t mask;
  const float *wavelet;     // device, one sample per step
  int it;
  int nt_w;                 // wavelet samples (batched launches check the step against it)
  float inv_scale[NFIELD];  // 1 / storage scale per field
};
struct RecordArgs {
  const void *f[NFIELD];
  const int8_t *lo[NFIELD];
  const int64_t *rec;       // device receiver cells
  int64_t nrec;
  int comp[NFIELD];
  int ncomp;
  float scale[NFIELD];      // storage scale per recorded component
  float *traces;            // device [comp][nt_max][nrec]
  double *traces64;         // same layout in double (FP64 handles), or null
  int *nonfinite;           // device: the first step with a non-finite sample (NONFINITE_NONE:
                            // none so far), or null
  int it, nt_max;
};
// the byte pattern a cleared non-finite flag holds (every byte 0x7F): larger than any step
constexpr int NONFINITE_NONE = 0x7F7F7F7F;
hipError_t launch_source(int prec, const SourceArgs &a, hipStream_t s);

// GPU-side set-up (biot_build.hip), bit-identical to the host builder.
struct MediumArgs {
  const double *Ks, *Kf, *Kb, *mu, *phi, *rhos, *rhof, *eta, *perm;
  double tao, porousm;
  float *out[9];            // M HH H2u C C1 C2 rho rhof mu
};
struct CoefRefArgs {
  const float *mod[9];      // as MediumArgs::out
  float *coef[23];          // the reference arrays (enum biot_coef order)
  int nx, ny, nz;
  float DT, Cv, Cp;
};
struct RefArgs {
  const float *ref[23];     // reference arrays on the device
  int64_t n;
  double H;                 // 1/dx as the kernels hold it
  uint32_t *stats_u32;      // 23 maxima of |ref| (float bits)
  unsigned long long *stats_u64;  // forced q response (double bits)
  void *kout;               // 22 kernel arrays of the storage type
  int scaled;
  double ratio[NKC], inv_ks[NKC];
};
hipError_t launch_moduli(const MediumArgs &m, int64_t n, hipStream_t s);
hipError_t launch_coef_ref(const CoefRefArgs &a, hipStream_t s);
hipError_t launch_ref_stats(const RefArgs &a, hipStream_t s);
hipError_t launch_ref_to_kernel(int prec, const RefArgs &a, hipStream_t s);

// Batched launches over many shots (one grid for all; device-resident argument tables).
struct BatchArgs {
  const StepArgs *shots;    // device, one per shot (gx, gy, gz, zchunk set for the shape)
  const int *tile_start;    // device, nshots + 1 prefix sums of the shots' tile counts
  int nshots, total_tiles;
  int *queue;               // device, 16 ints (zero between launches): per XCD group the tile
                            // counter and the count of finished blocks; null: one block per tile
};
// blocks: the persistent grid size (a multiple of 8)
hipError_t launch_step_batch(int prec, int order, bool vel, const BatchArgs &b, LaunchShape sh,
                             int blocks, hipStream_t s);
hipError_t launch_source_batch(int prec, const SourceArgs *tab, int nshots, const int *it,
                               hipStream_t s);
hipError_t launch_record_batch(int prec, const RecordArgs *tab, int nshots, int64_t max_nrec,
                               int max_ncomp, const int *it, hipStream_t s);
// *it += 1 / *it = v on the device (kernels, so they may sit in a captured graph)
hipError_t launch_advance(int *it, hipStream_t s);
hipError_t launch_set_counter(int *it, int v, hipStream_t s);
// multiply a stored array by f (lo: its residuals or null); carries use prec of the carry
// type, i.e. pass PREC_FP32 for FP16C/FP16T carries and PREC_FP16 for FP16 carries
hipError_t launch_rescale(int prec, void *p, int8_t *lo, int64_t n, double f, hipStream_t s);
// the same for an FP16M tau or p array of n values (lo: its tagged residual pointer)
hipError_t launch_rescale_pk(void *p, int8_t *lo, int64_t n, double f, hipStream_t s);
// element e of every 4-element record times f[e] (carry groups whose carries scale differently)
hipError_t launch_rescale4(int prec, void *p, int64_t nrec, const double f[4], hipStream_t s);
// the xz slice y = j of one field as float values (stored value times scale): out[k*nx + i]
hipError_t launch_slice_xz(int prec, const void *f, const int8_t *lo, bool comp, double scale, int j, int nx,
                           int ny, int nz, float *out, hipStream_t s);
// per-cell packed copies kv (K[0..13]) and kt (K[14..21]) of the 22 coefficient arrays
hipError_t launch_pack_coef(int prec, const void *soa, void *kv, void *kt, int64_t n, hipStream_t s);
hipError_t launch_record(int prec, const RecordArgs &a, hipStream_t s);

}  // namespace biot
