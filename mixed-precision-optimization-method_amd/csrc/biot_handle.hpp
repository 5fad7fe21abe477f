// The state behind a biot_handle: everything the reference kept in blockInfo
// (common.h:6-133) plus the time-loop state of callCUDAKernel (kernel.cu:1728-1760).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/biot.h"
#include "biot_adjoint.hpp"
#include "biot_device.hpp"

struct biot_handle {
  biot_grid g{};
  int prec = 0;
  size_t esz = 4;    // field / coefficient element bytes
  size_t csz = 4;    // C-PML carry element bytes
  int64_t n = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;

  // device memory
  void *d_fields = nullptr;          // 13 * n elements (FP16M: tau and p as their binary16 parts)
  int8_t *d_lo = nullptr;  // FP16C / FP16T residuals, 13 * n (compensated fields used); FP16M residual records, 4 * n
  void *d_coef = nullptr;            // 22 * n (FP16M: the medium records, 18 bytes per cell)
  void *d_kpack = nullptr;           // the coefficients packed per cell (velocity, then stress block)
  void *d_carry[3] = {nullptr, nullptr, nullptr};
  int *d_flag = nullptr;             // first step with a non-finite receiver sample
                                     // (biot::NONFINITE_NONE: none since the last clear)
  void *d_prof[3][4] = {};           // e_int, a_int, e_half, a_half per axis
  int64_t slab[3] = {0, 0, 0};
  int L[3] = {0, 0, 0}, R[3] = {0, 0, 0};
  int64_t device_bytes = 0;

  // host copies
  std::vector<float> coef_ref[BIOT_NCOEF];  // reference representation
  std::vector<float> prof[3][4];            // d, d2, e, e2 per axis (before the x0.01 pack)
  bool have_coef = false;
  float ks[biot::NKC];        // per-coefficient-array scale (binary16 storage)
  double kfac[biot::NKC];     // storage-unit coefficient / physical coefficient, per array
  float fscale[biot::NFIELD]; // per-field storage scale
  // FP16M: the 9 cell moduli the medium records are built from (M HH H2u C C1 C2 rho rhof mu,
  // float, as biot::Moduli) and the constants of the kernels' coefficient formulas
  std::vector<float> mod[9];
  float mc[biot::NMC] = {};

  // source
  std::vector<float> wavelet;
  void *d_wavelet = nullptr;
  int64_t src_cell = 0;
  uint32_t src_mask = 0;
  // receivers
  void *d_rec = nullptr;
  void *d_traces = nullptr;
  double *d_traces64 = nullptr;  // FP64 handles: the traces in double as well
  int64_t nrec = 0;
  uint32_t rec_mask = 0;
  int comp[biot::NFIELD] = {};
  int ncomp = 0;
  int nt_max = 0;

  int it = 0;
  double last_ms = 0.0;
  biot::LaunchShape shape{64, 4, 16};

  ~biot_handle();
  void *field_ptr(int f) const;
  // residual array of field f (FP16C / FP16T, stride 1), or for FP16M's tau and p the base of
  // the residual records (stride PK_RECORD) plus the field's nibble (the kernels' runtime
  // loaders read it from the pointer's low 3 bits), or null
  int8_t *lo_ptr(int f) const;
  // where a host copy of field f's residuals starts: FP16M's record base, else lo_ptr(f)
  int8_t *lo_copy_ptr(int f) const { return packed(f) ? d_lo : lo_ptr(f); }
  size_t lo_stride(int f) const { return packed(f) ? biot::PK_RECORD : 1; }
  size_t lo_bytes() const;
  // FP16M keeps tau and p as binary16 + a 4-bit residual: the binary16 part in the field array,
  // the residual in a per-cell record shared by the seven fields
  bool packed(int f) const { return prec == BIOT_FP16M && ((biot::PK_FIELDS >> f) & 1u); }
  size_t fields_bytes() const;
  size_t mt_offset() const { return ((size_t)8 * n + 255) & ~(size_t)255; }  // FP16M stress records
  size_t mu_offset() const { return (mt_offset() + (size_t)8 * n + 255) & ~(size_t)255; }  // FP16M mu
  bool half() const
  {
    return prec == BIOT_FP16 || prec == BIOT_FP16C || prec == BIOT_FP16T || prec == BIOT_FP16M;
  }
  void fill_args(biot::StepArgs &a) const;
  int allocate();
  int upload_profiles();
  int upload_coefficients();
  int pack_coefficients();           // d_coef -> d_kpack
  int build_adjoint_coefs();         // FP16M: d_kpack for the adjoint, from K_last
  int upload_kernel_coefs(const std::vector<double> (&K)[biot::NKC]);
  int upload_from_ref(const float *d_ref);      // 23 reference arrays on the device
  int upload_medium();                          // FP16M: mod -> medium records, mc
  void set_scales(const double amaxK[biot::NKC], double resp);
  int rescale_state(const float old_scale[biot::NFIELD]);

  // inversion parameters (biot_set_params) and the adjoint's buffers
  std::vector<double> par[8];                // lambda mu alpha M rho rhof C1 tphi
  std::vector<double> K_last[biot::NKC];     // kernel coefficients built from par
  std::vector<void *> ckpt;                  // checkpoint states, every ckpt_every steps
  int ckpt_every = 0, ckpt_nt = 0;
  // ckpt_every == 1: ckpt[0] is one block of ckpt_nt + 1 states, state it being step it's
  // state after its source (the last one the state after the run), which the adjoint reads
  // in place of segment recomputation
  void *d_slots = nullptr;                   // segment states (ckpt_every + 1)
  void *d_lam = nullptr;                     // adjoint fields, carries, derivative adjoints
  double *d_grad = nullptr;                  // 22 FP64 coefficient gradients
  float *d_gpart = nullptr;                  // 22 float per-segment partial sums (not FP64 handles)
  float *d_adjsrc = nullptr;                 // adjoint source traces
  size_t state_bytes() const;
  size_t state_off(int piece) const;  // 0 fields, 1..3 carries of x, y, z, 4 total
  size_t state_lo_off() const;        // the residual block (lo_bytes())
  void save_state(void *dst);
  void load_state(const void *src);
  void free_adjoint();
  void forget_model();  // drop parameter-built coefficients and every adjoint buffer
  // marks: null, or 3 events per step recorded before velocity, between, after stress;
  // states: null, or room for nsteps states, step s's state saved after its source
  int enqueue_steps(int nsteps, hipEvent_t *marks, char *states = nullptr);

  // argument blocks shared by the single-grid and batched drivers
  void source_args(biot::SourceArgs &sa) const;
  void record_args(biot::RecordArgs &ra) const;
  int clear_nonfinite();                     // enqueue a clear of d_flag on the stream
  int check_nonfinite();                     // synchronize; BIOT_ERR_NONFINITE if it was set
  size_t lam_bytes() const;                  // adjoint wavefields, carries, derivative adjoints
  int alloc_adjoint(int every);              // d_slots (every+1 states; none for 0), d_lam, d_grad, d_adjsrc
  void adj_args(biot::AdjArgs &A) const;     // after alloc_adjoint
  void inject_args(biot::InjectArgs &I) const;
};

biot::LaunchShape default_shape(int prec, int order);
// reference arrays -> unscaled kernel coefficients on the host (1/dx and 0.01 folded)
void host_kernel_coefs(double Hinv, const std::vector<float> (&ref)[BIOT_NCOEF],
                       std::vector<double> (&K)[biot::NKC]);
