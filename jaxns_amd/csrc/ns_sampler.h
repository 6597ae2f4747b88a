// ns_sampler.h -- launch interface of the fused sampling kernel (ns_sampler.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hipns.h"

namespace hipns {

enum SampleMode : int {
  MODE_SLICE = 0,  // UniDimSliceSampler chains of one shell (Philox draws)
  MODE_UNIF = 1,   // UniformSampler rejection chains of one shell
  MODE_INIT = 2,   // initial live points: uniform until logL > -inf
  MODE_INJ = 3,    // slice chains with injected draws and a recorded trace
  MODE_GSLICE = 4, // slice chains that use grad logL: in the fused kernel gradient_guided reflections and / or
                   // gradient_slice directions (run codes sampler 2-4), in the stepwise form gradient_guided
};

struct SampleArgs {
  const hipns_run_desc_t* runs;
  hipns_run_state_t* states;
  hipns_buffers_t B;
  int32_t max_steps;
  int32_t gmm_K;
  int32_t alpha_off;  // LDS float offset of the alpha table (after the model parameters)
  int32_t dir_ring;   // slice directions drawn ahead per chain (LDS ring depth; 0 = drawn at each slice start)
  // regression family: shared data of the batch
  const float* lr_r0;
  const uint16_t* lr_xh;
  const uint16_t* lr_xl;
  const float* lr_xf;
  int32_t lr_npad;
  int32_t lr_panels;
  int32_t lr_terms;
  int32_t ring_off;   // LDS float offset of the direction rings (16-byte aligned)
  // injected-draw mode
  int32_t inj_buf;
  int32_t inj_C;
  double inj_contour;
  const int32_t* inj_seed;
  const float* inj_dirs;
  const float* inj_tu;
  float* trace_t;
  double* trace_logL;
  int32_t* trace_steps;
  float* ph_U;
  double* ph_logL;
};

// regression-family view of the model: per-run vectors and exchange buffers in LDS, the shared data
// rows in HBM.  X is stored as `panels` column panels of 16: element (row, col) of the hi part is at
// xh[((col / 16) * npad + row) * 16 + col % 16].
struct LinRegLds {
  float* tf;           // [dps] 1 where the prior of the weight is Normal (q = ndtri(u))
  float* p;            // [dps] w = p + s (.) q(u)
  float* s;            // [dps]
  float* w0;           // [dps] centre of the residual GEMM
  double inv2s2;       // 1 / (2 sigma^2), held as FP64 so that logL is smooth in sigma
  void* bops;          // [G][NG][3][64] B operands (8 bf16 each) published by every wavefront
  double* partial;     // [2][G][G][CW] per-wavefront partial sums (|r|^2, then y^T r)
  const float* r0;     // [npad] r0 = y - X w0 (computed in FP64, zero on padding rows)
  const uint16_t* xh;  // [panels][npad][16] bf16 hi part of the column-scaled X
  const uint16_t* xl;  // [panels][npad][16] bf16 lo part (X - hi)
  const float* y;      // [npad] targets (zero on padding rows), read by the adjoint statistics only
  const float* xf;     // [panels][npad][16] FP32 column-scaled X (the plain FP32 form)
  const double* f64;   // [3][dps] p, s and w0 in FP64 (global memory), read by the refinement form
  int npad, panels;
  int dps;             // stride of the per-run vectors
  int terms;           // 1: plain FP32 MFMA; 3: BF16x3; 4: adds X hi times the third bf16 part of w - w0;
                       // 5: the plain FP32 product with w - w0 formed in FP64 (FP64 ndtri), for refinement
};

// register slots per lane of the dispatched kernel instance for runs of at most dmax dimensions
// (the launchers select the same instance)
__host__ __device__ inline int sampler_nj(int dmax, int lpc) {
  if (lpc == 8) return dmax <= 64 ? 8 : 16;  // the mixture family, dmax <= 128
  if (lpc == 2 && dmax <= 64) return dmax <= 16 ? 8 : (dmax <= 32 ? 16 : 32);
  return dmax <= 32 ? 8 : (dmax <= 64 ? 16 : 32);
}

// lanes per chain of a launch: 8 only for the mixture family (no MFMA tile ties its layout), and only
// where the U row stride covers every dimension slot of the 8-lane instance (8 slots: 64 dimensions,
// 16 slots: 128), since the kernels read and write all owned slots of a row; else 4 above 64
// dimensions (the widest instance), else the requested 2 or 4
__host__ __device__ inline int sampler_lpc(int family, int lanes_per_chain, int dps, int dmax) {
  if (lanes_per_chain == 8)
    return (family == HIPNS_FAM_GMM && dmax <= 128 && dps >= 8 * sampler_nj(dmax, 8)) ? 8 : 4;
  return dps > 64 ? 4 : lanes_per_chain;
}

// MVN: row stride of the transposed whitening matrix in LDS for the instance (NJ, LPC): the instance's
// full width NJ * LPC in blocks of CW = 64 / LPC output rows, + 1 against bank conflicts.  A
// compile-time stride lets the kernel address every A operand with an immediate offset.
__host__ __device__ constexpr int mvn_row_stride(int nj, int lpc) {
  return ((nj * lpc + 64 / lpc - 1) / (64 / lpc)) * (64 / lpc) + 1;
}

// rows of one wavefront's gradient scratch (MVN, d logL / dU): the output rows of the MFMA blocks of
// 64 / lpc rows that cover the nj * lpc dimension slots (32 for lpc = 2 and nj = 8, where one 32-row
// block covers 16 slots); the scratch is [rows][64 / lpc + 1] floats
__host__ __device__ constexpr int grad_rows(int nj, int lpc) {
  return ((nj * lpc + 64 / lpc - 1) / (64 / lpc)) * (64 / lpc);
}

// floats of dynamic LDS one sampler block needs (layout of load_params in ns_sampler_impl.h)
inline size_t sampler_lds_floats(int fam, int dps, int D, int K, int lpc, int G = 1) {
  if (fam == HIPNS_FAM_LINREG) {
    const size_t NG = (size_t)sampler_nj(D, lpc) / 8;
    const size_t CW = (size_t)(64 / lpc);
    return (size_t)4 * dps + 4 + (size_t)G * NG * 3 * 64 * 4 + (size_t)2 * G * G * CW * 2;
  }
  if (fam == HIPNS_FAM_MVN) {
    // W^T holds a (zero) row for every dimension slot of the instance, so the kernel reads all NJ slots
    const int nj = sampler_nj(D, lpc);
    const int rows = dps > nj * lpc ? dps : nj * lpc;
    return (size_t)2 * dps + (size_t)rows * mvn_row_stride(nj, lpc);
  }
  return (size_t)(3 + K) * dps + 2 * K;
}

// Depth of the per-chain ring of slice directions drawn ahead by an instance (0: none).  Every
// wavefront step in which some chain of the wavefront begins a slice runs the direction draw for all
// of its lanes, and ~43 % of chain steps begin one, so nearly every step paid for a draw.  With a ring
// the draw runs only when a chain about to begin a slice has no direction left, and then tops up every
// chain of the wavefront by one, so a pass serves several chains' later slices (0.96 -> ~0.6 passes per
// step at depth 3 for 16 chains).  One ring slot is 64 lanes x nj floats per wavefront.  The regression
// family (whose step is its data-row GEMM) and the 8-slot mixture instance (at 128 VGPRs, where the
// ring's registers would cost a wavefront per SIMD) keep the draw at the slice start.
__host__ __device__ constexpr int dir_ring_max(int mode, int fam, int nj) {
  return (mode != MODE_SLICE || fam == HIPNS_FAM_LINREG || (fam == HIPNS_FAM_GMM && nj <= 8))
             ? 0
             : (nj <= 8 ? 3 : (nj <= 16 ? 2 : 0));
}

hipError_t launch_ext_step(const hipns_run_desc_t* runs, hipns_run_state_t* states, const hipns_buffers_t& B,
                           const hipns_ext_state_t& E, int mode, int w, int max_steps, int R, int max_chains, int lpc,
                           int wpb, int dmax, hipStream_t s);
hipError_t launch_linreg_stats(const float* U, int64_t B, double* logL, double* rss, double* yr,
                               const hipns_run_desc_t* runs, const hipns_buffers_t& Bf, int lpc, int dmax,
                               size_t lds, hipStream_t s, const LinRegLds& lr);
hipError_t launch_loglik_grad(const float* U, int64_t B, double* out, float* grad, const hipns_run_desc_t* runs,
                              const hipns_buffers_t& Bf, const hipns_launch_t& L, int dmax, int gscr_off, size_t lds,
                              hipStream_t s);
hipError_t launch_loglik(const float* U, int64_t B, double* out, const hipns_run_desc_t* runs,
                         const hipns_buffers_t& Bf, const hipns_launch_t& L, int dmax, size_t lds, hipStream_t s,
                         const LinRegLds& lr);

}  // namespace hipns

namespace hipns {
hipError_t launch_sampler_slice(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                                size_t lds_bytes, hipStream_t s);
hipError_t launch_sampler_unif(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                               size_t lds_bytes, hipStream_t s);
hipError_t launch_sampler_init(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                               size_t lds_bytes, hipStream_t s);
hipError_t launch_sampler_inj(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                              size_t lds_bytes, hipStream_t s);
hipError_t launch_sampler_gslice(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                                 size_t lds_bytes, hipStream_t s);
hipError_t launch_seed_indices(uint32_t k0, uint32_t k1, uint32_t it32, int first, int N, int C, int32_t* idx,
                               double* u, hipStream_t s);
hipError_t launch_slice_dirs(uint32_t k0, uint32_t k1, uint32_t it32, int C, int num_slices, int D, int lpc,
                             float* out, hipStream_t s);
}  // namespace hipns
