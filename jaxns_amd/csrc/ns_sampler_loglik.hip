// ns_sampler_loglik.hip -- batched likelihood through the sampler's device code.
#include "ns_sampler_impl.h"

namespace hipns {

// ---------------------------------------------------------------------------
// batched likelihood (Model.forward over a batch) through the same device code:
// row i of U plays the role of chain i of a single wavefront tile.
// ---------------------------------------------------------------------------
template <int LPC, int NJ, int FAM>
__global__ void __launch_bounds__(256) k_loglik(const float* U, int64_t B, double* out, const hipns_run_desc_t* runs,
                                                hipns_buffers_t Bf, int K, LinRegLds lrd) {
  constexpr int CW = HIPNS_WAVE / LPC;
  extern __shared__ float lds[];
  const hipns_run_desc_t run = runs[0];
  const int dps = Bf.dps;
  const int D = run.D;
  const int DPr = ((D + 8 * LPC - 1) / (8 * LPC)) * (8 * LPC);
  const int nj = DPr / LPC;
  MvnLds mv;
  GmmLds gm;
  LinRegLds lr = lrd;
  const int wpb = blockDim.x / HIPNS_WAVE;
  load_params(FAM, Bf.params + run.param_off, dps, D, K, lds, mv, gm, LPC, &lr, wpb, NJ / 8);
  const int wave = threadIdx.x / HIPNS_WAVE;
  const int l = lane_id();
  const int c = l % CW, h = l / CW;
  const int64_t row = ((int64_t)blockIdx.x * wpb + wave) * CW + c;
  float up[NJ];
  const bool ok = row < B;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    int d = owned_dim<LPC>(h, j);
    up[j] = (ok && j < nj && d < D) ? U[row * dps + d] : 0.5f;
  }
  const uint32_t tfm = transform_mask<LPC, NJ>(FAM == HIPNS_FAM_MVN ? mv.tf : gm.tf, h, nj);
  double L;
  if constexpr (FAM == HIPNS_FAM_MVN)
    L = lik_mvn<LPC, NJ>(up, nj, D, dps, mv, run.log_const, tfm);
  else if constexpr (FAM == HIPNS_FAM_GMM)
    L = lik_gmm<LPC, NJ>(up, nj, D, gm, tfm);
  else
    L = lik_linreg<LPC, NJ, false, true>(up, nj, D, lr, tfm, wave, wpb, run.log_const);
  if (ok && h == 0) out[row] = L;
}

// the same with d logL / dU (MVN and mixture families), the device form the gradient-guided chains
// use: grad[row * dps + d] for d < D; the wavefront's gradient scratch follows the parameters in LDS
template <int LPC, int NJ, int FAM>
__global__ void __launch_bounds__(256) k_loglik_grad(const float* U, int64_t B, double* out, float* grad,
                                                     const hipns_run_desc_t* runs, hipns_buffers_t Bf, int K,
                                                     int gscr_off) {
  constexpr int CW = HIPNS_WAVE / LPC;
  extern __shared__ float lds[];
  const hipns_run_desc_t run = runs[0];
  const int dps = Bf.dps;
  const int D = run.D;
  const int DPr = ((D + 8 * LPC - 1) / (8 * LPC)) * (8 * LPC);
  const int nj = DPr / LPC;
  MvnLds mv;
  GmmLds gm;
  LinRegLds lr{};
  const int wpb = blockDim.x / HIPNS_WAVE;
  load_params(FAM, Bf.params + run.param_off, dps, D, K, lds, mv, gm, LPC, &lr, wpb, NJ / 8);
  const int wave = threadIdx.x / HIPNS_WAVE;
  const int l = lane_id();
  const int c = l % CW, h = l / CW;
  const int64_t row = ((int64_t)blockIdx.x * wpb + wave) * CW + c;
  float up[NJ], g[NJ];
  const bool ok = row < B;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    int d = owned_dim<LPC>(h, j);
    up[j] = (ok && j < nj && d < D) ? U[row * dps + d] : 0.5f;
  }
  const uint32_t tfm = transform_mask<LPC, NJ>(FAM == HIPNS_FAM_MVN ? mv.tf : gm.tf, h, nj);
  double L;
  if constexpr (FAM == HIPNS_FAM_MVN) {
    float* gscr = lds + gscr_off + (size_t)wave * grad_rows(NJ, LPC) * (CW + 1);
    L = lik_mvn<LPC, NJ, true>(up, nj, D, dps, mv, run.log_const, tfm, gscr, g, true);
  } else {
    L = lik_gmm<LPC, NJ, true>(up, nj, D, gm, tfm, g, true);
  }
  if (ok) {
    if (h == 0) out[row] = L;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int d = owned_dim<LPC>(h, j);
      if (j < nj && d < D) grad[row * dps + d] = g[j];
    }
  }
}

template <int LPC, int NJ, int FAM>
static hipError_t launch_loglik_grad_one(const float* U, int64_t B, double* out, float* grad,
                                         const hipns_run_desc_t* runs, const hipns_buffers_t& Bf, int K, int wpb,
                                         int gscr_off, size_t lds, hipStream_t s) {
  constexpr int CW = HIPNS_WAVE / LPC;
  int64_t groups = (B + CW - 1) / CW;
  dim3 grid((unsigned)((groups + wpb - 1) / wpb)), block(wpb * HIPNS_WAVE);
  hipLaunchKernelGGL((k_loglik_grad<LPC, NJ, FAM>), grid, block, lds, s, U, B, out, grad, runs, Bf, K, gscr_off);
  return hipGetLastError();
}

template <int FAM>
static hipError_t launch_loglik_grad_fam(const float* U, int64_t B, double* out, float* grad,
                                         const hipns_run_desc_t* runs, const hipns_buffers_t& Bf, int K, int lpc,
                                         int wpb, int dmax, int gscr_off, size_t lds, hipStream_t s) {
  if constexpr (FAM == HIPNS_FAM_GMM) {
    if (lpc == 8 && dmax <= 128) {
      if (dmax <= 64) return launch_loglik_grad_one<8, 8, FAM>(U, B, out, grad, runs, Bf, K, wpb, gscr_off, lds, s);
      return launch_loglik_grad_one<8, 16, FAM>(U, B, out, grad, runs, Bf, K, wpb, gscr_off, lds, s);
    }
  }
  if (lpc == 4 || lpc == 8 || dmax > 64) {
    if (dmax <= 32) return launch_loglik_grad_one<4, 8, FAM>(U, B, out, grad, runs, Bf, K, wpb, gscr_off, lds, s);
    if (dmax <= 64) return launch_loglik_grad_one<4, 16, FAM>(U, B, out, grad, runs, Bf, K, wpb, gscr_off, lds, s);
    return launch_loglik_grad_one<4, 32, FAM>(U, B, out, grad, runs, Bf, K, wpb, gscr_off, lds, s);
  }
  if (dmax <= 16) return launch_loglik_grad_one<2, 8, FAM>(U, B, out, grad, runs, Bf, K, wpb, gscr_off, lds, s);
  if (dmax <= 32) return launch_loglik_grad_one<2, 16, FAM>(U, B, out, grad, runs, Bf, K, wpb, gscr_off, lds, s);
  return launch_loglik_grad_one<2, 32, FAM>(U, B, out, grad, runs, Bf, K, wpb, gscr_off, lds, s);
}

template <int LPC, int NJ, int FAM>
static hipError_t launch_loglik_one(const float* U, int64_t B, double* out, const hipns_run_desc_t* runs,
                                   const hipns_buffers_t& Bf, int K, int wpb, size_t lds, hipStream_t s,
                                   const LinRegLds& lr) {
  constexpr int CW = HIPNS_WAVE / LPC;
  int64_t groups = (B + CW - 1) / CW;
  dim3 grid((unsigned)((groups + wpb - 1) / wpb)), block(wpb * HIPNS_WAVE);
  hipLaunchKernelGGL((k_loglik<LPC, NJ, FAM>), grid, block, lds, s, U, B, out, runs, Bf, K, lr);
  return hipGetLastError();
}

// the same instance choice as the sampler launchers (sampler_nj)
template <int FAM>
static hipError_t launch_loglik_fam(const float* U, int64_t B, double* out, const hipns_run_desc_t* runs,
                                   const hipns_buffers_t& Bf, int K, int lpc, int wpb, int dmax, size_t lds,
                                   hipStream_t s, const LinRegLds& lr) {
  if constexpr (FAM == HIPNS_FAM_GMM) {
    if (lpc == 8 && dmax <= 128) {
      if (dmax <= 64) return launch_loglik_one<8, 8, FAM>(U, B, out, runs, Bf, K, wpb, lds, s, lr);
      return launch_loglik_one<8, 16, FAM>(U, B, out, runs, Bf, K, wpb, lds, s, lr);
    }
  }
  if (lpc == 4 || lpc == 8 || dmax > 64) {
    if (dmax <= 32) return launch_loglik_one<4, 8, FAM>(U, B, out, runs, Bf, K, wpb, lds, s, lr);
    if (dmax <= 64) return launch_loglik_one<4, 16, FAM>(U, B, out, runs, Bf, K, wpb, lds, s, lr);
    return launch_loglik_one<4, 32, FAM>(U, B, out, runs, Bf, K, wpb, lds, s, lr);
  }
  if (dmax <= 16) return launch_loglik_one<2, 8, FAM>(U, B, out, runs, Bf, K, wpb, lds, s, lr);
  if (dmax <= 32) return launch_loglik_one<2, 16, FAM>(U, B, out, runs, Bf, K, wpb, lds, s, lr);
  return launch_loglik_one<2, 32, FAM>(U, B, out, runs, Bf, K, wpb, lds, s, lr);
}

// adjoint statistics of the regression likelihood over a batch of points: logL, |r|^2 and y^T r
// per row of U (one wavefront per chain group; the host derives d logL / d(log sigma, log tau))
template <int LPC, int NJ>
__global__ void __launch_bounds__(64) k_linreg_stats(const float* U, int64_t B, double* logL, double* rss, double* yr,
                                                     const hipns_run_desc_t* runs, hipns_buffers_t Bf, LinRegLds lrd) {
  constexpr int CW = HIPNS_WAVE / LPC;
  extern __shared__ float lds[];
  const hipns_run_desc_t run = runs[0];
  const int dps = Bf.dps;
  const int D = run.D;
  const int DPr = ((D + 8 * LPC - 1) / (8 * LPC)) * (8 * LPC);
  const int nj = DPr / LPC;
  MvnLds mv;
  GmmLds gm;
  LinRegLds lr = lrd;
  load_params(HIPNS_FAM_LINREG, Bf.params + run.param_off, dps, D, 1, lds, mv, gm, LPC, &lr, 1, NJ / 8);
  const int l = lane_id();
  const int c = l % CW, h = l / CW;
  const int64_t row = (int64_t)blockIdx.x * CW + c;
  float up[NJ];
  const bool ok = row < B;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    int d = owned_dim<LPC>(h, j);
    up[j] = (ok && j < nj && d < D) ? U[row * dps + d] : 0.5f;
  }
  const uint32_t tfm = transform_mask<LPC, NJ>(lr.tf, h, nj);
  double ss = 0.0, sy = 0.0;
  const double L = lik_linreg<LPC, NJ, true>(up, nj, D, lr, tfm, 0, 1, run.log_const, &ss, &sy);
  if (ok && h == 0) {
    logL[row] = L;
    rss[row] = ss;
    yr[row] = sy;
  }
}

hipError_t launch_linreg_stats(const float* U, int64_t B, double* logL, double* rss, double* yr,
                               const hipns_run_desc_t* runs, const hipns_buffers_t& Bf, int lpc, int dmax,
                               size_t lds, hipStream_t s, const LinRegLds& lr) {
  if (dmax > 64) lpc = 4;
  const int CW = HIPNS_WAVE / lpc;
  dim3 grid((unsigned)((B + CW - 1) / CW)), block(HIPNS_WAVE);
  if (lpc == 2) {
    if (dmax <= 16)
      hipLaunchKernelGGL((k_linreg_stats<2, 8>), grid, block, lds, s, U, B, logL, rss, yr, runs, Bf, lr);
    else if (dmax <= 32)
      hipLaunchKernelGGL((k_linreg_stats<2, 16>), grid, block, lds, s, U, B, logL, rss, yr, runs, Bf, lr);
    else
      hipLaunchKernelGGL((k_linreg_stats<2, 32>), grid, block, lds, s, U, B, logL, rss, yr, runs, Bf, lr);
  } else {
    if (dmax <= 32)
      hipLaunchKernelGGL((k_linreg_stats<4, 8>), grid, block, lds, s, U, B, logL, rss, yr, runs, Bf, lr);
    else if (dmax <= 64)
      hipLaunchKernelGGL((k_linreg_stats<4, 16>), grid, block, lds, s, U, B, logL, rss, yr, runs, Bf, lr);
    else
      hipLaunchKernelGGL((k_linreg_stats<4, 32>), grid, block, lds, s, U, B, logL, rss, yr, runs, Bf, lr);
  }
  return hipGetLastError();
}

hipError_t launch_loglik_grad(const float* U, int64_t B, double* out, float* grad, const hipns_run_desc_t* runs,
                              const hipns_buffers_t& Bf, const hipns_launch_t& L, int dmax, int gscr_off, size_t lds,
                              hipStream_t s) {
  const int lpc = L.lanes_per_chain, wpb = L.waves_per_block, K = L.gmm_K;
  if (L.family == HIPNS_FAM_GMM)
    return launch_loglik_grad_fam<HIPNS_FAM_GMM>(U, B, out, grad, runs, Bf, K, lpc, wpb, dmax, gscr_off, lds, s);
  if (L.family == HIPNS_FAM_MVN)
    return launch_loglik_grad_fam<HIPNS_FAM_MVN>(U, B, out, grad, runs, Bf, K, lpc, wpb, dmax, gscr_off, lds, s);
  return hipErrorInvalidValue;
}

hipError_t launch_loglik(const float* U, int64_t B, double* out, const hipns_run_desc_t* runs,
                         const hipns_buffers_t& Bf, const hipns_launch_t& L, int dmax, size_t lds, hipStream_t s,
                         const LinRegLds& lr) {
  const int lpc = L.lanes_per_chain, wpb = L.waves_per_block, K = L.gmm_K;
  if (L.family == HIPNS_FAM_GMM) return launch_loglik_fam<HIPNS_FAM_GMM>(U, B, out, runs, Bf, K, lpc, wpb, dmax, lds, s, lr);
  if (L.family == HIPNS_FAM_LINREG)
    return launch_loglik_fam<HIPNS_FAM_LINREG>(U, B, out, runs, Bf, K, lpc, wpb, dmax, lds, s, lr);
  return launch_loglik_fam<HIPNS_FAM_MVN>(U, B, out, runs, Bf, K, lpc, wpb, dmax, lds, s, lr);
}

}  // namespace hipns
