// ns_finals.hip -- the per-sample finals of _to_results (nested_samplers/sharded/sharded_static.py:706-745)
// on the device, after count_crossed_edges and the evidence scan:
//   log_dp_mean = log dZ normalised (normalise_log_space), H_mean (unstable form E_post[log L] - log Z,
//   with the stable form sum dp log X where the first is not finite), and, per sample in tree order,
//   U (FP64), X = Model.transform(U) (loc + S q(U)) and log_posterior_density = log L + log prior(X).
// The reductions run in a fixed grouping (tiles of 2048 samples, then one block over the tile partials),
// so results do not depend on scheduling.
#include "hipns_dev.h"

namespace hipns {

namespace {

constexpr int FIN_THREADS = 256;
constexpr int FIN_PER_THREAD = 8;
constexpr int FIN_TILE = FIN_THREADS * FIN_PER_THREAD;

template <typename T, typename Op>
__device__ T block_reduce(T v, Op op, T* sh) {
  const int l = lane_id(), w = threadIdx.x / HIPNS_WAVE, nw = blockDim.x / HIPNS_WAVE;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T other = __shfl_down(v, o, 64);
    if (l + o < HIPNS_WAVE) v = op(v, other);
  }
  if (l == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < nw; ++i) v = op(v, sh[i]);
    sh[0] = v;
  }
  __syncthreads();
  T r = sh[0];
  __syncthreads();
  return r;
}

struct MaxOp {
  __device__ double operator()(double a, double b) const {
    // jnp.max propagates NaN
    if (__builtin_isnan(a) || __builtin_isnan(b)) return __builtin_nan("");
    return a > b ? a : b;
  }
};
struct AddOp {
  __device__ double operator()(double a, double b) const { return a + b; }
};

// tile maxima of log dZ
__global__ void __launch_bounds__(FIN_THREADS) k_fin_max(const double* log_dZ, int64_t M, double* part) {
  __shared__ double sh[FIN_THREADS / HIPNS_WAVE];
  const int64_t base = (int64_t)blockIdx.x * FIN_TILE;
  double m = -__builtin_inf();
  for (int k = 0; k < FIN_PER_THREAD; ++k) {
    const int64_t i = base + (int64_t)k * FIN_THREADS + threadIdx.x;
    if (i < M) m = MaxOp()(m, log_dZ[i]);
  }
  m = block_reduce(m, MaxOp(), sh);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// one block: the maximum of the tile maxima -> scal[0]
__global__ void __launch_bounds__(1024) k_fin_max_all(const double* part, int nt, double* scal) {
  __shared__ double sh[1024 / HIPNS_WAVE];
  double m = -__builtin_inf();
  for (int i = threadIdx.x; i < nt; i += blockDim.x) m = MaxOp()(m, part[i]);
  m = block_reduce(m, MaxOp(), sh);
  if (threadIdx.x == 0) scal[0] = m;
}

// tile sums of exp(log dZ - max)
__global__ void __launch_bounds__(FIN_THREADS) k_fin_expsum(const double* log_dZ, int64_t M, const double* scal,
                                                            double* part) {
  __shared__ double sh[FIN_THREADS / HIPNS_WAVE];
  const double mx = scal[0];
  const bool fin = __builtin_isfinite(mx);
  const int64_t base = (int64_t)blockIdx.x * FIN_TILE;
  double s = 0.0;
  for (int k = 0; k < FIN_PER_THREAD; ++k) {
    const int64_t i = base + (int64_t)k * FIN_THREADS + threadIdx.x;
    if (i < M && fin) s += exp(log_dZ[i] - mx);
  }
  s = block_reduce(s, AddOp(), sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// one block: log of the normaliser, max + log(sum) -> scal[1]  (-inf for an all -inf input; NaN if any)
__global__ void __launch_bounds__(1024) k_fin_lnorm(const double* part, int nt, double* scal) {
  __shared__ double sh[1024 / HIPNS_WAVE];
  double s = 0.0;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) s += part[i];
  s = block_reduce(s, AddOp(), sh);
  if (threadIdx.x == 0) {
    const double mx = scal[0];
    scal[1] = __builtin_isfinite(mx) ? mx + log(s) : mx;
  }
}

// log dp = log dZ - log normaliser; tile sums of dp * L (L -> 0 where dp = 0) and dp * log X
__global__ void __launch_bounds__(FIN_THREADS) k_fin_apply(const double* log_dZ, const double* log_L,
                                                           const double* log_X, int64_t M, const double* scal,
                                                           double* log_dp, double* part) {
  __shared__ double sh[FIN_THREADS / HIPNS_WAVE];
  const double ln = scal[1];
  const int64_t base = (int64_t)blockIdx.x * FIN_TILE;
  double s1 = 0.0, s2 = 0.0;
  for (int k = 0; k < FIN_PER_THREAD; ++k) {
    const int64_t i = base + (int64_t)k * FIN_THREADS + threadIdx.x;
    if (i >= M) continue;
    // normalise_log_space: a -inf normaliser leaves every weight at -inf
    const double lp = (ln == -__builtin_inf()) ? -__builtin_inf() : log_dZ[i] - ln;
    log_dp[i] = lp;
    const double p = exp(lp);
    const double L = (lp == -__builtin_inf()) ? 0.0 : log_L[i];
    s1 += p * L;
    s2 += p * log_X[i];
  }
  s1 = block_reduce(s1, AddOp(), sh);
  s2 = block_reduce(s2, AddOp(), sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

// one block: H_out = {unstable, stable, chosen, log normaliser}
__global__ void __launch_bounds__(1024) k_fin_H(const double* part, int nt, const double* scal, double log_Z_mean,
                                                double* H_out) {
  __shared__ double sh[1024 / HIPNS_WAVE];
  double s1 = 0.0, s2 = 0.0;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) {
    s1 += part[2 * i];
    s2 += part[2 * i + 1];
  }
  s1 = block_reduce(s1, AddOp(), sh);
  s2 = block_reduce(s2, AddOp(), sh);
  if (threadIdx.x == 0) {
    const double h_unstable = -(s1 - log_Z_mean);
    // -(sum dp * (-log X)): the compression form
    const double h_stable = s2;
    H_out[0] = h_unstable;
    H_out[1] = h_stable;
    H_out[2] = __builtin_isfinite(h_unstable) ? h_unstable : h_stable;
    H_out[3] = scal[1];
  }
}

// Per sample b (row src = perm[b] of the collection, or b): U in FP64, X = loc + S q(U) with S lower
// triangular with a positive diagonal (DIAG: only its diagonal is read), and log L + log prior.  Rows
// at or past `valid` are the collection's placeholders (U = 0.5).
template <bool DIAG>
__global__ void __launch_bounds__(256) k_prior_finals(const float* samp_U, int ldu, const int64_t* perm,
                                                      int64_t valid, int64_t M, int D, const double* loc,
                                                      const double* S, const int32_t* qkind, const double* log_L,
                                                      double* U_out, double* X_out, double* log_post) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= M) return;
  const int64_t src = perm ? perm[b] : b;
  const bool placeholder = src >= valid;
  const float* u = samp_U + src * (int64_t)ldu;
  // log density of loc + S q: sum of the standard normal terms of the ndtri slots minus log|det S|
  double lp = 0.0;
  for (int i = 0; i < D; ++i) lp -= log(S[(int64_t)i * D + i]);
  constexpr double HALF_LOG_2PI = 0.91893853320467274178;
  for (int i = 0; i < D; ++i) {
    const double ui = placeholder ? 0.5 : (double)u[i];
    if (U_out) U_out[b * D + i] = ui;
    double q = ui;
    if (qkind[i]) {
      q = ndtri_f64(ui);
      lp += -0.5 * q * q - HALF_LOG_2PI;
    }
    if (DIAG) {
      if (X_out) X_out[b * D + i] = loc[i] + S[(int64_t)i * D + i] * q;
    } else if (X_out) {
      // X_out doubles as the store of q until row i is formed: X_i needs q_0..q_i
      X_out[b * D + i] = q;
    }
  }
  if (!DIAG && X_out) {
    for (int i = D - 1; i >= 0; --i) {
      double x = loc[i];
      for (int j = 0; j <= i; ++j) x = fma(S[(int64_t)i * D + j], X_out[b * D + j], x);
      X_out[b * D + i] = x;
    }
  }
  if (log_post) log_post[b] = log_L[b] + lp;
}

}  // namespace

size_t results_finals_ws(int64_t M) {
  const int64_t nt = (M + FIN_TILE - 1) / FIN_TILE;
  return (size_t)(4 * nt + 8) * sizeof(double);
}

hipError_t results_finals(const double* log_dZ, const double* log_L, const double* log_X, int64_t M,
                          double log_Z_mean, double* log_dp, double* H_out, void* ws, hipStream_t s) {
  const int nt = (int)((M + FIN_TILE - 1) / FIN_TILE);
  double* scal = (double*)ws;
  double* p1 = scal + 8;
  double* p2 = p1 + nt;
  double* p3 = p2 + nt;
  hipLaunchKernelGGL(k_fin_max, dim3(nt), dim3(FIN_THREADS), 0, s, log_dZ, M, p1);
  hipLaunchKernelGGL(k_fin_max_all, dim3(1), dim3(1024), 0, s, p1, nt, scal);
  hipLaunchKernelGGL(k_fin_expsum, dim3(nt), dim3(FIN_THREADS), 0, s, log_dZ, M, scal, p2);
  hipLaunchKernelGGL(k_fin_lnorm, dim3(1), dim3(1024), 0, s, p2, nt, scal);
  hipLaunchKernelGGL(k_fin_apply, dim3(nt), dim3(FIN_THREADS), 0, s, log_dZ, log_L, log_X, M, scal, log_dp, p3);
  hipLaunchKernelGGL(k_fin_H, dim3(1), dim3(1024), 0, s, p3, nt, scal, log_Z_mean, H_out);
  return hipGetLastError();
}

hipError_t prior_finals(const float* samp_U, int ldu, const int64_t* perm, int64_t valid, int64_t M, int D,
                        const double* loc, const double* S, const int32_t* qkind, int diag, const double* log_L,
                        double* U_out, double* X_out, double* log_post, hipStream_t s) {
  const dim3 grid((unsigned)((M + 255) / 256)), block(256);
  if (diag)
    hipLaunchKernelGGL(k_prior_finals<true>, grid, block, 0, s, samp_U, ldu, perm, valid, M, D, loc, S, qkind,
                       log_L, U_out, X_out, log_post);
  else
    hipLaunchKernelGGL(k_prior_finals<false>, grid, block, 0, s, samp_U, ldu, perm, valid, M, D, loc, S, qkind,
                       log_L, U_out, X_out, log_post);
  return hipGetLastError();
}

}  // namespace hipns
