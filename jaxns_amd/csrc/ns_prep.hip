// ns_prep.hip -- one-off FP64 preparation of the multivariate-normal likelihood on the device.
//
// log N(x; mu, C) on an affine prior x = loc + S q(u) is evaluated by the sampling kernel as
// c - |W q + b|^2 / 2 with W = L^-1 S, b = L^-1 (loc - mu), L = chol(C) and
// c = -D/2 log(2 pi) - sum log L_ii (tests/conftest.py:219-224 of the reference writes the same density
// through a triangular solve).  The factor and the two solves run once per likelihood, in FP64, in one
// workgroup; the FP32 parameter block the kernel reads is written directly.
#include "hipns_dev.h"

namespace hipns {

namespace {

constexpr int PREP_THREADS = 256;

__device__ double block_sum(double v, double* sh) {
  const int l = lane_id(), w = threadIdx.x / HIPNS_WAVE, nw = blockDim.x / HIPNS_WAVE;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double other = __shfl_down(v, o, 64);
    if (l + o < HIPNS_WAVE) v += other;
  }
  if (l == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < nw; ++i) v += sh[i];
    sh[0] = v;
  }
  __syncthreads();
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Lower Cholesky factor of the symmetric D x D matrix A (lower triangle read), column by column
// (left-looking): L_kk = sqrt(A_kk - sum_{j<k} L_kj^2), L_ik = (A_ik - sum_{j<k} L_ij L_kj) / L_kk.
// info = k + 1 at the first non-positive (or NaN) pivot, else 0.  Returns the info value to all threads.
__device__ int cholesky_block(const double* A, int D, double* L, int* info_out) {
  __shared__ double sh[PREP_THREADS / HIPNS_WAVE];
  __shared__ int bad;
  const int t = threadIdx.x;
  if (t == 0) bad = 0;
  for (int i = t; i < D * D; i += blockDim.x) L[i] = 0.0;
  __syncthreads();
  for (int k = 0; k < D; ++k) {
    double part = 0.0;
    for (int j = t; j < k; j += blockDim.x) part += L[k * D + j] * L[k * D + j];
    const double d = A[k * D + k] - block_sum(part, sh);
    if (!(d > 0.0)) {
      if (t == 0) bad = k + 1;
      __syncthreads();
      break;
    }
    const double lkk = sqrt(d);
    for (int i = k + 1 + t; i < D; i += blockDim.x) {
      double s = A[i * D + k];
      for (int j = 0; j < k; ++j) s -= L[i * D + j] * L[k * D + j];
      L[i * D + k] = s / lkk;
    }
    if (t == 0) L[k * D + k] = lkk;
    __syncthreads();
  }
  __syncthreads();
  const int r = bad;
  if (t == 0 && info_out) *info_out = r;
  return r;
}

__global__ void __launch_bounds__(PREP_THREADS) k_cholesky(const double* A, int D, double* L, int* info) {
  cholesky_block(A, D, L, info);
}

// W = L^-1 S (columns of S in parallel, one thread per column, forward substitution in FP64), b = L^-1
// (loc - mu) as column D, then the FP32 block [q-kind: dps][b: dps][W: dps x dps] (zero padding) and c.
__global__ void __launch_bounds__(PREP_THREADS) k_mvn_prepare(const double* C, const double* mu, const double* loc,
                                                              const double* S, const int32_t* qkind, int D, int dps,
                                                              double* L, double* Wd, float* prm, double* log_const,
                                                              int* info) {
  __shared__ double sh[PREP_THREADS / HIPNS_WAVE];
  // C == NULL: L is the given factor (a scale_tril), only the solves run
  if (C && cholesky_block(C, D, L, info)) return;
  if (!C && threadIdx.x == 0 && info) *info = 0;
  const int t = threadIdx.x;
  const int ncol = D + 1;
  for (int c = t; c < ncol; c += blockDim.x) {
    for (int i = 0; i < D; ++i) {
      double s = c < D ? S[i * D + c] : loc[i] - mu[i];
      for (int j = 0; j < i; ++j) s -= L[i * D + j] * Wd[j * ncol + c];
      Wd[i * ncol + c] = s / L[i * D + i];
    }
  }
  __syncthreads();
  for (int i = t; i < dps; i += blockDim.x) {
    prm[i] = i < D ? (float)qkind[i] : 0.0f;
    prm[dps + i] = i < D ? (float)Wd[i * ncol + D] : 0.0f;
  }
  for (int e = t; e < dps * dps; e += blockDim.x) {
    const int i = e / dps, j = e - i * dps;
    prm[2 * dps + e] = (i < D && j < D) ? (float)Wd[i * ncol + j] : 0.0f;
  }
  double part = 0.0;
  for (int i = t; i < D; i += blockDim.x) part += log(L[i * D + i]);
  const double ld = block_sum(part, sh);
  if (t == 0) *log_const = -0.5 * (double)D * 1.83787706640934548356 - ld;
}

}  // namespace

hipError_t cholesky_f64(const double* A, int D, double* L, int* info, hipStream_t s) {
  hipLaunchKernelGGL(k_cholesky, dim3(1), dim3(PREP_THREADS), 0, s, A, D, L, info);
  return hipGetLastError();
}

size_t mvn_prepare_ws(int D) { return (size_t)D * (D + 1) * sizeof(double); }

hipError_t mvn_prepare_f64(const double* C, const double* mu, const double* loc, const double* S,
                           const int32_t* qkind, int D, int dps, double* L, void* ws, float* prm, double* log_const,
                           int* info, hipStream_t s) {
  hipLaunchKernelGGL(k_mvn_prepare, dim3(1), dim3(PREP_THREADS), 0, s, C, mu, loc, S, qkind, D, dps, L, (double*)ws,
                     prm, log_const, info);
  return hipGetLastError();
}

}  // namespace hipns
