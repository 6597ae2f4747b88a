// ns_shell.h -- kernels of the shell loop (ns_shell.hip) and the to_results
// primitives (ns_results.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hipns.h"

namespace hipns {

__global__ void k_discard(const hipns_run_desc_t* runs, hipns_run_state_t* states, hipns_buffers_t B);
__global__ void k_replace(const hipns_run_desc_t* runs, hipns_run_state_t* states, hipns_buffers_t B, int init_mode);
__global__ void k_register(const hipns_run_desc_t* runs, hipns_run_state_t* states, hipns_buffers_t B,
                           hipns_term_cond_t term);
__global__ void k_check_term(const hipns_run_desc_t* runs, hipns_run_state_t* states, int R, hipns_term_cond_t term,
                             int set_no_seed, hipns_buffers_t B);
__global__ void k_init_state(hipns_run_state_t* states, int R);
__global__ void k_init_samples(const hipns_run_desc_t* runs, int R, hipns_buffers_t B);
__global__ void k_consume(const hipns_run_desc_t* runs, hipns_run_state_t* states, hipns_buffers_t B);
__global__ void k_consume_advance(const hipns_run_desc_t* runs, hipns_run_state_t* states, int R);

size_t replace_lds_bytes(int N);

// to_results primitives (ns_results.hip)
hipError_t evidence_scan(const double* logL, const double* n, int64_t M, int64_t stop, double* per8, double* fin8,
                         void* ws, size_t ws_bytes, hipStream_t s);
size_t evidence_scan_ws(int64_t M);
// the same scan with its workspace in a file-scope device array of HIPNS_SCAN_SCRATCH_BYTES (no allocation)
#define HIPNS_SCAN_SCRATCH_BYTES (32u << 20)
#define HIPNS_MAX_DEVICES 64
size_t evidence_scan_static_max();
hipError_t evidence_scan_static(const double* logL, const double* n, int64_t M, int64_t stop, double* per8,
                                double* fin8, hipStream_t s);
hipError_t argsort_stable(const double* keys, int64_t n, int32_t* perm, void* ws, size_t ws_bytes, hipStream_t s);
size_t argsort_ws(int64_t n);
hipError_t count_crossed_edges(const int32_t* sender, const double* logL, int64_t N, int64_t num_samples,
                               int64_t* samples_indices, int64_t* num_live, void* ws, size_t ws_bytes, hipStream_t s);
size_t count_crossed_edges_ws(int64_t N);
size_t batch_evidence_ws(int R, int64_t max_M);
hipError_t batch_evidence(const hipns_run_desc_t* runs, const hipns_run_state_t* states, int R, int64_t max_M,
                          const hipns_buffers_t& B, double* fin, int64_t* phantom, void* ws, hipStream_t s);

// one-off FP64 preparation of the multivariate-normal likelihood (ns_prep.hip)
hipError_t cholesky_f64(const double* A, int D, double* L, int* info, hipStream_t s);
size_t mvn_prepare_ws(int D);
hipError_t mvn_prepare_f64(const double* C, const double* mu, const double* loc, const double* S,
                           const int32_t* qkind, int D, int dps, double* L, void* ws, float* prm, double* log_const,
                           int* info, hipStream_t s);

// _to_results finals (ns_finals.hip)
size_t results_finals_ws(int64_t M);
hipError_t results_finals(const double* log_dZ, const double* log_L, const double* log_X, int64_t M,
                          double log_Z_mean, double* log_dp, double* H_out, void* ws, hipStream_t s);
hipError_t prior_finals(const float* samp_U, int ldu, const int64_t* perm, int64_t valid, int64_t M, int D,
                        const double* loc, const double* S, const int32_t* qkind, int diag, const double* log_L,
                        double* U_out, double* X_out, double* log_post, hipStream_t s);

}  // namespace hipns
