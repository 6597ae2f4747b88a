/*
 * hipns.h -- C ABI of the MI355X nested-sampling engine (libhipns.so).
 *
 * The library implements the sampling hot path of jaxns'
 * ShardedStaticNestedSampler + UniDimSliceSampler as HIP kernels for gfx950.
 * Every entry point below cites the reference function (jaxns 2.6.9,
 * src/jaxns/<file>:<line>) whose work it replaces.
 *
 * Conventions
 *  - Plain C: pointers, sizes and POD structs only.  Every pointer whose name
 *    ends in _dev (or that sits in a *_dev struct field) is device memory owned
 *    by the caller; the library never frees caller memory.
 *  - Every call is stream-ordered on `stream` (a hipStream_t passed as void*,
 *    NULL = default stream) and returns an int status: HIPNS_OK or a negative
 *    HIPNS_E* code; hipns_last_error() gives the message for the calling thread.
 *  - Scratch memory comes from the caller through (workspace, workspace_bytes);
 *    the matching *_workspace_size() reports how much is needed.
 *  - Calls are re-entrant for distinct buffers/workspaces.  No host threads.
 */
#ifndef HIPNS_H
#define HIPNS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIPNS_ABI_VERSION 1

#define HIPNS_OK 0
#define HIPNS_EINVAL (-1)  /* bad argument (null pointer, size, layout) */
#define HIPNS_EHIP (-2)    /* HIP runtime error */
#define HIPNS_ENOMEM (-3)  /* workspace too small */
#define HIPNS_ERANGE (-4)  /* size beyond a documented kernel limit */
#define HIPNS_ENOTSUP (-5) /* combination not supported */

/* likelihood families fused into the sampling kernels */
#define HIPNS_FAM_MVN 0    /* log N(x; mu, Sigma) on an affine prior: r = W q(u) + b, logL = c - |r|^2/2 */
#define HIPNS_FAM_GMM 1    /* log sum_k w_k N(x; m_k, s_k^2 I), x = p + s (.) q(u)                          */
#define HIPNS_FAM_LINREG 2 /* -N/2 log(2 pi s^2) - |y - X w|^2 / (2 s^2), w = p + s (.) q(u)                */

/* maximum sizes of the fused kernels */
#define HIPNS_MAX_DIM 128        /* padded U-space dimension */
#define HIPNS_MAX_LIVE 8192      /* live points per run */
#define HIPNS_MAX_GMM_K 8        /* mixture components */

/* ---------------------------------------------------------------------
 * Per-run description and mutable state (device arrays of these are
 * indexed by run).  Offsets are element offsets into the packed buffers.
 * ------------------------------------------------------------------- */
typedef struct hipns_run_desc {
  int32_t D;                /* U-space dimension */
  int32_t N;                /* num_live_points */
  int32_t shell;            /* shell size = int(N * shell_fraction) */
  int32_t num_slices;       /* slices per chain (s * D) */
  int32_t k_phantom;        /* phantom samples kept per chain (k) */
  int32_t midpoint_shrink;  /* 1: midpoint (alpha) shrink, 0: shrink to rejected point */
  int32_t sampler;          /* 0: UniDimSliceSampler, 1: UniformSampler (rejection), 2: UniDimSliceSampler
                             * with gradient-guided directions (uni_slice_sampler.py:255-271), 3: with
                             * gradient_slice (each slice along the unit gradient at its seed over [0, right],
                             * uni_slice_sampler.py:202-214), 4: both; for 2-4 d logL / dU is computed in the
                             * sampling kernel (MVN and mixture families only) */
  int32_t max_uniform_evals;/* UniformSampler cap (100 in the reference) */
  int64_t max_samples;      /* sample-collection length */
  int64_t live_off;         /* offset into the live-point arrays (points) */
  int64_t samp_off;         /* offset into the sample-collection arrays (samples) */
  int64_t chain_off;        /* offset into the new-sample arrays (chains) */
  int64_t param_off;        /* offset into the model parameter buffer (floats) */
  uint32_t seed_lo, seed_hi;/* Philox key of the run */
  double log_const;         /* additive constant of the likelihood */
} hipns_run_desc_t;

typedef struct hipns_run_state {
  int64_t next_idx;         /* NestedSamplerState.next_sample_idx */
  int64_t num_samples;      /* NestedSamplerState.num_samples */
  int64_t iteration;        /* shells collected */
  int64_t num_likelihood_evaluations; /* TerminationRegister fields ... */
  int64_t termination_reason;
  int64_t num_samples_used;
  double evidence_calc[8];  /* EvidenceCalculation: log_L, log_X, log_X2, log_Z, log_ZX, log_Z2, log_dZ, log_dZ2 */
  double evidence_calc_with_remaining[8];
  double log_L_contour;
  double efficiency;
  double relative_spread;
  double absolute_spread;
  double peak_log_XL;
  double contour;           /* contour of the shell being collected */
  int32_t plateau;
  int32_t no_seed_points;
  int32_t done;             /* 1 once the termination condition holds */
  int32_t buf;              /* which live buffer (0/1) holds the current sorted live set */
  int32_t first_seed;       /* index of the first live point with logL > contour */
  int32_t sender_node;      /* sender node index of the shell's new samples */
  int32_t flags;            /* diagnostic bits: 1 = a slice hit the step cap */
  int32_t pad0;
  int64_t cap_hits;         /* slices that hit the step cap (accepted the seed point) */
} hipns_run_state_t;

/* TerminationCondition (nested_samplers/common/types.py:26-54); a field is active when has_* != 0 */
typedef struct hipns_term_cond {
  int32_t has_ess, has_evidence_uncert, has_dlogZ, has_max_samples, has_max_num_likelihood_evaluations;
  int32_t has_log_L_contour, has_efficiency_threshold, has_rtol, has_atol, has_peak_XL_frac;
  double ess, evidence_uncert, dlogZ, max_samples, max_num_likelihood_evaluations;
  double log_L_contour, efficiency_threshold, rtol, atol, peak_XL_frac;
} hipns_term_cond_t;

/* Device buffers of a batch of runs.  U rows have stride `dps` floats
 * (dps = padded dimension, a multiple of 32, >= every run's D). */
typedef struct hipns_buffers {
  float* live_U[2];
  double* live_logL[2];
  double* live_cons[2];
  int32_t* live_sender[2];
  int32_t* live_evals[2];
  float* samp_U;
  double* samp_logL;
  int32_t* samp_sender;
  int32_t* samp_evals;
  uint8_t* samp_phantom;
  float* new_U;
  double* new_logL;
  int32_t* new_evals;
  const float* params;      /* packed model parameters */
  int32_t dps;
  int32_t pad0;
  const int32_t* run_order; /* optional [R]: the sampling launch takes runs in this order (blockIdx.y -> run),
                             * e.g. the longest chains first so that they do not start last; NULL = 0..R-1.
                             * It must be a permutation of 0..R-1 for the R runs of the call it is passed
                             * to: a call on a subset of a batch's runs (one slot, a gathered set) passes
                             * NULL, since the batch's order would index runs outside the subset. */
} hipns_buffers_t;

/* Kernel shape picked by the autotuner: lanes per chain (2: 32 chains per
 * wavefront with 32x32 MFMA tiles, 4: 16 chains per wavefront with 16x16 tiles,
 * 8: 8 chains per wavefront, mixture family only, used where the U row stride
 * covers the 8-lane instance (dps >= 64 for D <= 64, 128 for D <= 128), else 4),
 * wavefronts per workgroup, and the likelihood family. */
typedef struct hipns_launch {
  int32_t family;
  int32_t lanes_per_chain;
  int32_t waves_per_block;  /* chain groups (one per wavefront) per workgroup; LinReg multiplies each row
                             * tile against all of them */
  int32_t gmm_K;
  int32_t lr_ndata;         /* LinReg: data rows padded to a multiple of 32 */
  int32_t max_steps;        /* step cap per slice (safety bound on the shrink loop) */
  int32_t lr_panels;        /* LinReg: 16-column panels of X (even, 16 * lr_panels >= D rounded up to 32) */
  int32_t lr_terms;         /* LinReg products: 3 (or 0) = BF16x3 hi.hi + hi.lo + lo.hi; 4 adds X hi times the third
                             * bf16 part of w - w0; 1 = plain FP32 (FP32 X and w, FP32 MFMA); 5 = the plain FP32
                             * product with w - w0 formed in FP64 (FP64 ndtri and parameters; refinement),
                             * hipns_log_likelihood only (sampling launches and hipns_linreg_stats take 1, 3, 4) */
  int64_t lr_data_off;      /* LinReg: float offset in params of [r0 | X hi | X lo | y | X] shared by all runs;
                             * X parts are [lr_panels][lr_ndata][16], bf16 hi/lo then f32 */
} hipns_launch_t;

/* ---------------------------------------------------------------------
 * Library
 * ------------------------------------------------------------------- */
int hipns_abi_version(void);
/* copies the calling thread's last error message (NUL-terminated) */
int hipns_last_error(char* buf, size_t len);

/* ---------------------------------------------------------------------
 * (a)+(b)+(c)+(d): the shell loop.  R runs are advanced together; runs whose
 * state says done are skipped.
 * ------------------------------------------------------------------- */

/* create_init_state + create_init_termination_register
 * (nested_samplers/common/initialisation.py:20-108, uniform_sample.py:12-61):
 * draws each run's N live points uniformly with logL > -inf, sorts them
 * (stable argsort, initialisation.py:75-76) and clears the register. */
int hipns_create_init_state(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                            const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs,
                            const hipns_launch_t* launch, void* workspace, size_t workspace_bytes, void* stream);

/* One _collect_shell per run (nested_samplers/sharded/sharded_static.py:210-324) followed
 * by determine_termination (nested_samplers/common/termination.py:13-147):
 * discard the lowest `shell` live points into the sample collection, run one
 * sampler chain per discarded point from seeds above the contour
 * (UniDimSliceSampler.get_sample, samplers/uni_slice_sampler.py:114-441, or
 * UniformSampler, samplers/uniform_samplers.py:42-85), write the chains back,
 * stable-sort the live set, append phantoms, update the evidence register and
 * the termination flag. */
int hipns_collect_shell(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                        const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, const hipns_launch_t* launch,
                        const hipns_term_cond_t* term, void* workspace, size_t workspace_bytes, void* stream);

/* The four stages of hipns_collect_shell, exposed so that each can be timed:
 *  discard  : _add_samples_to_state(live[:shell]) + contour/seed range/sender (sharded_static.py:237-261)
 *  sample   : get_samples -> one UniDimSliceSampler (or UniformSampler) chain per discarded point
 *             (sharded_static.py:88-129, samplers/uni_slice_sampler.py:114-441, bases.py:63-74)
 *  replace  : live[:shell] = new; stable argsort; permute (sharded_static.py:262-279)
 *  register : evidence register + termination flag (sharded_static.py:281-323, termination.py:13-147) */
int hipns_shell_discard(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                        const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream);
int hipns_shell_sample(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                       const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, const hipns_launch_t* launch,
                       void* stream);
int hipns_shell_replace(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                        const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream);
int hipns_shell_register(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                         const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs,
                         const hipns_term_cond_t* term, void* stream);

/* Final `_add_samples_to_state(live_point_collection)` of _run
 * (sharded_static.py:834-838): appends the live set to the sample collection. */
int hipns_consume_live(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                       const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream);

/* determine_termination on each run's current register without collecting a
 * shell: the while_loop condition of _main_ns_thread (sharded_static.py:480-485)
 * evaluated before the first shell of a phase.  set_no_seed != 0 first sets
 * no_seed_points from the incoming live set (sharded_static.py:471-473). */
int hipns_check_termination(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                            const hipns_buffers_t* bufs, const hipns_term_cond_t* term, int32_t set_no_seed,
                            void* stream);

size_t hipns_collect_shell_workspace_size(int32_t R, const hipns_run_desc_t* runs_host);

/* ---------------------------------------------------------------------
 * (d) evidence statistics
 * ------------------------------------------------------------------- */

/* compute_evidence_stats (internals/shrinkage_statistics.py:131-157): the
 * _update_evidence_calc_op recurrence (:43-94) over M samples as a single-pass
 * log-space scan.  per_sample_dev: [8][M] SoA (field-major, may be NULL);
 * final_dev: [8].  stop_idx < 0 => static scan, else outputs past stop_idx hold
 * the initial state (cumulative_op_dynamic, internals/cumulative_ops.py:82-130). */
int hipns_evidence_scan(const double* log_L_dev, const double* num_live_points_dev, int64_t M, int64_t stop_idx,
                        double* per_sample_dev, double* final_dev, void* workspace, size_t workspace_bytes,
                        void* stream);
size_t hipns_evidence_scan_workspace_size(int64_t M);

/* hipns_evidence_scan for callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): no counterpart in the reference.  The same outputs, with the scan's
 * scratch in a fixed 32 MiB device array inside the library (one per GPU, held until the process ends);
 * M is limited to hipns_evidence_scan_noalloc_max() samples (HIPNS_ERANGE beyond it).  Nothing is
 * allocated.  Calls that overlap on one GPU share that array and race: serialise them (one stream, or
 * stream-ordered). */
int hipns_evidence_scan_noalloc(const double* log_L_dev, const double* num_live_points_dev, int64_t M,
                                int64_t stop_idx, double* per_sample_dev, double* final_dev, void* stream);
size_t hipns_evidence_scan_noalloc_max(void);

/* count_crossed_edges (internals/tree_structure.py:33-108): stable sort of the
 * N+1 node log-likelihoods (root -inf), out-degree by sender, prefix count.
 * num_samples < 0 => static form. */
int hipns_count_crossed_edges(const int32_t* sender_node_idx_dev, const double* log_L_dev, int64_t N,
                              int64_t num_samples, int64_t* samples_indices_dev, int64_t* num_live_points_dev,
                              void* workspace, size_t workspace_bytes, void* stream);
size_t hipns_count_crossed_edges_workspace_size(int64_t N);

/* The evidence statistics of _to_results (sharded_static.py:652-745: count_crossed_edges, then
 * compute_evidence_stats over the sorted samples) for R runs at once, in one set of launches.
 * Run r's samples are its sample-collection slice of M_r = min(num_samples, max_samples) entries,
 * read from states_dev on the device.  final_dev [R][8] receives the final EvidenceCalculation
 * (log_L, log_X_mean, log_X2_mean, log_Z_mean, log_ZX_mean, log_Z2_mean, log_dZ_mean,
 * log_dZ2_mean), phantom_dev [R] the number of phantom samples. */
int hipns_batch_evidence(const hipns_run_desc_t* runs_dev, const hipns_run_state_t* states_dev, int32_t R,
                         const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, double* final_dev,
                         int64_t* phantom_dev, void* workspace, size_t workspace_bytes, void* stream);
size_t hipns_batch_evidence_workspace_size(int32_t R, const hipns_run_desc_t* runs_host);

/* The per-sample finals of _to_results (sharded_static.py:706-745) over M samples in tree order, given
 * the evidence scan's per-sample log_dZ_mean and log_X_mean fields: log_dp_dev[i] = log dZ_i normalised
 * over all samples (normalise_log_space), and H_dev[4] = {H unstable = log Z - sum dp log L,
 * H stable = sum dp log X, H_mean (the unstable form where finite, else the stable one), log of the
 * normaliser}.  Deterministic reductions in a fixed grouping. */
int hipns_results_finals(const double* log_dZ_dev, const double* log_L_dev, const double* log_X_dev, int64_t M,
                         double log_Z_mean, double* log_dp_dev, double* H_dev, void* workspace, size_t workspace_bytes,
                         void* stream);
size_t hipns_results_finals_workspace_size(int64_t M);

/* Model.transform and log_posterior_density of _to_results (sharded_static.py:721-745, framework/model.py:
 * 155-185) for M samples: row b reads U row perm_dev[b] (or b when perm_dev is NULL) of samp_U_dev (FP32,
 * stride ldu); rows at or past `valid` are the collection's placeholders, U = 0.5.  With q_i = ndtri(U_i)
 * where qkind_dev[i] != 0, else U_i: U_dev [M][D] = U (FP64), X_dev [M][D] = loc + S q (S_dev [D][D]
 * lower triangular with a positive diagonal, row-major; diag != 0 reads its diagonal only), and
 * log_post_dev[b] = log_L_dev[b] + log prior density (sum of -q^2/2 - log(2 pi)/2 over the ndtri slots,
 * minus the sum of log S_ii).  Any of U_dev, X_dev, log_post_dev may be NULL. */
int hipns_prior_finals(const float* samp_U_dev, int32_t ldu, const int64_t* perm_dev, int64_t valid, int64_t M, int32_t D,
                       const double* loc_dev, const double* S_dev, const int32_t* qkind_dev, int32_t diag,
                       const double* log_L_dev, double* U_dev, double* X_dev, double* log_post_dev, void* stream);

/* Stable argsort of FP64 keys (jnp.argsort, stable=True; -inf < ... < +inf < NaN). */
int hipns_argsort_stable(const double* keys_dev, int64_t n, int64_t* perm_dev, void* workspace,
                         size_t workspace_bytes, void* stream);
size_t hipns_argsort_workspace_size(int64_t n);

/* ---------------------------------------------------------------------
 * Likelihoods evaluated by the caller (a torch callable): the same shell,
 * one step per call.  Chain state lives in caller-owned arrays of C rows
 * (C >= the sum of num_live_points; chain i of run r is row chain_off_r + i).
 * ------------------------------------------------------------------- */
typedef struct hipns_ext_state {
  float* U0;          /* [C][dps] current point of each chain */
  float* nv;          /* [C][dps] slice direction */
  float* prop_U;      /* [C][dps] proposal of the current step; rows of finished chains hold 0.5 */
  double* logL0;      /* [C] */
  double* prop_logL;  /* [C] written by the caller: Model.forward of prop_U (NaN is read as -inf) */
  float* left;        /* [C] bracket */
  float* right;
  float* t;           /* [C] last pick */
  int32_t* s;         /* [C] slice index */
  int32_t* k;         /* [C] shrinks of the current slice */
  int32_t* evals;     /* [C] likelihood evaluations of the chain */
  int32_t* flags;     /* [C] bit 0 running, bit 1 begins a slice, bit 2 shrinks, bit 3 asks for the
                       * gradient of logL at the prop_U row (mode 4) */
  float* grad;        /* [C][dps] written by the caller for rows with bit 3: d logL / dU (mode 4 only) */
} hipns_ext_state_t;

/* One step of every chain of the batch (UniDimSliceSampler.get_sample_from_seed,
 * uni_slice_sampler.py:360-441, UniformSampler uniform_samplers.py:42-85, or the initial
 * uniform draws, initialisation.py:20-84): step 0 seeds the chains and writes their first
 * proposals; step w > 0 first consumes prop_logL of step w-1 (acceptance, shrink, phantoms,
 * outputs into new_U/new_logL/new_evals), then writes the proposals of step w.  mode: 0 slice
 * shell, 1 uniform-sampler shell, 2 initial live points, 4 gradient-guided slice shell
 * (uni_slice_sampler.py:255-271: after each accepted slice the chain asks for d logL / dU at the
 * accepted point, counted as one likelihood evaluation, and reflects its direction about it;
 * a gradient of norm < 1e-10 or non-finite gives a random direction).  A shell is hipns_shell_discard, steps
 * until no chain runs (flags bit 0), hipns_shell_replace and hipns_shell_register; the initial
 * state is hipns_init_prepare, mode-2 steps, hipns_init_finalize. */
int hipns_ext_step(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                   const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, const hipns_launch_t* launch,
                   const hipns_ext_state_t* ext, int32_t mode, int32_t step, void* stream);

/* The parts of hipns_create_init_state around its sampling (initialisation.py:20-108): register
 * and sample-collection reset, then the initial sort of the live points. */
int hipns_init_prepare(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                       const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream);
int hipns_init_finalize(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                        const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream);

/* ---------------------------------------------------------------------
 * (b) batched prior transform and likelihood
 * ------------------------------------------------------------------- */

/* Lower Cholesky factor in FP64 of the symmetric positive-definite D x D matrix A_dev (row-major, lower
 * triangle read): L_dev [D][D] (zero above the diagonal).  info_dev (may be NULL) receives 0, or k + 1
 * when the leading minor of order k + 1 is not positive definite (L is then incomplete).  One
 * workgroup; for the one-off factor of a likelihood covariance (jaxns users pass it to
 * tfpd.MultivariateNormalFullCovariance / TriL, tests/conftest.py:219-224). */
int hipns_cholesky_f64(const double* A_dev, int32_t D, double* L_dev, int32_t* info_dev, void* stream);

/* The multivariate-normal family's parameters (HIPNS_FAM_MVN) for log N(x; mean, cov) on the affine
 * prior x = loc + S q(u), formed in FP64 on the device: L = chol(cov) (into L_dev), W = L^-1 S,
 * b = L^-1 (loc - mean), written as the FP32 block params_dev = [qkind: dps][b: dps][W: dps x dps]
 * (zero padding), and log_const_dev[0] = -D/2 log(2 pi) - sum log L_ii.  S_dev is D x D row-major;
 * info_dev as for hipns_cholesky_f64 (nothing else is written when it is non-zero).  cov_dev = NULL
 * takes L_dev as the given factor (a scale_tril) and runs the solves only. */
int hipns_mvn_prepare_f64(const double* cov_dev, const double* mean_dev, const double* loc_dev, const double* S_dev,
                          const int32_t* qkind_dev, int32_t D, int32_t dps, double* L_dev, float* params_dev,
                          double* log_const_dev, int32_t* info_dev, void* workspace, size_t workspace_bytes,
                          void* stream);
size_t hipns_mvn_prepare_workspace_size(int32_t D);

/* Model.forward over a batch (framework/model.py:167-176 -> framework/ops.py:302-326):
 * logL_dev[i] = likelihood(prior_transform(U_dev[i, :D])) of run 0 of `runs_dev`,
 * NaN -> -inf.  U rows have stride bufs->dps. */
int hipns_log_likelihood(const float* U_dev, int64_t B, double* logL_dev, const hipns_run_desc_t* runs_dev,
                         const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, const hipns_launch_t* launch,
                         void* stream);

/* The batched likelihood with its gradient d logL / dU (the gradient of Model.forward that the
 * gradient-guided and gradient-slice samplers ask for, uni_slice_sampler.py:204,258), computed as the
 * sampling kernel computes it for sampler codes 2-4: MVN, -(W^T r) (.) dq/du with r = W q(U) + b; mixture, the
 * softmax-weighted component gradients times the prior scale and dq/du.  grad_dev has row stride
 * bufs->dps (entries d < D written); FP32 products, as in the kernel.  MVN and mixture families. */
int hipns_log_likelihood_grad(const float* U_dev, int64_t B, double* logL_dev, float* grad_dev,
                              const hipns_run_desc_t* runs_dev, const hipns_run_desc_t* runs_host,
                              const hipns_buffers_t* bufs, const hipns_launch_t* launch, void* stream);

/* Adjoint statistics of the regression likelihood, for the M-step of evidence maximisation
 * (experimental/evidence_maximisation.py:151-160, log_evidence = logsumexp(forward(U; theta) +
 * log_weights), differentiated by jax there).  For each row b of U (stride bufs->dps), with
 * w = p + s (.) q(U_b) from run 0's prior and r = y - X w:  logL_dev[b], rss_dev[b] = |r|^2,
 * yr_dev[b] = y^T r.  Then d logL / d log sigma = -N + |r|^2 / sigma^2 and, for a zero-mean prior
 * scaled by tau, d logL / d log tau = w^T X^T r / sigma^2 = (y^T r - |r|^2) / sigma^2. */
int hipns_linreg_stats(const float* U_dev, int64_t B, double* logL_dev, double* rss_dev, double* yr_dev,
                       const hipns_run_desc_t* runs_dev, const hipns_run_desc_t* runs_host,
                       const hipns_buffers_t* bufs, const hipns_launch_t* launch, void* stream);

/* Prior.forward of every dimension (framework/bases.py:58-68 ->
 * framework/wrapped_tfp_distribution.py:77-84): X = loc + scale (.) q(U) for
 * per-dimension transform q in {identity, ndtri}, optional dense lower factor. */
int hipns_prior_transform(const float* U_dev, int64_t B, int32_t D, int32_t ldu, const float* loc_dev,
                          const float* scale_dev, const float* tril_dev, const int32_t* qkind_dev, float* X_dev,
                          int32_t ldx, void* stream);

/* ---------------------------------------------------------------------
 * Parity hooks
 * ------------------------------------------------------------------- */

/* UniDimSliceSampler chains with injected draws (no PRNG): chain c starts at
 * live point seed_idx_dev[c] of run 0 (live buffer `buf`), uses unit directions
 * dirs_dev [C][num_slices][dps] and pick uniforms tu_dev [C][num_slices][max_steps],
 * and records trace_t/trace_logL [C][num_slices][max_steps] and steps_dev
 * [C][num_slices].  Outputs go to bufs->new_* (chains) and, for k > 0, the
 * phantom arrays ph_U_dev [C][k][dps], ph_logL_dev [C][k]. */
int hipns_slice_chains_injected(const hipns_run_desc_t* runs_dev, const hipns_run_desc_t* runs_host,
                                const hipns_buffers_t* bufs, const hipns_launch_t* launch, int32_t buf,
                                int32_t C, double contour, const int32_t* seed_idx_dev, const float* dirs_dev,
                                const float* tu_dev, float* trace_t_dev, double* trace_logL_dev, int32_t* steps_dev,
                                float* ph_U_dev, double* ph_logL_dev, void* stream);

/* Philox4x32-10 blocks: out_dev[i] = philox(ctr_dev[i], key) (4 words each). */
int hipns_philox4x32(const uint32_t* ctr_dev, int64_t n, uint32_t key0, uint32_t key1, uint32_t* out_dev,
                     void* stream);

/* Device inverse normal CDF (the prior transform's q = ndtri), for accuracy tests. */
int hipns_ndtri_f32(const float* u_dev, int64_t n, float* z_dev, void* stream);

/* The slice bracket's quotient a / n (reciprocal + residual corrections, no operand scaling),
 * elementwise, for tests that it equals IEEE binary32 division over the bracket's operand range
 * (uni_slice_sampler.py:41-64 divides (0 - x) and (1 - x) by the direction).  a_dev = NULL: the
 * reciprocal y of n that the quotient starts from, for tests that it is 1 / n correctly rounded. */
int hipns_slice_quotients(const float* a_dev, const float* n_dev, int64_t count, float* q_dev, void* stream);

/* The seed point chain i of a shell picks in iteration `iteration` of a run keyed (key0, key1), i in
 * [0, C): idx_dev[i] and the uniform u_dev[i] it came from, for a sorted live set of N points whose
 * entries [first, N) lie above the contour (get_seed_point, samplers/uni_slice_sampler.py:343-358, via
 * resample_indicies, internals/random.py:34-84). */
int hipns_seed_indices(uint32_t key0, uint32_t key1, uint32_t iteration, int32_t first, int32_t N, int32_t C,
                       int32_t* idx_dev, double* u_dev, void* stream);

/* The unit direction of every slice of chains 0..C-1 in iteration `iteration` (_sample_direction,
 * samplers/uni_slice_sampler.py:23-38), computed in the fused kernel's lane layout with
 * lanes_per_chain lanes per chain: dirs_dev [C][num_slices][D]. */
int hipns_slice_directions(uint32_t key0, uint32_t key1, uint32_t iteration, int32_t C, int32_t num_slices, int32_t D,
                           int32_t lanes_per_chain, float* dirs_dev, void* stream);

#ifdef __cplusplus
}
#endif
#endif
