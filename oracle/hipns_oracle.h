/*
 * hipns_oracle.h -- CPU restatement of the jaxns nested-sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP kernels
 * in jaxns_amd/csrc.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the jaxns source tree, src/jaxns/).  Arithmetic is IEEE binary64 unless a
 * function name ends in _f32, in which case the slice arithmetic is binary32
 * (compiled with -ffp-contract=off so that every +,-,*,/ rounds once, exactly
 * like the device code built under `#pragma clang fp contract(off)`).
 */
#ifndef HIPNS_ORACLE_H
#define HIPNS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- log-space arithmetic (internals/log_semiring.py) ---------------- */
double orc_logaddexp(double a, double b);
void orc_signed_logaddexp(double a, double sa, double b, double sb, double* out_log, double* out_sign);
void orc_cumulative_logsumexp(const double* u, int64_t n, double* out);

/* ---- evidence recurrence (internals/shrinkage_statistics.py) ---------- */
/* state layout (8 doubles): log_L, log_X_mean, log_X2_mean, log_Z_mean,
 * log_ZX_mean, log_Z2_mean, log_dZ_mean, log_dZ2_mean  (types.py:12-23)   */
void orc_evidence_init(double* state8);
void orc_evidence_update(double* state8, double num_live_points, double log_L_next);
/* compute_evidence_stats: per_sample is [M][8] (may be NULL); stop_idx < 0 => static scan */
void orc_compute_evidence_stats(const double* log_L, const double* num_live_points, int64_t M,
                                int64_t stop_idx, double* final8, double* per_sample);

/* ---- stats (internals/stats.py) -------------------------------------- */
void orc_linear_to_log_stats(double log_f_mean, double log_f2_mean, double* mu, double* var);
double orc_ess_kish(double log_Z_mean, double log_dZ2_mean);

/* ---- sorting / tree (internals/tree_structure.py, jnp.argsort) ------- */
void orc_argsort_stable(const double* keys, int64_t n, int64_t* perm);
/* num_samples < 0 => static form */
void orc_count_crossed_edges(const int64_t* sender_node_idx, const double* log_L, int64_t N,
                             int64_t num_samples, int64_t* samples_indices, int64_t* num_live_points);
void orc_count_intervals_naive(const int64_t* sender_node_idx, const double* log_L, int64_t N,
                               int64_t* samples_indices, int64_t* num_live_points);

/* ---- seed selection (internals/random.py:34-84) ----------------------- */
/* first index replace_index writes an update of `count` entries starting at `start` into `size` */
int64_t orc_update_start(int64_t start, int64_t count, int64_t size);

int64_t orc_seed_index(const double* live_log_L, int64_t N, double contour, double u);

/* ---- special functions ------------------------------------------------ */
double orc_ndtri(double p);
double orc_ndtr(double x);

/* ---- Philox4x32-10 (Salmon et al., SC'11) ------------------------------ */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- model description (host memory) --------------------------------- */
enum { ORC_FAM_MVN = 0, ORC_FAM_GMM = 1, ORC_FAM_LINREG = 2, ORC_FAM_TABLE = 3 };
enum { ORC_PRIOR_NORMAL = 0, ORC_PRIOR_UNIFORM = 1, ORC_PRIOR_MVN = 2 };

typedef struct {
  int family;
  int D;
  /* prior: x = prior_loc + prior_scale (.) q(u)  with q = ndtri (normal/mvn) or identity (uniform)
   * prior_kind == ORC_PRIOR_MVN uses the dense lower-triangular prior_tril [D*D] */
  int prior_kind;
  const double* prior_loc;    /* [D] (uniform: low)            */
  const double* prior_scale;  /* [D] (uniform: high-low)       */
  const double* prior_tril;   /* [D*D] row-major, MVN prior    */
  /* MVN likelihood N(mu, L L^T) */
  const double* mvn_mu;       /* [D]   */
  const double* mvn_chol;     /* [D*D] lower, row-major */
  /* GMM: K components, means [K*D], sigmas [K], log weights [K] */
  int K;
  const double* gmm_means;
  const double* gmm_sigma;
  const double* gmm_logw;
  /* LinReg: y ~ N(X w, sigma^2) with X [Ndata*D] row-major */
  int64_t Ndata;
  const double* lr_X;
  const double* lr_y;
  double lr_sigma;
} orc_model_t;

void orc_prior_transform(const orc_model_t* m, const double* U, double* X);
double orc_log_likelihood(const orc_model_t* m, const double* U);
double orc_log_prob_prior(const orc_model_t* m, const double* U);

/* ---- slice sampler with injected draws (samplers/uni_slice_sampler.py) --
 * One chain of num_slices slices from seed (U0, logL0) under contour.
 * dirs:  [num_slices][D] unit directions
 * tu:    [num_slices][max_steps] uniforms for the pick at step k of slice s
 * logL_replay: if non-NULL, [num_slices][max_steps] likelihood values to use
 *              instead of evaluating the model (replays a device trace)
 * outputs: U_out [D], *logL_out, *evals_out, phantoms [k][D] + logL [k],
 *          trace_t [num_slices][max_steps], trace_logL, steps_per_slice [num_slices]
 * returns 0 on success, -1 if a slice needed more than max_steps picks. */
int orc_slice_chain_f64(const orc_model_t* m, const double* U0, double logL0, double contour,
                        int num_slices, int midpoint_shrink, int k_phantom,
                        const double* dirs, const double* tu, int max_steps, const double* logL_replay,
                        double* U_out, double* logL_out, int64_t* evals_out,
                        double* phantom_U, double* phantom_logL,
                        double* trace_t, double* trace_logL, int32_t* steps_per_slice);
int orc_slice_chain_f32(const orc_model_t* m, const float* U0, double logL0, double contour,
                        int num_slices, int midpoint_shrink, int k_phantom,
                        const float* dirs, const float* tu, int max_steps, const double* logL_replay,
                        float* U_out, double* logL_out, int64_t* evals_out,
                        float* phantom_U, double* phantom_logL,
                        float* trace_t, double* trace_logL, int32_t* steps_per_slice);
void orc_slice_bounds_f32(const float* U0, const float* n, int D, float* left, float* right);
void orc_slice_bounds_f64(const double* U0, const double* n, int D, double* left, double* right);

/* ---- termination (nested_samplers/common/termination.py:13-147) -------- */
typedef struct {
  /* a field is active when its has_ flag is nonzero */
  int has_ess; double ess;
  int has_evidence_uncert; double evidence_uncert;
  int has_dlogZ; double dlogZ;
  int has_max_samples; double max_samples;
  int has_max_num_likelihood_evaluations; double max_num_likelihood_evaluations;
  int has_log_L_contour; double log_L_contour;
  int has_efficiency_threshold; double efficiency_threshold;
  int has_rtol; double rtol;
  int has_atol; double atol;
  int has_peak_XL_frac; double peak_XL_frac;
} orc_term_cond_t;

typedef struct {
  int64_t num_samples_used;
  double evidence_calc[8];
  double evidence_calc_with_remaining[8];
  int64_t num_likelihood_evaluations;
  double log_L_contour;
  double efficiency;
  int32_t plateau;
  int32_t no_seed_points;
  double relative_spread;
  double absolute_spread;
  double peak_log_XL;
} orc_register_t;

void orc_init_register(orc_register_t* r);
int orc_determine_termination(const orc_term_cond_t* c, const orc_register_t* r, int64_t* reason);

/* ---- full run (nested_samplers/sharded/sharded_static.py) ------------- */
typedef struct {
  int num_live_points;
  int num_slices;
  int k_phantom;
  int midpoint_shrink;
  int64_t max_samples;       /* sample-buffer length */
  double init_efficiency_threshold;
  uint64_t seed;
  int num_threads;           /* OpenMP threads over the chains of a shell, <=0 => all */
  int64_t max_iterations;    /* stop early after this many shells (<0: none) -- for bounded timing samples */
  int dir16;                 /* 0: slice directions from FP64 Box-Muller per dimension; 1: from 16-bit
                                uniform pairs, eight dimensions per Philox block, as the device draws
                                them (a variant for comparing the two, not the reference's) */
} orc_run_cfg_t;

typedef struct {
  int64_t termination_reason;
  int64_t num_samples;       /* state.num_samples */
  int64_t num_iterations;
  int64_t total_likelihood_evaluations;
  double log_Z_mean, log_Z_uncert, ESS, H_mean;
  double log_efficiency;
  int64_t total_phantom_samples;
} orc_run_result_t;

/* Runs create_init_state + (uniform phase) + slice phase + to_results.
 * Sample collection buffers (caller-owned, length cfg->max_samples): may be NULL except U/log_L/sender/evals/phantom
 * which are required; returns 0 on success. */
int orc_ns_run(const orc_model_t* m, const orc_run_cfg_t* cfg, const orc_term_cond_t* term,
               double* samp_U, double* samp_logL, int64_t* samp_sender, int64_t* samp_evals, int8_t* samp_phantom,
               orc_run_result_t* res);

/* to_results statistics over a (trimmed) sample collection (sharded_static.py:652-773) */
int orc_to_results(const double* samp_logL, const int64_t* samp_sender, const int64_t* samp_evals,
                   const int8_t* samp_phantom, int64_t num_samples,
                   int64_t* samples_indices, int64_t* num_live_points, double* per_sample8,
                   orc_run_result_t* res);

#ifdef __cplusplus
}
#endif
#endif
