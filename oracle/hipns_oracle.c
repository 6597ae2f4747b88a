/*
 * hipns_oracle.c -- CPU restatement of the jaxns nested-sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels, never the thing
 * measured or shipped.  See hipns_oracle.h for the contract.
 *
 * Reference paths are relative to src/jaxns/ of jaxns 2.6.9.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include "hipns_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LOG2 0.69314718055994530942
#define LOG_HALF (-0.69314718055994530942)

/* ======================================================================
 * log-space arithmetic
 * ==================================================================== */

/* jnp.logaddexp as used by LogSpace.__add__ (internals/log_semiring.py:11-25,
 * :146-147): amax + log1p(exp(-|delta|)); NaN delta (same-sign infinities)
 * returns x1 + x2. */
double orc_logaddexp(double a, double b) {
  double amax = a > b ? a : b;
  double delta = a - b;
  if (isnan(delta)) return a + b;
  return amax + log1p(exp(-fabs(delta)));
}

/* signed_logaddexp (internals/log_semiring.py:28-48) */
void orc_signed_logaddexp(double a, double sa, double b, double sb, double* out_log, double* out_sign) {
  double amax = a > b ? a : b;
  double signmax = a > b ? sa : sb;
  double delta = -fabs(b - a);
  double sign = sa * sb;
  if (isnan(delta)) {
    *out_log = a + b;
  } else {
    *out_log = amax + log1p(sign * exp(delta));
  }
  *out_sign = signmax;
}

/* cumulative_logsumexp without signs (internals/log_semiring.py:51-92):
 * sequential lax.scan of logaddexp from -inf. */
void orc_cumulative_logsumexp(const double* u, int64_t n, double* out) {
  double acc = -INFINITY;
  for (int64_t i = 0; i < n; ++i) {
    acc = orc_logaddexp(acc, u[i]);
    out[i] = acc;
  }
}

/* ======================================================================
 * evidence recurrence
 * ==================================================================== */

/* create_init_evidence_calc (internals/shrinkage_statistics.py:112-128) */
void orc_evidence_init(double* s) {
  s[0] = -INFINITY; /* log_L */
  s[1] = 0.0;       /* log_X_mean */
  s[2] = 0.0;       /* log_X2_mean */
  s[3] = -INFINITY; /* log_Z_mean */
  s[4] = -INFINITY; /* log_ZX_mean */
  s[5] = -INFINITY; /* log_Z2_mean */
  s[6] = -INFINITY; /* log_dZ_mean */
  s[7] = -INFINITY; /* log_dZ2_mean */
}

/* _update_evidence_calc_op (internals/shrinkage_statistics.py:43-94).
 * LogSpace naked products are sums of logs, sums are logaddexp; the
 * expression structure (left-to-right grouping) follows the reference. */
void orc_evidence_update(double* s, double n, double log_L_next) {
  const double L_contour = s[0];
  const double X = s[1], X2 = s[2], Z = s[3], ZX = s[4], Z2 = s[5], dZ2 = s[7];
  double midL = LOG_HALF + orc_logaddexp(log_L_next, L_contour);
  double log_n = log(n);
  double log_n1 = log(n + 1.0);
  double log_n2 = log(n + 2.0);
  double T = -orc_logaddexp(0.0, -log_n);
  double t = -log_n1;
  double T2 = -orc_logaddexp(0.0, LOG2 - log_n);
  double t2 = LOG2 - log_n1 - log_n2;
  double tT = -orc_logaddexp(0.0, -log_n) - log_n2;
  double dZ = X + t + midL;
  double nX = X + T;
  double nX2 = X2 + T2;
  double nZ = orc_logaddexp(Z, dZ);
  double nZX = orc_logaddexp(ZX + T, X2 + tT + midL);
  double nZ2 = orc_logaddexp(orc_logaddexp(Z2, LOG2 + ZX + t + midL), X2 + t2 + 2.0 * midL);
  double ndZ2 = orc_logaddexp(dZ2, X2 + t2 + 2.0 * midL);
  s[0] = log_L_next;
  s[1] = nX;
  s[2] = nX2;
  s[3] = nZ;
  s[4] = nZX;
  s[5] = nZ2;
  s[6] = dZ;
  s[7] = ndZ2;
}

/* compute_evidence_stats (internals/shrinkage_statistics.py:131-157) with
 * cumulative_op_static / cumulative_op_dynamic (internals/cumulative_ops.py:50-130):
 * outputs past the dynamic stop hold the init value. */
void orc_compute_evidence_stats(const double* log_L, const double* num_live_points, int64_t M,
                                int64_t stop_idx, double* final8, double* per_sample) {
  double s[8], init[8];
  orc_evidence_init(s);
  orc_evidence_init(init);
  int64_t stop = (stop_idx < 0 || stop_idx > M) ? M : stop_idx;
  for (int64_t i = 0; i < M; ++i) {
    if (i < stop) {
      orc_evidence_update(s, num_live_points[i], log_L[i]);
      if (per_sample) memcpy(per_sample + 8 * i, s, sizeof(s));
    } else if (per_sample) {
      memcpy(per_sample + 8 * i, init, sizeof(init));
    }
  }
  if (final8) memcpy(final8, s, sizeof(s));
}

/* ======================================================================
 * stats
 * ==================================================================== */

/* linear_to_log_stats with log_f2_mean (internals/stats.py:55-74):
 * mu = f^2 / sqrt(f2), sigma2 = f2 / f^2 in log space; var floored at eps(f64). */
void orc_linear_to_log_stats(double log_f_mean, double log_f2_mean, double* mu, double* var) {
  double f_sq = log_f_mean + log_f_mean;
  *mu = f_sq - 0.5 * log_f2_mean;
  double s2 = log_f2_mean - f_sq;
  *var = s2 > DBL_EPSILON ? s2 : DBL_EPSILON;
  if (isnan(s2)) *var = s2; /* jnp.maximum propagates NaN */
}

/* effective_sample_size_kish (internals/stats.py:77-86) */
double orc_ess_kish(double log_Z_mean, double log_dZ2_mean) {
  return exp((log_Z_mean + log_Z_mean) - log_dZ2_mean);
}

/* ======================================================================
 * stable argsort (jnp.argsort / lax.sort_key_val with is_stable=True)
 * total order: -inf < finite < +inf < NaN; ties keep input order.
 * ==================================================================== */
static int key_less(double a, double b) {
  if (isnan(a)) return 0;
  if (isnan(b)) return 1;
  return a < b;
}

static void merge_sort_rec(const double* keys, int64_t* idx, int64_t* tmp, int64_t n) {
  if (n < 2) return;
  int64_t h = n / 2;
  merge_sort_rec(keys, idx, tmp, h);
  merge_sort_rec(keys, idx + h, tmp, n - h);
  int64_t i = 0, j = h, o = 0;
  while (i < h && j < n) {
    /* take from the right only if strictly less: stability */
    if (key_less(keys[idx[j]], keys[idx[i]])) tmp[o++] = idx[j++];
    else tmp[o++] = idx[i++];
  }
  while (i < h) tmp[o++] = idx[i++];
  while (j < n) tmp[o++] = idx[j++];
  memcpy(idx, tmp, (size_t)n * sizeof(int64_t));
}

void orc_argsort_stable(const double* keys, int64_t n, int64_t* perm) {
  for (int64_t i = 0; i < n; ++i) perm[i] = i;
  if (n < 2) return;
  int64_t* tmp = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  merge_sort_rec(keys, perm, tmp, n);
  free(tmp);
}

/* ======================================================================
 * tree counting
 * ==================================================================== */

/* count_crossed_edges (internals/tree_structure.py:33-108) */
void orc_count_crossed_edges(const int64_t* sender, const double* log_L, int64_t N, int64_t num_samples,
                             int64_t* samples_indices, int64_t* num_live_points) {
  double* nodes = (double*)malloc((size_t)(N + 1) * sizeof(double));
  int64_t* sort_idx = (int64_t*)malloc((size_t)(N + 1) * sizeof(int64_t));
  int32_t* out_degree = (int32_t*)calloc((size_t)(N + 1), sizeof(int32_t));
  int64_t* crossed = (int64_t*)malloc((size_t)(N + 1) * sizeof(int64_t));
  nodes[0] = -INFINITY;
  for (int64_t i = 0; i < N; ++i) nodes[i + 1] = log_L[i];
  orc_argsort_stable(nodes, N + 1, sort_idx);
  for (int64_t i = 0; i < N; ++i) {
    int64_t s = sender[i] > 0 ? sender[i] : 0; /* lax.max(sender, 0) */
    out_degree[s] += 1;
  }
  int32_t fake = (num_samples < 0) ? 0 : (int32_t)(N - num_samples);
  int64_t stop = (num_samples < 0) ? N + 1 : num_samples;
  int32_t acc = 1;
  for (int64_t i = 0; i < N + 1; ++i) {
    if (i < stop) {
      acc += out_degree[sort_idx[i]] - 1;
      crossed[i] = acc;
    } else {
      crossed[i] = fake; /* empty_fill of cumulative_op_dynamic */
    }
  }
  if (num_samples >= 0)
    for (int64_t i = 0; i < N + 1; ++i) crossed[i] -= fake;
  for (int64_t i = 0; i < N; ++i) {
    samples_indices[i] = sort_idx[i + 1] - 1;
    num_live_points[i] = crossed[i];
  }
  free(nodes);
  free(sort_idx);
  free(out_degree);
  free(crossed);
}

/* count_intervals_naive (internals/tree_structure.py:136-154): O(N^2)
 * brute force used only to pin count_crossed_edges on tiny trees. */
void orc_count_intervals_naive(const int64_t* sender, const double* log_L, int64_t N,
                               int64_t* samples_indices, int64_t* num_live_points) {
  double* nodes = (double*)malloc((size_t)(N + 1) * sizeof(double));
  double* cons = (double*)malloc((size_t)N * sizeof(double));
  int64_t* sort_idx = (int64_t*)malloc((size_t)N * sizeof(int64_t));
  char* avail = (char*)malloc((size_t)N);
  nodes[0] = -INFINITY;
  for (int64_t i = 0; i < N; ++i) nodes[i + 1] = log_L[i];
  for (int64_t i = 0; i < N; ++i) cons[i] = nodes[sender[i]];
  orc_argsort_stable(log_L, N, sort_idx);
  memset(avail, 1, (size_t)N);
  double contour = nodes[0];
  for (int64_t i = 0; i < N; ++i) {
    int64_t cnt = 0;
    for (int64_t j = 0; j < N; ++j) {
      int64_t s = sort_idx[j];
      if (cons[s] <= contour && log_L[s] > contour && avail[s]) cnt++;
    }
    num_live_points[i] = cnt;
    contour = nodes[sort_idx[i] + 1];
    avail[sort_idx[i]] = 0;
  }
  for (int64_t i = 0; i < N; ++i) samples_indices[i] = sort_idx[i];
  free(nodes);
  free(cons);
  free(sort_idx);
  free(avail);
}

/* ======================================================================
 * seed selection
 * ==================================================================== */

/* get_seed_point (samplers/uni_slice_sampler.py:343-358) via
 * sample_uniformly_masked / resample_indicies with replacement
 * (internals/random.py:34-84): log weights 0 where logL > contour else -inf,
 * log_p_cuml = cumulative_logsumexp, log_r = log_p_cuml[-1] + log(1-u),
 * idx = searchsorted(log_p_cuml, log_r, side='left') (clamped like a JAX gather). */
int64_t orc_seed_index(const double* live_log_L, int64_t N, double contour, double u) {
  double* w = (double*)malloc((size_t)N * sizeof(double));
  double* c = (double*)malloc((size_t)N * sizeof(double));
  for (int64_t i = 0; i < N; ++i) w[i] = live_log_L[i] > contour ? 0.0 : -INFINITY;
  orc_cumulative_logsumexp(w, N, c);
  double log_r = c[N - 1] + log(1.0 - u);
  int64_t lo = 0, hi = N; /* first i with c[i] >= log_r (c non-decreasing) */
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (c[mid] < log_r) lo = mid + 1;
    else hi = mid;
  }
  free(w);
  free(c);
  if (lo > N - 1) lo = N - 1;
  return lo;
}

/* ======================================================================
 * special functions
 * ==================================================================== */

/* Standard normal quantile: Wichura, "Algorithm AS 241: The percentage
 * points of the normal distribution", Applied Statistics 37 (1988) 477-484,
 * routine PPND16 (about 1e-16 relative accuracy).  Pinned against
 * scipy.special.ndtri in tests/test_oracle.py. */
double orc_ndtri(double p) {
  static const double a[8] = {3.3871328727963666080e0, 1.3314166789178437745e+2, 1.9715909503065514427e+3,
                              1.3731693765509461125e+4, 4.5921953931549871457e+4, 6.7265770927008700853e+4,
                              3.3430575583588128105e+4, 2.5090809287301226727e+3};
  static const double b[8] = {1.0, 4.2313330701600911252e+1, 6.8718700749205790830e+2, 5.3941960214247511077e+3,
                              2.1213794301586595867e+4, 3.9307895800092710610e+4, 2.8729085735721942674e+4,
                              5.2264952788528545610e+3};
  static const double c[8] = {1.42343711074968357734e0, 4.63033784615654529590e0, 5.76949722146069140550e0,
                              3.64784832476320460504e0, 1.27045825245236838258e0, 2.41780725177450611770e-1,
                              2.27238449892691845833e-2, 7.74545014278341407640e-4};
  static const double d[8] = {1.0, 2.05319162663775882187e0, 1.67638483018380384940e0, 6.89767334985100004550e-1,
                              1.48103976427480074590e-1, 1.51986665636164571966e-2, 5.47593808499534494600e-4,
                              1.05075007164441684324e-9};
  static const double e[8] = {6.65790464350110377720e0, 5.46378491116411436990e0, 1.78482653991729133580e0,
                              2.96560571828504891230e-1, 2.65321895265761230930e-2, 1.24266094738807843860e-3,
                              2.71155556874348757815e-5, 2.01033439929228813265e-7};
  static const double f[8] = {1.0, 5.99832206555887937690e-1, 1.36929880922735805310e-1, 1.48753612908506148525e-2,
                              7.86869131145613259100e-4, 1.84631831751005468180e-5, 1.42151175831644588870e-7,
                              2.04426310338993978564e-15};
  if (isnan(p)) return NAN;
  if (p <= 0.0) return p == 0.0 ? -INFINITY : NAN;
  if (p >= 1.0) return p == 1.0 ? INFINITY : NAN;
  double q = p - 0.5;
  if (fabs(q) <= 0.425) {
    double r = 0.180625 - q * q;
    double num = a[7], den = b[7];
    for (int i = 6; i >= 0; --i) {
      num = num * r + a[i];
      den = den * r + b[i];
    }
    return q * num / den;
  }
  double r = q < 0.0 ? p : 1.0 - p;
  r = sqrt(-log(r));
  double val;
  if (r <= 5.0) {
    r -= 1.6;
    double num = c[7], den = d[7];
    for (int i = 6; i >= 0; --i) {
      num = num * r + c[i];
      den = den * r + d[i];
    }
    val = num / den;
  } else {
    r -= 5.0;
    double num = e[7], den = f[7];
    for (int i = 6; i >= 0; --i) {
      num = num * r + e[i];
      den = den * r + f[i];
    }
    val = num / den;
  }
  return q < 0.0 ? -val : val;
}

double orc_ndtr(double x) { return 0.5 * erfc(-x / sqrt(2.0)); }

/* ======================================================================
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Pinned by the published known-answer vectors.
 * ==================================================================== */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* ======================================================================
 * model: prior transform and user likelihoods of the benchmark configs
 * ==================================================================== */

/* Prior.forward through TFP quantile + bijectors (framework/bases.py:58-68,
 * framework/wrapped_tfp_distribution.py:77-84): Normal -> loc + scale*ndtri(u),
 * Uniform -> low + (high-low)*u, MultivariateNormalTriL -> loc + L ndtri(u). */
void orc_prior_transform(const orc_model_t* m, const double* U, double* X) {
  int D = m->D;
  if (m->prior_kind == ORC_PRIOR_UNIFORM) {
    for (int i = 0; i < D; ++i) X[i] = m->prior_loc[i] + m->prior_scale[i] * U[i];
  } else if (m->prior_kind == ORC_PRIOR_NORMAL) {
    for (int i = 0; i < D; ++i) X[i] = m->prior_loc[i] + m->prior_scale[i] * orc_ndtri(U[i]);
  } else {
    double z[512];
    for (int i = 0; i < D; ++i) z[i] = orc_ndtri(U[i]);
    for (int i = 0; i < D; ++i) {
      double s = m->prior_loc[i];
      for (int j = 0; j <= i; ++j) s += m->prior_tril[i * D + j] * z[j];
      X[i] = s;
    }
  }
}

/* log N(x; mu, L L^T) the way tests/conftest.py:219-224 writes it:
 * -D/2 log(2 pi) - sum log diag(L) - 1/2 |L^-1 (x - mu)|^2 (forward substitution). */
static double mvn_loglik(int D, const double* x, const double* mu, const double* L) {
  double y[512];
  double q = 0.0, ld = 0.0;
  for (int i = 0; i < D; ++i) {
    double s = x[i] - mu[i];
    for (int j = 0; j < i; ++j) s -= L[i * D + j] * y[j];
    y[i] = s / L[i * D + i];
    q += y[i] * y[i];
    ld += log(L[i * D + i]);
  }
  return -0.5 * D * log(2.0 * M_PI) - ld - 0.5 * q;
}

/* Equal/arbitrary-weight isotropic Gaussian mixture: log sum_k w_k N(x; m_k, s_k^2 I). */
static double gmm_loglik(int D, int K, const double* x, const double* means, const double* sig, const double* logw) {
  double acc = -INFINITY;
  for (int k = 0; k < K; ++k) {
    double q = 0.0;
    for (int i = 0; i < D; ++i) {
      double r = x[i] - means[k * D + i];
      q += r * r;
    }
    double s2 = sig[k] * sig[k];
    double lk = logw[k] - 0.5 * D * log(2.0 * M_PI * s2) - 0.5 * q / s2;
    acc = orc_logaddexp(acc, lk);
  }
  return acc;
}

/* Gaussian linear regression: -N/2 log(2 pi sigma^2) - |y - X w|^2 / (2 sigma^2). */
static double linreg_loglik(int D, int64_t N, const double* w, const double* X, const double* y, double sigma) {
  double q = 0.0;
  for (int64_t n = 0; n < N; ++n) {
    double r = y[n];
    const double* row = X + n * D;
    for (int i = 0; i < D; ++i) r -= row[i] * w[i];
    q += r * r;
  }
  return -0.5 * (double)N * log(2.0 * M_PI * sigma * sigma) - 0.5 * q / (sigma * sigma);
}

/* Model.forward (framework/model.py:167-176) -> compute_log_likelihood
 * (framework/ops.py:302-326): logL = user_fn(prepare_input(U)); NaN -> -inf. */
double orc_log_likelihood(const orc_model_t* m, const double* U) {
  double x[512];
  orc_prior_transform(m, U, x);
  double v;
  switch (m->family) {
    case ORC_FAM_MVN: v = mvn_loglik(m->D, x, m->mvn_mu, m->mvn_chol); break;
    case ORC_FAM_GMM: v = gmm_loglik(m->D, m->K, x, m->gmm_means, m->gmm_sigma, m->gmm_logw); break;
    case ORC_FAM_LINREG: v = linreg_loglik(m->D, m->Ndata, x, m->lr_X, m->lr_y, m->lr_sigma); break;
    default: v = NAN;
  }
  if (isnan(v)) v = -INFINITY;
  return v;
}

/* Model.log_prob_prior (framework/model.py:178-185): sum of the priors' log densities at X(U). */
double orc_log_prob_prior(const orc_model_t* m, const double* U) {
  int D = m->D;
  double lp = 0.0;
  if (m->prior_kind == ORC_PRIOR_UNIFORM) {
    for (int i = 0; i < D; ++i) lp += -log(m->prior_scale[i]);
  } else if (m->prior_kind == ORC_PRIOR_NORMAL) {
    for (int i = 0; i < D; ++i) {
      double z = orc_ndtri(U[i]);
      lp += -0.5 * z * z - 0.5 * log(2.0 * M_PI) - log(m->prior_scale[i]);
    }
  } else {
    for (int i = 0; i < D; ++i) {
      double z = orc_ndtri(U[i]);
      lp += -0.5 * z * z - 0.5 * log(2.0 * M_PI) - log(m->prior_tril[i * D + i]);
    }
  }
  return lp;
}

/* ======================================================================
 * slice sampler (samplers/uni_slice_sampler.py), binary32 and binary64
 * ==================================================================== */

/* The template below is instantiated twice.  REAL is the arithmetic type of
 * the U-space state (the reference uses float64; the device plain form uses
 * float32).  Likelihood values are always binary64. */
#define SLICE_TEMPLATE(REAL, SFX)                                                                             \
  /* _slice_bounds (uni_slice_sampler.py:41-64) */                                                          \
  void orc_slice_bounds_##SFX(const REAL* U0, const REAL* n, int D, REAL* left, REAL* right) {              \
    REAL t1r = (REAL)INFINITY, t1l = (REAL)-INFINITY, t0r = (REAL)INFINITY, t0l = (REAL)-INFINITY;          \
    for (int i = 0; i < D; ++i) {                                                                           \
      REAL t1 = ((REAL)1 - U0[i]) / n[i];                                                                   \
      REAL t0 = -U0[i] / n[i];                                                                              \
      REAL c1r = t1 >= (REAL)0 ? t1 : (REAL)INFINITY;                                                       \
      REAL c1l = t1 <= (REAL)0 ? t1 : (REAL)-INFINITY;                                                      \
      REAL c0r = t0 >= (REAL)0 ? t0 : (REAL)INFINITY;                                                       \
      REAL c0l = t0 <= (REAL)0 ? t0 : (REAL)-INFINITY;                                                      \
      if (c1r < t1r) t1r = c1r;                                                                             \
      if (c1l > t1l) t1l = c1l;                                                                             \
      if (c0r < t0r) t0r = c0r;                                                                             \
      if (c0l > t0l) t0l = c0l;                                                                             \
    }                                                                                                       \
    *right = t0r < t1r ? t0r : t1r;                                                                         \
    *left = t0l > t1l ? t0l : t1l;                                                                          \
  }                                                                                                         \
                                                                                                            \
  static double loglik_##SFX(const orc_model_t* m, const REAL* U) {                                         \
    double u64[512];                                                                                        \
    for (int i = 0; i < m->D; ++i) u64[i] = (double)U[i];                                                   \
    return orc_log_likelihood(m, u64);                                                                      \
  }                                                                                                         \
                                                                                                            \
  /* get_sample_from_seed (uni_slice_sampler.py:360-441) over _new_proposal (:114-273):                     \
   * per slice: perfect bracket, uniform pick, while-not-accepted shrink (midpoint: moved                   \
   * bound becomes alpha*t, alpha = linspace(0.5, 1, num_slices)), acceptance                               \
   * logL > contour or (logL0 == contour and logL == contour). */                                           \
  int orc_slice_chain_##SFX(const orc_model_t* m, const REAL* U0in, double logL0, double contour,          \
                            int num_slices, int midpoint_shrink, int k_phantom, const REAL* dirs,           \
                            const REAL* tu, int max_steps, const double* logL_replay, REAL* U_out,          \
                            double* logL_out, int64_t* evals_out, REAL* phantom_U, double* phantom_logL,    \
                            REAL* trace_t, double* trace_logL, int32_t* steps_per_slice) {                 \
    int D = m->D;                                                                                           \
    REAL U0[512], U[512];                                                                                   \
    memcpy(U0, U0in, sizeof(REAL) * (size_t)D);                                                             \
    int64_t evals_total = 0;                                                                                \
    double cur_logL = logL0;                                                                                \
    for (int s = 0; s < num_slices; ++s) {                                                                  \
      double alpha64 = num_slices > 1 ? 0.5 + 0.5 * (double)s / (double)(num_slices - 1) : 0.5;            \
      REAL alpha = (REAL)alpha64;                                                                           \
      const REAL* n = dirs + (size_t)s * D;                                                                 \
      REAL left, right;                                                                                     \
      orc_slice_bounds_##SFX(U0, n, D, &left, &right);                                                      \
      int k = 0;                                                                                            \
      REAL t = left + tu[(size_t)s * max_steps + 0] * (right - left);                                       \
      for (int i = 0; i < D; ++i) U[i] = U0[i] + t * n[i];                                                  \
      double lL = logL_replay ? logL_replay[(size_t)s * max_steps + 0] : loglik_##SFX(m, U);                \
      if (trace_t) trace_t[(size_t)s * max_steps + 0] = t;                                                  \
      if (trace_logL) trace_logL[(size_t)s * max_steps + 0] = lL;                                           \
      int64_t ev = 1;                                                                                       \
      while (!((lL > contour) || (cur_logL == contour && lL == contour))) {                                 \
        ++k;                                                                                                \
        if (k >= max_steps) return -1;                                                                      \
        if (t < (REAL)0) left = t;                                                                          \
        if (t > (REAL)0) right = t;                                                                         \
        if (midpoint_shrink) {                                                                              \
          if (t < (REAL)0) left = alpha * left;                                                             \
          if (t > (REAL)0) right = alpha * right;                                                           \
        }                                                                                                   \
        t = left + tu[(size_t)s * max_steps + k] * (right - left);                                          \
        for (int i = 0; i < D; ++i) U[i] = U0[i] + t * n[i];                                                \
        lL = logL_replay ? logL_replay[(size_t)s * max_steps + k] : loglik_##SFX(m, U);                     \
        if (trace_t) trace_t[(size_t)s * max_steps + k] = t;                                                \
        if (trace_logL) trace_logL[(size_t)s * max_steps + k] = lL;                                         \
        ++ev;                                                                                               \
      }                                                                                                     \
      if (steps_per_slice) steps_per_slice[s] = k + 1;                                                      \
      evals_total += ev;                                                                                    \
      memcpy(U0, U, sizeof(REAL) * (size_t)D);                                                              \
      cur_logL = lL;                                                                                        \
      /* phantoms: x[-(k+1):-1] of the per-slice samples (uni_slice_sampler.py:432) */                     \
      int pidx = s - (num_slices - 1 - k_phantom);                                                          \
      if (k_phantom > 0 && pidx >= 0 && s < num_slices - 1) {                                               \
        if (phantom_U) memcpy(phantom_U + (size_t)pidx * D, U, sizeof(REAL) * (size_t)D);                   \
        if (phantom_logL) phantom_logL[pidx] = lL;                                                          \
      }                                                                                                     \
    }                                                                                                       \
    memcpy(U_out, U0, sizeof(REAL) * (size_t)D);                                                            \
    *logL_out = cur_logL;                                                                                   \
    *evals_out = evals_total;                                                                               \
    return 0;                                                                                               \
  }

SLICE_TEMPLATE(double, f64)
SLICE_TEMPLATE(float, f32)

/* ======================================================================
 * termination
 * ==================================================================== */

/* create_init_termination_register (nested_samplers/common/initialisation.py:87-108) */
void orc_init_register(orc_register_t* r) {
  memset(r, 0, sizeof(*r));
  orc_evidence_init(r->evidence_calc);
  orc_evidence_init(r->evidence_calc_with_remaining);
  r->log_L_contour = -INFINITY;
  r->efficiency = 0.0;
  r->relative_spread = INFINITY;
  r->absolute_spread = INFINITY;
  r->peak_log_XL = -INFINITY;
}

/* determine_termination for a single TerminationCondition
 * (nested_samplers/common/termination.py:67-147). Returns done; writes the bit mask. */
int orc_determine_termination(const orc_term_cond_t* c, const orc_register_t* r, int64_t* reason) {
  int done = 0;
  int64_t why = 0;
#define SETBIT(cond, bit)        \
  do {                           \
    if (cond) {                  \
      done = 1;                  \
      why += ((int64_t)1 << bit); \
    }                            \
  } while (0)
  if (c->has_max_samples) SETBIT((double)r->num_samples_used >= c->max_samples, 0);
  if (c->has_evidence_uncert) {
    double mu, var;
    orc_linear_to_log_stats(r->evidence_calc_with_remaining[3], r->evidence_calc_with_remaining[5], &mu, &var);
    SETBIT(var <= c->evidence_uncert * c->evidence_uncert, 1);
  }
  if (c->has_dlogZ) {
    double mu1, v1, mu0, v0;
    orc_linear_to_log_stats(r->evidence_calc_with_remaining[3], r->evidence_calc_with_remaining[5], &mu1, &v1);
    orc_linear_to_log_stats(r->evidence_calc[3], r->evidence_calc[5], &mu0, &v0);
    SETBIT((mu1 - mu0) < c->dlogZ, 2);
  }
  if (c->has_ess) {
    double ess = orc_ess_kish(r->evidence_calc_with_remaining[3], r->evidence_calc_with_remaining[7]);
    SETBIT(ess >= c->ess, 3);
  }
  if (c->has_max_num_likelihood_evaluations)
    SETBIT((double)r->num_likelihood_evaluations >= c->max_num_likelihood_evaluations, 4);
  if (c->has_log_L_contour) SETBIT(r->log_L_contour >= c->log_L_contour, 5);
  if (c->has_efficiency_threshold) SETBIT(r->efficiency < c->efficiency_threshold, 6);
  SETBIT(r->plateau, 7);
  if (c->has_rtol) SETBIT(r->relative_spread < c->rtol, 8);
  if (c->has_atol) SETBIT(r->absolute_spread < c->atol, 9);
  SETBIT(r->no_seed_points, 10);
  if (c->has_peak_XL_frac) {
    double log_XL = r->evidence_calc[1] + r->evidence_calc[0];
    SETBIT(log_XL < r->peak_log_XL + log(c->peak_XL_frac), 11);
  }
#undef SETBIT
  *reason = why;
  return done;
}

/* ======================================================================
 * full run (CPU baseline): sharded_static.py restated on host cores
 * ==================================================================== */

typedef struct {
  uint32_t key[2];
  int dir16;
} orc_rng_t;

static double u01_from(uint32_t a, uint32_t b) { /* 53-bit uniform in [0,1) */
  uint64_t v = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
  v &= ((uint64_t)1 << 53) - 1;
  return (double)v * (1.0 / 9007199254740992.0);
}

static void rng_block(const orc_rng_t* g, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4]) {
  uint32_t ctr[4] = {c0, c1, c2, c3};
  orc_philox4x32(ctr, g->key, out);
}

static double rng_uniform(const orc_rng_t* g, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
  uint32_t o[4];
  rng_block(g, c0, c1, c2, c3, o);
  return u01_from(o[0], o[1]);
}

static double rng_normal(const orc_rng_t* g, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
  uint32_t o[4];
  rng_block(g, c0, c1, c2, c3, o);
  double u1 = 1.0 - u01_from(o[0], o[1]); /* (0, 1] */
  double u2 = u01_from(o[2], o[3]);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* _sample_direction (uni_slice_sampler.py:23-38) */
static void sample_direction(const orc_rng_t* g, int D, uint32_t chain, uint32_t it, uint32_t slice, double* n) {
  if (D == 1) {
    n[0] = 1.0;
    return;
  }
  double ss = 0.0;
  if (g->dir16) {
    /* the device's form: each 32-bit word of a Philox block gives u1 = (low 16 bits + 1) / 2^16 in
     * (0, 1] and u2 = (high 16 bits) / 2^16, and the pair r cos(2 pi u2), r sin(2 pi u2) */
    for (int i0 = 0; i0 < D; i0 += 8) {
      uint32_t w[4];
      rng_block(g, chain, it, slice, (3u << 24) | (uint32_t)(i0 >> 3), w);
      for (int q = 0; q < 4; ++q) {
        double u1 = (double)((w[q] & 0xFFFFu) + 1u) / 65536.0;
        double u2 = (double)(w[q] >> 16) / 65536.0;
        double r = sqrt(-2.0 * log(u1));
        if (i0 + 2 * q < D) n[i0 + 2 * q] = r * cos(2.0 * M_PI * u2);
        if (i0 + 2 * q + 1 < D) n[i0 + 2 * q + 1] = r * sin(2.0 * M_PI * u2);
      }
    }
    for (int i = 0; i < D; ++i) ss += n[i] * n[i];
  } else {
    for (int i = 0; i < D; ++i) {
      n[i] = rng_normal(g, chain, it, slice, (1u << 24) | (uint32_t)i);
      ss += n[i] * n[i];
    }
  }
  double nrm = sqrt(ss);
  for (int i = 0; i < D; ++i) n[i] /= nrm;
}

/* one UniDimSliceSampler chain with on-the-fly draws (same logic as orc_slice_chain_f64) */
static int64_t slice_chain_rng(const orc_model_t* m, const orc_rng_t* g, uint32_t chain, uint32_t it, double* U0,
                               double* logL0, double contour, int num_slices, int midpoint, int k_phantom,
                               double* ph_U, double* ph_logL) {
  int D = m->D;
  double n[512], U[512];
  int64_t evals_total = 0;
  for (int s = 0; s < num_slices; ++s) {
    double alpha = num_slices > 1 ? 0.5 + 0.5 * (double)s / (double)(num_slices - 1) : 0.5;
    sample_direction(g, D, chain, it, (uint32_t)s, n);
    double left, right;
    orc_slice_bounds_f64(U0, n, D, &left, &right);
    uint32_t k = 0;
    double t = left + rng_uniform(g, chain, it, (uint32_t)s, (2u << 24) | k) * (right - left);
    for (int i = 0; i < D; ++i) U[i] = U0[i] + t * n[i];
    double lL = orc_log_likelihood(m, U);
    int64_t ev = 1;
    while (!((lL > contour) || (*logL0 == contour && lL == contour))) {
      ++k;
      if (t < 0.0) left = t;
      if (t > 0.0) right = t;
      if (midpoint) {
        if (t < 0.0) left = alpha * left;
        if (t > 0.0) right = alpha * right;
      }
      t = left + rng_uniform(g, chain, it, (uint32_t)s, (2u << 24) | k) * (right - left);
      for (int i = 0; i < D; ++i) U[i] = U0[i] + t * n[i];
      lL = orc_log_likelihood(m, U);
      ++ev;
      if (k > 100000) break; /* cannot happen for a continuous likelihood; keeps the baseline bounded */
    }
    evals_total += ev;
    memcpy(U0, U, sizeof(double) * (size_t)D);
    *logL0 = lL;
    int pidx = s - (num_slices - 1 - k_phantom);
    if (k_phantom > 0 && pidx >= 0 && s < num_slices - 1) {
      memcpy(ph_U + (size_t)pidx * D, U, sizeof(double) * (size_t)D);
      ph_logL[pidx] = lL;
    }
  }
  return evals_total;
}

typedef struct {
  int64_t* sender;
  double* U;
  double* cons;
  double* logL;
  int64_t* evals;
} live_t;

static void live_alloc(live_t* L, int N, int D) {
  L->sender = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
  L->U = (double*)malloc(sizeof(double) * (size_t)N * D);
  L->cons = (double*)malloc(sizeof(double) * (size_t)N);
  L->logL = (double*)malloc(sizeof(double) * (size_t)N);
  L->evals = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
}
static void live_free(live_t* L) {
  free(L->sender);
  free(L->U);
  free(L->cons);
  free(L->logL);
  free(L->evals);
}
static void live_permute(live_t* dst, const live_t* src, const int64_t* perm, int N, int D) {
  for (int i = 0; i < N; ++i) {
    int64_t p = perm[i];
    dst->sender[i] = src->sender[p];
    dst->cons[i] = src->cons[p];
    dst->logL[i] = src->logL[p];
    dst->evals[i] = src->evals[p];
    memcpy(dst->U + (size_t)i * D, src->U + (size_t)p * D, sizeof(double) * (size_t)D);
  }
}

/* replace_index (internals/maps.py:15-25) writes through lax.dynamic_update_slice, which clamps the
 * start index so that the update stays inside the operand: the first index written. */
int64_t orc_update_start(int64_t start, int64_t count, int64_t size) {
  int64_t last = size - count;
  if (last < 0) last = 0;
  if (start < 0) start = 0;
  return start < last ? start : last;
}

/* _add_samples_to_state (sharded_static.py:40-85): write at next_sample_idx through replace_index (the
 * clamped start), then next_sample_idx advances modulo max_samples */
static void add_samples(int64_t* next_idx, int64_t* num_samples, int64_t max_samples, int D, int n,
                        const int64_t* sender, const double* U, const double* logL, const int64_t* evals,
                        int phantom, double* sU, double* sL, int64_t* sS, int64_t* sE, int8_t* sP) {
  const int64_t start = orc_update_start(*next_idx, n, max_samples);
  for (int i = 0; i < n; ++i) {
    int64_t j = start + i;
    sS[j] = sender[i];
    sL[j] = logL[i];
    sE[j] = evals ? evals[i] : 0;
    sP[j] = (int8_t)phantom;
    memcpy(sU + (size_t)j * D, U + (size_t)i * D, sizeof(double) * (size_t)D);
  }
  *next_idx = (*next_idx + n) % max_samples;
  *num_samples += n;
}

int orc_to_results(const double* samp_logL, const int64_t* samp_sender, const int64_t* samp_evals,
                   const int8_t* samp_phantom, int64_t M, int64_t* si, int64_t* nl, double* per8,
                   orc_run_result_t* res) {
  /* sharded_static.py:652-745, trim=True */
  orc_count_crossed_edges(samp_sender, samp_logL, M, -1, si, nl);
  double* lL = (double*)malloc(sizeof(double) * (size_t)M);
  double* nn = (double*)malloc(sizeof(double) * (size_t)M);
  int64_t tot_evals = 0, tot_ph = 0;
  for (int64_t i = 0; i < M; ++i) {
    lL[i] = samp_logL[si[i]];
    nn[i] = (double)nl[i];
    tot_evals += samp_evals[si[i]];
    tot_ph += samp_phantom[i] ? 1 : 0;
  }
  double fin[8];
  orc_compute_evidence_stats(lL, nn, M, -1, fin, per8);
  double mu, var;
  orc_linear_to_log_stats(fin[3], fin[5], &mu, &var);
  double unc = sqrt(var);
  double pf = (double)tot_ph / (double)M;
  double k = pf / (1.0 - pf);
  unc *= sqrt(1.0 + k);
  double ess = orc_ess_kish(fin[3], fin[7]) / (1.0 + k);
  /* H: normalise_log_space(log_dZ) then -(sum p logL - log Z), fallback -(sum p (-log X)) */
  double mx = -INFINITY;
  for (int64_t i = 0; i < M; ++i) {
    double v = per8[8 * i + 6];
    if (v > mx) mx = v;
  }
  double ssum = 0.0;
  for (int64_t i = 0; i < M; ++i) ssum += exp(per8[8 * i + 6] - mx);
  double lnorm = mx + log(ssum);
  double h1 = 0.0, h2 = 0.0;
  for (int64_t i = 0; i < M; ++i) {
    double ldp = isinf(lnorm) && lnorm < 0 ? -INFINITY : per8[8 * i + 6] - lnorm;
    double p = exp(ldp);
    double v = (isinf(ldp) && ldp < 0) ? 0.0 : lL[i];
    h1 += p * v;
    h2 += p * (-per8[8 * i + 1]);
  }
  double H1 = -(h1 - mu);
  double H2 = -h2;
  res->log_Z_mean = mu;
  res->log_Z_uncert = unc;
  res->ESS = ess;
  res->H_mean = isfinite(H1) ? H1 : H2;
  res->total_likelihood_evaluations = tot_evals;
  res->total_phantom_samples = tot_ph;
  res->log_efficiency = log((double)M) - log((double)tot_evals);
  res->num_samples = M;
  free(lL);
  free(nn);
  return 0;
}

int orc_ns_run(const orc_model_t* m, const orc_run_cfg_t* cfg, const orc_term_cond_t* term_in, double* sU,
               double* sL, int64_t* sS, int64_t* sE, int8_t* sP, orc_run_result_t* res) {
  const int D = m->D;
  const int N = cfg->num_live_points;
  const int kph = cfg->k_phantom;
  const int64_t max_samples = cfg->max_samples;
  orc_rng_t g;
  g.key[0] = (uint32_t)cfg->seed;
  g.key[1] = (uint32_t)(cfg->seed >> 32);
  g.dir16 = cfg->dir16;
#ifdef _OPENMP
  int nthr = cfg->num_threads > 0 ? cfg->num_threads : omp_get_max_threads();
#endif
  /* ShardedStaticNestedSampler.__post_init__ (sharded_static.py:625-647), one device */
  double shell_fraction = 0.5;
  if (1.0 / N > shell_fraction) shell_fraction = 1.0 / N;
  const int shell = (int)(N * shell_fraction);

  for (int64_t j = 0; j < max_samples; ++j) { /* create_init_state buffers (initialisation.py:39-45) */
    sS[j] = 0;
    sL[j] = INFINITY;
    sE[j] = 0;
    sP[j] = 0;
    for (int i = 0; i < D; ++i) sU[j * D + i] = 0.5;
  }
  live_t live, tmp;
  live_alloc(&live, N, D);
  live_alloc(&tmp, N, D);
  int64_t next_idx = 0, num_samples = 0, total_evals_init = 0;

  /* create_init_state: _single_uniform_sample per live point (uniform_sample.py:12-61) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthr) schedule(dynamic, 4)
#endif
  for (int c = 0; c < N; ++c) {
    double U[512];
    uint32_t att = 0;
    double lL;
    int64_t ev = 0;
    do {
      for (int i = 0; i < D; ++i) U[i] = rng_uniform(&g, (uint32_t)c, 0xFFFFFFFFu, att, (3u << 24) | (uint32_t)i);
      lL = orc_log_likelihood(m, U);
      ++ev;
      ++att;
    } while (lL <= -INFINITY && att < 1000000u);
    memcpy(tmp.U + (size_t)c * D, U, sizeof(double) * (size_t)D);
    tmp.logL[c] = lL;
    tmp.cons[c] = -INFINITY;
    tmp.sender[c] = 0;
    tmp.evals[c] = ev;
  }
  for (int c = 0; c < N; ++c) total_evals_init += tmp.evals[c];
  int64_t* perm = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
  orc_argsort_stable(tmp.logL, N, perm);
  live_permute(&live, &tmp, perm, N, D);

  orc_register_t reg;
  orc_init_register(&reg);

  /* phase 1 (uniform sampling to init_efficiency_threshold, sharded_static.py:793-815):
   * its termination is tested on the initial register first (efficiency 0), exactly as
   * the reference's while_loop does, so it runs zero shells for any threshold > 0. */
  if (cfg->init_efficiency_threshold > 0.0) {
    orc_term_cond_t uc;
    memset(&uc, 0, sizeof(uc));
    uc.has_efficiency_threshold = 1;
    uc.efficiency_threshold = cfg->init_efficiency_threshold;
    uc.has_dlogZ = 1;
    uc.dlogZ = 0.0;
    uc.has_max_samples = 1;
    uc.max_samples = (double)max_samples;
    int64_t r0;
    reg.no_seed_points = live.logL[shell - 1] >= live.logL[N - 1];
    int done0 = orc_determine_termination(&uc, &reg, &r0);
    (void)done0; /* with done0 == 1 the uniform phase body never executes */
  }

  /* phase 2: _main_ns_thread with the slice sampler (sharded_static.py:427-574) */
  orc_term_cond_t term = *term_in;
  double space_needed = (double)shell * (1 + kph);
  double cap = (double)max_samples - space_needed;
  if (term.has_max_samples) {
    if (cap < term.max_samples) term.max_samples = cap;
  }
  reg.no_seed_points = live.logL[shell - 1] >= live.logL[N - 1];

  double* new_U = (double*)malloc(sizeof(double) * (size_t)shell * D);
  double* new_logL = (double*)malloc(sizeof(double) * (size_t)shell);
  int64_t* new_evals = (int64_t*)malloc(sizeof(int64_t) * (size_t)shell);
  double* ph_U = (double*)malloc(sizeof(double) * (size_t)shell * (kph > 0 ? kph : 1) * D);
  double* ph_logL = (double*)malloc(sizeof(double) * (size_t)shell * (kph > 0 ? kph : 1));
  int64_t* ph_sender = (int64_t*)malloc(sizeof(int64_t) * (size_t)shell * (kph > 0 ? kph : 1));
  double* disc_logL = (double*)malloc(sizeof(double) * (size_t)shell);
  int64_t reason = 0;
  int64_t it = 0;
  while (1) {
    if (orc_determine_termination(&term, &reg, &reason)) break;
    if (cfg->max_iterations >= 0 && it >= cfg->max_iterations) break;
    /* _collect_shell (sharded_static.py:210-324) */
    add_samples(&next_idx, &num_samples, max_samples, D, shell, live.sender, live.U, live.logL, live.evals, 0, sU,
                sL, sS, sE, sP);
    memcpy(disc_logL, live.logL, sizeof(double) * (size_t)shell);
    double contour = live.logL[shell - 1];
    int64_t sender_node = next_idx - 1; /* sharded_static.py:261 */
    /* seed points chosen from the pre-replacement live set (pre_process returns it) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthr) schedule(dynamic, 1)
#endif
    for (int c = 0; c < shell; ++c) {
      double u = rng_uniform(&g, (uint32_t)c, (uint32_t)it, 0xFFFFFFFFu, 0u);
      int64_t idx = orc_seed_index(live.logL, N, contour, u);
      double U0[512];
      memcpy(U0, live.U + (size_t)idx * D, sizeof(double) * (size_t)D);
      double l0 = live.logL[idx];
      int64_t ev = slice_chain_rng(m, &g, (uint32_t)c, (uint32_t)it, U0, &l0, contour, cfg->num_slices,
                                   cfg->midpoint_shrink, kph, ph_U + (size_t)c * kph * D, ph_logL + (size_t)c * kph);
      memcpy(new_U + (size_t)c * D, U0, sizeof(double) * (size_t)D);
      new_logL[c] = l0;
      new_evals[c] = ev;
    }
    /* live[:shell] = new; stable argsort; permute (sharded_static.py:269-275) */
    int64_t new_sum_evals = 0;
    for (int c = 0; c < shell; ++c) {
      live.sender[c] = sender_node;
      live.cons[c] = contour;
      live.logL[c] = new_logL[c];
      live.evals[c] = new_evals[c];
      memcpy(live.U + (size_t)c * D, new_U + (size_t)c * D, sizeof(double) * (size_t)D);
      new_sum_evals += new_evals[c];
    }
    orc_argsort_stable(live.logL, N, perm);
    live_permute(&tmp, &live, perm, N, D);
    {
      live_t sw = live;
      live = tmp;
      tmp = sw;
    }
    if (kph > 0) {
      for (int i = 0; i < shell * kph; ++i) ph_sender[i] = sender_node;
      add_samples(&next_idx, &num_samples, max_samples, D, shell * kph, ph_sender, ph_U, ph_logL, NULL, 1, sU, sL,
                  sS, sE, sP);
    }
    /* register (sharded_static.py:284-323) */
    double ec[8], ecr[8];
    memcpy(ec, reg.evidence_calc, sizeof(ec));
    for (int i = 0; i < shell; ++i) orc_evidence_update(ec, (double)N, disc_logL[i]);
    memcpy(ecr, ec, sizeof(ec));
    for (int i = 0; i < N; ++i) orc_evidence_update(ecr, (double)(N - i), live.logL[i]);
    int64_t live_evals = 0;
    for (int i = 0; i < N; ++i) live_evals += live.evals[i];
    reg.num_samples_used = num_samples;
    memcpy(reg.evidence_calc, ec, sizeof(ec));
    memcpy(reg.evidence_calc_with_remaining, ecr, sizeof(ecr));
    reg.num_likelihood_evaluations += new_sum_evals;
    reg.log_L_contour = contour;
    reg.efficiency = (double)N / (double)live_evals;
    int plateau = 1;
    for (int i = 0; i < N; ++i)
      if (live.logL[i] != live.logL[0]) plateau = 0;
    reg.plateau = plateau;
    reg.absolute_spread = fabs(live.logL[N - 1] - live.logL[0]);
    reg.relative_spread = 2.0 * reg.absolute_spread / fabs(live.logL[0] + live.logL[N - 1]);
    reg.no_seed_points = live.logL[shell - 1] >= live.logL[N - 1];
    double pxl = ec[1] + ec[0];
    reg.peak_log_XL = reg.peak_log_XL > pxl ? reg.peak_log_XL : pxl;
    ++it;
  }
  orc_determine_termination(&term, &reg, &reason);
  /* consume live points (sharded_static.py:834-838) */
  add_samples(&next_idx, &num_samples, max_samples, D, N, live.sender, live.U, live.logL, live.evals, 0, sU, sL, sS,
              sE, sP);

  int64_t M = num_samples < max_samples ? num_samples : max_samples;
  int64_t* si = (int64_t*)malloc(sizeof(int64_t) * (size_t)M);
  int64_t* nl = (int64_t*)malloc(sizeof(int64_t) * (size_t)M);
  double* per8 = (double*)malloc(sizeof(double) * 8 * (size_t)M);
  orc_to_results(sL, sS, sE, sP, M, si, nl, per8, res);
  res->termination_reason = reason;
  res->num_iterations = it;
  res->num_samples = num_samples;
  (void)total_evals_init;
  free(si);
  free(nl);
  free(per8);
  free(new_U);
  free(new_logL);
  free(new_evals);
  free(ph_U);
  free(ph_logL);
  free(ph_sender);
  free(disc_logL);
  free(perm);
  live_free(&live);
  live_free(&tmp);
  return 0;
}
