// ns_shell.hip -- per-shell bookkeeping kernels of the nested-sampling loop.
//
// One shell (nested_samplers/sharded/sharded_static.py:210-324) is
//   k_discard  -> k_sample (ns_sampler_*.hip) -> k_replace -> k_register
// with every kernel covering all runs of a batch (grid.y or grid.x = run) and
// skipping runs whose state says done.  The live set is double-buffered; the
// run state holds which buffer is current, so a captured sequence of shells
// replays correctly.
#include "hipns_dev.h"
#include "ns_shell.h"

namespace hipns {

// ---------------------------------------------------------------------------
// _add_samples_to_state(discarded) (sharded_static.py:40-85, 240-245) plus the
// contour, seed range and sender of this shell (sharded_static.py:250-261).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_discard(const hipns_run_desc_t* runs, hipns_run_state_t* states,
                                                 hipns_buffers_t B) {
  const int r = blockIdx.y;
  hipns_run_state_t* st = states + r;
  if (st->done) return;
  const hipns_run_desc_t run = runs[r];
  const int S = run.shell, N = run.N, dps = B.dps, cur = st->buf;
  const int64_t next = st->next_idx;
  const int64_t start = append_start(next, S, run.max_samples);
  // one float4 of a row per thread (dps is a multiple of 32)
  const uint32_t q = (uint32_t)dps >> 2;
  const float4* srcU = (const float4*)(B.live_U[cur] + run.live_off * dps);
  float4* dstU = (float4*)B.samp_U;
  const uint32_t total = (uint32_t)S * q;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    const uint32_t i = e / q, c = e - i * q;
    const int64_t j = run.samp_off + start + i;
    dstU[j * q + c] = srcU[e];
    if (c == 0) {
      B.samp_logL[j] = B.live_logL[cur][run.live_off + i];
      B.samp_sender[j] = B.live_sender[cur][run.live_off + i];
      B.samp_evals[j] = B.live_evals[cur][run.live_off + i];
      B.samp_phantom[j] = 0;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double* L = B.live_logL[cur] + run.live_off;
    double contour = L[S - 1];
    // first index with logL > contour (the live set is sorted ascending)
    int lo = 0, hi = N;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (kless(contour, L[mid])) hi = mid;
      else lo = mid + 1;
    }
    st->contour = contour;
    st->first_seed = lo;
    // sender_node_idx = next_sample_idx - 1 after the discard (sharded_static.py:261)
    st->sender_node = (int32_t)((next + S) % run.max_samples) - 1;
  }
}

// ---------------------------------------------------------------------------
// Replacement: live[:S] = new samples, then stable argsort of the whole live
// set by logL and permutation of every field (sharded_static.py:262-275; at
// initialisation S = N, initialisation.py:68-76).  The surviving live[S:N] are
// already sorted, so the new samples are bitonic-sorted by (logL, index) in LDS
// and merged: a new sample precedes a survivor with an equal key, exactly the
// order jnp.argsort(stable=True) gives to the concatenation [new, survivors].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_replace(const hipns_run_desc_t* runs, hipns_run_state_t* states,
                                                  hipns_buffers_t B, int init_mode) {
  extern __shared__ unsigned char smem[];
  const int r = blockIdx.x;
  hipns_run_state_t* st = states + r;
  if (!init_mode && st->done) return;
  const hipns_run_desc_t run = runs[r];
  const int N = run.N;
  const int S = init_mode ? N : run.shell;
  const int dps = B.dps;
  const int cur = init_mode ? 0 : st->buf;
  const int nxt = init_mode ? 0 : 1 - cur;
  int P = 1;
  while (P < S) P <<= 1;
  double* keys = (double*)smem;
  int* idx = (int*)(keys + P);
  int* perm = idx + P;
  const int tid = threadIdx.x, nt = blockDim.x;
  // padding slots hold NaN with indices >= S: they sort after every real key, NaN included (NaN keys
  // tie-break on the index)
  for (int i = tid; i < P; i += nt) {
    keys[i] = i < S ? B.new_logL[run.chain_off + i] : __builtin_nan("");
    idx[i] = i;
  }
  __syncthreads();
  // bitonic sort of (key, index)
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += nt) {
        int p = i ^ j;
        if (p > i) {
          bool up = (i & k) == 0;
          double ki = keys[i], kp = keys[p];
          int ii = idx[i], ip = idx[p];
          bool sw = up ? key_lt(kp, ip, ki, ii) : key_lt(ki, ii, kp, ip);
          if (sw) {
            keys[i] = kp;
            keys[p] = ki;
            idx[i] = ip;
            idx[p] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
  const double* surv = B.live_logL[cur] + run.live_off + S;  // survivors, sorted
  const int NS = N - S;
  for (int i = tid; i < S; i += nt) {
    double kk = keys[i];
    int lo = 0, hi = NS;  // survivors strictly less than kk
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (kless(surv[mid], kk)) lo = mid + 1;
      else hi = mid;
    }
    perm[i + lo] = idx[i];
  }
  for (int j = tid; j < NS; j += nt) {
    double kk = surv[j];
    int lo = 0, hi = S;  // new samples with key <= kk
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (!kless(kk, keys[mid])) lo = mid + 1;
      else hi = mid;
    }
    perm[j + lo] = S + j;
  }
  __syncthreads();
  const double contour = init_mode ? -__builtin_inf() : st->contour;
  const int sender = init_mode ? 0 : st->sender_node;
  const int64_t lo_ = run.live_off;
  for (int p = tid; p < N; p += nt) {
    int src = perm[p];
    if (src < S) {
      B.live_logL[nxt][lo_ + p] = B.new_logL[run.chain_off + src];
      B.live_cons[nxt][lo_ + p] = contour;
      B.live_sender[nxt][lo_ + p] = sender;
      B.live_evals[nxt][lo_ + p] = B.new_evals[run.chain_off + src];
    } else {
      B.live_logL[nxt][lo_ + p] = B.live_logL[cur][lo_ + src];
      B.live_cons[nxt][lo_ + p] = B.live_cons[cur][lo_ + src];
      B.live_sender[nxt][lo_ + p] = B.live_sender[cur][lo_ + src];
      B.live_evals[nxt][lo_ + p] = B.live_evals[cur][lo_ + src];
    }
  }
  // rows moved one float4 per thread (dps is a multiple of 32)
  const uint32_t q = (uint32_t)dps >> 2;
  const float4* newU = (const float4*)B.new_U;
  const float4* curU = (const float4*)B.live_U[cur];
  float4* nxtU = (float4*)B.live_U[nxt];
  const uint32_t tot = (uint32_t)N * q;
  for (uint32_t e = tid; e < tot; e += nt) {
    const uint32_t p = e / q, c = e - p * q;
    const int src = perm[p];
    nxtU[(lo_ + p) * q + c] = src < S ? newU[(run.chain_off + src) * q + c] : curU[(lo_ + src) * q + c];
  }
}

size_t replace_lds_bytes(int N) {
  int P = 1;
  while (P < N) P <<= 1;
  return (size_t)P * 8 + (size_t)P * 4 + (size_t)N * 4;
}

// Aggregate maps of two element sequences at once: waves [0, w1) of the block cover the first
// (n_i = N, the discarded shell), waves [w1, nw) the second (n_i = N - i, the remaining live points).
// Each thread composes a contiguous run of its sequence, each wavefront reduces its lanes in order,
// and a pairwise tree over the wavefront maps of each sequence (both trees in the same passes)
// finishes.  The serial depth is a few compositions per thread plus log2 of the lane and wave counts.
// Returns through lds_maps[0] (first sequence) and lds_maps[w1] (second).
struct SeqSpec {
  const double* L;
  int n;
  double nconst;  // > 0: every n_i = nconst; else n_i = ndesc - i
  int ndesc;
  double L0;      // Lprev of element 0
};

__device__ EvMap thread_map(const SeqSpec& q, int tt, int nthreads) {
  const int per = (q.n + nthreads - 1) / nthreads;
  const int b = tt * per;
  const int e = min(q.n, b + per);
  EvMap m = evmap_identity();
  EvCoef cc = evcoef_empty();
  for (int i = b; i < e; ++i) {
    double Lp = i == 0 ? q.L0 : q.L[i - 1];
    double ni = q.nconst > 0.0 ? q.nconst : (double)(q.ndesc - i);
    double lt;
    m = evmap_then(m, evmap_step_c(evcoef_cached(cc, ni), Lp, q.L[i], &lt));
  }
  return m;
}

__device__ void two_sequence_maps(const SeqSpec& q1, const SeqSpec& q2, int w1, EvMap* lds_maps) {
  const int l = lane_id();
  const int w = threadIdx.x / HIPNS_WAVE;
  const int nw = blockDim.x / HIPNS_WAVE;
  EvMap m = (w < w1) ? thread_map(q1, threadIdx.x, w1 * HIPNS_WAVE)
                     : thread_map(q2, threadIdx.x - w1 * HIPNS_WAVE, (nw - w1) * HIPNS_WAVE);
#pragma unroll
  for (int o = 1; o < HIPNS_WAVE; o <<= 1) {
    EvMap other = {__shfl_down(m.a, o, 64), __shfl_down(m.b, o, 64), __shfl_down(m.g, o, 64),
                   __shfl_down(m.e, o, 64), __shfl_down(m.z, o, 64), __shfl_down(m.h, o, 64),
                   __shfl_down(m.th, o, 64)};
    if (l + o < HIPNS_WAVE) m = evmap_then(m, other);
  }
  if (l == 0) lds_maps[w] = m;
  __syncthreads();
  for (int span = 1; span < nw; span <<= 1) {
    const int t = threadIdx.x;
    if (t < nw) {
      const int base = t < w1 ? 0 : w1, cnt = t < w1 ? w1 : nw - w1, i = t - base;
      if ((i % (2 * span)) == 0 && i + span < cnt) {
        const EvMap a = lds_maps[t], b = lds_maps[t + span];
        lds_maps[t] = evmap_then(a, b);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Termination register (sharded_static.py:281-323) and determine_termination.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_register(const hipns_run_desc_t* runs, hipns_run_state_t* states,
                                                  hipns_buffers_t B, hipns_term_cond_t term) {
  __shared__ EvMap lds_maps[16];
  __shared__ long long red_i[32];
  __shared__ int red_p[16];
  const int r = blockIdx.x;
  hipns_run_state_t* st = states + r;
  if (st->done) return;
  const hipns_run_desc_t run = runs[r];
  const int N = run.N, S = run.shell;
  const int cur = st->buf, nxt = 1 - cur;
  const double* disc = B.live_logL[cur] + run.live_off;
  const double* live = B.live_logL[nxt] + run.live_off;
  const EvState s0 = evstate_from(st->evidence_calc);
  // discarded shell with n = N (sharded_static.py:284-291) and the remaining live points with
  // n = N, N-1, ..., 1 (sharded_static.py:292-299), wavefronts split in proportion to their lengths
  const int nw = blockDim.x / HIPNS_WAVE;
  const int w1 = max(1, min(nw - 1, (int)(((int64_t)nw * S + (S + N) / 2) / (S + N))));
  two_sequence_maps(SeqSpec{disc, S, (double)N, 0, s0.L}, SeqSpec{live, N, 0.0, N, disc[S - 1]}, w1, lds_maps);
  const EvMap A1 = lds_maps[0], A2 = lds_maps[w1];
  // live evaluation counts, new evaluation counts, plateau
  long long le = 0, ne = 0;
  int plat = 1;
  const double L0 = live[0];
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    le += B.live_evals[nxt][run.live_off + i];
    if (live[i] != L0) plat = 0;
  }
  for (int i = threadIdx.x; i < S; i += blockDim.x) ne += B.new_evals[run.chain_off + i];
  for (int o = 32; o > 0; o >>= 1) {
    le += __shfl_down(le, o, 64);
    ne += __shfl_down(ne, o, 64);
    plat &= __shfl_down(plat, o, 64);
  }
  const int w = threadIdx.x / 64;
  if (lane_id() == 0) {
    red_i[w] = le;
    red_i[16 + w] = ne;
    red_p[w] = plat;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < nw; ++i) {
      le += red_i[i];
      ne += red_i[16 + i];
      plat &= red_p[i];
    }
    double lt;
    EvState s1 = evmap_apply(A1, s0);
    s1.L = disc[S - 1];
    {
      double Lp = S >= 2 ? disc[S - 2] : s0.L;
      EvMap f = evmap_step((double)N, Lp, disc[S - 1], &lt);
      s1.dZ = lt + (s1.X - f.a);
    }
    EvState s2 = evmap_apply(A2, s1);
    s2.L = live[N - 1];
    {
      double Lp = N >= 2 ? live[N - 2] : s1.L;
      EvMap f = evmap_step(1.0, Lp, live[N - 1], &lt);
      s2.dZ = lt + (s2.X - f.a);
    }
    // the uniform sampler keeps no phantom samples (UniformSampler.num_phantom() = 0,
    // samplers/uniform_samplers.py), so its shells advance the collection by S only
    const int kph = run.sampler == 1 ? 0 : run.k_phantom;
    const int64_t added = (int64_t)S + (int64_t)S * kph;
    st->num_samples += added;
    st->next_idx = (st->next_idx + added) % run.max_samples;
    st->num_samples_used = st->num_samples;
    evstate_to(s1, st->evidence_calc);
    evstate_to(s2, st->evidence_calc_with_remaining);
    st->num_likelihood_evaluations += ne;
    st->log_L_contour = st->contour;
    st->efficiency = (double)N / (double)le;
    st->plateau = plat;
    st->absolute_spread = fabs(live[N - 1] - live[0]);
    st->relative_spread = 2.0 * st->absolute_spread / fabs(live[0] + live[N - 1]);
    st->no_seed_points = live[S - 1] >= live[N - 1];
    // jnp.maximum propagates NaN
    double pxl = s1.X + s1.L;
    double prev = st->peak_log_XL;
    st->peak_log_XL = (__builtin_isnan(prev) || __builtin_isnan(pxl)) ? __builtin_nan("") : fmax(prev, pxl);
    st->iteration += 1;
    st->buf = nxt;
    // max_samples is capped to leave room for one more shell (sharded_static.py:464-470)
    hipns_term_cond_t c = term;
    if (c.has_max_samples) {
      double cap = (double)run.max_samples - (double)added;
      c.max_samples = fmin(c.max_samples, cap);
    }
    int64_t why;
    int done = determine_termination(c, *st, &why);
    st->termination_reason = why;
    st->done = done;
  }
}

// determine_termination on the current register (the while_loop condition,
// sharded_static.py:480-485, also evaluated before the first shell)
__global__ void k_check_term(const hipns_run_desc_t* runs, hipns_run_state_t* states, int R, hipns_term_cond_t term,
                             int set_no_seed, hipns_buffers_t B) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  hipns_run_state_t* st = states + r;
  const hipns_run_desc_t run = runs[r];
  if (set_no_seed) {
    // no_seed_points of the incoming live set (sharded_static.py:471-473)
    const double* L = B.live_logL[st->buf] + run.live_off;
    st->no_seed_points = L[run.shell - 1] >= L[run.N - 1];
  }
  hipns_term_cond_t c = term;
  if (c.has_max_samples) {
    // space_needed_per_iteration = shell * (1 + sampler.num_phantom()) (sharded_static.py:461-470)
    const int kph = run.sampler == 1 ? 0 : run.k_phantom;
    double cap = (double)run.max_samples - (double)run.shell * (1.0 + kph);
    c.max_samples = fmin(c.max_samples, cap);
  }
  int64_t why;
  int done = determine_termination(c, *st, &why);
  st->termination_reason = why;
  st->done = done;
}

// create_init_termination_register (initialisation.py:87-108) + state counters
__global__ void k_init_state(hipns_run_state_t* states, int R) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  hipns_run_state_t* st = states + r;
  const double ninf = -__builtin_inf();
  st->next_idx = 0;
  st->num_samples = 0;
  st->iteration = 0;
  st->num_likelihood_evaluations = 0;
  st->termination_reason = 0;
  st->num_samples_used = 0;
  double init[8] = {ninf, 0.0, 0.0, ninf, ninf, ninf, ninf, ninf};
  for (int i = 0; i < 8; ++i) {
    st->evidence_calc[i] = init[i];
    st->evidence_calc_with_remaining[i] = init[i];
  }
  st->log_L_contour = ninf;
  st->efficiency = 0.0;
  st->relative_spread = __builtin_inf();
  st->absolute_spread = __builtin_inf();
  st->peak_log_XL = ninf;
  st->contour = ninf;
  st->plateau = 0;
  st->no_seed_points = 0;
  st->done = 0;
  st->buf = 0;
  st->first_seed = 0;
  st->sender_node = 0;
  st->flags = 0;
  st->pad0 = 0;
  st->cap_hits = 0;
}

// sample-collection defaults (initialisation.py:39-45): logL +inf, U 0.5, sender 0
// The sample collection of create_init_state (initialisation.py:20-84): log L = +inf and sender,
// evals and phantom 0 in every entry.  The U rows are not written: entries past num_samples are never
// read (results take the first num_samples rows), so the reference's U placeholder has no reader.
__global__ void k_init_samples(const hipns_run_desc_t* runs, int R, hipns_buffers_t B) {
  const int r = blockIdx.y;
  const hipns_run_desc_t run = runs[r];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < run.max_samples;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = run.samp_off + e;
    B.samp_logL[j] = __builtin_inf();
    B.samp_sender[j] = 0;
    B.samp_evals[j] = 0;
    B.samp_phantom[j] = 0;
  }
}

// live_point_collection appended at the end of _run (sharded_static.py:834-838)
__global__ void __launch_bounds__(256) k_consume(const hipns_run_desc_t* runs, hipns_run_state_t* states,
                                                 hipns_buffers_t B) {
  const int r = blockIdx.y;
  hipns_run_state_t* st = states + r;
  const hipns_run_desc_t run = runs[r];
  const int N = run.N, dps = B.dps, cur = st->buf;
  const int64_t start = append_start(st->next_idx, N, run.max_samples);
  const uint32_t q = (uint32_t)dps >> 2;
  const float4* srcU = (const float4*)(B.live_U[cur] + run.live_off * dps);
  float4* dstU = (float4*)B.samp_U;
  const uint32_t total = (uint32_t)N * q;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    const uint32_t i = e / q, c = e - i * q;
    const int64_t j = run.samp_off + start + i;
    dstU[j * q + c] = srcU[e];
    if (c == 0) {
      B.samp_logL[j] = B.live_logL[cur][run.live_off + i];
      B.samp_sender[j] = B.live_sender[cur][run.live_off + i];
      B.samp_evals[j] = B.live_evals[cur][run.live_off + i];
      B.samp_phantom[j] = 0;
    }
  }
}

__global__ void k_consume_advance(const hipns_run_desc_t* runs, hipns_run_state_t* states, int R) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  hipns_run_state_t* st = states + r;
  const hipns_run_desc_t run = runs[r];
  st->num_samples += run.N;
  st->next_idx = (st->next_idx + run.N) % run.max_samples;
}

}  // namespace hipns
