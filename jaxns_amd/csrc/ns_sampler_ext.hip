// ns_sampler_ext.hip -- the sampler for a likelihood evaluated outside the library (a torch callable).
//
// The fused kernel keeps every chain in registers for a whole shell and evaluates the likelihood
// inline.  Here one launch advances every chain of the batch by one step: it consumes the caller's
// likelihood values of the previous step's proposals (acceptance, shrink, phantoms, outputs), then
// begins/shrinks/picks and writes the proposals of this step for the caller to evaluate.  The draws
// are keyed exactly as in the fused kernel (the pick of step w is word w mod 4 of the chain's Philox
// block w / 4), so with the same likelihood values both forms produce the same chains.
#include "ns_sampler_impl.h"

namespace hipns {

struct ExtArgs {
  const hipns_run_desc_t* runs;
  hipns_run_state_t* states;
  hipns_buffers_t B;
  hipns_ext_state_t E;
  int32_t max_steps;
  int32_t w;  // step index
};

template <int LPC, int NJ, int MODE>
__global__ void __launch_bounds__(256) k_ext_step(ExtArgs A) {
  constexpr int CW = HIPNS_WAVE / LPC;
  constexpr bool SL = (MODE == MODE_SLICE || MODE == MODE_GSLICE);  // slice chains
  constexpr bool GG = (MODE == MODE_GSLICE);                        // gradient-guided directions
  const int r = blockIdx.y;
  const hipns_run_desc_t run = A.runs[r];
  hipns_run_state_t* st = A.states + r;
  if (st->done) return;
  const int wpb = blockDim.x / HIPNS_WAVE;
  const int wave = threadIdx.x / HIPNS_WAVE;
  const int l = lane_id();
  const int c = l % CW;
  const int h = l / CW;
  const int nchains = (MODE == MODE_INIT) ? run.N : run.shell;
  if (blockIdx.x * wpb * CW >= nchains) return;
  const int chain = (blockIdx.x * wpb + wave) * CW + c;
  const bool valid = chain < nchains;
  const int64_t o = run.chain_off + (valid ? chain : 0);
  const int dps = A.B.dps;
  const int D = run.D;
  const int DPr = ((D + 8 * LPC - 1) / (8 * LPC)) * (8 * LPC);
  const int nj = DPr / LPC;
  const int64_t iter = (MODE == MODE_INIT) ? 0xFFFFFFFFll : st->iteration;
  const uint32_t it32 = (uint32_t)iter, k0 = run.seed_lo, k1 = run.seed_hi;
  const double contour = (MODE == MODE_INIT) ? -__builtin_inf() : st->contour;
  const int N = run.N, S = run.shell, num_slices = run.num_slices, kph = run.k_phantom;
  const bool midpoint = run.midpoint_shrink != 0;
  const int max_uniform = (MODE == MODE_INIT) ? 1000000 : run.max_uniform_evals;
  const hipns_ext_state_t& E = A.E;
  const int w = A.w;

  float U0[NJ], nv[NJ], up[NJ];
  double logL0;
  float left, right, t;
  int s, k, evals;
  bool alive, need_begin, need_shrink, need_grad, dir_ready;

  if (w == 0) {
    alive = valid;
    s = 0;
    k = 0;
    evals = 0;
    left = right = t = 0.0f;
    need_begin = SL && alive;
    need_shrink = need_grad = dir_ready = false;
    logL0 = SL ? 0.0 : -__builtin_inf();
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      U0[j] = 0.5f;
      nv[j] = 0.0f;
      up[j] = 0.5f;
    }
    if (SL && alive) {
      const int cur = st->buf;
      const int idx = seed_index_dev(chain, it32, k0, k1, st->first_seed, N);
      const float* src = A.B.live_U[cur] + (run.live_off + idx) * (int64_t)dps;
#pragma unroll
      for (int j = 0; j < NJ; ++j) U0[j] = (j < nj) ? src[owned_dim<LPC>(h, j)] : 0.5f;
      logL0 = A.B.live_logL[cur][run.live_off + idx];
    }
  } else {
    const int fl = valid ? E.flags[o] : 0;
    alive = (fl & 1) != 0;
    need_begin = (fl & 2) != 0;
    need_shrink = (fl & 4) != 0;
    need_grad = (fl & 8) != 0;
    dir_ready = (fl & 16) != 0;
    logL0 = E.logL0[o];
    left = E.left[o];
    right = E.right[o];
    t = E.t[o];
    s = E.s[o];
    k = E.k[o];
    evals = E.evals[o];
    const float* u0p = E.U0 + o * dps;
    const float* nvp = E.nv + o * dps;
    const float* upp = E.prop_U + o * dps;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int d = owned_dim<LPC>(h, j);
      U0[j] = (j < nj) ? u0p[d] : 0.5f;
      nv[j] = (j < nj) ? nvp[d] : 0.0f;
      up[j] = (j < nj) ? upp[d] : 0.5f;
    }
    if (alive && GG && need_grad) {
      // the gradient at the accepted point (one likelihood evaluation), then the Householder
      // reflection of the slice's direction about it (uni_slice_sampler.py:255-271)
      ++evals;
      need_grad = false;
      if (s == num_slices) {
        alive = false;
        float* dst = A.B.new_U + o * dps;
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          if (j < nj) dst[owned_dim<LPC>(h, j)] = U0[j];
        if (h == 0) {
          A.B.new_logL[o] = logL0;
          A.B.new_evals[o] = evals;
        }
      } else {
        const float* gp = E.grad + o * dps;
        float g[NJ];
        float gg2 = 0.0f;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int d = owned_dim<LPC>(h, j);
          g[j] = (j < nj && d < D) ? gp[d] : 0.0f;
          gg2 = fmaf(g[j], g[j], gg2);
        }
        const float gn = __builtin_sqrtf(chain_sum<LPC>(gg2));
        if (!(gn >= 1e-10f) || !__builtin_isfinite(gn)) {
          slice_direction<LPC, NJ>(nv, chain, it32, s, D, nj, h, k0, k1);
        } else {
          float dot = 0.0f;
#pragma unroll
          for (int j = 0; j < NJ; ++j) {
            g[j] = g[j] / gn;
            dot = fmaf(nv[j], g[j], dot);
          }
          dot = chain_sum<LPC>(dot);
          float rn2 = 0.0f;
#pragma unroll
          for (int j = 0; j < NJ; ++j) {
            nv[j] = nv[j] - 2.0f * dot * g[j];
            rn2 = fmaf(nv[j], nv[j], rn2);
          }
          const float inv = 1.0f / __builtin_sqrtf(chain_sum<LPC>(rn2));
#pragma unroll
          for (int j = 0; j < NJ; ++j) nv[j] = nv[j] * inv;
        }
        need_begin = true;
        dir_ready = true;
      }
    } else if (alive) {
      // the caller's Model.forward of the previous proposal; NaN -> -inf (framework/ops.py:302-326)
      double L = E.prop_logL[o];
      if (__builtin_isnan(L)) L = -__builtin_inf();
      ++evals;
      if (MODE == MODE_INIT || MODE == MODE_UNIF) {
        // uniform_sample.py:35-36 (logL > -inf) / uniform_samplers.py:44-50
        bool acc = (MODE == MODE_INIT) ? (L > -__builtin_inf()) : ((L > contour) || (L == contour));
        if (!acc && evals >= max_uniform) acc = true;
        if (acc) {
#pragma unroll
          for (int j = 0; j < NJ; ++j) U0[j] = up[j];
          logL0 = L;
          alive = false;
          float* dst = A.B.new_U + o * dps;
#pragma unroll
          for (int j = 0; j < NJ; ++j)
            if (j < nj) dst[owned_dim<LPC>(h, j)] = U0[j];
          if (h == 0) {
            A.B.new_logL[o] = logL0;
            A.B.new_evals[o] = evals;
          }
        }
      } else {
        // _new_proposal acceptance (uni_slice_sampler.py:160-166)
        bool acc = (L > contour) || (logL0 == contour && L == contour);
        if (!acc && k + 1 >= A.max_steps) {
          // step cap: keep the seed point (t = 0 lies inside every bracket)
          acc = true;
          L = logL0;
#pragma unroll
          for (int j = 0; j < NJ; ++j) up[j] = U0[j];
          if (h == 0) {
            atomicAdd((unsigned long long*)&st->cap_hits, 1ull);
            atomicOr(&st->flags, 1);
          }
        }
        if (acc) {
#pragma unroll
          for (int j = 0; j < NJ; ++j) U0[j] = up[j];
          logL0 = L;
          // phantom samples: x[-(k+1):-1] (uni_slice_sampler.py:432-440), evals 0
          const int pidx = s - (num_slices - 1 - kph);
          if (kph > 0 && pidx >= 0 && s < num_slices - 1) {
            const int64_t phantom_base =
                append_start((st->next_idx + S) % run.max_samples, (int64_t)S * kph, run.max_samples);
            const int64_t si = run.samp_off + phantom_base + (int64_t)chain * kph + pidx;
            float* pd = A.B.samp_U + si * dps;
#pragma unroll
            for (int j = 0; j < NJ; ++j)
              if (j < nj) pd[owned_dim<LPC>(h, j)] = U0[j];
            if (h == 0) {
              A.B.samp_logL[si] = L;
              A.B.samp_sender[si] = st->sender_node;
              A.B.samp_evals[si] = 0;
              A.B.samp_phantom[si] = 1;
            }
          }
          ++s;
          if (GG) {
            need_grad = true;  // every accepted slice asks for the gradient, the last one included
          } else if (s == num_slices) {
            alive = false;
            float* dst = A.B.new_U + o * dps;
#pragma unroll
            for (int j = 0; j < NJ; ++j)
              if (j < nj) dst[owned_dim<LPC>(h, j)] = U0[j];
            if (h == 0) {
              A.B.new_logL[o] = logL0;
              A.B.new_evals[o] = evals;
            }
          } else {
            need_begin = true;
          }
        } else {
          need_shrink = true;
        }
      }
    }
  }

  // ---- the proposal of step w (or the point whose gradient is asked for)
  if (alive && GG && need_grad) {
#pragma unroll
    for (int j = 0; j < NJ; ++j) up[j] = U0[j];
  } else if (alive) {
    if (SL) {
      if (need_begin) {
        k = 0;
        if (!dir_ready) slice_direction<LPC, NJ>(nv, chain, it32, s, D, nj, h, k0, k1);
        dir_ready = false;
        slice_bracket<LPC, NJ>(U0, nv, left, right);
        need_begin = false;
      } else if (need_shrink) {
        // _shrink_interval (uni_slice_sampler.py:92-111); alpha = linspace(0.5, 1, num_slices)[s]
        const float alpha = (float)(num_slices > 1 ? 0.5 + 0.5 * (double)s / (double)(num_slices - 1) : 0.5);
        shrink(t, alpha, midpoint, left, right);
        ++k;
      }
      need_shrink = false;
      const u32x4 pb = philox4x32((uint32_t)chain, it32, (uint32_t)(w >> 2), RNG_PICK, k0, k1);
      const int kw = w & 3;
      const float u = u01f(kw == 0 ? pb.x : (kw == 1 ? pb.y : (kw == 2 ? pb.z : pb.w)));
      // _pick_point_in_interval (uni_slice_sampler.py:67-89)
      t = pick_t(left, right, u);
#pragma unroll
      for (int j = 0; j < NJ; ++j) up[j] = point_on_line(U0[j], t, nv[j]);
    } else {
      uniform_point<LPC, NJ>(up, chain, it32, evals, D, nj, h, k0, k1);
    }
  }
  if (!valid) return;
  float* u0p = E.U0 + o * dps;
  float* nvp = E.nv + o * dps;
  float* upp = E.prop_U + o * dps;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    if (j < nj) {
      const int d = owned_dim<LPC>(h, j);
      u0p[d] = U0[j];
      nvp[d] = nv[j];
      upp[d] = alive ? up[j] : 0.5f;
    }
  }
  if (h == 0) {
    E.flags[o] = (alive ? 1 : 0) | (need_begin ? 2 : 0) | (need_shrink ? 4 : 0) | (need_grad ? 8 : 0) |
                 (dir_ready ? 16 : 0);
    E.logL0[o] = logL0;
    E.left[o] = left;
    E.right[o] = right;
    E.t[o] = t;
    E.s[o] = s;
    E.k[o] = k;
    E.evals[o] = evals;
  }
}

template <int LPC, int NJ, int MODE>
static hipError_t launch_ext_one(const ExtArgs& A, int R, int max_chains, int wpb, hipStream_t s) {
  constexpr int CW = HIPNS_WAVE / LPC;
  int groups = (max_chains + CW - 1) / CW;
  int bx = (groups + wpb - 1) / wpb;
  if (bx < 1) bx = 1;
  hipLaunchKernelGGL((k_ext_step<LPC, NJ, MODE>), dim3(bx, R), dim3(wpb * HIPNS_WAVE), 0, s, A);
  return hipGetLastError();
}

// the same instance choice as the fused launchers (sampler_nj)
template <int MODE>
static hipError_t launch_ext_mode(const ExtArgs& A, int R, int max_chains, int lpc, int wpb, int dmax, hipStream_t s) {
  if (lpc == 4 || dmax > 64) {
    if (dmax <= 32) return launch_ext_one<4, 8, MODE>(A, R, max_chains, wpb, s);
    if (dmax <= 64) return launch_ext_one<4, 16, MODE>(A, R, max_chains, wpb, s);
    return launch_ext_one<4, 32, MODE>(A, R, max_chains, wpb, s);
  }
  if (dmax <= 16) return launch_ext_one<2, 8, MODE>(A, R, max_chains, wpb, s);
  if (dmax <= 32) return launch_ext_one<2, 16, MODE>(A, R, max_chains, wpb, s);
  return launch_ext_one<2, 32, MODE>(A, R, max_chains, wpb, s);
}

hipError_t launch_ext_step(const hipns_run_desc_t* runs, hipns_run_state_t* states, const hipns_buffers_t& B,
                           const hipns_ext_state_t& E, int mode, int w, int max_steps, int R, int max_chains, int lpc,
                           int wpb, int dmax, hipStream_t s) {
  ExtArgs A;
  A.runs = runs;
  A.states = states;
  A.B = B;
  A.E = E;
  A.max_steps = max_steps;
  A.w = w;
  if (mode == MODE_INIT) return launch_ext_mode<MODE_INIT>(A, R, max_chains, lpc, wpb, dmax, s);
  if (mode == MODE_UNIF) return launch_ext_mode<MODE_UNIF>(A, R, max_chains, lpc, wpb, dmax, s);
  if (mode == MODE_GSLICE) return launch_ext_mode<MODE_GSLICE>(A, R, max_chains, lpc, wpb, dmax, s);
  return launch_ext_mode<MODE_SLICE>(A, R, max_chains, lpc, wpb, dmax, s);
}

}  // namespace hipns
