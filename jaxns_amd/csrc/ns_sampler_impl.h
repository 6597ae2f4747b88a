// ns_sampler_impl.h -- the fused sampling kernel of libhipns (gfx950).
//
// One wavefront advances CW = 64/LPC Markov chains in lock-step "evaluation
// steps".  Lane l serves chain (l % CW) and part h = l / CW of that chain's
// U-space vector: it owns the dimensions d(h, j) = 8*LPC*(j/8) + 8h + (j%8),
// j < DP/LPC.  Each step every live chain proposes one point, the wavefront
// evaluates all CW likelihoods at once (for the Gaussian family as an MFMA
// GEMM whose B operand is exactly this layout), and each chain then accepts,
// shrinks or starts its next slice.  Chains never wait for each other except
// at the very end of the launch.
//
// Reference semantics: samplers/uni_slice_sampler.py:23-441 (slice chain),
// samplers/uniform_samplers.py:42-85 (rejection sampler),
// nested_samplers/common/uniform_sample.py:12-61 (initial live points),
// samplers/bases.py:63-74 + internals/random.py:34-84 (seed choice).
#pragma once
#include "hipns_dev.h"
#include "ns_sampler.h"

namespace hipns {

// ---------------------------------------------------------------------------
// slice arithmetic, every operation one IEEE rounding (matches the oracle)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bound_candidates(float u0, float n, float& rmin, float& lmax) {
#pragma clang fp contract(off)
  float t1 = (1.0f - u0) / n;
  float t0 = -u0 / n;
  float c1r = t1 >= 0.0f ? t1 : __builtin_inff();
  float c1l = t1 <= 0.0f ? t1 : -__builtin_inff();
  float c0r = t0 >= 0.0f ? t0 : __builtin_inff();
  float c0l = t0 <= 0.0f ? t0 : -__builtin_inff();
  // min over {t0 right, t1 right}, max over {t0 left, t1 left}; fminf/fmaxf are exact
  rmin = fminf(rmin, fminf(c0r, c1r));
  lmax = fmaxf(lmax, fmaxf(c0l, c1l));
}

__device__ __forceinline__ float pick_t(float left, float right, float u) {
#pragma clang fp contract(off)
  return left + u * (right - left);
}

__device__ __forceinline__ float point_on_line(float u0, float t, float n) {
#pragma clang fp contract(off)
  return u0 + t * n;
}

__device__ __forceinline__ void shrink(float t, float alpha, bool midpoint, float& left, float& right) {
#pragma clang fp contract(off)
  if (t < 0.0f) left = t;
  if (t > 0.0f) right = t;
  if (midpoint) {
    if (t < 0.0f) left = alpha * left;
    if (t > 0.0f) right = alpha * right;
  }
}

// owned dimension of lane-part h at register slot j
template <int LPC>
__device__ __forceinline__ int owned_dim(int h, int j) {
  return 8 * LPC * (j >> 3) + 8 * h + (j & 7);
}

// ---------------------------------------------------------------------------
// likelihood families in the chain-tile layout.  Parameters live in LDS.
// ---------------------------------------------------------------------------
struct MvnLds {
  float* WT;  // [DP][RS]: WT[d*RS + m] = W[m][d]
  float* b;   // [DP]
  float* tf;  // [DP]: 1 -> ndtri, 0 -> identity
  int RS;
};

struct GmmLds {
  float* tf;     // [DP]
  float* p;      // [DP]
  float* q;      // [DP]
  float* mean;   // [K][DP]
  float* inv2s2; // [K]
  float* logc;   // [K]
  int K;
  int dps;
};


typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ double chain_sum_d_swap16(double v) {
  float lo = __uint_as_float((uint32_t)__double_as_longlong(v)), hi = __uint_as_float((uint32_t)(__double_as_longlong(v) >> 32));
  float a0, b0, a1, b1;
  swap16(lo, a0, b0);
  swap16(hi, a1, b1);
  double A = __longlong_as_double(((long long)__float_as_uint(a1) << 32) | __float_as_uint(a0));
  double B = __longlong_as_double(((long long)__float_as_uint(b1) << 32) | __float_as_uint(b0));
  return A + B;
}
__device__ __forceinline__ double chain_sum_d_swap32(double v) {
  float lo = __uint_as_float((uint32_t)__double_as_longlong(v)), hi = __uint_as_float((uint32_t)(__double_as_longlong(v) >> 32));
  float a0, b0, a1, b1;
  swap32(lo, a0, b0);
  swap32(hi, a1, b1);
  double A = __longlong_as_double(((long long)__float_as_uint(a1) << 32) | __float_as_uint(a0));
  double B = __longlong_as_double(((long long)__float_as_uint(b1) << 32) | __float_as_uint(b0));
  return A + B;
}
template <int LPC>
__device__ __forceinline__ double chain_sum_d(double v) {
  if constexpr (LPC == 4) v = chain_sum_d_swap16(v);
  return chain_sum_d_swap32(v);
}

// Gaussian linear regression: r = r0 - X (w - w0) with w = p + s (.) q(u); logL = c - |r|^2 / (2 sigma^2).
// The MFMA chain of a tile starts from C = r0 with the weight operands negated, so it ends on r.
// The G wavefronts of a workgroup each own one group of CW chains.  Every step they publish the
// B operands of their chains (w - w0 split into bf16 parts) in LDS, then each wavefront takes 1/G of
// the data-row tiles and multiplies every row tile it loads against all G groups' operands, so X is
// read once per step per workgroup instead of once per chain group.  Each product is BF16x3
// (hi.hi + hi.lo + lo.hi, FP32 accumulate; terms == 4 adds hi times the third bf16 part of w - w0,
// which removes the error of the 16-bit weight split that every row shares): LPC=2 uses
// v_mfma_f32_32x32x16_bf16 over 32-row tiles, LPC=4 v_mfma_f32_16x16x32_bf16 over 16-row tiles.
// X is stored in 16-column panels ([panel][row][16] bf16) so both layouts read dense rows.  Squares
// of a tile are summed in FP32 and accumulated in FP64; the G partial sums of a group meet in LDS
// in a fixed order.  All G wavefronts must call this together (two workgroup barriers).
constexpr int LR_MAX_G = 4;

template <int LPC>
__device__ __forceinline__ const uint16_t* lr_panel_ptr(const uint16_t* x, int npad, int gg, int row, int h) {
  if constexpr (LPC == 2) return x + ((int64_t)gg * npad + row) * 16 + 8 * h;
  return x + ((int64_t)(2 * gg + (h >> 1)) * npad + row) * 16 + 8 * (h & 1);
}

// q(u) of the owned slots for the prior transform: ndtri where bit j of tfm is set, the identity
// elsewhere.  Straight-line over the slots; the binary64 far-tail path and the identity selects run
// only when some lane of the wavefront needs them.
template <int NJ>
__device__ __forceinline__ void prior_q(const float (&up)[NJ], float (&q)[NJ], uint32_t tfm) {
  if (!__any(tfm != 0u)) {
#pragma unroll
    for (int j = 0; j < NJ; ++j) q[j] = up[j];
    return;
  }
  // the central polynomial for every slot, then the tail polynomial (w >= 5, u beyond about 0.0017 of
  // either end) behind one wavefront-uniform branch for all slots: the same values as ndtri_giles.  The
  // largest w of the lane also gates the far tails: min(u, 1 - u) < 1e-7 means w > 14.7, so no lane
  // with a slot there (ndtri_far) has w <= 14.  A u outside [0, 1] gives w = NaN, which fmaxf skips:
  // its q is NaN and the proposal's logL -inf, as with the far-tail routine (AS 241 of u outside (0, 1)).
  float wmax = 0.0f;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const float u = up[j];
    const float w = giles_w(u);
    q[j] = giles_central(w) * giles_x(u);
    wmax = fmaxf(wmax, w);
  }
  if (__any(wmax >= 5.0f)) {
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const float w = giles_w(up[j]);
      if (w >= 5.0f) q[j] = giles_tail(w) * giles_x(up[j]);
    }
  }
  if (__any(wmax > 14.0f)) {
#pragma unroll
    for (int j = 0; j < NJ; ++j)
      if (((tfm >> j) & 1u) && ndtri_far(up[j])) q[j] = ndtri_tail_f32(up[j]);
  }
  constexpr uint32_t all = NJ >= 32 ? 0xFFFFFFFFu : ((1u << NJ) - 1u);
  if (__any(tfm != all)) {
#pragma unroll
    for (int j = 0; j < NJ; ++j)
      if (!((tfm >> j) & 1u)) q[j] = up[j];
  }
}

template <int LPC>
struct LrAcc;
template <>
struct LrAcc<2> {
  typedef f32x16 type;  // v_mfma_f32_32x32x16_bf16: 32 rows x 32 chains, 16 rows per lane
};
template <>
struct LrAcc<4> {
  typedef f32x4 type;  // v_mfma_f32_16x16x32_bf16: 16 rows x 16 chains, 4 rows per lane
};

template <int LPC>
__device__ __forceinline__ typename LrAcc<LPC>::type lr_mfma(const bf16x8_t& a, const bf16x8_t& b,
                                                              const typename LrAcc<LPC>::type& c) {
  if constexpr (LPC == 2)
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// The BF16 residual tiles [t0, t1) of one wavefront against the B operands of GG chain groups (GG
// known at compile time, so the groups' MFMA chains are interleaved and nothing branches per group).
// The tile loop has wavefront-uniform bounds; at LPC=4 the tiles are double-buffered (the next tile's
// X panels and r0 load while this one multiplies).  With few (group, k-group) pairs the step's B
// operands are held in registers for all tiles; otherwise they are read from LDS per tile.
template <int LPC, int NJ, int GG, bool STATS>
__device__ __forceinline__ void lr_bf16_tiles(const LinRegLds& P, const bf16x8_t* Bs, int t0, int t1, bool t4,
                                              double (&ss)[LR_MAX_G], double (&sy)[LR_MAX_G]) {
  constexpr int CW = HIPNS_WAVE / LPC;
  constexpr int NG = NJ / 8;
  constexpr int NR = (LPC == 2) ? 16 : 4;
  constexpr bool BREG = LPC == 4 && GG * NG <= 4;
  typedef typename LrAcc<LPC>::type Acc;
  const int l = lane_id();
  const int m = l % CW;
  const int h = l / CW;
  auto bop = [&](int q, int gg, int part) -> bf16x8_t { return Bs[((q * NG + gg) * 3 + part) * HIPNS_WAVE + l]; };
  bf16x8_t bh[BREG ? GG : 1][NG], bl[BREG ? GG : 1][NG];
  if constexpr (BREG) {
#pragma unroll
    for (int q = 0; q < GG; ++q)
#pragma unroll
      for (int gg = 0; gg < NG; ++gg) {
        bh[q][gg] = bop(q, gg, 0);
        bl[q][gg] = bop(q, gg, 1);
      }
  }
  auto tile_row = [&](int tl, int i) { return (LPC == 2) ? tl * 32 + 8 * (i >> 2) + 4 * h + (i & 3) : tl * 16 + 4 * h + i; };
  auto load = [&](int tl, bf16x8_t (&xh)[NG], bf16x8_t (&xl)[NG], float (&rv)[NR]) {
    const int row = tl * CW + m;
#pragma unroll
    for (int gg = 0; gg < NG; ++gg) {
      xh[gg] = __builtin_bit_cast(bf16x8_t, *(const uint4*)lr_panel_ptr<LPC>(P.xh, P.npad, gg, row, h));
      xl[gg] = __builtin_bit_cast(bf16x8_t, *(const uint4*)lr_panel_ptr<LPC>(P.xl, P.npad, gg, row, h));
    }
    // the lane's rows of the tile come in aligned runs of four (tile_row): r0 in 16-byte loads
    if constexpr (LPC == 2) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 v = *(const float4*)(P.r0 + tl * 32 + 8 * q + 4 * h);
        rv[4 * q] = v.x;
        rv[4 * q + 1] = v.y;
        rv[4 * q + 2] = v.z;
        rv[4 * q + 3] = v.w;
      }
    } else {
      const float4 v = *(const float4*)(P.r0 + tl * 16 + 4 * h);
      rv[0] = v.x;
      rv[1] = v.y;
      rv[2] = v.z;
      rv[3] = v.w;
    }
  };
  // groups per pass over a tile: LPC=2 accumulators are 16 VGPRs each, so two at a time
  constexpr int QB = (LPC == 2 && GG > 2) ? 2 : GG;
  auto compute = [&](int tl, const bf16x8_t (&xh)[NG], const bf16x8_t (&xl)[NG], const float (&rv)[NR]) {
    float yv[NR];
    if (STATS) {
#pragma unroll
      for (int i = 0; i < NR; ++i) yv[i] = P.y[tile_row(tl, i)];
    }
#pragma unroll
    for (int q0 = 0; q0 < GG; q0 += QB) {
      constexpr int QN = QB;
      Acc acc[QN];
#pragma unroll
      for (int qq = 0; qq < QN; ++qq)
#pragma unroll
        for (int i = 0; i < NR; ++i) acc[qq][i] = rv[i];
#pragma unroll
      for (int gg = 0; gg < NG; ++gg) {
        // hi.hi, hi.lo, lo.hi of every group of the pass in turn: consecutive MFMAs are independent
#pragma unroll
        for (int qq = 0; qq < QN; ++qq)
          if (q0 + qq < GG)
            acc[qq] = lr_mfma<LPC>(xh[gg], BREG ? bh[q0 + qq][gg] : bop(q0 + qq, gg, 0), acc[qq]);
#pragma unroll
        for (int qq = 0; qq < QN; ++qq)
          if (q0 + qq < GG)
            acc[qq] = lr_mfma<LPC>(xh[gg], BREG ? bl[q0 + qq][gg] : bop(q0 + qq, gg, 1), acc[qq]);
#pragma unroll
        for (int qq = 0; qq < QN; ++qq)
          if (q0 + qq < GG)
            acc[qq] = lr_mfma<LPC>(xl[gg], BREG ? bh[q0 + qq][gg] : bop(q0 + qq, gg, 0), acc[qq]);
        if (t4) {
#pragma unroll
          for (int qq = 0; qq < QN; ++qq)
            if (q0 + qq < GG) acc[qq] = lr_mfma<LPC>(xh[gg], bop(q0 + qq, gg, 2), acc[qq]);
        }
      }
#pragma unroll
      for (int qq = 0; qq < QN; ++qq) {
        if (q0 + qq < GG) {
          // four independent partial sums of the squared residuals (no serial FMA chain)
          float pp[4] = {0.0f, 0.0f, 0.0f, 0.0f}, py = 0.0f;
#pragma unroll
          for (int i = 0; i < NR; ++i) {
            const float r = acc[qq][i];
            pp[i & 3] = fmaf(r, r, pp[i & 3]);
            if (STATS) py = fmaf(yv[i], r, py);
          }
          ss[q0 + qq] += (double)((pp[0] + pp[1]) + (pp[2] + pp[3]));
          if (STATS) sy[q0 + qq] += (double)py;
        }
      }
    }
  };
  bf16x8_t xh0[NG], xl0[NG];
  float r00[NR];
  if constexpr (LPC == 4) {
    // double-buffered: the next tile's X panels and r0 load while this one multiplies
    bf16x8_t xh1[NG], xl1[NG];
    float r01[NR];
    if (t0 < t1) load(t0, xh0, xl0, r00);
    for (int tl = t0; tl < t1; tl += 2) {
      if (tl + 1 < t1) load(tl + 1, xh1, xl1, r01);
      compute(tl, xh0, xl0, r00);
      if (tl + 1 >= t1) break;
      if (tl + 2 < t1) load(tl + 2, xh0, xl0, r00);
      compute(tl + 1, xh1, xl1, r01);
    }
  } else {
    // LPC=2: the second buffer (24 VGPRs) would drop the kernel to one wavefront per SIMD, which
    // costs more than the exposed load latency (4096 x 16, 128 runs: 17.1 ms with it, 14.6 without)
    for (int tl = t0; tl < t1; ++tl) {
      load(tl, xh0, xl0, r00);
      compute(tl, xh0, xl0, r00);
    }
  }
}

// F64T: the instance also carries the refinement form (terms 5: the FP64 transform).  Only the batched
// likelihood kernel needs it; in the sampling kernel its FP64 registers would cost a wavefront per SIMD
// (256 VGPRs against 214 at 8 slots and LPC=2), so there terms 5 is rejected by the host.
template <int LPC, int NJ, bool STATS = false, bool F64T = false>
__device__ __forceinline__ double lik_linreg(const float (&up)[NJ], int nj, int D, const LinRegLds& P, uint32_t tfm,
                                             int g, int G, double log_const, double* rss = nullptr,
                                             double* yr = nullptr) {
  constexpr int CW = HIPNS_WAVE / LPC;
  constexpr int NG = NJ / 8;  // k-groups of 8 owned dims
  const int l = lane_id();
  const int m = l % CW;
  const int h = l / CW;
  const bool t4 = P.terms == 4;
  const bool fp32 = P.terms == 1 || (F64T && P.terms == 5);
  float dwv[NJ], qv[NJ];
  if (F64T && P.terms == 5) {
    // refinement: w - w0 from the FP64 transform and FP64 parameters, rounded once to FP32 (the FP32
    // ndtri's ~1e-7 relative error would otherwise shift w by far more than the posterior width of
    // strongly determined weights)
    const double* p64 = P.f64;
    const double* s64 = P.f64 + P.dps;
    const double* w64 = P.f64 + 2 * P.dps;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int d = owned_dim<LPC>(h, j);
      float dw = 0.0f;
      if (j < nj && d < D) {
        const double u = (double)up[j];
        const double q = ((tfm >> j) & 1u) ? ndtri_f64(u) : u;
        dw = (float)(fma(s64[d], q, p64[d]) - w64[d]);
      }
      dwv[j] = dw;
    }
  } else {
    prior_q<NJ>(up, qv, tfm);
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int d = owned_dim<LPC>(h, j);
      float dw = 0.0f;
      if (j < nj && d < D) dw = fmaf(P.s[d], qv[j], P.p[d]) - P.w0[d];
      dwv[j] = dw;
    }
  }
  // publish this wavefront's B operands, negated so that the MFMA chain started from C = r0 yields
  // r = r0 - X (w - w0): BF16 forms [G][NG][3][64] vectors of 8 bf16, plain FP32 [G][NJ][64] floats
  bf16x8_t* Bs = (bf16x8_t*)P.bops;
  float* Bf = (float*)P.bops;
  if (fp32) {
#pragma unroll
    for (int j = 0; j < NJ; ++j) Bf[(g * NJ + j) * HIPNS_WAVE + l] = -dwv[j];
  } else {
#pragma unroll
    for (int gg = 0; gg < NG; ++gg) {
      bf16x8_t bh, bl, bt;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const float dw = dwv[gg * 8 + i];
        const __bf16 hi = (__bf16)dw;
        const float rem = dw - (float)hi;
        const __bf16 lo = (__bf16)rem;
        bh[i] = -hi;
        bl[i] = -lo;
        bt[i] = -(__bf16)(rem - (float)lo);
      }
      Bs[((g * NG + gg) * 3 + 0) * HIPNS_WAVE + l] = bh;
      Bs[((g * NG + gg) * 3 + 1) * HIPNS_WAVE + l] = bl;
      Bs[((g * NG + gg) * 3 + 2) * HIPNS_WAVE + l] = bt;
    }
  }
  __syncthreads();
  const int tiles = P.npad / CW;
  const int t0 = (int)(((long long)tiles * g) / G), t1 = (int)(((long long)tiles * (g + 1)) / G);
  double ss[LR_MAX_G], sy[LR_MAX_G];
#pragma unroll
  for (int q = 0; q < LR_MAX_G; ++q) {
    ss[q] = 0.0;
    sy[q] = 0.0;
  }
  constexpr int NR = (LPC == 2) ? 16 : 4;  // accumulator rows per lane
  // rows of tile tl held by this lane (the MFMA output layout)
  auto tile_row = [&](int tl, int i) { return (LPC == 2) ? tl * 32 + 8 * (i >> 2) + 4 * h + (i & 3) : tl * 16 + 4 * h + i; };
  if (fp32) {
    // plain FP32 form: v_mfma_f32_32x32x2_f32 (LPC=2) / v_mfma_f32_16x16x4_f32 (LPC=4), one MFMA per
    // owned dimension slot, X read in FP32 from its own panels
    for (int tl = t0; tl < t1; ++tl) {
      const int row = tl * CW + m;
      float xa[NJ];
#pragma unroll
      for (int gg = 0; gg < NG; ++gg) {
        const float* xp = (LPC == 2) ? P.xf + ((int64_t)gg * P.npad + row) * 16 + 8 * h
                                     : P.xf + ((int64_t)(2 * gg + (h >> 1)) * P.npad + row) * 16 + 8 * (h & 1);
        const float4 v0 = *(const float4*)xp, v1 = *(const float4*)(xp + 4);
        xa[gg * 8 + 0] = v0.x;
        xa[gg * 8 + 1] = v0.y;
        xa[gg * 8 + 2] = v0.z;
        xa[gg * 8 + 3] = v0.w;
        xa[gg * 8 + 4] = v1.x;
        xa[gg * 8 + 5] = v1.y;
        xa[gg * 8 + 6] = v1.z;
        xa[gg * 8 + 7] = v1.w;
      }
      float rv[NR], yv[NR];
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        rv[i] = P.r0[tile_row(tl, i)];
        if (STATS) yv[i] = P.y[tile_row(tl, i)];
      }
#pragma unroll
      for (int q = 0; q < LR_MAX_G; ++q) {
        if (q < G) {
          const float* B = Bf + q * NJ * HIPNS_WAVE + l;
          float part = 0.0f, ypart = 0.0f;
          if constexpr (LPC == 2) {
            f32x16 acc;
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[i] = rv[i];
#pragma unroll
            for (int j = 0; j < NJ; ++j)
              acc = __builtin_amdgcn_mfma_f32_32x32x2f32(xa[j], B[j * HIPNS_WAVE], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
              part = fmaf(acc[i], acc[i], part);
              if (STATS) ypart = fmaf(yv[i], acc[i], ypart);
            }
          } else {
            f32x4 acc;
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] = rv[i];
#pragma unroll
            for (int j = 0; j < NJ; ++j)
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[j], B[j * HIPNS_WAVE], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              part = fmaf(acc[i], acc[i], part);
              if (STATS) ypart = fmaf(yv[i], acc[i], ypart);
            }
          }
          ss[q] += (double)part;
          if (STATS) sy[q] += (double)ypart;
        }
      }
    }
  } else {
    // wavefront-uniform tile range (scalar loop control)
    const int u0 = __builtin_amdgcn_readfirstlane(t0), u1 = __builtin_amdgcn_readfirstlane(t1);
    if (G == 4)
      lr_bf16_tiles<LPC, NJ, 4, STATS>(P, Bs, u0, u1, t4, ss, sy);
    else if (G == 2)
      lr_bf16_tiles<LPC, NJ, 2, STATS>(P, Bs, u0, u1, t4, ss, sy);
    else if (G == 1)
      lr_bf16_tiles<LPC, NJ, 1, STATS>(P, Bs, u0, u1, t4, ss, sy);
    else
      lr_bf16_tiles<LPC, NJ, 3, STATS>(P, Bs, u0, u1, t4, ss, sy);
  }
  // partial sums of every group from this wavefront's rows -> LDS [group][wavefront][chain]
#pragma unroll
  for (int q = 0; q < LR_MAX_G; ++q) {
    if (q < G) {
      const double v = chain_sum_d<LPC>(ss[q]);
      const double vy = STATS ? chain_sum_d<LPC>(sy[q]) : 0.0;
      if (h == 0) {
        P.partial[(q * G + g) * CW + m] = v;
        if (STATS) P.partial[(G * G + q * G + g) * CW + m] = vy;
      }
    }
  }
  __syncthreads();
  double tot = 0.0, toty = 0.0;
  for (int k = 0; k < G; ++k) {
    tot += P.partial[(g * G + k) * CW + m];
    if (STATS) toty += P.partial[(G * G + g * G + k) * CW + m];
  }
  if constexpr (STATS) {
    *rss = tot;
    *yr = toty;
  }
  double L = log_const - tot * P.inv2s2;
  return __builtin_isnan(L) ? -__builtin_inf() : L;
}

// dq/du of the prior transform at a slot: 1 / phi(q) = sqrt(2 pi) exp(q^2 / 2) where q = ndtri(u), 1 for
// the identity slots
__device__ __forceinline__ float dq_du(float q, bool normal) {
  return normal ? 2.50662827463100050f * __expf(0.5f * q * q) : 1.0f;
}

// r = W q(u) + b ; logL = c - |r|^2 / 2 ; W rows (outputs) tiled by 32 (LPC=2) or 16 (LPC=4).
// GRAD: when some lane of the wavefront asks (want), also g = d logL / du = -(W^T r) (.) dq/du of the
// owned slots.  W^T r is a second MFMA GEMM whose B operands are the lanes' own accumulators of r (the
// K index of that product runs over r's output rows in the accumulator layout, and the A operands are
// read from W^T at the matching rows), so r never moves; its output goes through the wavefront's LDS
// scratch gscr ([grad_rows(NJ, LPC)][CW + 1]) into the owned-slot layout.
template <int LPC, int NJ, bool GRAD = false>
__device__ __forceinline__ double lik_mvn(const float (&up)[NJ], int nj, int D, int dps, const MvnLds& P,
                                          double log_const, uint32_t tfm, float* gscr = nullptr,
                                          float* g = nullptr, bool want = false) {
  constexpr int CW = HIPNS_WAVE / LPC;
  const int l = lane_id();
  const int m = l % CW;
  const int h = l / CW;
  float z[NJ];
  prior_q<NJ>(up, z, tfm);
#pragma unroll
  for (int j = 0; j < NJ; ++j)
    if (!(j < nj && owned_dim<LPC>(h, j) < D)) z[j] = 0.0f;
  // A operands W[row][owned dim] at compile-time offsets from the lane's base (row m of block mb, owned
  // dim 8h); the accumulator of each block starts from this lane's output rows of b.  W^T has a row for
  // every slot of the instance (zero beyond the model), and z is zero on slots beyond D, so all NJ
  // slots are multiplied unconditionally.
  constexpr int RS = mvn_row_stride(NJ, LPC);
  constexpr int MBI = (NJ * LPC + CW - 1) / CW;
  const float* wt = P.WT + 8 * h * RS + m;
  float ss = 0.0f;
  const int MB = (D + CW - 1) / CW;
  constexpr int V = (LPC == 2) ? 16 : 4;  // accumulator values per lane per output block
  float rv[GRAD ? MBI * V : 1];
#pragma unroll
  for (int mb = 0; mb < MBI; ++mb) {
    if (mb >= MB) break;
    if constexpr (LPC == 2) {
      f32x16 acc0, acc1 = {0};
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const f32x4 bb = *(const f32x4*)(P.b + mb * 32 + 8 * g + 4 * h);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc0[4 * g + i] = bb[i];
      }
#pragma unroll
      for (int j = 0; j < NJ; j += 2) {
        const float a0 = wt[(8 * LPC * (j >> 3) + (j & 7)) * RS + mb * CW];
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, z[j], acc0, 0, 0, 0);
        const float a1 = wt[(8 * LPC * ((j + 1) >> 3) + ((j + 1) & 7)) * RS + mb * CW];
        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, z[j + 1], acc1, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const float v = acc0[i] + acc1[i];
        if constexpr (GRAD) rv[mb * V + i] = v;
        ss = fmaf(v, v, ss);
      }
    } else {
      f32x4 acc0 = *(const f32x4*)(P.b + mb * 16 + 4 * h), acc1 = {0};
#pragma unroll
      for (int j = 0; j < NJ; j += 2) {
        const float a0 = wt[(8 * LPC * (j >> 3) + (j & 7)) * RS + mb * CW];
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, z[j], acc0, 0, 0, 0);
        const float a1 = wt[(8 * LPC * ((j + 1) >> 3) + ((j + 1) & 7)) * RS + mb * CW];
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, z[j + 1], acc1, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float v = acc0[i] + acc1[i];
        if constexpr (GRAD) rv[mb * V + i] = v;
        ss = fmaf(v, v, ss);
      }
    }
  }
  if constexpr (GRAD) {
    if (__any(want)) {
      // (W^T r)[d] for the output rows d = db * CW + (this lane's row), K over r's rows o: lane part h
      // holds o = 16 mb + 4 h + i (LPC=4) or 32 mb + 8 (i / 4) + 4 h + i % 4 (LPC=2) in value i
      constexpr int CWP = CW + 1;
#pragma unroll
      for (int db = 0; db < MBI; ++db) {
        if (db >= MB) break;
        const float* wd = P.WT + (db * CW + m) * RS;
        if constexpr (LPC == 2) {
          f32x16 acc = {0};
#pragma unroll
          for (int mb = 0; mb < MBI; ++mb) {
            if (mb >= MB) break;
#pragma unroll
            for (int i = 0; i < 16; ++i)
              acc = __builtin_amdgcn_mfma_f32_32x32x2f32(wd[32 * mb + 8 * (i / 4) + 4 * h + (i % 4)], rv[mb * V + i],
                                                        acc, 0, 0, 0);
          }
#pragma unroll
          for (int i = 0; i < 16; ++i) gscr[(32 * db + 8 * (i / 4) + 4 * h + (i % 4)) * CWP + m] = acc[i];
        } else {
          f32x4 acc = {0};
#pragma unroll
          for (int mb = 0; mb < MBI; ++mb) {
            if (mb >= MB) break;
#pragma unroll
            for (int i = 0; i < 4; ++i)
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(wd[16 * mb + 4 * h + i], rv[mb * V + i], acc, 0, 0, 0);
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) gscr[(16 * db + 4 * h + i) * CWP + m] = acc[i];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int d = owned_dim<LPC>(h, j);
        const float wr = (j < nj && d < D) ? gscr[d * CWP + m] : 0.0f;
        g[j] = -wr * dq_du(z[j], (tfm >> j) & 1u);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  (void)dps;
  ss = chain_sum<LPC>(ss);
  double L = log_const - 0.5 * (double)ss;
  return __builtin_isnan(L) ? -__builtin_inf() : L;
}

// GRAD: when some lane of the wavefront asks (want), also g = d logL / du of the owned slots:
// d logL / dx_d = sum_k w_k (-2 inv2s2_k (x_d - m_kd)) with w_k the components' softmax weights, and
// dx_d / du_d = q_d dq/du (x = p + q (.) q(u))
template <int LPC, int NJ, bool GRAD = false>
__device__ __forceinline__ double lik_gmm(const float (&up)[NJ], int nj, int D, const GmmLds& P, uint32_t tfm,
                                          float* g = nullptr, bool want = false) {
  constexpr int CW = HIPNS_WAVE / LPC;
  const int h = lane_id() / CW;
  float part[HIPNS_MAX_GMM_K];
#pragma unroll
  for (int k = 0; k < HIPNS_MAX_GMM_K; ++k) part[k] = 0.0f;
  float qv[NJ];
  prior_q<NJ>(up, qv, tfm);
  // Slots 4i..4i+3 of a lane are four consecutive dimensions (owned_dim), so the prior scale, offset and
  // component means come in 16-byte LDS reads, all issued before the arithmetic that waits on them.
  // The reads cover the run's padding dimensions (below the row stride, so inside the LDS block);
  // those slots contribute nothing.
  float xv[NJ];
#pragma unroll
  for (int j = 0; j < NJ; j += 4) {
    if (j < nj) {  // uniform per run
      const int d0 = owned_dim<LPC>(h, j);
      const float4 pp = *reinterpret_cast<const float4*>(P.p + d0);
      const float4 qq = *reinterpret_cast<const float4*>(P.q + d0);
      xv[j] = fmaf(qq.x, qv[j], pp.x);
      xv[j + 1] = fmaf(qq.y, qv[j + 1], pp.y);
      xv[j + 2] = fmaf(qq.z, qv[j + 2], pp.z);
      xv[j + 3] = fmaf(qq.w, qv[j + 3], pp.w);
    }
  }
#pragma unroll
  for (int j = 0; j < NJ; j += 4) {
    if (j < nj) {
      const int d0 = owned_dim<LPC>(h, j);
      // every component's four means of these slots, read before any is used (components past K
      // read component K - 1 again, a valid row, and are not used)
      float4 m[HIPNS_MAX_GMM_K];
#pragma unroll
      for (int k = 0; k < HIPNS_MAX_GMM_K; ++k)
        m[k] = *reinterpret_cast<const float4*>(P.mean + (k < P.K ? k : P.K - 1) * P.dps + d0);
#pragma unroll
      for (int k = 0; k < HIPNS_MAX_GMM_K; ++k) {
        if (k < P.K) {
          const float mv[4] = {m[k].x, m[k].y, m[k].z, m[k].w};
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float r = d0 + i < D ? xv[j + i] - mv[i] : 0.0f;  // padding slots contribute nothing
            part[k] = fmaf(r, r, part[k]);
          }
        }
      }
    }
  }
  // log sum_k exp(c_k), c_k = logc_k - |x - m_k|^2 / (2 s_k^2), as c_top + log sum_k exp(c_k - c_top)
  // about the largest component: c_top is assembled in FP64 from its FP32 square sum (so points near
  // a peak keep distinct values), the correction (>= 0, <= log K) in FP32 with the hardware exp2 / log2
  float mx = -__builtin_inff(), ss_top = 0.0f, i_top = 0.0f, c_top = 0.0f;
#pragma unroll
  for (int k = 0; k < HIPNS_MAX_GMM_K; ++k) {
    if (k < P.K) {
      const float ss = chain_sum<LPC>(part[k]);
      part[k] = fmaf(-ss, P.inv2s2[k], P.logc[k]);  // component log-density, FP32
      if (!(part[k] <= mx)) {  // first maximum (a NaN component is taken, and propagates)
        mx = part[k];
        ss_top = ss;
        i_top = P.inv2s2[k];
        c_top = P.logc[k];
      }
    }
  }
  float sum = 0.0f;
#pragma unroll
  for (int k = 0; k < HIPNS_MAX_GMM_K; ++k)
    if (k < P.K) sum += __builtin_amdgcn_exp2f((part[k] - mx) * 1.44269504088896341f);
  if constexpr (GRAD) {
    if (__any(want)) {
      float wk[HIPNS_MAX_GMM_K];
      const float inv_sum = 1.0f / sum;
#pragma unroll
      for (int k = 0; k < HIPNS_MAX_GMM_K; ++k)
        wk[k] = k < P.K ? -2.0f * P.inv2s2[k] * __builtin_amdgcn_exp2f((part[k] - mx) * 1.44269504088896341f) *
                              inv_sum
                        : 0.0f;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int d = owned_dim<LPC>(h, j);
        const bool in = j < nj && d < D;
        float dx = 0.0f;
        if (in) {
          const float x = xv[j];
#pragma unroll
          for (int k = 0; k < HIPNS_MAX_GMM_K; ++k)
            if (k < P.K) dx = fmaf(wk[k], x - P.mean[k * P.dps + d], dx);
        }
        g[j] = in ? dx * P.q[d] * dq_du(qv[j], (tfm >> j) & 1u) : 0.0f;
      }
    }
  }
  const double L = fma(-(double)ss_top, (double)i_top, (double)c_top) +
                   (double)(0.693147180559945309f * __builtin_amdgcn_logf(sum));
  return __builtin_isnan(L) ? -__builtin_inf() : L;
}

// bit j set when owned dimension j of lane-part h uses q = ndtri
template <int LPC, int NJ>
__device__ __forceinline__ uint32_t transform_mask(const float* tf, int h, int nj) {
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < NJ; ++j)
    if (j < nj && tf[owned_dim<LPC>(h, j)] != 0.0f) m |= 1u << j;
  return m;
}

// ---------------------------------------------------------------------------
// the sampling kernel
// ---------------------------------------------------------------------------


// dynamic LDS layout of one block (one run per block)
__device__ __forceinline__ void load_params(int fam, const float* prm, int dps, int D, int K, float* lds, MvnLds& mv,
                                            GmmLds& gm, int LPCv, LinRegLds* lr = nullptr, int G = 1, int NG = 1) {
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  if (fam == HIPNS_FAM_LINREG) {
    lr->tf = lds;
    lr->p = lds + dps;
    lr->s = lds + 2 * dps;
    lr->w0 = lds + 3 * dps;
    lr->bops = (void*)(lds + 4 * dps + 4);
    lr->partial = (double*)(lds + 4 * dps + 4 + G * NG * 3 * 64 * 4);
    for (int i = tid; i < 4 * dps + 4; i += nt) lds[i] = prm[i];
    // FP64 copies of p, s and w0 follow the FP32 block (8-byte aligned: 4 dps + 4 floats)
    lr->f64 = (const double*)(prm + 4 * dps + 4);
    lr->dps = dps;
    __syncthreads();
    // 1/(2 sigma^2) as the two 32-bit words of an FP64 value (little-endian)
    lr->inv2s2 = __longlong_as_double(((long long)__float_as_uint(lds[4 * dps + 1]) << 32) |
                                      (long long)__float_as_uint(lds[4 * dps]));
    gm.tf = lr->tf;
    return;
  }
  if (fam == HIPNS_FAM_MVN) {
    const int RS = mvn_row_stride(8 * NG, LPCv);
    mv.tf = lds;
    mv.b = lds + dps;
    mv.WT = lds + 2 * dps;
    mv.RS = RS;
    for (int i = tid; i < dps; i += nt) {
      mv.tf[i] = prm[i];
    }
    for (int i = tid; i < dps; i += nt) mv.b[i] = prm[dps + i];
    const float* W = prm + 2 * dps;
    const int rows = dps > 8 * NG * LPCv ? dps : 8 * NG * LPCv;
    for (int i = tid; i < rows * RS; i += nt) {
      int d = i / RS, m = i % RS;
      mv.WT[i] = (d < dps && m < dps && m < RS - 1) ? W[m * dps + d] : 0.0f;
    }
  } else {
    gm.tf = lds;
    gm.p = lds + dps;
    gm.q = lds + 2 * dps;
    gm.mean = lds + 3 * dps;
    gm.inv2s2 = lds + 3 * dps + K * dps;
    gm.logc = gm.inv2s2 + K;
    gm.K = K;
    gm.dps = dps;
    for (int i = tid; i < (3 + K) * dps + 2 * K; i += nt) lds[i] = prm[i];
  }
  __syncthreads();
}


// ---------------------------------------------------------------------------
// slice-sampler steps shared by the fused kernel and the external-likelihood steps
// ---------------------------------------------------------------------------

// get_seed_point (uni_slice_sampler.py:343-358): uniform choice among live points [first, N) above the
// contour, the first i with (i - first + 1) >= M (1 - u), i.e. first + ceil(M (1 - u)) - 1
// (random.py:57-60)
__device__ __forceinline__ int seed_index_dev(int chain, uint32_t it32, uint32_t k0, uint32_t k1, int first, int N) {
  u32x4 w = philox4x32((uint32_t)chain, it32, 0xFFFFFFFFu, RNG_SEED, k0, k1);
  double u = u01d(w.x, w.y);
  int M = N - first;
  int idx = M > 0 ? first + (int)ceil((double)M * (1.0 - u)) - 1 : 0;
  return idx < 0 ? 0 : (idx > N - 1 ? N - 1 : idx);
}

// _sample_direction (uni_slice_sampler.py:23-38): z ~ N(0, I), n = z / |z|.  Two Philox blocks feed
// the 8 owned dimensions of a group through Box-Muller: word q of the first block gives the radius
// uniform u1 = (top 24 bits + 1) 2^-24 in (0, 1] and word q of the second the angle uniform
// u2 = (top 24 bits) 2^-24, for dimensions 2q and 2q + 1.  (16-bit halves of one block were cheaper,
// but their radius levels stop at 4.71 sigma and shifted the mixture batch's log Z by about +0.03
// nat.)  Only the direction matters (t scales with 1/|n|), so the hardware rsq normalises it.
template <int LPC, int NJ>
__device__ __forceinline__ void slice_direction(float (&nv)[NJ], int chain, uint32_t it32, int si, int D, int nj,
                                                int h, uint32_t k0, uint32_t k1) {
  if (D == 1) {
#pragma unroll
    for (int j = 0; j < NJ; ++j) nv[j] = (owned_dim<LPC>(h, j) == 0) ? 1.0f : 0.0f;
    return;
  }
  float ss = 0.0f;
#pragma unroll
  for (int j = 0; j < NJ; j += 8) {
    if (j < nj) {
      const int d0 = owned_dim<LPC>(h, j);
      const u32x4 w = philox4x32((uint32_t)chain, it32, (uint32_t)si, RNG_DIR | (uint32_t)(d0 >> 3), k0, k1);
      const u32x4 v = philox4x32((uint32_t)chain, it32, (uint32_t)si, RNG_DIR2 | (uint32_t)(d0 >> 3), k0, k1);
      const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
      const uint32_t vs[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float u1 = (float)((ws[q] >> 8) + 1u) * 5.9604644775390625e-8f;  // (0, 1]
        float u2 = (float)(vs[q] >> 8) * 5.9604644775390625e-8f;          // [0, 1)
        // sqrt(-2 ln u1) with the hardware log2; sin / cos of 2 pi u2 take the angle in revolutions
        float rr = __builtin_amdgcn_sqrtf(-1.38629436111989061f * __builtin_amdgcn_logf(u1));
        float sn = __builtin_amdgcn_sinf(u2), cs = __builtin_amdgcn_cosf(u2);
        float z0 = d0 + 2 * q < D ? rr * cs : 0.0f;
        float z1 = d0 + 2 * q + 1 < D ? rr * sn : 0.0f;
        nv[j + 2 * q] = z0;
        nv[j + 2 * q + 1] = z1;
        ss = fmaf(z0, z0, ss);
        ss = fmaf(z1, z1, ss);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) nv[j + q] = 0.0f;
    }
  }
  ss = chain_sum<LPC>(ss);
  const float inv = __builtin_amdgcn_rsqf(ss);
#pragma unroll
  for (int j = 0; j < NJ; ++j) nv[j] = nv[j] * inv;
}

// _slice_bounds (uni_slice_sampler.py:41-64).  Padding slots (beyond D or nj) hold U0 = 0.5 and n = +0,
// whose quotients are +-inf and never the bound, so every slot is taken unconditionally.
// With every coordinate inside (2^-60, 1 - 2^-60) the numerators 1 - u0 > 0 > -u0 are nonzero, so the
// sign of n alone says which quotient bounds which side, and both quotients share one reciprocal.  A
// unit direction has a component of size >= 1/sqrt(D), which bounds the bracket by sqrt(D); slots with
// |n| < 2^-94 (quotients beyond 2^34) can therefore never be the bound.  A wavefront holding a
// coordinate on or outside that range takes the general candidate rule with full divisions.
template <int LPC, int NJ>
__device__ __forceinline__ void slice_bracket(const float (&U0)[NJ], const float (&nv)[NJ], float& left, float& right) {
#pragma clang fp contract(off)
  float rmin = __builtin_inff(), lmax = -__builtin_inff();
  // smallest numerator, compared as signed bit patterns (the order of non-negative floats; a negative
  // numerator is below every positive one)
  int amin = 0x3F800000;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const float u0 = U0[j], n = nv[j];
    const float a1 = 1.0f - u0, a0 = -u0;
    amin = min(amin, min(__float_as_int(u0), __float_as_int(a1)));
    const float y = recip_rn(n);
    const float t1 = quot_rn(a1, n, y), t0 = quot_rn(a0, n, y);
    // a1 >= 0 >= a0, so the quotient of the sign of n is the right candidate and the other the left one
    // (n = +-0: both NaN, which fminf / fmaxf skip)
    rmin = fminf(rmin, fmaxf(t0, t1));
    lmax = fmaxf(lmax, fminf(t0, t1));
  }
  if (__any(amin < 0x21800000)) {  // 0x21800000 = 2^-60
    rmin = __builtin_inff();
    lmax = -__builtin_inff();
#pragma unroll
    for (int j = 0; j < NJ; ++j) bound_candidates(U0[j], nv[j], rmin, lmax);
  }
  right = chain_min<LPC>(rmin);
  left = chain_max<LPC>(lmax);
}

// gradient-guided direction (uni_slice_sampler.py:255-271): the Householder reflection of the last
// direction nv about the unit gradient, n - 2 (n . g) g, renormalised; a gradient of norm below 1e-10
// or not finite gives a new random direction (the slice's own draw)
template <int LPC, int NJ>
__device__ __forceinline__ void reflect_direction(float (&nv)[NJ], float (&g)[NJ], int chain, uint32_t it32, int si,
                                                  int D, int nj, int h, uint32_t k0, uint32_t k1) {
  float gg2 = 0.0f;
#pragma unroll
  for (int j = 0; j < NJ; ++j) gg2 = fmaf(g[j], g[j], gg2);
  const float gn = __builtin_sqrtf(chain_sum<LPC>(gg2));
  if (!(gn >= 1e-10f) || !__builtin_isfinite(gn)) {
    slice_direction<LPC, NJ>(nv, chain, it32, si, D, nj, h, k0, k1);
    return;
  }
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

// gradient slice direction (uni_slice_sampler.py:202-214): the unit gradient g / |g|, unless |g| is 0 or
// not finite (then false, and the slice keeps its own direction and full bracket)
template <int LPC, int NJ>
__device__ __forceinline__ bool gradient_direction(float (&nv)[NJ], const float (&g)[NJ]) {
  float gg2 = 0.0f;
#pragma unroll
  for (int j = 0; j < NJ; ++j) gg2 = fmaf(g[j], g[j], gg2);
  const float gn = __builtin_sqrtf(chain_sum<LPC>(gg2));
  if (!(gn > 0.0f) || !__builtin_isfinite(gn)) return false;
#pragma unroll
  for (int j = 0; j < NJ; ++j) nv[j] = g[j] / gn;
  return true;
}

// a uniform point of the unit cube for attempt `attempt` of a chain (Model.sample_U; padding 0.5)
template <int LPC, int NJ>
__device__ __forceinline__ void uniform_point(float (&up)[NJ], int chain, uint32_t it32, int attempt, int D, int nj,
                                              int h, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int j = 0; j < NJ; j += 4) {
    int d0 = owned_dim<LPC>(h, j);
    u32x4 w = philox4x32((uint32_t)chain, it32, (uint32_t)attempt, RNG_UNIF | (uint32_t)(d0 >> 2), k0, k1);
    float uu[4] = {u01f(w.x), u01f(w.y), u01f(w.z), u01f(w.w)};
#pragma unroll
    for (int q = 0; q < 4; ++q) up[j + q] = (j < nj && d0 + q < D) ? uu[q] : 0.5f;
  }
}

template <int LPC, int NJ, int FAM, int MODE>
__global__ void __launch_bounds__(256) k_sample(SampleArgs A) {
  constexpr int CW = HIPNS_WAVE / LPC;
  // the regression family multiplies the data rows against every chain group of the workgroup
  // cooperatively, so its wavefronts step together (workgroup-uniform loop)
  constexpr bool COOP = (FAM == HIPNS_FAM_LINREG);
  // gradient slice chains, run code sampler 2, 3 or 4.  gradient_guided (2 and 4,
  // uni_slice_sampler.py:255-271): after each accepted slice the chain spends one step (one likelihood
  // evaluation) on d logL / dU at the accepted point, and the next direction is the reflection of the
  // last one about it.  gradient_slice (3 and 4, uni_slice_sampler.py:202-214): every slice runs along
  // the unit gradient at its seed point over the half bracket [0, right]; the gradient step before slice
  // s > 0 is the step after slice s - 1, at the same point.
  constexpr bool GG = (MODE == MODE_GSLICE);
  constexpr bool SL = (MODE == MODE_SLICE || MODE == MODE_INJ || GG);  // slice chains
  constexpr bool DRAWN = (MODE == MODE_SLICE || GG);                   // slice chains on Philox draws
  static_assert(!GG || FAM != HIPNS_FAM_LINREG, "gradient-guided chains need the MVN or mixture family");
  extern __shared__ float lds[];
  // run of this block row: the batch's dispatch order (longest chains first), or the identity
  const int r = (MODE != MODE_INJ && A.B.run_order) ? A.B.run_order[blockIdx.y] : (int)blockIdx.y;
  const hipns_run_desc_t run = A.runs[r];
  hipns_run_state_t* st = A.states + r;
  if (MODE != MODE_INJ && st->done) return;  // uniform across the block
  const int wpb = blockDim.x / HIPNS_WAVE;
  const int wave = threadIdx.x / HIPNS_WAVE;
  const int G = COOP ? wpb : 1;
  const int l = lane_id();
  const int c = l % CW;
  const int h = l / CW;
  const int nchains = (MODE == MODE_INIT) ? run.N : ((MODE == MODE_INJ) ? A.inj_C : run.shell);
  const int chain0 = (blockIdx.x * wpb + wave) * CW;
  if (blockIdx.x * wpb * CW >= nchains) return;  // whole block idle
  const int dps = A.B.dps;
  const int D = run.D;
  const int DPr = ((D + 8 * LPC - 1) / (8 * LPC)) * (8 * LPC);
  const int nj = DPr / LPC;

  MvnLds mv;
  GmmLds gm;
  LinRegLds lr;
  load_params(FAM, A.B.params + run.param_off, dps, D, A.gmm_K, lds, mv, gm, LPC, &lr, G, NJ / 8);
  if (COOP) {
    lr.r0 = A.lr_r0;
    lr.xh = A.lr_xh;
    lr.xl = A.lr_xl;
    lr.xf = A.lr_xf;
    lr.y = nullptr;
    lr.npad = A.lr_npad;
    lr.panels = A.lr_panels;
    lr.terms = A.lr_terms;
  }
  const uint32_t tfm = transform_mask<LPC, NJ>(FAM == HIPNS_FAM_MVN ? mv.tf : gm.tf, h, nj);

  const int chain = chain0 + c;
  bool alive = chain < nchains;
  const int64_t iter = (MODE == MODE_INIT) ? 0xFFFFFFFFll : ((MODE == MODE_INJ) ? 0 : st->iteration);
  const uint32_t k0 = run.seed_lo, k1 = run.seed_hi;
  const uint32_t it32 = (uint32_t)iter;
  const int cur = (MODE == MODE_INJ) ? A.inj_buf : st->buf;
  const double contour = (MODE == MODE_INJ) ? A.inj_contour : ((MODE == MODE_INIT) ? -__builtin_inf() : st->contour);
  const int N = run.N;
  const int S = run.shell;
  const int num_slices = run.num_slices;
  const int kph = run.k_phantom;
  const bool midpoint = run.midpoint_shrink != 0;

  float U0[NJ], nv[NJ], up[NJ];
  double logL0 = -__builtin_inf();
  int64_t evals = 0;
  int s = 0, k = 0;
  float left = 0.0f, right = 0.0f, t = 0.0f, alpha = 0.5f;
  int cap_hits = 0;

  // ---- seed point (get_seed_point, uni_slice_sampler.py:343-358)
  if (SL) {
    int idx = 0;
    if (alive) {
      if (MODE == MODE_INJ) {
        idx = A.inj_seed[chain];
      } else {
        idx = seed_index_dev(chain, it32, k0, k1, st->first_seed, N);
      }
    }
    const float* src = A.B.live_U[cur] + (run.live_off + idx) * (int64_t)dps;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      int d = owned_dim<LPC>(h, j);
      U0[j] = (alive && j < nj) ? src[d] : 0.5f;
    }
    logL0 = alive ? A.B.live_logL[cur][run.live_off + idx] : 0.0;
  } else {
#pragma unroll
    for (int j = 0; j < NJ; ++j) U0[j] = 0.5f;
  }

  // alpha = linspace(0.5, 1, num_slices)[s] (uni_slice_sampler.py:422), one table per block
  float* alpha_tab = lds + A.alpha_off;
  if (SL) {
    for (int i = threadIdx.x; i < num_slices; i += blockDim.x)
      alpha_tab[i] = (float)(num_slices > 1 ? 0.5 + 0.5 * (double)i / (double)(num_slices - 1) : 0.5);
    __syncthreads();
  }

  // ---- ring of slice directions drawn ahead (MODE_SLICE): slot b of this wavefront holds, for every
  // lane, the nj owned components of one direction, as nj / 4 float4 rows of 64 lanes.  A chain's ring
  // holds the directions of its slices [s, gen) (gen - s <= ring), in slot (slice index mod ring).
  constexpr int RB = dir_ring_max(MODE, FAM, NJ);
  constexpr int RSLOT = (NJ / 4) * HIPNS_WAVE;  // float4s of one ring slot
  const int ring = RB > 0 ? A.dir_ring : 0;
  float4* ring_base = nullptr;
  if (RB > 0) ring_base = (float4*)(lds + A.ring_off) + (size_t)wave * ring * RSLOT + l;
  int gen = 0, gslot = 0;  // next slice to draw and its slot
  // gradient slice chains: the wavefront's LDS scratch for d logL / dU, and the state of the gradient step
  float* gscr = GG ? lds + A.ring_off + (size_t)wave * grad_rows(NJ, LPC) * (CW + 1) : nullptr;
  const bool gs_on = GG && run.sampler >= 3;  // gradient_slice
  const bool gg_on = GG && run.sampler != 3;  // gradient_guided
  bool need_grad = false, dir_ready = false, half = false;
  float gv[NJ];

  // ---- begin a slice: direction and perfect bracket (the pick follows in the step)
  auto begin_slice = [&](int si) {
    k = 0;
    alpha = alpha_tab[si];
    if (GG && dir_ready) {
      dir_ready = false;  // the direction of the gradient step (reflected, or the unit gradient)
    } else if (RB > 0 && ring > 0) {
      // slice si = s is in the ring (gen > s), gen - s slots behind the next free one
      int rslot = gslot - (gen - si);
      if (rslot < 0) rslot += ring;
      const float4* src = ring_base + (size_t)rslot * RSLOT;
#pragma unroll
      for (int j4 = 0; j4 < NJ / 4; ++j4) {
        const float4 v = src[j4 * HIPNS_WAVE];
        nv[4 * j4] = v.x;
        nv[4 * j4 + 1] = v.y;
        nv[4 * j4 + 2] = v.z;
        nv[4 * j4 + 3] = v.w;
      }
    } else if (MODE == MODE_INJ) {
      const float* dsrc = A.inj_dirs + ((int64_t)chain * num_slices + si) * dps;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int d = owned_dim<LPC>(h, j);
        nv[j] = (j < nj && d < D) ? dsrc[d] : 0.0f;
      }
    } else {
      slice_direction<LPC, NJ>(nv, chain, it32, si, D, nj, h, k0, k1);
    }
    slice_bracket<LPC, NJ>(U0, nv, left, right);
    if (GG && half) {
      left = 0.0f;  // climb the gradient: the bracket's negative side is dropped (uni_slice_sampler.py:212)
      half = false;
    }
  };

  auto draw_uniform_point = [&](int attempt) { uniform_point<LPC, NJ>(up, chain, it32, attempt, D, nj, h, k0, k1); };

  bool need_begin = SL && alive;
  bool need_shrink = false;
  need_grad = gs_on && alive;  // gradient_slice: the first slice needs the gradient at the seed point

  // the shell's phantoms are appended after its discarded points (add_phantom_samples_to_state,
  // sharded_static.py:181-207), chain-major, k per chain
  const int64_t phantom_base =
      (DRAWN && kph > 0) ? append_start((st->next_idx + S) % run.max_samples, (int64_t)S * kph, run.max_samples)
                         : 0;
  const int sender = DRAWN ? st->sender_node : 0;
  const int max_uniform = (MODE == MODE_INIT) ? 1000000 : run.max_uniform_evals;
  u32x4 pickblk = {0u, 0u, 0u, 0u};

  // Every live chain evaluates exactly once per step, so the step index w equals each live
  // chain's evaluation count: the pick uniform of a chain's w-th evaluation is word (w mod 4)
  // of its Philox block w/4.  The LPC lanes of a chain draw LPC consecutive blocks at once (part h
  // block w/4 + h, every 4 LPC steps), and each step the lane part holding block w/4 passes its
  // word to the others.
  for (int w = 0; COOP ? __syncthreads_or(alive) : __any(alive); ++w) {
    uint32_t pickw = 0u;
    if (DRAWN) {
      if ((w & (4 * LPC - 1)) == 0)
        pickblk = philox4x32((uint32_t)chain, it32, (uint32_t)((w >> 2) + h), RNG_PICK, k0, k1);
      const int kw = w & 3;
      const uint32_t own = kw == 0 ? pickblk.x : (kw == 1 ? pickblk.y : (kw == 2 ? pickblk.z : pickblk.w));
      pickw = part_broadcast<LPC>(own, (w >> 2) & (LPC - 1));
    }
    if (RB > 0 && ring > 0) {
      // a chain about to begin slice s with an empty ring: one draw for every live chain whose ring has
      // room (the same Philox counters as a draw at the slice start, so the directions are unchanged)
      if (__any(alive && need_begin && gen == s)) {
        if (alive && gen < num_slices && gen - s < ring) {
          float dn[NJ];
          slice_direction<LPC, NJ>(dn, chain, it32, gen, D, nj, h, k0, k1);
          float4* dst = ring_base + (size_t)gslot * RSLOT;
#pragma unroll
          for (int j4 = 0; j4 < NJ / 4; ++j4)
            dst[j4 * HIPNS_WAVE] = make_float4(dn[4 * j4], dn[4 * j4 + 1], dn[4 * j4 + 2], dn[4 * j4 + 3]);
          ++gen;
          gslot = gslot + 1 == ring ? 0 : gslot + 1;
        }
      }
    }
    if (SL) {
      if (alive && !(GG && need_grad)) {
        if (need_begin) {
          begin_slice(s);
          need_begin = false;
        } else if (need_shrink) {
          // _shrink_interval (uni_slice_sampler.py:92-111)
          shrink(t, alpha, midpoint, left, right);
          ++k;
        }
        need_shrink = false;
        float u;
        if (MODE == MODE_INJ) {
          u = A.inj_tu[((int64_t)chain * num_slices + s) * A.max_steps + k];
        } else {
          u = u01f(pickw);
        }
        // _pick_point_in_interval (uni_slice_sampler.py:67-89)
        t = pick_t(left, right, u);
      }
    }
    // ---- proposal (a gradient step evaluates at the accepted point)
    if (SL) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) up[j] = (GG && need_grad) ? U0[j] : point_on_line(U0[j], t, nv[j]);
    } else {
      draw_uniform_point((int)evals);
    }
    // ---- batched likelihood of the wavefront's CW proposals (Model.forward)
    double L;
    if constexpr (GG) {
      if constexpr (FAM == HIPNS_FAM_MVN)
        L = lik_mvn<LPC, NJ, true>(up, nj, D, dps, mv, run.log_const, tfm, gscr, gv, alive && need_grad);
      else
        L = lik_gmm<LPC, NJ, true>(up, nj, D, gm, tfm, gv, alive && need_grad);
    } else if constexpr (FAM == HIPNS_FAM_MVN) {
      L = lik_mvn<LPC, NJ>(up, nj, D, dps, mv, run.log_const, tfm);
    } else if constexpr (FAM == HIPNS_FAM_GMM) {
      L = lik_gmm<LPC, NJ>(up, nj, D, gm, tfm);
    } else {
      L = lik_linreg<LPC, NJ>(up, nj, D, lr, tfm, wave, G, run.log_const);
    }

    if (alive) {
      ++evals;
      if (MODE == MODE_INJ) {
        int64_t o = ((int64_t)chain * num_slices + s) * A.max_steps + k;
        if (h == 0) {
          A.trace_t[o] = t;
          A.trace_logL[o] = L;
        }
      }
      if (GG && need_grad) {
        // the gradient step, counted as one evaluation per grad_fn call of the reference: with both
        // gradient_guided and gradient_slice, the step between two slices stands for both calls at the
        // accepted point.  Finish after the last slice; else, gradient_slice, the next slice runs along
        // the unit gradient (or its own direction where the gradient is 0 or not finite:
        // uni_slice_sampler.py:202-214); gradient_guided, reflect the direction about the gradient, or
        // draw a new one where the gradient is below 1e-10 or not finite (uni_slice_sampler.py:255-271)
        need_grad = false;
        if (gs_on && gg_on && s > 0 && s < num_slices) ++evals;
        if (s == num_slices) {
          alive = false;
          const int64_t o = run.chain_off + chain;
          float* dst = A.B.new_U + o * dps;
#pragma unroll
          for (int j = 0; j < NJ; ++j)
            if (j < nj) dst[owned_dim<LPC>(h, j)] = U0[j];
          if (h == 0) {
            A.B.new_logL[o] = logL0;
            A.B.new_evals[o] = (int32_t)evals;
          }
        } else if (gs_on) {
          // a gradient below 1e-10 but not 0 is reflected about by gradient_guided, and that direction
          // then gives way to the unit gradient; a zero or non-finite one leaves the slice's own draw
          dir_ready = gradient_direction<LPC, NJ>(nv, gv);
          half = dir_ready;
          need_begin = true;
        } else {
          reflect_direction<LPC, NJ>(nv, gv, chain, it32, s, D, nj, h, k0, k1);
          dir_ready = true;
          need_begin = true;
        }
      } else if (MODE == MODE_INIT || MODE == MODE_UNIF) {
        // uniform_sample.py:35-36 (logL > -inf) / uniform_samplers.py:44-50
        bool acc = (MODE == MODE_INIT) ? (L > -__builtin_inf()) : ((L > contour) || (L == contour));
        if (!acc && evals >= max_uniform) acc = true;
        if (acc) {
#pragma unroll
          for (int j = 0; j < NJ; ++j) U0[j] = up[j];
          logL0 = L;
          alive = false;
          const int64_t o = run.chain_off + chain;
          float* dst = A.B.new_U + o * dps;
#pragma unroll
          for (int j = 0; j < NJ; ++j)
            if (j < nj) dst[owned_dim<LPC>(h, j)] = U0[j];
          if (h == 0) {
            A.B.new_logL[o] = logL0;
            A.B.new_evals[o] = (int32_t)evals;
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
          ++cap_hits;
        }
        if (acc) {
          if (MODE == MODE_INJ && h == 0) A.trace_steps[(int64_t)chain * num_slices + s] = k + 1;
#pragma unroll
          for (int j = 0; j < NJ; ++j) U0[j] = up[j];
          logL0 = L;
          // phantom samples: x[-(k+1):-1] (uni_slice_sampler.py:432-440), evals 0
          int pidx = s - (num_slices - 1 - kph);
          if (kph > 0 && pidx >= 0 && s < num_slices - 1) {
            if (MODE == MODE_INJ) {
              float* pd = A.ph_U + ((int64_t)chain * kph + pidx) * dps;
#pragma unroll
              for (int j = 0; j < NJ; ++j)
                if (j < nj) pd[owned_dim<LPC>(h, j)] = U0[j];
              if (h == 0) A.ph_logL[(int64_t)chain * kph + pidx] = L;
            } else {
              const int64_t si = run.samp_off + phantom_base + (int64_t)chain * kph + pidx;
              float* pd = A.B.samp_U + si * dps;
#pragma unroll
              for (int j = 0; j < NJ; ++j)
                if (j < nj) pd[owned_dim<LPC>(h, j)] = U0[j];
              if (h == 0) {
                A.B.samp_logL[si] = L;
                A.B.samp_sender[si] = sender;
                A.B.samp_evals[si] = 0;
                A.B.samp_phantom[si] = 1;
              }
            }
          }
          ++s;
          if (GG && (gg_on || s < num_slices)) {
            // gradient_guided: every accepted slice asks for the gradient, the last one included;
            // gradient_slice: every slice but the first starts with the gradient at its seed
            need_grad = true;
          } else if (s == num_slices) {
            alive = false;
            const int64_t o = run.chain_off + chain;
            float* dst = A.B.new_U + o * dps;
#pragma unroll
            for (int j = 0; j < NJ; ++j)
              if (j < nj) dst[owned_dim<LPC>(h, j)] = U0[j];
            if (h == 0) {
              A.B.new_logL[o] = logL0;
              A.B.new_evals[o] = (int32_t)evals;
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
  if (cap_hits > 0 && h == 0 && MODE != MODE_INJ) {
    atomicAdd((unsigned long long*)&st->cap_hits, (unsigned long long)cap_hits);
    atomicOr(&st->flags, 1);
  }
}

}  // namespace hipns
