// hipns_dev.h -- device helpers shared by the gfx950 kernels of libhipns.
//
// Wavefront-64 idioms, the Philox4x32-10 counter PRNG, the inverse normal
// CDF used by the prior transform, FP64 log-space arithmetic and the
// composition algebra of the evidence recurrence.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hipns.h"

#define HIPNS_WAVE 64

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace hipns {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// ---------------------------------------------------------------------------
// cross-lane reductions among the LPC lanes that share one chain.  Lane l
// serves chain (l % CW) with part h = l / CW, CW = 64 / LPC, so partners are
// at xor distances CW, 2 CW, ...
// ---------------------------------------------------------------------------
// gfx950 lane swaps: permlane16_swap exchanges the odd 16-lane rows of its
// first operand with the even rows of its second, permlane32_swap the upper
// half of the first with the lower half of the second.  Called with the same
// value twice, the two results hold the two partners of every lane, so
// r[0] (op) r[1] is the reduction over lanes l and l^16 (resp. l^32), with the
// operands in the same order on both partners.
__device__ __forceinline__ void swap16(float v, float& a, float& b) {
  auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap32(float v, float& a, float& b) {
  auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}

// the value of lane l ^ 8 (LPC = 8: the partner 8 lanes away, inside the same 16-lane row): a DPP
// row rotation by 8
__device__ __forceinline__ uint32_t xor8(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false);  // row_ror:8
}
__device__ __forceinline__ float xor8f(float v) { return __uint_as_float(xor8(__float_as_uint(v))); }

// value of lane part `part` (wavefront-uniform) of this lane's chain, on every part of the chain:
// r[0] / r[1] of a swap are the values of the lower / upper partner
template <int LPC>
__device__ __forceinline__ uint32_t part_broadcast(uint32_t v, int part) {
  auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  if constexpr (LPC == 8) {
    v = (part & 4) ? r[1] : r[0];
    auto q = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    v = (part & 2) ? q[1] : q[0];
    // bit 0 of the part is bit 3 of the lane: keep this lane's value where they agree, else the
    // partner's
    const uint32_t x = xor8(v);
    return (((lane_id() >> 3) & 1) == (part & 1)) ? v : x;
  } else if constexpr (LPC == 4) {
    v = (part & 2) ? r[1] : r[0];
    auto q = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (part & 1) ? q[1] : q[0];
  }
  return part ? r[1] : r[0];
}

// reductions over the LPC lanes of a chain; with a commutative operation the partners compute the
// same value (a + b = b + a in IEEE arithmetic)
template <int LPC>
__device__ __forceinline__ float chain_sum(float v) {
  float a, b;
  if constexpr (LPC == 8) v = v + xor8f(v);
  if constexpr (LPC >= 4) {
    swap16(v, a, b);
    v = a + b;
  }
  swap32(v, a, b);
  return a + b;
}
template <int LPC>
__device__ __forceinline__ float chain_min(float v) {
  float a, b;
  if constexpr (LPC == 8) v = fminf(v, xor8f(v));
  if constexpr (LPC >= 4) {
    swap16(v, a, b);
    v = fminf(a, b);
  }
  swap32(v, a, b);
  return fminf(a, b);
}
template <int LPC>
__device__ __forceinline__ float chain_max(float v) {
  float a, b;
  if constexpr (LPC == 8) v = fmaxf(v, xor8f(v));
  if constexpr (LPC >= 4) {
    swap16(v, a, b);
    v = fmaxf(a, b);
  }
  swap32(v, a, b);
  return fmaxf(a, b);
}

// ---------------------------------------------------------------------------
// a / n rounded to nearest without the operand scaling of the general IEEE division: y = RN(1/n) from
// the hardware reciprocal and one Newton step, then one residual correction of q = a y.  With y the
// correctly rounded reciprocal and q within an ulp of a / n, r = a - n q is exact and RN(q + r y) =
// RN(a / n) (P. Markstein, IBM J. Res. Dev. 34(1), 1990).  y = RN(1/n) holds for every |n| in
// [2^-94, 1] (checked exhaustively on the device); for |a| in [2^-60, 2] nothing under- or overflows,
// so the quotient is bit-identical to a / n.  n = +-0 yields NaN (where a / n is +-inf): both say "no
// bound from this slot".
__device__ __forceinline__ float recip_rn(float n) {
  const float y = __builtin_amdgcn_rcpf(n);
  const float e = fmaf(-n, y, 1.0f);
  return fmaf(e, y, y);
}
__device__ __forceinline__ float quot_rn(float a, float n, float y) {
  const float q = a * y;
  const float r = fmaf(-n, q, a);
  return fmaf(r, y, q);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of the
// multiply-xor bijection with Weyl key schedule.
// ---------------------------------------------------------------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                            uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32 -> 64-bit product per multiplier gives both halves of mulhilo
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return {c0, c1, c2, c3};
}

// uniform in [0,1) with 24 random bits
__device__ __forceinline__ float u01f(uint32_t w) { return (float)(w >> 8) * 5.9604644775390625e-8f; }
// uniform in (0,1]
__device__ __forceinline__ float u01f_open0(uint32_t w) { return (float)((w >> 8) + 1u) * 5.9604644775390625e-8f; }
// uniform in [0,1) with 53 random bits
__device__ __forceinline__ double u01d(uint32_t a, uint32_t b) {
  uint64_t v = (((uint64_t)a << 21) ^ (uint64_t)(b >> 11)) & ((1ull << 53) - 1);
  return (double)v * (1.0 / 9007199254740992.0);
}

// counter-domain tags (word 3 of the Philox counter)
enum : uint32_t {
  RNG_SEED = 0u << 24,
  RNG_DIR = 1u << 24,
  RNG_PICK = 2u << 24,
  RNG_UNIF = 3u << 24,
  RNG_DIR2 = 4u << 24,  // the angle uniforms of the slice directions
};

// ---------------------------------------------------------------------------
// Inverse standard-normal CDF.
//  - binary64: Wichura, Algorithm AS 241 (Applied Statistics 37, 1988), PPND16.
//  - binary32: M. Giles, "Approximating the erfinv function" (GPU Computing
//    Gems Jade, 2011), single-precision polynomials in w = -log(4u(1-u)),
//    evaluated directly in u so that both tails keep relative accuracy;
//    |u - 1/2| > 1/2 - 1e-7 falls back to PPND16.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ndtri_f64(double p) {
  if (!(p > 0.0 && p < 1.0)) {
    if (p == 0.0) return -__builtin_inf();
    if (p == 1.0) return __builtin_inf();
    return __builtin_nan("");
  }
  double q = p - 0.5;
  if (fabs(q) <= 0.425) {
    double r = 0.180625 - q * q;
    double num = 2.5090809287301226727e+3;
    num = num * r + 3.3430575583588128105e+4;
    num = num * r + 6.7265770927008700853e+4;
    num = num * r + 4.5921953931549871457e+4;
    num = num * r + 1.3731693765509461125e+4;
    num = num * r + 1.9715909503065514427e+3;
    num = num * r + 1.3314166789178437745e+2;
    num = num * r + 3.3871328727963666080e0;
    double den = 5.2264952788528545610e+3;
    den = den * r + 2.8729085735721942674e+4;
    den = den * r + 3.9307895800092710610e+4;
    den = den * r + 2.1213794301586595867e+4;
    den = den * r + 5.3941960214247511077e+3;
    den = den * r + 6.8718700749205790830e+2;
    den = den * r + 4.2313330701600911252e+1;
    den = den * r + 1.0;
    return q * num / den;
  }
  double r = q < 0.0 ? p : 1.0 - p;
  r = sqrt(-log(r));
  double num, den;
  if (r <= 5.0) {
    r -= 1.6;
    num = 7.74545014278341407640e-4;
    num = num * r + 2.27238449892691845833e-2;
    num = num * r + 2.41780725177450611770e-1;
    num = num * r + 1.27045825245236838258e0;
    num = num * r + 3.64784832476320460504e0;
    num = num * r + 5.76949722146069140550e0;
    num = num * r + 4.63033784615654529590e0;
    num = num * r + 1.42343711074968357734e0;
    den = 1.05075007164441684324e-9;
    den = den * r + 5.47593808499534494600e-4;
    den = den * r + 1.51986665636164571966e-2;
    den = den * r + 1.48103976427480074590e-1;
    den = den * r + 6.89767334985100004550e-1;
    den = den * r + 1.67638483018380384940e0;
    den = den * r + 2.05319162663775882187e0;
    den = den * r + 1.0;
  } else {
    r -= 5.0;
    num = 2.01033439929228813265e-7;
    num = num * r + 2.71155556874348757815e-5;
    num = num * r + 1.24266094738807843860e-3;
    num = num * r + 2.65321895265761230930e-2;
    num = num * r + 2.96560571828504891230e-1;
    num = num * r + 1.78482653991729133580e0;
    num = num * r + 5.46378491116411436990e0;
    num = num * r + 6.65790464350110377720e0;
    den = 2.04426310338993978564e-15;
    den = den * r + 1.42151175831644588870e-7;
    den = den * r + 1.84631831751005468180e-5;
    den = den * r + 7.86869131145613259100e-4;
    den = den * r + 1.48753612908506148525e-2;
    den = den * r + 1.36929880922735805310e-1;
    den = den * r + 5.99832206555887937690e-1;
    den = den * r + 1.0;
  }
  double v = num / den;
  return q < 0.0 ? -v : v;
}

__device__ __noinline__ float ndtri_tail_f32(float u) { return (float)ndtri_f64((double)u); }

// Giles' body for u in [1e-7, 1 - 1e-7].  w = -ln(4u(1-u)) comes from the hardware log2 (one
// rounding more than a full-precision log: about 1e-7 relative in w, below the polynomials' own error).
// The w >= 5 polynomial runs only when some lane of the wavefront needs it.
__device__ __forceinline__ float giles_w(float u) {
  return -0.693147180559945309f * __builtin_amdgcn_logf(4.0f * u * (1.0f - u));
}
__device__ __forceinline__ float giles_x(float u) {
  return fmaf(u, 2.82842712474619010f, -1.41421356237309505f);  // sqrt(2) (2u - 1), one rounding
}
// the central polynomial (w < 5)
__device__ __forceinline__ float giles_central(float w) {
  const float z = w - 2.5f;
  float p = 2.81022636e-08f;
  p = fmaf(p, z, 3.43273939e-07f);
  p = fmaf(p, z, -3.5233877e-06f);
  p = fmaf(p, z, -4.39150654e-06f);
  p = fmaf(p, z, 0.00021858087f);
  p = fmaf(p, z, -0.00125372503f);
  p = fmaf(p, z, -0.00417768164f);
  p = fmaf(p, z, 0.246640727f);
  p = fmaf(p, z, 1.50140941f);
  return p;
}
// the tail polynomial (w >= 5)
__device__ __forceinline__ float giles_tail(float w) {
  const float v = __builtin_amdgcn_sqrtf(w) - 3.0f;
  float pt = -0.000200214257f;
  pt = fmaf(pt, v, 0.000100950558f);
  pt = fmaf(pt, v, 0.00134934322f);
  pt = fmaf(pt, v, -0.00367342844f);
  pt = fmaf(pt, v, 0.00573950773f);
  pt = fmaf(pt, v, -0.0076224613f);
  pt = fmaf(pt, v, 0.00943887047f);
  pt = fmaf(pt, v, 1.00167406f);
  pt = fmaf(pt, v, 2.83297682f);
  return pt;
}

__device__ __forceinline__ float ndtri_giles(float u) {
  const float w = giles_w(u);
  const float x = giles_x(u);
  float p = giles_central(w);
  const bool tl = w >= 5.0f;
  if (__any(tl)) p = tl ? giles_tail(w) : p;
  return p * x;
}

// u far in the tails (min(u, 1 - u) < 1e-7) or outside (0, 1): binary64 AS 241
__device__ __forceinline__ bool ndtri_far(float u) { return !(u >= 1e-7f && u <= 0.9999999f); }

__device__ __forceinline__ float ndtri_f32(float u) {
  float z = ndtri_giles(u);
  if (ndtri_far(u)) z = ndtri_tail_f32(u);
  return z;
}

// ---------------------------------------------------------------------------
// FP64 log-space arithmetic.  lae(a, b) = log(e^a + e^b) with the jnp.logaddexp
// convention for equal infinities (returns a + b).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lae(double a, double b) {
  double m = fmax(a, b);
  double d = a - b;
  if (__builtin_isnan(d)) return a + b;
  return m + log1p(exp(-fabs(d)));
}

// ---------------------------------------------------------------------------
// Evidence recurrence as composable maps.
//
// One _update_evidence_calc_op step (internals/shrinkage_statistics.py:43-94)
// acts on S = (X, X2, Z, ZX, Z2, dZ2) (all >= 0, kept as logs) linearly:
//   X'  = T X                 X2' = T2 X2
//   Z'  = Z  + (t m) X        ZX' = T ZX + (tT m) X2
//   Z2' = Z2 + (2 t m) ZX + (t2 m^2) X2
//   dZ2'= dZ2 + (t2 m^2) X2
// with T=n/(n+1), t=1/(n+1), T2=n/(n+2), t2=2/((n+1)(n+2)), tT=n/((n+1)(n+2)),
// m = (L_prev + L)/2.  Maps of the form
//   X'=aX, X2'=bX2, Z'=Z+gX, ZX'=aZX+eX2, Z2'=Z2+zZX+hX2, dZ2'=dZ2+thX2
// are closed under composition, so the scan over samples is a prefix scan
// of 7-coefficient maps in log space.
// ---------------------------------------------------------------------------
struct EvMap {
  double a, b, g, e, z, h, th;
};

struct EvState {
  double L, X, X2, Z, ZX, Z2, dZ, dZ2;
};

__device__ __forceinline__ EvMap evmap_identity() {
  const double ninf = -__builtin_inf();
  return {0.0, 0.0, ninf, ninf, ninf, ninf, ninf};
}

// the single-step map of sample (n, logL) preceded by log-likelihood Lprev; the
// per-factor expressions follow shrinkage_statistics.py:47-72.
// the coefficients of one step that depend on the live-point count n alone (T, T2, t, t2, tT)
struct EvCoef {
  double n, T, T2, t, t2, tT;
};
__device__ __forceinline__ EvCoef evcoef(double n) {
  const double LOG2 = 0.69314718055994530942;
  const double log_n = log(n);
  const double log_n1 = log(n + 1.0);
  const double log_n2 = log(n + 2.0);
  EvCoef c;
  c.n = n;
  c.T = -lae(0.0, -log_n);
  c.t = -log_n1;
  c.T2 = -lae(0.0, LOG2 - log_n);
  c.t2 = LOG2 - log_n1 - log_n2;
  c.tT = c.T - log_n2;
  return c;
}
// the coefficients of n, formed again only when n differs from the cached step's (runs of equal n
// are the rule: a shell's discarded samples all have n = N); the values are those of evcoef(n)
__device__ __forceinline__ const EvCoef& evcoef_cached(EvCoef& c, double n) {
  if (!(c.n == n)) c = evcoef(n);
  return c;
}
__device__ __forceinline__ EvCoef evcoef_empty() {
  EvCoef c;
  c.n = __builtin_nan("");
  c.T = c.T2 = c.t = c.t2 = c.tT = 0.0;
  return c;
}

__device__ __forceinline__ EvMap evmap_step_c(const EvCoef& c, double Lprev, double L, double* log_tm) {
  const double LOG2 = 0.69314718055994530942;
  double midL = -LOG2 + lae(L, Lprev);
  EvMap m;
  m.a = c.T;
  m.b = c.T2;
  m.g = c.t + midL;
  m.e = c.tT + midL;
  m.z = LOG2 + c.t + midL;
  m.h = c.t2 + 2.0 * midL;
  m.th = m.h;
  *log_tm = m.g;
  return m;
}

__device__ __forceinline__ EvMap evmap_step(double n, double Lprev, double L, double* log_tm) {
  return evmap_step_c(evcoef(n), Lprev, L, log_tm);
}

// f then g
__device__ __forceinline__ EvMap evmap_then(const EvMap& f, const EvMap& g) {
  EvMap r;
  r.a = g.a + f.a;
  r.b = g.b + f.b;
  r.g = lae(f.g, g.g + f.a);
  r.e = lae(g.a + f.e, g.e + f.b);
  r.z = lae(f.z, g.z + f.a);
  r.h = lae(lae(f.h, g.z + f.e), g.h + f.b);
  r.th = lae(f.th, g.th + f.b);
  return r;
}

__device__ __forceinline__ EvState evmap_apply(const EvMap& f, const EvState& s) {
  EvState o;
  o.L = s.L;
  o.X = f.a + s.X;
  o.X2 = f.b + s.X2;
  o.Z = lae(s.Z, f.g + s.X);
  o.ZX = lae(f.a + s.ZX, f.e + s.X2);
  o.Z2 = lae(lae(s.Z2, f.z + s.ZX), f.h + s.X2);
  o.dZ = s.dZ;
  o.dZ2 = lae(s.dZ2, f.th + s.X2);
  return o;
}

// one sequential recurrence step (the reference op itself); returns the new state
__device__ __forceinline__ EvState ev_step_c(const EvState& s, const EvCoef& c, double L) {
  double log_tm;
  EvMap f = evmap_step_c(c, s.L, L, &log_tm);
  EvState o = evmap_apply(f, s);
  o.L = L;
  o.dZ = log_tm + s.X;
  return o;
}
__device__ __forceinline__ EvState ev_step(const EvState& s, double n, double L) {
  double log_tm;
  EvMap f = evmap_step(n, s.L, L, &log_tm);
  EvState o = evmap_apply(f, s);
  o.L = L;
  o.dZ = log_tm + s.X;
  return o;
}

__device__ __forceinline__ EvState evstate_from(const double* s8) {
  return {s8[0], s8[1], s8[2], s8[3], s8[4], s8[5], s8[6], s8[7]};
}
__device__ __forceinline__ void evstate_to(const EvState& s, double* s8) {
  s8[0] = s.L;
  s8[1] = s.X;
  s8[2] = s.X2;
  s8[3] = s.Z;
  s8[4] = s.ZX;
  s8[5] = s.Z2;
  s8[6] = s.dZ;
  s8[7] = s.dZ2;
}

__device__ __forceinline__ EvMap shfl_xor_map(const EvMap& m, int o) {
  return {__shfl_xor(m.a, o, 64), __shfl_xor(m.b, o, 64), __shfl_xor(m.g, o, 64), __shfl_xor(m.e, o, 64),
          __shfl_xor(m.z, o, 64), __shfl_xor(m.h, o, 64), __shfl_xor(m.th, o, 64)};
}
__device__ __forceinline__ EvMap shfl_down_map(const EvMap& m, int o) {
  return {__shfl_down(m.a, o, 64), __shfl_down(m.b, o, 64), __shfl_down(m.g, o, 64), __shfl_down(m.e, o, 64),
          __shfl_down(m.z, o, 64), __shfl_down(m.h, o, 64), __shfl_down(m.th, o, 64)};
}
__device__ __forceinline__ EvMap shfl_up_map(const EvMap& m, int o) {
  return {__shfl_up(m.a, o, 64), __shfl_up(m.b, o, 64), __shfl_up(m.g, o, 64), __shfl_up(m.e, o, 64),
          __shfl_up(m.z, o, 64), __shfl_up(m.h, o, 64), __shfl_up(m.th, o, 64)};
}

// linear_to_log_stats (internals/stats.py:55-74)
__device__ __forceinline__ void linear_to_log_stats(double log_f, double log_f2, double* mu, double* var) {
  double fsq = log_f + log_f;
  *mu = fsq - 0.5 * log_f2;
  double s2 = log_f2 - fsq;
  const double eps = 2.220446049250313e-16;
  *var = __builtin_isnan(s2) ? s2 : fmax(s2, eps);
}

// determine_termination for one TerminationCondition (termination.py:67-147)
__device__ __forceinline__ int determine_termination(const hipns_term_cond_t& c, const hipns_run_state_t& r,
                                                     int64_t* reason) {
  int done = 0;
  int64_t why = 0;
  auto set = [&](bool cond, int bit) {
    if (cond) {
      done = 1;
      why += (int64_t)1 << bit;
    }
  };
  if (c.has_max_samples) set((double)r.num_samples_used >= c.max_samples, 0);
  if (c.has_evidence_uncert) {
    double mu, var;
    linear_to_log_stats(r.evidence_calc_with_remaining[3], r.evidence_calc_with_remaining[5], &mu, &var);
    set(var <= c.evidence_uncert * c.evidence_uncert, 1);
  }
  if (c.has_dlogZ) {
    double mu1, v1, mu0, v0;
    linear_to_log_stats(r.evidence_calc_with_remaining[3], r.evidence_calc_with_remaining[5], &mu1, &v1);
    linear_to_log_stats(r.evidence_calc[3], r.evidence_calc[5], &mu0, &v0);
    set((mu1 - mu0) < c.dlogZ, 2);
  }
  if (c.has_ess) {
    double ess = exp((r.evidence_calc_with_remaining[3] + r.evidence_calc_with_remaining[3]) -
                     r.evidence_calc_with_remaining[7]);
    set(ess >= c.ess, 3);
  }
  if (c.has_max_num_likelihood_evaluations)
    set((double)r.num_likelihood_evaluations >= c.max_num_likelihood_evaluations, 4);
  if (c.has_log_L_contour) set(r.log_L_contour >= c.log_L_contour, 5);
  if (c.has_efficiency_threshold) set(r.efficiency < c.efficiency_threshold, 6);
  set(r.plateau != 0, 7);
  if (c.has_rtol) set(r.relative_spread < c.rtol, 8);
  if (c.has_atol) set(r.absolute_spread < c.atol, 9);
  set(r.no_seed_points != 0, 10);
  if (c.has_peak_XL_frac) {
    double log_XL = r.evidence_calc[1] + r.evidence_calc[0];
    set(log_XL < r.peak_log_XL + log(c.peak_XL_frac), 11);
  }
  *reason = why;
  return done;
}

// First index of `count` entries appended to a sample collection of `size` entries at `next`
// (_add_samples_to_state, sharded_static.py:40-85, through replace_index, internals/maps.py:15-25):
// lax.dynamic_update_slice clamps its start so that the update stays in bounds, so an append that
// would run past the end is shifted back to end at `size` (it overwrites the entries before it) rather
// than wrapped.  `next` itself wraps modulo `size` afterwards.
__host__ __device__ __forceinline__ int64_t append_start(int64_t next, int64_t count, int64_t size) {
  const int64_t last = size - count;
  return next < last ? next : (last > 0 ? last : 0);
}

// key order of a stable FP64 sort: (key, index) lexicographic with NaN last
__device__ __forceinline__ bool key_lt(double ka, int ia, double kb, int ib) {
  bool na = __builtin_isnan(ka), nb = __builtin_isnan(kb);
  if (na || nb) {
    if (na && nb) return ia < ib;
    return nb;  // a < NaN
  }
  return (ka < kb) || (ka == kb && ia < ib);
}
// strict key order ignoring indices (NaN last)
__device__ __forceinline__ bool kless(double a, double b) {
  if (__builtin_isnan(a)) return false;
  if (__builtin_isnan(b)) return true;
  return a < b;
}

}  // namespace hipns
