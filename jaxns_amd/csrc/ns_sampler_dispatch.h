// ns_sampler_dispatch.h -- host-side dispatch of the fused sampling kernel over its template grid
// (lanes per chain, register slots, likelihood family).  Each sampling mode instantiates the grid in
// a translation unit of its own (ns_sampler_{slice,unif,init,inj}.hip), so the four compile in parallel.
#pragma once
#include "ns_sampler_impl.h"

namespace hipns {

// ---------------------------------------------------------------------------
// host-side dispatch over the template grid
// ---------------------------------------------------------------------------
template <int LPC, int NJ, int FAM, int MODE>
inline hipError_t launch_one(const SampleArgs& A, int R, int max_chains, int wpb, size_t lds_bytes, hipStream_t s) {
  constexpr int CW = HIPNS_WAVE / LPC;
  int groups = (max_chains + CW - 1) / CW;
  int bx = (groups + wpb - 1) / wpb;
  if (bx < 1) bx = 1;
  dim3 grid(bx, R), block(wpb * HIPNS_WAVE);
  hipLaunchKernelGGL((k_sample<LPC, NJ, FAM, MODE>), grid, block, lds_bytes, s, A);
  return hipGetLastError();
}

// register slots per lane: the smallest instance whose NJ * LPC covers the widest run (D rounded up
// to 8 * LPC), which keeps VGPRs (and so occupancy) in step with the model dimension
template <int FAM, int MODE>
inline hipError_t launch_lpc(const SampleArgs& A, int R, int max_chains, int lpc, int wpb, int dmax, size_t lds,
                             hipStream_t s) {
  if constexpr (FAM == HIPNS_FAM_GMM) {
    // the mixture family also runs 8 lanes (8 slots) per chain: 8 chains per wavefront, up to 64 dims
    // in 8 slots, and about half the registers of the 16-slot instance of 4 lanes per chain
    if (lpc == 8 && dmax <= 128) {
      if (dmax <= 64) return launch_one<8, 8, FAM, MODE>(A, R, max_chains, wpb, lds, s);
      return launch_one<8, 16, FAM, MODE>(A, R, max_chains, wpb, lds, s);
    }
  }
  if (lpc == 4 || lpc == 8 || dmax > 64) {
    if (dmax <= 32) return launch_one<4, 8, FAM, MODE>(A, R, max_chains, wpb, lds, s);
    if (dmax <= 64) return launch_one<4, 16, FAM, MODE>(A, R, max_chains, wpb, lds, s);
    return launch_one<4, 32, FAM, MODE>(A, R, max_chains, wpb, lds, s);
  }
  if (dmax <= 16) return launch_one<2, 8, FAM, MODE>(A, R, max_chains, wpb, lds, s);
  if (dmax <= 32) return launch_one<2, 16, FAM, MODE>(A, R, max_chains, wpb, lds, s);
  return launch_one<2, 32, FAM, MODE>(A, R, max_chains, wpb, lds, s);
}

template <int MODE>
inline hipError_t launch_fam(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                             size_t lds, hipStream_t s) {
  if (L.family == HIPNS_FAM_GMM)
    return launch_lpc<HIPNS_FAM_GMM, MODE>(A, R, max_chains, L.lanes_per_chain, L.waves_per_block, dmax, lds, s);
  if (L.family == HIPNS_FAM_LINREG)
    return launch_lpc<HIPNS_FAM_LINREG, MODE>(A, R, max_chains, L.lanes_per_chain, L.waves_per_block, dmax, lds, s);
  return launch_lpc<HIPNS_FAM_MVN, MODE>(A, R, max_chains, L.lanes_per_chain, L.waves_per_block, dmax, lds, s);
}

}  // namespace hipns
