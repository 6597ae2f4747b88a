// ns_sampler_inj.hip -- instantiations of the sampling kernel for MODE_INJ.
#include "ns_sampler_impl.h"

namespace hipns {

// ---------------------------------------------------------------------------
// host-side dispatch over the template grid
// ---------------------------------------------------------------------------
template <int LPC, int NJ, int FAM, int MODE>
static hipError_t launch_one(const SampleArgs& A, int R, int max_chains, int wpb, size_t lds_bytes, hipStream_t s) {
  constexpr int CW = HIPNS_WAVE / LPC;
  int groups = (max_chains + CW - 1) / CW;
  int bx = (FAM == HIPNS_FAM_LINREG) ? groups : (groups + wpb - 1) / wpb;
  if (bx < 1) bx = 1;
  dim3 grid(bx, R), block(wpb * HIPNS_WAVE);
  hipLaunchKernelGGL((k_sample<LPC, NJ, FAM, MODE>), grid, block, lds_bytes, s, A);
  return hipGetLastError();
}

template <int LPC, int FAM, int MODE>
static hipError_t launch_dp(const SampleArgs& A, int R, int max_chains, int wpb, int dpmax, size_t lds, hipStream_t s) {
  if (dpmax <= 32) return launch_one<LPC, 32 / LPC, FAM, MODE>(A, R, max_chains, wpb, lds, s);
  if (dpmax <= 64) return launch_one<LPC, 64 / LPC, FAM, MODE>(A, R, max_chains, wpb, lds, s);
  return launch_one<LPC, 128 / LPC, FAM, MODE>(A, R, max_chains, wpb, lds, s);
}

template <int FAM, int MODE>
static hipError_t launch_lpc(const SampleArgs& A, int R, int max_chains, int lpc, int wpb, int dpmax, size_t lds,
                             hipStream_t s) {
  if (lpc == 4 || dpmax > 64) return launch_dp<4, FAM, MODE>(A, R, max_chains, wpb, dpmax, lds, s);
  if (dpmax <= 32) return launch_one<2, 16, FAM, MODE>(A, R, max_chains, wpb, lds, s);
  return launch_one<2, 32, FAM, MODE>(A, R, max_chains, wpb, lds, s);
}

template <int MODE>
static hipError_t launch_fam(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dpmax,
                             size_t lds, hipStream_t s) {
  if (L.family == HIPNS_FAM_GMM)
    return launch_lpc<HIPNS_FAM_GMM, MODE>(A, R, max_chains, L.lanes_per_chain, L.waves_per_block, dpmax, lds, s);
  if (L.family == HIPNS_FAM_LINREG)
    return launch_lpc<HIPNS_FAM_LINREG, MODE>(A, R, max_chains, L.lanes_per_chain, L.waves_per_block, dpmax, lds, s);
  return launch_lpc<HIPNS_FAM_MVN, MODE>(A, R, max_chains, L.lanes_per_chain, L.waves_per_block, dpmax, lds, s);
}

hipError_t launch_sampler_inj(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dpmax,
                          size_t lds_bytes, hipStream_t s) {
  return launch_fam<MODE_INJ>(A, R, max_chains, L, dpmax, lds_bytes, s);
}

}  // namespace hipns
