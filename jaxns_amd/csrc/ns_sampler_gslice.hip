// ns_sampler_gslice.hip -- the sampling kernel's MODE_GSLICE instances: gradient-guided UniDimSliceSampler
// chains of a shell, with d logL / dU computed inside the kernel (MVN and mixture families).
#include "ns_sampler_dispatch.h"

namespace hipns {

hipError_t launch_sampler_gslice(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                                 size_t lds_bytes, hipStream_t s) {
  if (L.family == HIPNS_FAM_GMM)
    return launch_lpc<HIPNS_FAM_GMM, MODE_GSLICE>(A, R, max_chains, L.lanes_per_chain, L.waves_per_block, dmax,
                                                  lds_bytes, s);
  if (L.family == HIPNS_FAM_MVN)
    return launch_lpc<HIPNS_FAM_MVN, MODE_GSLICE>(A, R, max_chains, L.lanes_per_chain, L.waves_per_block, dmax,
                                                  lds_bytes, s);
  return hipErrorInvalidValue;
}

}  // namespace hipns
