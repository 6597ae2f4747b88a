// ns_sampler_unif.hip -- the sampling kernel's MODE_UNIF instances: UniformSampler chains of a shell.
#include "ns_sampler_dispatch.h"

namespace hipns {

hipError_t launch_sampler_unif(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                                size_t lds_bytes, hipStream_t s) {
  return launch_fam<MODE_UNIF>(A, R, max_chains, L, dmax, lds_bytes, s);
}

}  // namespace hipns
