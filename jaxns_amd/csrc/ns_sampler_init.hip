// ns_sampler_init.hip -- the sampling kernel's MODE_INIT instances: initial live points.
#include "ns_sampler_dispatch.h"

namespace hipns {

hipError_t launch_sampler_init(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                                size_t lds_bytes, hipStream_t s) {
  return launch_fam<MODE_INIT>(A, R, max_chains, L, dmax, lds_bytes, s);
}

}  // namespace hipns
