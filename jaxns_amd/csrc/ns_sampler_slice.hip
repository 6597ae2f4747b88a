// ns_sampler_slice.hip -- the sampling kernel's MODE_SLICE instances: UniDimSliceSampler chains of a shell.
#include "ns_sampler_dispatch.h"

namespace hipns {

hipError_t launch_sampler_slice(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                                size_t lds_bytes, hipStream_t s) {
  return launch_fam<MODE_SLICE>(A, R, max_chains, L, dmax, lds_bytes, s);
}

}  // namespace hipns
