// ns_sampler_inj.hip -- the sampling kernel's MODE_INJ instances: slice chains with injected draws (parity hook).
#include "ns_sampler_dispatch.h"

namespace hipns {

hipError_t launch_sampler_inj(const SampleArgs& A, int R, int max_chains, const hipns_launch_t& L, int dmax,
                                size_t lds_bytes, hipStream_t s) {
  return launch_fam<MODE_INJ>(A, R, max_chains, L, dmax, lds_bytes, s);
}

}  // namespace hipns

// ---------------------------------------------------------------------------
// parity hooks on the draws of the fused kernel
// ---------------------------------------------------------------------------
namespace hipns {

// the seed index (and its uniform) that chain i of a shell draws in iteration `it32`
__global__ void k_seed_indices(uint32_t k0, uint32_t k1, uint32_t it32, int first, int N, int C, int32_t* idx,
                               double* u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  idx[i] = seed_index_dev(i, it32, k0, k1, first, N);
  const u32x4 w = philox4x32((uint32_t)i, it32, 0xFFFFFFFFu, RNG_SEED, k0, k1);
  u[i] = u01d(w.x, w.y);
}

// the direction of slice si (grid.y) of chains 0..C-1, computed in the lane layout of the fused kernel
// (LPC lanes per chain, NJ owned slots per lane); every lane of a wavefront takes part in the
// cross-lane sums, chains past C only skip the store
template <int LPC, int NJ>
__global__ void __launch_bounds__(256) k_slice_dirs(uint32_t k0, uint32_t k1, uint32_t it32, int C, int D,
                                                    float* out) {
  constexpr int CW = HIPNS_WAVE / LPC;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / HIPNS_WAVE);
  const int l = lane_id(), c = l % CW, h = l / CW;
  const int chain = wave * CW + c;
  const int si = blockIdx.y;
  const int nj = ((D + 8 * LPC - 1) / (8 * LPC)) * 8;
  float nv[NJ];
  slice_direction<LPC, NJ>(nv, chain, it32, si, D, nj, h, k0, k1);
  if (chain >= C) return;
  float* o = out + ((int64_t)chain * gridDim.y + si) * D;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int d = owned_dim<LPC>(h, j);
    if (j < nj && d < D) o[d] = nv[j];
  }
}

hipError_t launch_seed_indices(uint32_t k0, uint32_t k1, uint32_t it32, int first, int N, int C, int32_t* idx,
                               double* u, hipStream_t s) {
  hipLaunchKernelGGL(k_seed_indices, dim3((C + 255) / 256), dim3(256), 0, s, k0, k1, it32, first, N, C, idx, u);
  return hipGetLastError();
}

template <int LPC, int NJ>
static hipError_t launch_dirs_one(uint32_t k0, uint32_t k1, uint32_t it32, int C, int num_slices, int D, float* out,
                                  hipStream_t s) {
  constexpr int CW = HIPNS_WAVE / LPC;
  const int waves = (C + CW - 1) / CW;
  dim3 grid((waves + 3) / 4, num_slices), block(256);
  hipLaunchKernelGGL((k_slice_dirs<LPC, NJ>), grid, block, 0, s, k0, k1, it32, C, D, out);
  return hipGetLastError();
}

hipError_t launch_slice_dirs(uint32_t k0, uint32_t k1, uint32_t it32, int C, int num_slices, int D, int lpc,
                             float* out, hipStream_t s) {
  // the fused kernel's instance for this dimension (launch_lpc: beyond 64 dimensions, 4 lanes per chain)
  if (D > 64) lpc = 4;
  const int nj = sampler_nj(D, lpc);
  if (lpc == 4) {
    if (nj == 8) return launch_dirs_one<4, 8>(k0, k1, it32, C, num_slices, D, out, s);
    if (nj == 16) return launch_dirs_one<4, 16>(k0, k1, it32, C, num_slices, D, out, s);
    return launch_dirs_one<4, 32>(k0, k1, it32, C, num_slices, D, out, s);
  }
  if (nj == 8) return launch_dirs_one<2, 8>(k0, k1, it32, C, num_slices, D, out, s);
  if (nj == 16) return launch_dirs_one<2, 16>(k0, k1, it32, C, num_slices, D, out, s);
  return launch_dirs_one<2, 32>(k0, k1, it32, C, num_slices, D, out, s);
}

}  // namespace hipns
