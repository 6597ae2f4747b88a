// ns_results.hip -- to_results primitives: the evidence scan over all samples,
// a stable FP64 argsort, and count_crossed_edges.
#include <utility>

#include "hipns_dev.h"
#include "ns_shell.h"

namespace hipns {

// ===========================================================================
// Evidence scan: compute_evidence_stats (internals/shrinkage_statistics.py:131-157)
// as a single-pass decoupled look-back scan over 7-coefficient log-space maps.
// Each 256-thread block owns a tile of 256*ITEMS consecutive samples; tiles
// are claimed in order from an atomic counter so that look-back only waits on
// tiles that are already running.
// ===========================================================================
constexpr int EV_THREADS = 256;
constexpr int EV_ITEMS = 8;
constexpr int EV_TILE = EV_THREADS * EV_ITEMS;

struct TileStatus {
  EvMap agg;
  EvMap incl;
  int flag;  // 0 = empty, 1 = aggregate published, 2 = inclusive prefix published
  int pad;
};

__device__ __forceinline__ double ld_relaxed(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ EvMap load_map(const EvMap* m) {
  const double* p = (const double*)m;
  return {ld_relaxed(p + 0), ld_relaxed(p + 1), ld_relaxed(p + 2), ld_relaxed(p + 3),
          ld_relaxed(p + 4), ld_relaxed(p + 5), ld_relaxed(p + 6)};
}
__device__ __forceinline__ void store_map(EvMap* m, const EvMap& v) {
  double* p = (double*)m;
  st_relaxed(p + 0, v.a);
  st_relaxed(p + 1, v.b);
  st_relaxed(p + 2, v.g);
  st_relaxed(p + 3, v.e);
  st_relaxed(p + 4, v.z);
  st_relaxed(p + 5, v.h);
  st_relaxed(p + 6, v.th);
}

__global__ void k_fill_zero_u32(uint32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// sources of the scan: (logL, n) at sorted position i
struct EvSrcFlat {
  const double* logL;
  const double* n;
  __device__ double L(int64_t i) const { return logL[i]; }
  __device__ double N(int64_t i) const { return n[i]; }
};
// sorted through a node permutation: position i holds sample perm[i + 1] - 1 (node 0 is the root)
struct EvSrcSorted {
  const double* logL;
  const int32_t* perm;
  const int64_t* nl;
  __device__ double L(int64_t i) const { return logL[perm[i + 1] - 1]; }
  __device__ double N(int64_t i) const { return (double)nl[i]; }
};

template <class Src>
__device__ __forceinline__ void evidence_scan_tile(const Src& src, int64_t M, int64_t stop, double* per8, double* fin8,
                                                   TileStatus* status, uint32_t* tile_counter) {
  __shared__ EvMap wave_aggs[EV_THREADS / 64];
  __shared__ EvMap s_excl;
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  if (tile * EV_TILE >= M) return;  // block-uniform: no tile left for this block
  const int tid = threadIdx.x;
  const int l = lane_id(), w = tid / 64;
  const int64_t base = tile * EV_TILE + (int64_t)tid * EV_ITEMS;
  const double ninf = -__builtin_inf();

  // thread aggregate of its ITEMS consecutive elements
  EvMap m = evmap_identity();
  EvMap item[EV_ITEMS];
  double item_lt[EV_ITEMS];
  double item_L[EV_ITEMS];
#pragma unroll
  for (int k = 0; k < EV_ITEMS; ++k) {
    int64_t i = base + k;
    EvMap f = evmap_identity();
    double lt = ninf, Li = ninf;
    if (i < M && i < stop) {
      double Lp = i == 0 ? ninf : src.L(i - 1);
      Li = src.L(i);
      f = evmap_step(src.N(i), Lp, Li, &lt);
    }
    item[k] = f;
    item_lt[k] = lt;
    item_L[k] = Li;
    m = evmap_then(m, f);
  }
  // inclusive scan across the wavefront (ordered)
  EvMap incl = m;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    EvMap up = shfl_up_map(incl, o);
    if (l >= o) incl = evmap_then(up, incl);
  }
  if (l == 63) wave_aggs[w] = incl;
  __syncthreads();
  // exclusive prefix of this thread within the tile
  EvMap wave_prefix = evmap_identity();
  for (int i = 0; i < w; ++i) wave_prefix = evmap_then(wave_prefix, wave_aggs[i]);
  EvMap lane_excl = shfl_up_map(incl, 1);
  if (l == 0) lane_excl = evmap_identity();
  EvMap thr_excl = evmap_then(wave_prefix, lane_excl);

  // tile aggregate and decoupled look-back (one thread)
  if (tid == 0) {
    EvMap tile_agg = evmap_identity();
    for (int i = 0; i < EV_THREADS / 64; ++i) tile_agg = evmap_then(tile_agg, wave_aggs[i]);
    EvMap excl = evmap_identity();
    if (tile == 0) {
      store_map(&status[0].incl, tile_agg);
      __hip_atomic_store(&status[0].flag, 2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      store_map(&status[tile].agg, tile_agg);
      __hip_atomic_store(&status[tile].flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      EvMap run = evmap_identity();
      int64_t t = tile - 1;
      while (true) {
        int f;
        do {
          f = __hip_atomic_load(&status[t].flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        } while (f == 0);
        if (f == 2) {
          run = evmap_then(load_map(&status[t].incl), run);
          break;
        }
        run = evmap_then(load_map(&status[t].agg), run);
        --t;
      }
      excl = run;
      store_map(&status[tile].incl, evmap_then(excl, tile_agg));
      __hip_atomic_store(&status[tile].flag, 2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_excl = excl;
  }
  __syncthreads();
  const EvMap pre = evmap_then(s_excl, thr_excl);
  // state before this thread's first element
  EvState s;
  const EvState s_init = {ninf, 0.0, 0.0, ninf, ninf, ninf, ninf, ninf};
  s = evmap_apply(pre, s_init);
  s.L = (base == 0) ? ninf : (base - 1 < stop ? src.L(base - 1) : ninf);
  s.dZ = ninf;
#pragma unroll
  for (int k = 0; k < EV_ITEMS; ++k) {
    int64_t i = base + k;
    if (i < M) {
      EvState o;
      if (i < stop) {
        o = evmap_apply(item[k], s);
        o.L = item_L[k];
        o.dZ = item_lt[k] + s.X;
        s = o;
      } else {
        o = s_init;
      }
      if (per8) {
        per8[0 * M + i] = o.L;
        per8[1 * M + i] = o.X;
        per8[2 * M + i] = o.X2;
        per8[3 * M + i] = o.Z;
        per8[4 * M + i] = o.ZX;
        per8[5 * M + i] = o.Z2;
        per8[6 * M + i] = o.dZ;
        per8[7 * M + i] = o.dZ2;
      }
      if (fin8 && i == (stop < M ? stop : M) - 1) {
        fin8[0] = o.L;
        fin8[1] = o.X;
        fin8[2] = o.X2;
        fin8[3] = o.Z;
        fin8[4] = o.ZX;
        fin8[5] = o.Z2;
        fin8[6] = o.dZ;
        fin8[7] = o.dZ2;
      }
    }
  }
}

__global__ void __launch_bounds__(EV_THREADS) k_evidence_scan(const double* logL, const double* nlive, int64_t M,
                                                              int64_t stop, double* per8, double* fin8,
                                                              TileStatus* status, uint32_t* tile_counter) {
  evidence_scan_tile(EvSrcFlat{logL, nlive}, M, stop, per8, fin8, status, tile_counter);
}

__global__ void k_fin_init(double* fin8) {
  const double ninf = -__builtin_inf();
  double v[8] = {ninf, 0.0, 0.0, ninf, ninf, ninf, ninf, ninf};
  for (int i = 0; i < 8; ++i) fin8[i] = v[i];
}

size_t evidence_scan_ws(int64_t M) {
  int64_t tiles = (M + EV_TILE - 1) / EV_TILE;
  if (tiles < 1) tiles = 1;
  return 256 + (size_t)tiles * sizeof(TileStatus);
}

hipError_t evidence_scan(const double* logL, const double* n, int64_t M, int64_t stop, double* per8, double* fin8,
                         void* ws, size_t ws_bytes, hipStream_t s) {
  if (stop < 0 || stop > M) stop = M;
  int64_t tiles = (M + EV_TILE - 1) / EV_TILE;
  if (fin8 && (stop == 0 || M == 0)) hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(1), 0, s, fin8);
  if (M == 0) return hipGetLastError();
  uint32_t* counter = (uint32_t*)ws;
  TileStatus* status = (TileStatus*)((char*)ws + 256);
  int64_t words = (256 + tiles * (int64_t)sizeof(TileStatus)) / 4;
  hipLaunchKernelGGL(k_fill_zero_u32, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, (uint32_t*)ws, words);
  hipLaunchKernelGGL(k_evidence_scan, dim3((unsigned)tiles), dim3(EV_THREADS), 0, s, logL, n, M, stop, per8,
                     stop > 0 ? fin8 : nullptr, status, counter);
  return hipGetLastError();
}

// ===========================================================================
// Stable argsort of FP64 keys: bitonic sort of 2048-element tiles in LDS by
// (key, index), then merge passes where each element finds its output slot by
// binary search in the partner run (ties: left run first).
// ===========================================================================
constexpr int SORT_THREADS = 256;
constexpr int SORT_TILE = 2048;

__global__ void __launch_bounds__(SORT_THREADS) k_tile_sort(const double* keys_in, int64_t n, double* keys_out,
                                                            int32_t* idx_out) {
  __shared__ double k[SORT_TILE];
  __shared__ int32_t v[SORT_TILE];
  const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
  for (int i = threadIdx.x; i < SORT_TILE; i += SORT_THREADS) {
    int64_t g = base + i;
    k[i] = g < n ? keys_in[g] : __builtin_nan("");
    v[i] = g < n ? (int32_t)g : (int32_t)0x7fffffff;
  }
  __syncthreads();
  for (int kk = 2; kk <= SORT_TILE; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < SORT_TILE; i += SORT_THREADS) {
        int p = i ^ j;
        if (p > i) {
          bool up = (i & kk) == 0;
          double ki = k[i], kp = k[p];
          int32_t ii = v[i], ip = v[p];
          bool sw = up ? key_lt(kp, ip, ki, ii) : key_lt(ki, ii, kp, ip);
          if (sw) {
            k[i] = kp;
            k[p] = ki;
            v[i] = ip;
            v[p] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < SORT_TILE; i += SORT_THREADS) {
    int64_t g = base + i;
    if (g < n) {
      keys_out[g] = k[i];
      idx_out[g] = v[i];
    }
  }
}

// merge runs of length `run` pairwise: element of the left run goes before equal keys of the right run
__global__ void k_merge_pass(const double* kin, const int32_t* vin, int64_t n, int64_t run, double* kout,
                             int32_t* vout) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  int64_t pair = g / (2 * run);
  int64_t lbeg = pair * 2 * run;
  int64_t rbeg = lbeg + run;
  int64_t lend = rbeg < n ? rbeg : n;
  int64_t rend = rbeg + run < n ? rbeg + run : n;
  double key = kin[g];
  int64_t pos;
  if (g < rbeg) {
    // left element: count right elements strictly less
    int64_t lo = rbeg, hi = rend;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (kless(kin[mid], key)) lo = mid + 1;
      else hi = mid;
    }
    pos = lbeg + (g - lbeg) + (lo - rbeg);
  } else {
    // right element: count left elements less or equal
    int64_t lo = lbeg, hi = lend;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (!kless(key, kin[mid])) lo = mid + 1;
      else hi = mid;
    }
    pos = lbeg + (g - rbeg) + (lo - lbeg);
  }
  kout[pos] = key;
  vout[pos] = vin[g];
}

size_t argsort_ws(int64_t n) { return (size_t)n * (2 * sizeof(double) + sizeof(int32_t)) + 512; }

// perm receives int32 indices; workspace holds two key buffers and one index buffer
hipError_t argsort_stable(const double* keys, int64_t n, int32_t* perm, void* ws, size_t ws_bytes, hipStream_t s) {
  if (n == 0) return hipSuccess;
  double* kA = (double*)ws;
  double* kB = kA + n;
  int32_t* vB = (int32_t*)(kB + n);
  int64_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
  // ping-pong between (kA, perm) and (kB, vB); arrange the final pass to land in perm
  int passes = 0;
  for (int64_t r = SORT_TILE; r < n; r <<= 1) ++passes;
  double* k0 = (passes % 2 == 0) ? kA : kB;
  int32_t* v0 = (passes % 2 == 0) ? perm : vB;
  double* k1 = (passes % 2 == 0) ? kB : kA;
  int32_t* v1 = (passes % 2 == 0) ? vB : perm;
  hipLaunchKernelGGL(k_tile_sort, dim3((unsigned)tiles), dim3(SORT_THREADS), 0, s, keys, n, k0, v0);
  for (int64_t r = SORT_TILE; r < n; r <<= 1) {
    hipLaunchKernelGGL(k_merge_pass, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, k0, v0, n, r, k1, v1);
    double* tk = k0;
    k0 = k1;
    k1 = tk;
    int32_t* tv = v0;
    v0 = v1;
    v1 = tv;
  }
  return hipGetLastError();
}

// ===========================================================================
// count_crossed_edges (internals/tree_structure.py:33-108)
// ===========================================================================
__global__ void k_nodes(const double* logL, int64_t N, double* nodes) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > N) return;
  nodes[i] = i == 0 ? -__builtin_inf() : logL[i - 1];
}

__global__ void k_out_degree(const int32_t* sender, int64_t N, int32_t* deg) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int32_t s = sender[i] > 0 ? sender[i] : 0;  // lax.max(sender, 0)
  if (s <= N) atomicAdd(&deg[s], 1);           // out-of-range scatter updates are dropped
}

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int64_t crossed_term(const int32_t* deg, const int32_t* perm, int64_t i) {
  return (int64_t)deg[perm[i]] - 1;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const int32_t* deg, const int32_t* perm, int64_t n,
                                                             int64_t stop, int64_t* tile_sums) {
  __shared__ int64_t red[SCAN_THREADS / 64];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t acc = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + k;
    if (i < n && i < stop) acc += crossed_term(deg, perm, i);
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if (lane_id() == 0) red[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int i = 0; i < SCAN_THREADS / 64; ++i) t += red[i];
    tile_sums[blockIdx.x] = t;
  }
}

// exclusive scan of tile sums in one block (sequential chunks)
__global__ void k_scan_sums(int64_t* tile_sums, int64_t T) {
  __shared__ int64_t buf[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < T; c0 += 1024) {
    int64_t i = c0 + threadIdx.x;
    int64_t v = i < T ? tile_sums[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int64_t add = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    int64_t incl = buf[threadIdx.x];
    if (i < T) tile_sums[i] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const int32_t* deg, const int32_t* perm, int64_t n,
                                                             int64_t stop, int64_t fake, const int64_t* tile_sums,
                                                             int64_t* num_live, int64_t* samples_indices) {
  __shared__ int64_t tsum[SCAN_THREADS];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t local[SCAN_ITEMS];
  int64_t acc = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + k;
    int64_t v = (i < n && i < stop) ? crossed_term(deg, perm, i) : 0;
    acc += v;
    local[k] = acc;
  }
  tsum[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 1; o < SCAN_THREADS; o <<= 1) {
    int64_t add = threadIdx.x >= o ? tsum[threadIdx.x - o] : 0;
    __syncthreads();
    tsum[threadIdx.x] += add;
    __syncthreads();
  }
  int64_t pre = tile_sums[blockIdx.x] + tsum[threadIdx.x] - acc;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + k;
    if (i >= n) break;
    // crossed_edges_sorted[i]; the last node (i = n-1 = N) is dropped
    int64_t crossed = (i < stop) ? (1 + pre + local[k]) : fake;
    crossed -= fake;
    if (i < n - 1) {
      num_live[i] = crossed;
      samples_indices[i] = (int64_t)perm[i + 1] - 1;
    }
  }
}

size_t count_crossed_edges_ws(int64_t N) {
  int64_t n = N + 1;
  int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  return (size_t)n * sizeof(double) + (size_t)n * sizeof(int32_t) * 2 + (size_t)tiles * sizeof(int64_t) +
         argsort_ws(n) + 1024;
}

hipError_t count_crossed_edges(const int32_t* sender, const double* logL, int64_t N, int64_t num_samples,
                               int64_t* samples_indices, int64_t* num_live, void* ws, size_t ws_bytes, hipStream_t s) {
  const int64_t n = N + 1;
  char* p = (char*)ws;
  double* nodes = (double*)p;
  p += n * sizeof(double);
  int32_t* perm = (int32_t*)p;
  p += n * sizeof(int32_t);
  int32_t* deg = (int32_t*)p;
  p += n * sizeof(int32_t);
  p = (char*)(((uintptr_t)p + 15) & ~(uintptr_t)15);
  int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  int64_t* tsums = (int64_t*)p;
  p += tiles * sizeof(int64_t);
  p = (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
  void* sort_ws = p;
  hipLaunchKernelGGL(k_nodes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, logL, N, nodes);
  hipError_t e = argsort_stable(nodes, n, perm, sort_ws, argsort_ws(n), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (uint32_t*)deg, n);
  if (N > 0)
    hipLaunchKernelGGL(k_out_degree, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, sender, N, deg);
  const int64_t stop = num_samples < 0 ? n : num_samples;
  const int64_t fake = num_samples < 0 ? 0 : (N - num_samples);
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, deg, perm, n, stop, tsums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, tsums, tiles);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, deg, perm, n, stop, fake, tsums,
                     num_live, samples_indices);
  return hipGetLastError();
}

// ===========================================================================
// _to_results evidence of R runs at once: count_crossed_edges and the evidence
// scan of every run of a batch in one set of launches (grid.y = run).  Run r's
// samples are samp_*[samp_off_r, samp_off_r + M_r) with M_r = min(num_samples,
// max_samples) read on the device; per-run segments of the workspace have a
// common stride S >= max_r(M_r) + 1.
// ===========================================================================
__device__ __forceinline__ int64_t batch_M(const hipns_run_desc_t* runs, const hipns_run_state_t* states, int r) {
  const int64_t ns = states[r].num_samples, ms = runs[r].max_samples;
  return ns < ms ? ns : ms;
}

__global__ void k_b_nodes(const hipns_run_desc_t* runs, const hipns_run_state_t* states, const double* samp_logL,
                          int64_t S, double* nodes) {
  const int r = blockIdx.y;
  const int64_t M = batch_M(runs, states, r);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > M) return;
  nodes[r * S + i] = i == 0 ? -__builtin_inf() : samp_logL[runs[r].samp_off + i - 1];
}

__global__ void __launch_bounds__(SORT_THREADS) k_b_tile_sort(const hipns_run_desc_t* runs,
                                                              const hipns_run_state_t* states, const double* keys_in,
                                                              int64_t S, double* keys_out, int32_t* idx_out) {
  __shared__ double k[SORT_TILE];
  __shared__ int32_t v[SORT_TILE];
  const int r = blockIdx.y;
  const int64_t n = batch_M(runs, states, r) + 1;
  const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
  if (base >= n) return;  // block-uniform
  const double* kin = keys_in + r * S;
  for (int i = threadIdx.x; i < SORT_TILE; i += SORT_THREADS) {
    int64_t g = base + i;
    k[i] = g < n ? kin[g] : __builtin_nan("");
    v[i] = g < n ? (int32_t)g : (int32_t)0x7fffffff;
  }
  __syncthreads();
  for (int kk = 2; kk <= SORT_TILE; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < SORT_TILE; i += SORT_THREADS) {
        int p = i ^ j;
        if (p > i) {
          bool up = (i & kk) == 0;
          double ki = k[i], kp = k[p];
          int32_t ii = v[i], ip = v[p];
          bool sw = up ? key_lt(kp, ip, ki, ii) : key_lt(ki, ii, kp, ip);
          if (sw) {
            k[i] = kp;
            k[p] = ki;
            v[i] = ip;
            v[p] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < SORT_TILE; i += SORT_THREADS) {
    int64_t g = base + i;
    if (g < n) {
      keys_out[r * S + g] = k[i];
      idx_out[r * S + g] = v[i];
    }
  }
}

__global__ void k_b_merge_pass(const hipns_run_desc_t* runs, const hipns_run_state_t* states, const double* kin_all,
                               const int32_t* vin_all, int64_t S, int64_t run, double* kout_all, int32_t* vout_all) {
  const int r = blockIdx.y;
  const int64_t n = batch_M(runs, states, r) + 1;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double* kin = kin_all + r * S;
  int64_t pair = g / (2 * run);
  int64_t lbeg = pair * 2 * run;
  int64_t rbeg = lbeg + run;
  int64_t lend = rbeg < n ? rbeg : n;
  int64_t rend = rbeg + run < n ? rbeg + run : n;
  double key = kin[g];
  int64_t pos;
  if (g < rbeg) {
    int64_t lo = rbeg, hi = rend;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (kless(kin[mid], key)) lo = mid + 1;
      else hi = mid;
    }
    pos = lbeg + (g - lbeg) + (lo - rbeg);
  } else {
    int64_t lo = lbeg, hi = lend;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (!kless(key, kin[mid])) lo = mid + 1;
      else hi = mid;
    }
    pos = lbeg + (g - rbeg) + (lo - lbeg);
  }
  kout_all[r * S + pos] = key;
  vout_all[r * S + pos] = vin_all[r * S + g];
}

// out-degree of every node by sender (tree_structure.py:33-108) and the phantom count of each run
__global__ void __launch_bounds__(256) k_b_out_degree(const hipns_run_desc_t* runs, const hipns_run_state_t* states,
                                                      hipns_buffers_t B, int64_t S, int32_t* deg,
                                                      unsigned long long* phantom) {
  __shared__ unsigned long long red[4];
  const int r = blockIdx.y;
  const int64_t M = batch_M(runs, states, r);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t off = runs[r].samp_off;
  unsigned long long ph = 0;
  if (i < M) {
    int32_t snd = B.samp_sender[off + i];
    int64_t sn = snd > 0 ? snd : 0;                // lax.max(sender, 0)
    if (sn <= M) atomicAdd(&deg[r * S + sn], 1);  // out-of-range scatter updates are dropped
    ph = B.samp_phantom[off + i] ? 1ull : 0ull;
  }
  for (int o = 32; o > 0; o >>= 1) ph += __shfl_down(ph, o, 64);
  if (lane_id() == 0) red[threadIdx.x / 64] = ph;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = red[0] + red[1] + red[2] + red[3];
    if (t) atomicAdd(&phantom[r], t);
  }
}

__global__ void __launch_bounds__(SCAN_THREADS) k_b_scan_tiles(const hipns_run_desc_t* runs,
                                                               const hipns_run_state_t* states, const int32_t* deg,
                                                               const int32_t* perm, int64_t S, int64_t T,
                                                               int64_t* tile_sums) {
  __shared__ int64_t red[SCAN_THREADS / 64];
  const int r = blockIdx.y;
  const int64_t n = batch_M(runs, states, r) + 1;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t acc = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + k;
    if (i < n) acc += crossed_term(deg + r * S, perm + r * S, i);
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if (lane_id() == 0) red[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int i = 0; i < SCAN_THREADS / 64; ++i) t += red[i];
    tile_sums[r * T + blockIdx.x] = t;
  }
}

__global__ void k_b_scan_sums(int64_t* tile_sums_all, int64_t T) {
  __shared__ int64_t buf[1024];
  __shared__ int64_t carry;
  int64_t* tile_sums = tile_sums_all + (int64_t)blockIdx.y * T;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < T; c0 += 1024) {
    int64_t i = c0 + threadIdx.x;
    int64_t v = i < T ? tile_sums[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int64_t add = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    int64_t incl = buf[threadIdx.x];
    if (i < T) tile_sums[i] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
}

// num_live_points of sorted sample i (the static form: every node counts)
__global__ void __launch_bounds__(SCAN_THREADS) k_b_scan_apply(const hipns_run_desc_t* runs,
                                                               const hipns_run_state_t* states, const int32_t* deg,
                                                               const int32_t* perm, int64_t S, int64_t T,
                                                               const int64_t* tile_sums, int64_t* num_live) {
  __shared__ int64_t tsum[SCAN_THREADS];
  const int r = blockIdx.y;
  const int64_t n = batch_M(runs, states, r) + 1;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t local[SCAN_ITEMS];
  int64_t acc = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + k;
    int64_t v = i < n ? crossed_term(deg + r * S, perm + r * S, i) : 0;
    acc += v;
    local[k] = acc;
  }
  tsum[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 1; o < SCAN_THREADS; o <<= 1) {
    int64_t add = threadIdx.x >= o ? tsum[threadIdx.x - o] : 0;
    __syncthreads();
    tsum[threadIdx.x] += add;
    __syncthreads();
  }
  const int64_t pre = tile_sums[r * T + blockIdx.x] + tsum[threadIdx.x] - acc;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + k;
    if (i >= n - 1) break;  // the last node is dropped
    num_live[r * S + i] = 1 + pre + local[k];
  }
}

__global__ void __launch_bounds__(EV_THREADS) k_b_evidence_scan(const hipns_run_desc_t* runs,
                                                                const hipns_run_state_t* states, const double* samp_logL,
                                                                int64_t S, const int32_t* perm, const int64_t* num_live,
                                                                TileStatus* status, int64_t Tev, uint32_t* counters,
                                                                double* fin) {
  const int r = blockIdx.y;
  const int64_t M = batch_M(runs, states, r);
  double* fin8 = fin + 8 * r;
  if (M == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const double ninf = -__builtin_inf();
      const double v[8] = {ninf, 0.0, 0.0, ninf, ninf, ninf, ninf, ninf};
      for (int i = 0; i < 8; ++i) fin8[i] = v[i];
    }
    return;
  }
  const EvSrcSorted src{samp_logL + runs[r].samp_off, perm + r * S, num_live + r * S};
  evidence_scan_tile(src, M, M, nullptr, fin8, status + r * Tev, counters + r);
}

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct BatchEvLayout {
  int64_t S, Tscan, Tev;
  size_t zero_bytes, total;
  size_t off_status, off_counters, off_kA, off_kB, off_vA, off_vB, off_nl, off_ts;
};

static BatchEvLayout batch_ev_layout(int R, int64_t max_M) {
  BatchEvLayout L{};
  L.S = round_up(max_M + 1, SORT_TILE);
  L.Tscan = (L.S + SCAN_TILE - 1) / SCAN_TILE;
  L.Tev = (L.S + EV_TILE - 1) / EV_TILE;
  size_t o = 0;
  o += (size_t)R * L.S * sizeof(int32_t);  // deg
  o = round_up(o, 256);
  L.off_status = o;
  o += (size_t)R * L.Tev * sizeof(TileStatus);
  o = round_up(o, 256);
  L.off_counters = o;
  o += (size_t)R * sizeof(uint32_t);
  o = round_up(o, 256);
  L.zero_bytes = o;
  L.off_kA = o;
  o += (size_t)R * L.S * sizeof(double);
  L.off_kB = o;
  o += (size_t)R * L.S * sizeof(double);
  L.off_vA = o;
  o += (size_t)R * L.S * sizeof(int32_t);
  L.off_vB = o;
  o += (size_t)R * L.S * sizeof(int32_t);
  o = round_up(o, 256);
  L.off_nl = o;
  o += (size_t)R * L.S * sizeof(int64_t);
  L.off_ts = o;
  o += (size_t)R * L.Tscan * sizeof(int64_t);
  L.total = round_up(o, 256);
  return L;
}

size_t batch_evidence_ws(int R, int64_t max_M) { return batch_ev_layout(R, max_M).total; }

hipError_t batch_evidence(const hipns_run_desc_t* runs, const hipns_run_state_t* states, int R, int64_t max_M,
                          const hipns_buffers_t& B, double* fin, int64_t* phantom, void* ws, hipStream_t s) {
  const BatchEvLayout L = batch_ev_layout(R, max_M);
  char* base = (char*)ws;
  int32_t* deg = (int32_t*)base;
  TileStatus* status = (TileStatus*)(base + L.off_status);
  uint32_t* counters = (uint32_t*)(base + L.off_counters);
  double* kA = (double*)(base + L.off_kA);
  double* kB = (double*)(base + L.off_kB);
  int32_t* vA = (int32_t*)(base + L.off_vA);
  int32_t* vB = (int32_t*)(base + L.off_vB);
  int64_t* nl = (int64_t*)(base + L.off_nl);
  int64_t* ts = (int64_t*)(base + L.off_ts);
  const int64_t words = (int64_t)(L.zero_bytes / 4);
  hipLaunchKernelGGL(k_fill_zero_u32, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, (uint32_t*)base, words);
  hipLaunchKernelGGL(k_fill_zero_u32, dim3((unsigned)((2 * R + 255) / 256)), dim3(256), 0, s, (uint32_t*)phantom,
                     (int64_t)2 * R);
  // nodes into kA; sort: tile pass writes (k0, v0), merges ping-pong; the final buffers hold the permutation
  hipLaunchKernelGGL(k_b_nodes, dim3((unsigned)((L.S + 255) / 256), R), dim3(256), 0, s, runs, states, B.samp_logL,
                     L.S, kA);
  const int64_t tiles = L.S / SORT_TILE;
  hipLaunchKernelGGL(k_b_tile_sort, dim3((unsigned)tiles, R), dim3(SORT_THREADS), 0, s, runs, states, kA, L.S, kB, vB);
  double* k0 = kB;
  int32_t* v0 = vB;
  double* k1 = kA;
  int32_t* v1 = vA;
  for (int64_t run = SORT_TILE; run < L.S; run <<= 1) {
    hipLaunchKernelGGL(k_b_merge_pass, dim3((unsigned)((L.S + 255) / 256), R), dim3(256), 0, s, runs, states, k0, v0,
                       L.S, run, k1, v1);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  const int32_t* perm = v0;
  hipLaunchKernelGGL(k_b_out_degree, dim3((unsigned)((L.S + 255) / 256), R), dim3(256), 0, s, runs, states, B, L.S,
                     deg, (unsigned long long*)phantom);
  hipLaunchKernelGGL(k_b_scan_tiles, dim3((unsigned)L.Tscan, R), dim3(SCAN_THREADS), 0, s, runs, states, deg, perm,
                     L.S, L.Tscan, ts);
  hipLaunchKernelGGL(k_b_scan_sums, dim3(1, R), dim3(1024), 0, s, ts, L.Tscan);
  hipLaunchKernelGGL(k_b_scan_apply, dim3((unsigned)L.Tscan, R), dim3(SCAN_THREADS), 0, s, runs, states, deg, perm,
                     L.S, L.Tscan, ts, nl);
  hipLaunchKernelGGL(k_b_evidence_scan, dim3((unsigned)L.Tev, R), dim3(EV_THREADS), 0, s, runs, states, B.samp_logL,
                     L.S, perm, nl, status, L.Tev, counters, fin);
  return hipGetLastError();
}

}  // namespace hipns
