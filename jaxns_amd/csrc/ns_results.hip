// ns_results.hip -- to_results primitives: the evidence scan over all samples,
// a stable FP64 argsort, and count_crossed_edges.
#include <mutex>
#include <type_traits>
#include <utility>

#include "hipns_dev.h"
#include "ns_shell.h"

namespace hipns {

// ===========================================================================
// Evidence scan: compute_evidence_stats (internals/shrinkage_statistics.py:131-157)
// as a reduce-then-scan over 7-coefficient log-space maps, in three launches
// (grid.y = segment, so one set of launches serves every run of a batch):
//   1. k_ev_reduce: each 256-thread block reduces a tile of 256*ITEMS
//      consecutive samples (each thread composes its ITEMS step maps, then an
//      ordered scan across the block) and stores the tile's aggregate and
//      every thread's exclusive prefix within the tile;
//   2. k_ev_scan_tiles: one block per segment scans the tile aggregates
//      (exclusive, in place);
//   3. k_ev_apply: each thread applies its prefix to the initial state and
//      steps through its samples, forming each step map again.
// The grouping of the compositions depends only on the sample count, so a
// segment's results do not depend on the other segments or on timing.
// ===========================================================================
constexpr int EV_THREADS = 256;
constexpr int EV_ITEMS = 8;
constexpr int EV_TILE = EV_THREADS * EV_ITEMS;
constexpr int EV_SCAN_THREADS = 1024;

__global__ void k_fill_zero_u32(uint32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__device__ __forceinline__ int64_t batch_M(const hipns_run_desc_t* runs, const hipns_run_state_t* states, int r) {
  const int64_t ns = states[r].num_samples, ms = runs[r].max_samples;
  return ns < ms ? ns : ms;
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

// segments: one flat array (samples past `stop` keep the initial state), or the sorted samples of
// every run of a batch
struct EvSegFlat {
  const double* logL;
  const double* n;
  int64_t M, stop;
  __device__ int64_t len(int) const { return M; }
  __device__ int64_t stop_at(int) const { return stop; }
  __device__ EvSrcFlat src(int) const { return {logL, n}; }
};
struct EvSegBatch {
  const hipns_run_desc_t* runs;
  const hipns_run_state_t* states;
  const double* samp_logL;
  const int32_t* perm;
  const int64_t* num_live;
  int64_t S;
  __device__ int64_t len(int r) const { return batch_M(runs, states, r); }
  __device__ int64_t stop_at(int r) const { return batch_M(runs, states, r); }
  __device__ EvSrcSorted src(int r) const {
    return {samp_logL + runs[r].samp_off, perm + (int64_t)r * S, num_live + (int64_t)r * S};
  }
};

// ordered inclusive scan of one map per lane across the wavefront
__device__ __forceinline__ EvMap wave_incl_scan(EvMap x) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const EvMap up = shfl_up_map(x, o);
    if (l >= o) x = evmap_then(up, x);
  }
  return x;
}

template <class Seg>
__global__ void __launch_bounds__(EV_THREADS) k_ev_reduce(Seg seg, int64_t T, EvMap* tile_agg, EvMap* thr_excl) {
  __shared__ EvMap wave_aggs[EV_THREADS / 64];
  const int r = blockIdx.y;
  const int64_t M = seg.len(r), stop = seg.stop_at(r);
  const int64_t tile = blockIdx.x;
  if (tile * EV_TILE >= M) return;  // block-uniform
  const auto src = seg.src(r);
  const int tid = threadIdx.x, l = lane_id(), w = tid / 64;
  const int64_t base = tile * EV_TILE + (int64_t)tid * EV_ITEMS;
  const double ninf = -__builtin_inf();
  EvMap m = evmap_identity();
  EvCoef cc = evcoef_empty();
  for (int k = 0; k < EV_ITEMS; ++k) {
    const int64_t i = base + k;
    if (i < M && i < stop) {
      double lt;
      const double Lp = i == 0 ? ninf : src.L(i - 1);
      m = evmap_then(m, evmap_step_c(evcoef_cached(cc, src.N(i)), Lp, src.L(i), &lt));
    }
  }
  const EvMap incl = wave_incl_scan(m);
  if (l == 63) wave_aggs[w] = incl;
  __syncthreads();
  EvMap wave_prefix = evmap_identity();
  for (int i = 0; i < w; ++i) wave_prefix = evmap_then(wave_prefix, wave_aggs[i]);
  EvMap lane_excl = shfl_up_map(incl, 1);
  if (l == 0) lane_excl = evmap_identity();
  thr_excl[((int64_t)r * T + tile) * EV_THREADS + tid] = evmap_then(wave_prefix, lane_excl);
  if (tid == 0) {
    EvMap agg = evmap_identity();
    for (int i = 0; i < EV_THREADS / 64; ++i) agg = evmap_then(agg, wave_aggs[i]);
    tile_agg[(int64_t)r * T + tile] = agg;
  }
}

// exclusive ordered scan of segment r's tile aggregates, in place: each thread composes a contiguous
// chunk, the chunk totals are scanned across the block, then each thread walks its chunk again
template <class Seg>
__global__ void __launch_bounds__(EV_SCAN_THREADS) k_ev_scan_tiles(Seg seg, int64_t T, EvMap* tile_agg) {
  constexpr int NW = EV_SCAN_THREADS / 64;
  __shared__ EvMap wave_tot[NW];
  const int r = blockIdx.y;
  const int64_t Tr = (seg.len(r) + EV_TILE - 1) / EV_TILE;
  if (Tr == 0) return;
  EvMap* a = tile_agg + (int64_t)r * T;
  const int tid = threadIdx.x, l = lane_id(), w = tid / 64;
  const int64_t per = (Tr + EV_SCAN_THREADS - 1) / EV_SCAN_THREADS;
  const int64_t j0 = tid * per;
  const int64_t j1 = j0 + per < Tr ? j0 + per : Tr;
  EvMap m = evmap_identity();
  for (int64_t j = j0; j < j1; ++j) m = evmap_then(m, a[j]);
  const EvMap incl = wave_incl_scan(m);
  if (l == 63) wave_tot[w] = incl;
  __syncthreads();
  if (w == 0) {
    EvMap x = l < NW ? wave_tot[l] : evmap_identity();
#pragma unroll
    for (int o = 1; o < NW; o <<= 1) {
      const EvMap up = shfl_up_map(x, o);
      if (l >= o) x = evmap_then(up, x);
    }
    EvMap ex = shfl_up_map(x, 1);
    if (l == 0) ex = evmap_identity();
    __syncwarp();
    if (l < NW) wave_tot[l] = ex;
  }
  __syncthreads();
  EvMap lane_excl = shfl_up_map(incl, 1);
  if (l == 0) lane_excl = evmap_identity();
  EvMap run = evmap_then(wave_tot[w], lane_excl);
  for (int64_t j = j0; j < j1; ++j) {
    const EvMap v = a[j];
    a[j] = run;
    run = evmap_then(run, v);
  }
}

__device__ __forceinline__ void write_fin8(double* fin8, const EvState& o) {
  fin8[0] = o.L;
  fin8[1] = o.X;
  fin8[2] = o.X2;
  fin8[3] = o.Z;
  fin8[4] = o.ZX;
  fin8[5] = o.Z2;
  fin8[6] = o.dZ;
  fin8[7] = o.dZ2;
}

// per8 (PER8, flat segment only): [8][M] field-major; fin: 8 doubles per segment (may be NULL)
template <bool PER8, class Seg>
__global__ void __launch_bounds__(EV_THREADS) k_ev_apply(Seg seg, int64_t T, const EvMap* tile_excl,
                                                         const EvMap* thr_excl, double* per8, double* fin) {
  __shared__ double s_stage[PER8 ? EV_TILE : 1];
  const int r = blockIdx.y;
  const int64_t M = seg.len(r), stop = seg.stop_at(r);
  const int64_t tile = blockIdx.x;
  const int tid = threadIdx.x, l = lane_id(), w = tid / 64;
  const double ninf = -__builtin_inf();
  const EvState s_init = {ninf, 0.0, 0.0, ninf, ninf, ninf, ninf, ninf};
  double* fin8 = fin ? fin + 8 * (int64_t)r : nullptr;
  if (M == 0) {
    if (fin8 && tile == 0 && tid == 0) write_fin8(fin8, s_init);
    return;
  }
  if (tile * EV_TILE >= M) return;  // block-uniform
  const auto src = seg.src(r);
  const int64_t base = tile * EV_TILE + (int64_t)tid * EV_ITEMS;
  const EvMap pre = evmap_then(tile_excl[(int64_t)r * T + tile], thr_excl[((int64_t)r * T + tile) * EV_THREADS + tid]);
  EvState s = evmap_apply(pre, s_init);
  s.L = (base == 0) ? ninf : (base - 1 < stop ? src.L(base - 1) : ninf);
  s.dZ = ninf;
  const int64_t last = (stop < M ? stop : M) - 1;
  EvState out[PER8 ? EV_ITEMS : 1];
  EvCoef cc = evcoef_empty();
#pragma unroll
  for (int k = 0; k < EV_ITEMS; ++k) {
    const int64_t i = base + k;
    EvState o = s_init;
    if (i < M && i < stop) {
      o = ev_step_c(s, evcoef_cached(cc, src.N(i)), src.L(i));
      s = o;
    }
    if constexpr (PER8) out[k] = o;
    if (fin8 && i == last) write_fin8(fin8, o);
  }
  if constexpr (PER8) {
    // field by field through a per-wavefront LDS stage, so that each store writes 64 consecutive
    // samples (the thread's own samples are consecutive, 8 doubles apart between lanes)
    double* stage = s_stage + w * 64 * EV_ITEMS;
    const int64_t wbase = tile * EV_TILE + (int64_t)w * 64 * EV_ITEMS;
#pragma unroll
    for (int f = 0; f < 8; ++f) {
#pragma unroll
      for (int k = 0; k < EV_ITEMS; ++k) {
        const EvState& o = out[k];
        const double v = f == 0 ? o.L : f == 1 ? o.X : f == 2 ? o.X2 : f == 3 ? o.Z : f == 4 ? o.ZX
                       : f == 5 ? o.Z2 : f == 6 ? o.dZ : o.dZ2;
        stage[l * EV_ITEMS + k] = v;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < EV_ITEMS; ++k) {
        const int64_t i = wbase + k * 64 + l;
        if (i < M) per8[(int64_t)f * M + i] = stage[k * 64 + l];
      }
      __syncthreads();
    }
  }
}

// workspace of the scan of R segments of at most Mmax samples: tile aggregates / prefixes, then the
// threads' prefixes within their tiles
static int64_t ev_tiles(int64_t Mmax) { return Mmax > 0 ? (Mmax + EV_TILE - 1) / EV_TILE : 1; }
static size_t ev_ws_bytes(int R, int64_t Mmax) {
  const int64_t T = ev_tiles(Mmax);
  return (size_t)R * T * sizeof(EvMap) * (1 + EV_THREADS) + 256;
}

template <class Seg>
static void ev_launch(const Seg& seg, int R, int64_t Mmax, void* ws, double* per8, double* fin, hipStream_t s) {
  const int64_t T = ev_tiles(Mmax);
  EvMap* tile_agg = (EvMap*)ws;
  EvMap* thr_excl = tile_agg + (size_t)R * T;
  const dim3 grid((unsigned)T, (unsigned)R);
  hipLaunchKernelGGL(k_ev_reduce<Seg>, grid, dim3(EV_THREADS), 0, s, seg, T, tile_agg, thr_excl);
  hipLaunchKernelGGL(k_ev_scan_tiles<Seg>, dim3(1, (unsigned)R), dim3(EV_SCAN_THREADS), 0, s, seg, T, tile_agg);
  if (per8)
    hipLaunchKernelGGL((k_ev_apply<true, Seg>), grid, dim3(EV_THREADS), 0, s, seg, T, tile_agg, thr_excl, per8, fin);
  else
    hipLaunchKernelGGL((k_ev_apply<false, Seg>), grid, dim3(EV_THREADS), 0, s, seg, T, tile_agg, thr_excl, per8, fin);
}

__global__ void k_fin_init(double* fin8) {
  const double ninf = -__builtin_inf();
  double v[8] = {ninf, 0.0, 0.0, ninf, ninf, ninf, ninf, ninf};
  for (int i = 0; i < 8; ++i) fin8[i] = v[i];
}

size_t evidence_scan_ws(int64_t M) { return ev_ws_bytes(1, M); }

hipError_t evidence_scan(const double* logL, const double* n, int64_t M, int64_t stop, double* per8, double* fin8,
                         void* ws, size_t ws_bytes, hipStream_t s) {
  (void)ws_bytes;
  if (stop < 0 || stop > M) stop = M;
  if (fin8 && (stop == 0 || M == 0)) hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(1), 0, s, fin8);
  if (M == 0) return hipGetLastError();
  ev_launch(EvSegFlat{logL, n, M, stop}, 1, M, ws, per8, stop > 0 ? fin8 : nullptr, s);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The evidence scan without a caller workspace: its tile aggregates and per-thread prefixes live in a
// file-scope device array, one copy per GPU for the life of the process.  Every element the scan
// reads there is written by the same call first (k_ev_reduce writes every aggregate and prefix,
// k_ev_scan_tiles rewrites the aggregates in place), so nothing needs clearing between calls.  Calls
// that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[HIPNS_SCAN_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_dev[HIPNS_MAX_DEVICES] = {};
}  // namespace

size_t evidence_scan_static_max() {
  // the largest M whose workspace fits the array
  const size_t per_tile = sizeof(EvMap) * (1 + EV_THREADS);
  return (size_t)((HIPNS_SCAN_SCRATCH_BYTES - 256) / per_tile) * EV_TILE;
}

hipError_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= HIPNS_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

hipError_t evidence_scan_static(const double* logL, const double* n, int64_t M, int64_t stop, double* per8,
                                double* fin8, hipStream_t s) {
  void* ws = nullptr;
  hipError_t e = scratch_address(&ws);
  if (e != hipSuccess) return e;
  return evidence_scan(logL, n, M, stop, per8, fin8, ws, HIPNS_SCAN_SCRATCH_BYTES, s);
}

// ===========================================================================
// Stable argsort of FP64 keys (jnp.argsort(stable=True) order: -inf < ... < +inf < NaN, -0 == +0,
// equal keys keep their input order) as an LSD radix sort of order-preserving 64-bit codes: eight
// passes over 8-bit digits.  Each pass is three launches: per-tile digit histograms, a scan of every
// (segment, digit) row of tile counts, and a stable scatter that ranks each tile's elements by digit
// in input order, stages the tile in LDS in digit order and writes each digit's run contiguously.
// A launch may sort R segments of a common stride S at once (grid.y = segment), each of its own
// length; nothing waits on another workgroup.
// ===========================================================================
constexpr int RS_THREADS = 256;  // == RS_BINS: thread d owns digit d in the per-digit steps
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_BINS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;

// order-preserving code: negative -> bit complement, non-negative -> sign bit set; -0 as +0; every NaN
// one code above +inf.  Codes are equal exactly when the keys compare equal (or are both NaN).
__device__ __forceinline__ uint64_t sort_code(double x) {
  if (__builtin_isnan(x)) return ~0ull;
  if (x == 0.0) x = 0.0;
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

template <bool FIRST>
__device__ __forceinline__ uint64_t rs_key(const void* kin, int64_t g) {
  if constexpr (FIRST) return sort_code(((const double*)kin)[g]);
  else return ((const uint64_t*)kin)[g];
}

// lanes of the wavefront that are valid and hold the same digit as this lane
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// exclusive scan of one value per thread over a 256-thread block; *total gets the sum (all threads)
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wsum, uint32_t* total) {
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < RS_WAVES; ++q) {
    if (q < w) pre += s_wsum[q];
    tot += s_wsum[q];
  }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}

// segment lengths: one array of n keys, or run r's nodes (M_r + 1) of a batch
struct SegFixed {
  int64_t n;
  __device__ int64_t len(int) const { return n; }
};

struct SegBatchNodes {
  const hipns_run_desc_t* runs;
  const hipns_run_state_t* states;
  __device__ int64_t len(int r) const { return batch_M(runs, states, r) + 1; }
};

// counts[(seg * 256 + digit) * T + tile] = elements of the tile with that digit
template <bool FIRST, class Seg>
__global__ void __launch_bounds__(RS_THREADS) k_rs_count(const void* kin, int64_t S, int T, Seg seg, int shift,
                                                         uint32_t* counts) {
  __shared__ uint32_t hist[RS_BINS];
  const int r = blockIdx.y, t = blockIdx.x, tid = threadIdx.x;
  const int64_t n = seg.len(r);
  const int64_t base = (int64_t)t * RS_TILE;
  if (base >= n) return;  // block-uniform: the scans read only the segment's own tiles
  hist[tid] = 0u;
  __syncthreads();
  {
    const uint64_t lt = (1ull << lane_id()) - 1ull;
    // all loads first (clamped in-bounds index, so they issue together), then the histogram
    uint64_t kr[RS_ITEMS];
#pragma unroll
    for (int k = 0; k < RS_ITEMS; ++k) {
      const int64_t i = base + k * RS_THREADS + tid;
      kr[k] = rs_key<FIRST>(kin, (int64_t)r * S + (i < n ? i : n - 1));
    }
#pragma unroll
    for (int k = 0; k < RS_ITEMS; ++k) {
      const bool valid = base + k * RS_THREADS + tid < n;
      const uint32_t d = (uint32_t)(kr[k] >> shift) & 255u;
      const uint64_t peers = digit_peers(d, valid);
      if (valid && (peers & lt) == 0ull) atomicAdd(&hist[d], (uint32_t)__popcll(peers));
    }
  }
  __syncthreads();
  counts[((int64_t)r * RS_BINS + tid) * T + t] = hist[tid];
}

// each (digit, segment) row of T tile counts -> exclusive offsets in place; totals[seg * 256 + digit]
__global__ void __launch_bounds__(RS_THREADS) k_rs_scan(uint32_t* counts, int T, uint32_t* totals) {
  __shared__ uint32_t s_wsum[RS_WAVES];
  const int d = blockIdx.x, r = blockIdx.y;
  uint32_t* c = counts + ((int64_t)r * RS_BINS + d) * T;
  const int per = (T + RS_THREADS - 1) / RS_THREADS;
  const int j0 = threadIdx.x * per;
  const int j1 = j0 + per < T ? j0 + per : T;
  uint32_t sum = 0u;
  for (int j = j0; j < j1; ++j) sum += c[j];
  uint32_t tot;
  uint32_t run = block_excl_scan(sum, s_wsum, &tot);
  for (int j = j0; j < j1; ++j) {
    const uint32_t v = c[j];
    c[j] = run;
    run += v;
  }
  if (threadIdx.x == 0) totals[r * RS_BINS + d] = tot;
}

// the same per segment for batches: one block per segment scans only that segment's Tr tiles (the
// count launch writes nothing for empty tiles), staging 16 columns of all 256 rows in LDS at a time;
// thread d scans digit d's row
constexpr int RS_SCAN_COLS = 16;
template <class Seg>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scan_seg(Seg seg, uint32_t* counts, int T, uint32_t* totals) {
  __shared__ uint32_t sc[RS_BINS * RS_SCAN_COLS];
  const int r = blockIdx.x, d = threadIdx.x;
  const int Tr = (int)((seg.len(r) + RS_TILE - 1) / RS_TILE);
  uint32_t* c = counts + (int64_t)r * RS_BINS * T;
  uint32_t run = 0u;
  for (int c0 = 0; c0 < Tr; c0 += RS_SCAN_COLS) {
    const int w = Tr - c0 < RS_SCAN_COLS ? Tr - c0 : RS_SCAN_COLS;
    for (int i = d; i < RS_BINS * w; i += RS_THREADS) {
      const int row = i / w, col = i - row * w;
      sc[i] = c[(int64_t)row * T + c0 + col];
    }
    __syncthreads();
    for (int j = 0; j < w; ++j) {
      const uint32_t v = sc[d * w + j];
      sc[d * w + j] = run;
      run += v;
    }
    __syncthreads();
    for (int i = d; i < RS_BINS * w; i += RS_THREADS) {
      const int row = i / w, col = i - row * w;
      c[(int64_t)row * T + c0 + col] = sc[i];
    }
    __syncthreads();
  }
  totals[r * RS_BINS + d] = run;
}

// stable scatter of one pass; the first pass reads FP64 keys (codes formed here) and takes the
// element's position as its value, the last writes values only
template <bool FIRST, bool LAST, class Seg>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const void* kin, const int32_t* vin, int64_t S, int T,
                                                           Seg seg, int shift, const uint32_t* offs,
                                                           const uint32_t* totals, uint64_t* kout, int32_t* vout) {
  __shared__ uint64_t s_k[RS_TILE];
  __shared__ int32_t s_v[RS_TILE];
  __shared__ uint32_t s_wcnt[2][RS_WAVES][RS_BINS];
  __shared__ uint32_t s_run[RS_BINS];
  __shared__ uint32_t s_toff[RS_BINS];
  __shared__ uint32_t s_gbase[RS_BINS];
  __shared__ uint32_t s_wsum[RS_WAVES];
  const int r = blockIdx.y, t = blockIdx.x, tid = threadIdx.x;
  const int64_t n = seg.len(r);
  const int64_t base = (int64_t)t * RS_TILE;
  if (base >= n) return;  // block-uniform: an empty tile places nothing
  const int w = tid >> 6;
  const uint64_t lt = (1ull << lane_id()) - 1ull;
  const int64_t so = (int64_t)r * S;
  {
    uint32_t tot;
    const uint32_t ex = block_excl_scan(totals[r * RS_BINS + tid], s_wsum, &tot);
    s_gbase[tid] = ex + offs[((int64_t)r * RS_BINS + tid) * T + t];
    s_run[tid] = 0u;
#pragma unroll
    for (int q = 0; q < RS_WAVES; ++q) {
      s_wcnt[0][q][tid] = 0u;
      s_wcnt[1][q][tid] = 0u;
    }
  }
  __syncthreads();
  uint64_t kr[RS_ITEMS];
  int32_t vr[RS_ITEMS];
  uint32_t dr[RS_ITEMS];  // digit | (rank among the tile's elements of that digit) << 8; ~0 = none
  // all loads first (clamped in-bounds index, so they issue together), then the ranking rounds
#pragma unroll
  for (int k = 0; k < RS_ITEMS; ++k) {
    const int64_t i = base + k * RS_THREADS + tid;
    const int64_t ic = i < n ? i : n - 1;
    kr[k] = rs_key<FIRST>(kin, so + ic);
    vr[k] = FIRST ? (int32_t)i : vin[so + ic];
  }
#pragma unroll
  for (int k = 0; k < RS_ITEMS; ++k) {
    const int buf = k & 1;
    const int64_t i = base + k * RS_THREADS + tid;
    const bool valid = i < n;
    const uint64_t key = kr[k];
    const int32_t val = vr[k];
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    const uint64_t peers = digit_peers(d, valid);
    const uint32_t rw = (uint32_t)__popcll(peers & lt);
    if (valid && rw == 0u) s_wcnt[buf][w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    uint32_t pre = s_run[d];
#pragma unroll
    for (int q = 0; q < RS_WAVES; ++q)
      if (q < w) pre += s_wcnt[buf][q][d];
    kr[k] = key;
    vr[k] = val;
    dr[k] = valid ? (d | ((pre + rw) << 8)) : ~0u;
    __syncthreads();
    // thread tid folds digit tid's counts of this round and clears them for the round after next
    uint32_t add = 0u;
#pragma unroll
    for (int q = 0; q < RS_WAVES; ++q) {
      add += s_wcnt[buf][q][tid];
      s_wcnt[buf][q][tid] = 0u;
    }
    s_run[tid] += add;
  }
  __syncthreads();
  {
    uint32_t tot;
    s_toff[tid] = block_excl_scan(s_run[tid], s_wsum, &tot);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < RS_ITEMS; ++k) {
    if (dr[k] != ~0u) {
      const uint32_t pos = s_toff[dr[k] & 255u] + (dr[k] >> 8);
      s_k[pos] = kr[k];
      s_v[pos] = vr[k];
    }
  }
  __syncthreads();
  const int cnt = (int)(n - base < RS_TILE ? n - base : RS_TILE);
  for (int j = tid; j < cnt; j += RS_THREADS) {
    const uint64_t key = s_k[j];
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    const int64_t o = so + s_gbase[d] + (uint32_t)j - s_toff[d];
    if constexpr (!LAST) kout[o] = key;
    vout[o] = s_v[j];
  }
}

// workspace of a radix sort of R segments of stride S: two key buffers, one value buffer, the tile
// counts and the digit totals
struct RadixWs {
  uint64_t* kA;
  uint64_t* kB;
  int32_t* vB;
  uint32_t* counts;
  uint32_t* totals;
};

static size_t radix_ws_bytes(int R, int64_t S) {
  const int64_t T = (S + RS_TILE - 1) / RS_TILE;
  size_t o = 0;
  o += (size_t)R * S * 8 * 2;
  o = (o + 255) & ~(size_t)255;
  o += (size_t)R * S * 4;
  o = (o + 255) & ~(size_t)255;
  o += (size_t)R * RS_BINS * T * 4;
  o = (o + 255) & ~(size_t)255;
  o += (size_t)R * RS_BINS * 4 + 256;
  return o;
}

static RadixWs radix_ws(void* ws, int R, int64_t S) {
  const int64_t T = (S + RS_TILE - 1) / RS_TILE;
  char* p = (char*)ws;
  RadixWs w;
  w.kA = (uint64_t*)p;
  w.kB = w.kA + (size_t)R * S;
  size_t o = (size_t)R * S * 8 * 2;
  o = (o + 255) & ~(size_t)255;
  w.vB = (int32_t*)(p + o);
  o += (size_t)R * S * 4;
  o = (o + 255) & ~(size_t)255;
  w.counts = (uint32_t*)(p + o);
  o += (size_t)R * RS_BINS * T * 4;
  o = (o + 255) & ~(size_t)255;
  w.totals = (uint32_t*)(p + o);
  return w;
}

// sorts every segment r (keys[r * S + i], i < seg.len(r)) and writes its permutation to perm[r * S + i];
// keys are read, not written; perm also serves as the value buffer of the odd passes
template <class Seg>
static hipError_t radix_argsort(const double* keys, int R, int64_t S, Seg seg, int32_t* perm, const RadixWs& w,
                                hipStream_t s) {
  const int T = (int)((S + RS_TILE - 1) / RS_TILE);
  const dim3 grid((unsigned)T, (unsigned)R);
  // pass p writes keys to kB (p even) / kA (p odd) and values to vB (p even) / perm (p odd); the last
  // pass (odd) reads kB / vB and writes values to perm
  for (int p = 0; p < 8; ++p) {
    const int shift = 8 * p;
    const void* kin = p == 0 ? (const void*)keys : (const void*)((p & 1) ? w.kB : w.kA);
    const int32_t* vin = (p & 1) ? w.vB : perm;
    uint64_t* kout = (p & 1) ? w.kA : w.kB;
    int32_t* vout = (p & 1) ? perm : w.vB;
    if (p == 0)
      hipLaunchKernelGGL((k_rs_count<true, Seg>), grid, dim3(RS_THREADS), 0, s, kin, S, T, seg, shift, w.counts);
    else
      hipLaunchKernelGGL((k_rs_count<false, Seg>), grid, dim3(RS_THREADS), 0, s, kin, S, T, seg, shift, w.counts);
    // the digit-parallel scan reads every tile of a row, so only for one flat array (all its tiles
    // hold keys); a batch segment's tiles past its length hold no counts
    if (std::is_same<Seg, SegFixed>::value && T > RS_SCAN_COLS)
      hipLaunchKernelGGL(k_rs_scan, dim3(RS_BINS, 1), dim3(RS_THREADS), 0, s, w.counts, T, w.totals);
    else
      hipLaunchKernelGGL((k_rs_scan_seg<Seg>), dim3((unsigned)R), dim3(RS_THREADS), 0, s, seg, w.counts, T, w.totals);
    if (p == 0)
      hipLaunchKernelGGL((k_rs_scatter<true, false, Seg>), grid, dim3(RS_THREADS), 0, s, kin, vin, S, T, seg, shift,
                         w.counts, w.totals, kout, vout);
    else if (p < 7)
      hipLaunchKernelGGL((k_rs_scatter<false, false, Seg>), grid, dim3(RS_THREADS), 0, s, kin, vin, S, T, seg, shift,
                         w.counts, w.totals, kout, vout);
    else
      hipLaunchKernelGGL((k_rs_scatter<false, true, Seg>), grid, dim3(RS_THREADS), 0, s, kin, vin, S, T, seg, shift,
                         w.counts, w.totals, kout, vout);
  }
  return hipGetLastError();
}

size_t argsort_ws(int64_t n) { return radix_ws_bytes(1, n < 1 ? 1 : n); }

// perm receives int32 indices
hipError_t argsort_stable(const double* keys, int64_t n, int32_t* perm, void* ws, size_t ws_bytes, hipStream_t s) {
  (void)ws_bytes;
  if (n == 0) return hipSuccess;
  return radix_argsort(keys, 1, n, SegFixed{n}, perm, radix_ws(ws, 1, n), s);
}

// ===========================================================================
// count_crossed_edges (internals/tree_structure.py:33-108)
// ===========================================================================
// deg[node] += 1 for each lane's node (node < 0: nothing), one atomic per run of equal nodes in
// consecutive lanes: the samples of one shell share their sender, so runs are long
__device__ __forceinline__ void add_degree_runs(int32_t* deg, int64_t node) {
  const int l = lane_id();
  const int64_t prev = __shfl_up(node, 1, 64);
  const bool start = l == 0 || node != prev;
  const uint64_t starts = __ballot(start);
  const uint64_t later = l == 63 ? 0ull : starts & (~0ull << (l + 1));
  const int end = later ? __ffsll((unsigned long long)later) - 1 : 64;
  if (start && node >= 0) atomicAdd(&deg[node], end - l);
}

__global__ void k_nodes(const double* logL, int64_t N, double* nodes) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > N) return;
  nodes[i] = i == 0 ? -__builtin_inf() : logL[i - 1];
}

__global__ void k_out_degree(const int32_t* sender, int64_t N, int32_t* deg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t node = -1;
  if (i < N) {
    const int32_t s = sender[i] > 0 ? sender[i] : 0;  // lax.max(sender, 0)
    if (s <= N) node = s;                             // out-of-range scatter updates are dropped
  }
  add_degree_runs(deg, node);
}

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int64_t crossed_term(const int32_t* deg, const int32_t* perm, int64_t i) {
  return (int64_t)deg[perm[i]] - 1;
}

// gathers the sorted crossed-edge terms deg[perm[i]] - 1 once (zero from `stop` on) into term[] and sums
// each tile (strided, coalesced: a sum does not need the element order)
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const int32_t* deg, const int32_t* perm, int64_t n,
                                                             int64_t stop, int64_t* tile_sums, int32_t* term) {
  __shared__ int64_t red[SCAN_THREADS / 64];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t acc = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + (int64_t)k * SCAN_THREADS + threadIdx.x;
    if (i < n) {
      const int32_t v = i < stop ? deg[perm[i]] - 1 : 0;
      term[i] = v;
      acc += v;
    }
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

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const int32_t* term, const int32_t* perm, int64_t n,
                                                             int64_t stop, int64_t fake, const int64_t* tile_sums,
                                                             int64_t* num_live, int64_t* samples_indices) {
  __shared__ int64_t tsum[SCAN_THREADS];
  __shared__ int64_t s_out[SCAN_TILE];
  const int64_t tbase = (int64_t)blockIdx.x * SCAN_TILE;
  // samples_indices[i] = perm[i + 1] - 1 (tree_structure.py:104), written directly (coalesced)
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t i = tbase + (int64_t)k * SCAN_THREADS + threadIdx.x;
    if (i < n - 1) samples_indices[i] = (int64_t)perm[i + 1] - 1;
  }
  int64_t base = tbase + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t local[SCAN_ITEMS];
  int64_t acc = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + k;
    int64_t v = i < n ? term[i] : 0;
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
  // crossed_edges_sorted[i] through an LDS stage, so that each store writes consecutive entries
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t i = base + k;
    int64_t crossed = (i < stop) ? (1 + pre + local[k]) : fake;
    s_out[threadIdx.x * SCAN_ITEMS + k] = crossed - fake;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int j = k * SCAN_THREADS + threadIdx.x;
    const int64_t i = tbase + j;
    if (i < n - 1) num_live[i] = s_out[j];  // the last node (i = n - 1 = N) is dropped
  }
}

size_t count_crossed_edges_ws(int64_t N) {
  int64_t n = N + 1;
  int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  return (size_t)n * sizeof(double) + (size_t)n * sizeof(int32_t) * 3 + (size_t)tiles * sizeof(int64_t) +
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
  int32_t* term = (int32_t*)p;
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
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, deg, perm, n, stop, tsums, term);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, tsums, tiles);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)tiles), dim3(SCAN_THREADS), 0, s, term, perm, n, stop, fake, tsums,
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
__global__ void k_b_nodes(const hipns_run_desc_t* runs, const hipns_run_state_t* states, const double* samp_logL,
                          int64_t S, double* nodes) {
  const int r = blockIdx.y;
  const int64_t M = batch_M(runs, states, r);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > M) return;
  nodes[r * S + i] = i == 0 ? -__builtin_inf() : samp_logL[runs[r].samp_off + i - 1];
}

// out-degree of every node by sender (tree_structure.py:33-108) and the phantom count of each run
constexpr int OUT_DEGREE_ITEMS = 8;

__global__ void __launch_bounds__(256) k_b_out_degree(const hipns_run_desc_t* runs, const hipns_run_state_t* states,
                                                      hipns_buffers_t B, int64_t S, int32_t* deg,
                                                      unsigned long long* phantom) {
  __shared__ unsigned long long red[4];
  const int r = blockIdx.y;
  const int64_t M = batch_M(runs, states, r);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x * OUT_DEGREE_ITEMS + threadIdx.x;
  const int64_t off = runs[r].samp_off;
  unsigned long long ph = 0;
  // OUT_DEGREE_ITEMS samples per thread, 256 apart (each wavefront covers 64 consecutive samples)
  for (int k = 0; k < OUT_DEGREE_ITEMS; ++k) {
    const int64_t ik = i + (int64_t)k * blockDim.x;
    int64_t node = -1;
    if (ik < M) {
      int32_t snd = B.samp_sender[off + ik];
      int64_t sn = snd > 0 ? snd : 0;  // lax.max(sender, 0)
      if (sn <= M) node = sn;          // out-of-range scatter updates are dropped
      ph += B.samp_phantom[off + ik] ? 1ull : 0ull;
    }
    add_degree_runs(deg + r * S, node);
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

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct BatchEvLayout {
  int64_t S, Tscan;
  size_t zero_bytes, total;
  size_t off_perm, off_radix, off_nl, off_ts;
};

static BatchEvLayout batch_ev_layout(int R, int64_t max_M) {
  BatchEvLayout L{};
  L.S = round_up(max_M + 1, RS_TILE);
  L.Tscan = (L.S + SCAN_TILE - 1) / SCAN_TILE;
  size_t o = 0;
  o += (size_t)R * L.S * sizeof(int32_t);  // deg
  o = round_up(o, 256);
  L.zero_bytes = o;
  L.off_perm = o;
  o += (size_t)R * L.S * sizeof(int32_t);
  o = round_up(o, 256);
  // the node keys go into the radix sort's first key buffer; once the sort is done the region holds
  // the evidence scan's tile and thread prefixes
  L.off_radix = o;
  const size_t rb = radix_ws_bytes(R, L.S), eb = ev_ws_bytes(R, L.S);
  o += rb > eb ? rb : eb;
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
  int32_t* perm = (int32_t*)(base + L.off_perm);
  const RadixWs rw = radix_ws(base + L.off_radix, R, L.S);
  double* nodes = (double*)rw.kA;
  int64_t* nl = (int64_t*)(base + L.off_nl);
  int64_t* ts = (int64_t*)(base + L.off_ts);
  const int64_t words = (int64_t)(L.zero_bytes / 4);
  hipLaunchKernelGGL(k_fill_zero_u32, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, (uint32_t*)base, words);
  hipLaunchKernelGGL(k_fill_zero_u32, dim3((unsigned)((2 * R + 255) / 256)), dim3(256), 0, s, (uint32_t*)phantom,
                     (int64_t)2 * R);
  // node keys (root -inf, then the samples), sorted per run by the radix sort into perm
  hipLaunchKernelGGL(k_b_nodes, dim3((unsigned)((L.S + 255) / 256), R), dim3(256), 0, s, runs, states, B.samp_logL,
                     L.S, nodes);
  hipError_t e = radix_argsort(nodes, R, L.S, SegBatchNodes{runs, states}, perm, rw, s);
  if (e != hipSuccess) return e;
  const unsigned deg_blocks = (unsigned)((L.S + 256 * OUT_DEGREE_ITEMS - 1) / (256 * OUT_DEGREE_ITEMS));
  hipLaunchKernelGGL(k_b_out_degree, dim3(deg_blocks, R), dim3(256), 0, s, runs, states, B, L.S, deg,
                     (unsigned long long*)phantom);
  hipLaunchKernelGGL(k_b_scan_tiles, dim3((unsigned)L.Tscan, R), dim3(SCAN_THREADS), 0, s, runs, states, deg, perm,
                     L.S, L.Tscan, ts);
  hipLaunchKernelGGL(k_b_scan_sums, dim3(1, R), dim3(1024), 0, s, ts, L.Tscan);
  hipLaunchKernelGGL(k_b_scan_apply, dim3((unsigned)L.Tscan, R), dim3(SCAN_THREADS), 0, s, runs, states, deg, perm,
                     L.S, L.Tscan, ts, nl);
  ev_launch(EvSegBatch{runs, states, B.samp_logL, perm, nl, L.S}, R, L.S, base + L.off_radix, nullptr, fin, s);
  return hipGetLastError();
}

}  // namespace hipns
