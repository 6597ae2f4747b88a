// api.hip -- the extern "C" entry points of libhipns (declared in include/hipns.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>

#include "hipns_dev.h"
#include "ns_sampler.h"
#include "ns_shell.h"

// the ABI structs have the sizes the ctypes mirrors (jaxns_amd/_lib.py) assume
static_assert(sizeof(hipns_run_desc_t) == 88, "hipns_run_desc_t layout");
static_assert(sizeof(hipns_run_state_t) == 264, "hipns_run_state_t layout");
static_assert(sizeof(hipns_launch_t) == 40, "hipns_launch_t layout");
static_assert(sizeof(hipns_buffers_t) == 168, "hipns_buffers_t layout");

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(HIPNS_EHIP, "%s: %s", where, hipGetErrorString(e));
}

#define HIPNS_CHECK_LAUNCH(where)                \
  do {                                           \
    hipError_t _e = hipGetLastError();           \
    if (_e != hipSuccess) return hip_fail(_e, where); \
  } while (0)

hipStream_t as_stream(void* s) { return (hipStream_t)s; }

int check_buffers(const hipns_buffers_t* b) {
  if (!b) return fail(HIPNS_EINVAL, "null buffers");
  if (b->dps <= 0 || b->dps % 32 != 0 || b->dps > HIPNS_MAX_DIM)
    return fail(HIPNS_EINVAL, "dps=%d must be a positive multiple of 32 and <= %d", b->dps, HIPNS_MAX_DIM);
  return HIPNS_OK;
}

int check_runs(const hipns_run_desc_t* runs_host, int32_t R, const hipns_buffers_t* b, int* max_D, int* max_N,
               int* max_shell) {
  if (!runs_host || R <= 0) return fail(HIPNS_EINVAL, "need R > 0 run descriptors (host copy)");
  *max_D = 0;
  *max_N = 0;
  *max_shell = 0;
  for (int r = 0; r < R; ++r) {
    const hipns_run_desc_t& d = runs_host[r];
    if (d.D < 1 || d.D > b->dps) return fail(HIPNS_EINVAL, "run %d: D=%d outside [1, dps=%d]", r, d.D, b->dps);
    if (d.N < 2 || d.N > HIPNS_MAX_LIVE) return fail(HIPNS_ERANGE, "run %d: N=%d outside [2, %d]", r, d.N, HIPNS_MAX_LIVE);
    if (d.shell < 1 || d.shell > d.N) return fail(HIPNS_EINVAL, "run %d: shell=%d outside [1, N]", r, d.shell);
    if (d.num_slices < 1) return fail(HIPNS_EINVAL, "run %d: num_slices must be >= 1", r);
    if (d.k_phantom < 0 || d.k_phantom >= d.num_slices)
      return fail(HIPNS_EINVAL, "run %d: need 0 <= k < num_slices", r);
    if (d.max_samples < (int64_t)d.shell * (1 + d.k_phantom) + d.N)
      return fail(HIPNS_EINVAL, "run %d: max_samples=%lld too small", r, (long long)d.max_samples);
    if (d.max_samples > (int64_t)0x7fffffff) return fail(HIPNS_ERANGE, "run %d: max_samples beyond int32", r);
    if (d.D > *max_D) *max_D = d.D;
    if (d.N > *max_N) *max_N = d.N;
    if (d.shell > *max_shell) *max_shell = d.shell;
  }
  return HIPNS_OK;
}

int check_launch(const hipns_launch_t* L) {
  if (!L) return fail(HIPNS_EINVAL, "null launch config");
  if (L->family != HIPNS_FAM_MVN && L->family != HIPNS_FAM_GMM && L->family != HIPNS_FAM_LINREG)
    return fail(HIPNS_ENOTSUP, "family %d not supported by the fused sampler", L->family);
  if (L->family == HIPNS_FAM_LINREG) {
    if (L->lr_ndata < 32 || L->lr_ndata % 32) return fail(HIPNS_EINVAL, "lr_ndata must be a positive multiple of 32");
    if (L->lr_panels < 2 || L->lr_panels % 2 || L->lr_panels > HIPNS_MAX_DIM / 16)
      return fail(HIPNS_EINVAL, "lr_panels must be even, 2..%d", HIPNS_MAX_DIM / 16);
    if (L->lr_data_off < 0 || L->lr_data_off % 4) return fail(HIPNS_EINVAL, "lr_data_off must be a multiple of 4");
    if (L->lr_terms != 0 && L->lr_terms != 1 && L->lr_terms != 3 && L->lr_terms != 4 && L->lr_terms != 5)
      return fail(HIPNS_EINVAL, "lr_terms must be 1, 3, 4 or 5");
  }
  if (L->lanes_per_chain != 2 && L->lanes_per_chain != 4 && !(L->lanes_per_chain == 8 && L->family == HIPNS_FAM_GMM))
    return fail(HIPNS_EINVAL, "lanes_per_chain must be 2 or 4 (or 8 for the mixture family)");
  if (L->waves_per_block < 1 || L->waves_per_block > 4) return fail(HIPNS_EINVAL, "waves_per_block must be 1..4");
  if (L->family == HIPNS_FAM_GMM && (L->gmm_K < 1 || L->gmm_K > HIPNS_MAX_GMM_K))
    return fail(HIPNS_EINVAL, "gmm_K must be 1..%d", HIPNS_MAX_GMM_K);
  if (L->max_steps < 2) return fail(HIPNS_EINVAL, "max_steps must be >= 2");
  return HIPNS_OK;
}

// device view of the shared regression rows: params + lr_data_off = [r0: npad]
// [X hi: panels x npad x 16 bf16][X lo: same][y: npad][X: panels x npad x 16 f32]
hipns::LinRegLds linreg_view(const hipns_buffers_t* bufs, const hipns_launch_t* L) {
  hipns::LinRegLds lr{};
  const float* base = bufs->params + L->lr_data_off;
  const int64_t npad = L->lr_ndata, xe = (int64_t)L->lr_panels * npad * 16;
  lr.r0 = base;
  lr.xh = (const uint16_t*)(base + npad);
  lr.xl = lr.xh + xe;
  lr.y = base + npad + xe;
  lr.xf = base + 2 * npad + xe;
  lr.npad = L->lr_ndata;
  lr.panels = L->lr_panels;
  lr.terms = (L->lr_terms == 4 || L->lr_terms == 1 || L->lr_terms == 5) ? L->lr_terms : 3;
  return lr;
}

// every column a kernel instance reads (NJ * LPC of the instance chosen for maxD) must lie inside the
// panels
int check_panels(const hipns_launch_t* L, int maxD, int lpc) {
  const int cols = hipns::sampler_nj(maxD, lpc) * lpc;
  if (16 * L->lr_panels < cols)
    return fail(HIPNS_EINVAL, "lr_panels=%d cannot hold the %d columns read for D=%d", L->lr_panels, cols, maxD);
  return HIPNS_OK;
}

int max_lds_bytes() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 65536;
  return v;
}

int sampler_launch(int mode, const hipns_launch_t* L, const hipns::SampleArgs& A0, int R, int max_chains, int maxD,
                   int dps, int max_slices, hipStream_t s) {
  int lpc = hipns::sampler_lpc(L->family, L->lanes_per_chain, dps, maxD);
  hipns::SampleArgs A = A0;
  if (L->family == HIPNS_FAM_LINREG) {
    if (L->lr_terms == 5)
      return fail(HIPNS_EINVAL, "lr_terms 5 (the refinement form) is for hipns_log_likelihood; sampling takes 1, 3 or 4");
    int rc = check_panels(L, maxD, lpc);
    if (rc) return rc;
    const hipns::LinRegLds lr = linreg_view(&A.B, L);
    A.lr_r0 = lr.r0;
    A.lr_xh = lr.xh;
    A.lr_xl = lr.xl;
    A.lr_xf = lr.xf;
    A.lr_npad = lr.npad;
    A.lr_panels = lr.panels;
    A.lr_terms = lr.terms;
  }
  A.alpha_off = (int32_t)hipns::sampler_lds_floats(L->family, dps, maxD, L->gmm_K, lpc, L->waves_per_block);
  const size_t base_f = (size_t)A.alpha_off + (size_t)max_slices;
  size_t lds = base_f * sizeof(float);
  if ((int)lds > max_lds_bytes())
    return fail(HIPNS_ERANGE, "model parameters need %zu B of LDS, device allows %d", lds, max_lds_bytes());
  // direction rings after the alpha table (16-byte aligned): depth 3 for the 8-slot instances, 2 with one
  // wavefront per block (where the per-block parameters repeat per wavefront and LDS would cap residency)
  // and for 16 slots; none when the block's LDS would exceed the device limit
  const int nj = hipns::sampler_nj(maxD, lpc);
  int ring = hipns::dir_ring_max(mode, L->family, nj);
  if (ring > 2 && L->waves_per_block < 2) ring = 2;
  A.ring_off = (int32_t)((base_f + 3) / 4 * 4);
  const size_t ring_lds =
      ((size_t)A.ring_off + (size_t)ring * HIPNS_WAVE * nj * L->waves_per_block) * sizeof(float);
  if (ring > 0 && (int)ring_lds <= max_lds_bytes()) {
    lds = ring_lds;
  } else {
    ring = 0;
  }
  A.dir_ring = ring;
  if (mode == hipns::MODE_GSLICE) {
    // the gradient's scratch at the same offset: [grad_rows][64 / lpc + 1] floats per wavefront
    lds = ((size_t)A.ring_off + (size_t)hipns::grad_rows(nj, lpc) * (HIPNS_WAVE / lpc + 1) * L->waves_per_block) *
          sizeof(float);
    if ((int)lds > max_lds_bytes())
      return fail(HIPNS_ERANGE, "gradient-guided sampling needs %zu B of LDS, device allows %d", lds,
                  max_lds_bytes());
  }
  hipns_launch_t L2 = *L;
  L2.lanes_per_chain = lpc;
  hipError_t e;
  switch (mode) {
    case hipns::MODE_SLICE: e = hipns::launch_sampler_slice(A, R, max_chains, L2, maxD, lds, s); break;
    case hipns::MODE_UNIF: e = hipns::launch_sampler_unif(A, R, max_chains, L2, maxD, lds, s); break;
    case hipns::MODE_INIT: e = hipns::launch_sampler_init(A, R, max_chains, L2, maxD, lds, s); break;
    case hipns::MODE_GSLICE: e = hipns::launch_sampler_gslice(A, R, max_chains, L2, maxD, lds, s); break;
    default: e = hipns::launch_sampler_inj(A, R, max_chains, L2, maxD, lds, s); break;
  }
  if (e != hipSuccess) return hip_fail(e, "sampler launch");
  return HIPNS_OK;
}

int rows_blocks(int64_t rows, int dps) {
  int64_t b = (rows * dps + 255) / 256;
  if (b > 1024) b = 1024;
  return (int)(b < 1 ? 1 : b);
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels local to the API (test hooks and the prior transform)
// ---------------------------------------------------------------------------
__global__ void k_philox_fill(const uint32_t* ctr, int64_t n, uint32_t k0, uint32_t k1, uint32_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  hipns::u32x4 w = hipns::philox4x32(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], k0, k1);
  out[4 * i] = w.x;
  out[4 * i + 1] = w.y;
  out[4 * i + 2] = w.z;
  out[4 * i + 3] = w.w;
}

__global__ void k_ndtri(const float* u, int64_t n, float* z) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = hipns::ndtri_f32(u[i]);
}

__global__ void k_slice_quotients(const float* a, const float* n, int64_t count, float* q) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    const float y = hipns::recip_rn(n[i]);
    q[i] = a ? hipns::quot_rn(a[i], n[i], y) : y;
  }
}

// X[b, i] = loc[i] + sum_{j<=i} tril[i, j] q_j(U[b, j]) when tril != NULL, else loc[i] + scale[i] q_i(U[b, i])
__global__ void k_prior_transform(const float* U, int64_t B, int D, int ldu, const float* loc, const float* scale,
                                  const float* tril, const int32_t* qkind, float* X, int ldx) {
  int64_t b = (int64_t)blockIdx.x * blockDim.y + threadIdx.y;
  if (b >= B) return;
  for (int i = threadIdx.x; i < D; i += blockDim.x) {
    float acc = loc[i];
    if (tril) {
      for (int j = 0; j <= i; ++j) {
        float u = U[b * ldu + j];
        float q = qkind[j] ? hipns::ndtri_f32(u) : u;
        acc = fmaf(tril[(int64_t)i * D + j], q, acc);
      }
    } else {
      float u = U[b * ldu + i];
      float q = qkind[i] ? hipns::ndtri_f32(u) : u;
      acc = fmaf(scale[i], q, acc);
    }
    X[b * ldx + i] = acc;
  }
}

__global__ void k_widen_i32(const int32_t* in, int64_t n, int64_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// ---------------------------------------------------------------------------
extern "C" {

int hipns_abi_version(void) { return HIPNS_ABI_VERSION; }

int hipns_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return HIPNS_EINVAL;
  size_t n = g_err.size() < len - 1 ? g_err.size() : len - 1;
  memcpy(buf, g_err.data(), n);
  buf[n] = 0;
  return HIPNS_OK;
}

int hipns_create_init_state(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                            const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs,
                            const hipns_launch_t* launch, void* workspace, size_t workspace_bytes, void* stream) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  int maxD, maxN, maxS;
  if ((rc = check_runs(runs_host, R, bufs, &maxD, &maxN, &maxS))) return rc;
  if ((rc = check_launch(launch))) return rc;
  if (!runs_dev || !states_dev) return fail(HIPNS_EINVAL, "null device run/state arrays");
  hipStream_t s = as_stream(stream);
  if ((rc = hipns_init_prepare(runs_dev, states_dev, R, runs_host, bufs, stream))) return rc;
  hipns::SampleArgs A = {};
  A.runs = runs_dev;
  A.states = states_dev;
  A.B = *bufs;
  A.max_steps = launch->max_steps;
  A.gmm_K = launch->gmm_K;
  if ((rc = sampler_launch(hipns::MODE_INIT, launch, A, R, maxN, maxD, bufs->dps, 0, s))) return rc;
  (void)workspace;
  (void)workspace_bytes;
  return hipns_init_finalize(runs_dev, states_dev, R, runs_host, bufs, stream);
}

int hipns_init_prepare(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                       const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  int maxD, maxN, maxS;
  if ((rc = check_runs(runs_host, R, bufs, &maxD, &maxN, &maxS))) return rc;
  if (!runs_dev || !states_dev) return fail(HIPNS_EINVAL, "null device run/state arrays");
  hipStream_t s = as_stream(stream);
  int64_t max_ms = 0;
  for (int r = 0; r < R; ++r) max_ms = runs_host[r].max_samples > max_ms ? runs_host[r].max_samples : max_ms;
  hipLaunchKernelGGL(hipns::k_init_state, dim3((R + 63) / 64), dim3(64), 0, s, states_dev, R);
  HIPNS_CHECK_LAUNCH("k_init_state");
  hipLaunchKernelGGL(hipns::k_init_samples, dim3(rows_blocks(max_ms, 1), R), dim3(256), 0, s, runs_dev, R,
                     *bufs);
  HIPNS_CHECK_LAUNCH("k_init_samples");
  return HIPNS_OK;
}

int hipns_init_finalize(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                        const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  int maxD, maxN, maxS;
  if ((rc = check_runs(runs_host, R, bufs, &maxD, &maxN, &maxS))) return rc;
  if (!runs_dev || !states_dev) return fail(HIPNS_EINVAL, "null device run/state arrays");
  size_t lds = hipns::replace_lds_bytes(maxN);
  if ((int)lds > max_lds_bytes()) return fail(HIPNS_ERANGE, "N=%d needs %zu B LDS for the sort", maxN, lds);
  hipLaunchKernelGGL(hipns::k_replace, dim3(R), dim3(1024), lds, as_stream(stream), runs_dev, states_dev, *bufs, 1);
  HIPNS_CHECK_LAUNCH("k_replace(init)");
  return HIPNS_OK;
}

int hipns_ext_step(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                   const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, const hipns_launch_t* launch,
                   const hipns_ext_state_t* ext, int32_t mode, int32_t step, void* stream) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  int maxD, maxN, maxS;
  if ((rc = check_runs(runs_host, R, bufs, &maxD, &maxN, &maxS))) return rc;
  if (!runs_dev || !states_dev || !launch || !ext) return fail(HIPNS_EINVAL, "null pointer");
  if (R > 65535) return fail(HIPNS_ERANGE, "at most 65535 runs per launch");
  if (launch->lanes_per_chain != 2 && launch->lanes_per_chain != 4)
    return fail(HIPNS_EINVAL, "lanes_per_chain must be 2 or 4");
  if (launch->waves_per_block < 1 || launch->waves_per_block > 4)
    return fail(HIPNS_EINVAL, "waves_per_block must be 1..4");
  if (launch->max_steps < 2) return fail(HIPNS_EINVAL, "max_steps must be >= 2");
  if (mode < 0 || mode > 4 || mode == 3)
    return fail(HIPNS_EINVAL, "mode must be 0 (slice), 1 (uniform), 2 (initial) or 4 (gradient-guided slice)");
  if (mode == hipns::MODE_GSLICE && !ext->grad) return fail(HIPNS_EINVAL, "mode 4 needs ext->grad");
  if (step < 0) return fail(HIPNS_EINVAL, "step < 0");
  if (!ext->U0 || !ext->nv || !ext->prop_U || !ext->logL0 || !ext->prop_logL || !ext->left || !ext->right ||
      !ext->t || !ext->s || !ext->k || !ext->evals || !ext->flags)
    return fail(HIPNS_EINVAL, "null ext state array");
  const int lpc = bufs->dps > 64 ? 4 : launch->lanes_per_chain;
  const int maxC = mode == hipns::MODE_INIT ? maxN : maxS;
  hipError_t e = hipns::launch_ext_step(runs_dev, states_dev, *bufs, *ext, mode, step, launch->max_steps, R, maxC, lpc,
                                        launch->waves_per_block, maxD, as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "ext step launch");
  return HIPNS_OK;
}

size_t hipns_collect_shell_workspace_size(int32_t R, const hipns_run_desc_t* runs_host) {
  (void)R;
  (void)runs_host;
  return 0;
}

static int shell_common(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                        const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, const hipns_launch_t* launch,
                        int* maxD, int* maxN, int* maxS) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  if ((rc = check_runs(runs_host, R, bufs, maxD, maxN, maxS))) return rc;
  if (launch && (rc = check_launch(launch))) return rc;
  if (!runs_dev || !states_dev) return fail(HIPNS_EINVAL, "null device run/state arrays");
  return HIPNS_OK;
}

int hipns_shell_discard(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                        const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream) {
  int rc, maxD, maxN, maxS;
  if ((rc = shell_common(runs_dev, states_dev, R, runs_host, bufs, nullptr, &maxD, &maxN, &maxS))) return rc;
  hipLaunchKernelGGL(hipns::k_discard, dim3(rows_blocks(maxS, bufs->dps / 4), R), dim3(256), 0, as_stream(stream),
                     runs_dev, states_dev, *bufs);
  HIPNS_CHECK_LAUNCH("k_discard");
  return HIPNS_OK;
}

int hipns_shell_sample(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                       const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, const hipns_launch_t* launch,
                       void* stream) {
  int rc, maxD, maxN, maxS;
  if ((rc = shell_common(runs_dev, states_dev, R, runs_host, bufs, launch, &maxD, &maxN, &maxS))) return rc;
  const int sampler = runs_host[0].sampler;
  for (int r = 1; r < R; ++r)
    if (runs_host[r].sampler != sampler) return fail(HIPNS_EINVAL, "all runs of a batch must use the same sampler");
  hipns::SampleArgs A = {};
  A.runs = runs_dev;
  A.states = states_dev;
  A.B = *bufs;
  A.max_steps = launch->max_steps;
  A.gmm_K = launch->gmm_K;
  int max_slices = 0;
  for (int r = 0; r < R; ++r) max_slices = runs_host[r].num_slices > max_slices ? runs_host[r].num_slices : max_slices;
  if (sampler < 0 || sampler > 4) return fail(HIPNS_EINVAL, "sampler must be 0, 1, 2, 3 or 4");
  if (sampler >= 2 && launch->family == HIPNS_FAM_LINREG)
    return fail(HIPNS_EINVAL, "gradient slice shells (sampler %d) run the MVN and mixture families; gradient-guided "
                              "regression shells step through hipns_ext_step", sampler);
  const int mode = sampler == 1 ? hipns::MODE_UNIF : (sampler >= 2 ? hipns::MODE_GSLICE : hipns::MODE_SLICE);
  return sampler_launch(mode, launch, A, R, maxS, maxD, bufs->dps, max_slices, as_stream(stream));
}

int hipns_shell_replace(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                        const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream) {
  int rc, maxD, maxN, maxS;
  if ((rc = shell_common(runs_dev, states_dev, R, runs_host, bufs, nullptr, &maxD, &maxN, &maxS))) return rc;
  size_t lds = hipns::replace_lds_bytes(maxS) + (size_t)maxN * 4;
  if ((int)lds > max_lds_bytes()) return fail(HIPNS_ERANGE, "N=%d needs %zu B LDS for the sort", maxN, lds);
  hipLaunchKernelGGL(hipns::k_replace, dim3(R), dim3(1024), lds, as_stream(stream), runs_dev, states_dev, *bufs, 0);
  HIPNS_CHECK_LAUNCH("k_replace");
  return HIPNS_OK;
}

int hipns_shell_register(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                         const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs,
                         const hipns_term_cond_t* term, void* stream) {
  int rc, maxD, maxN, maxS;
  if ((rc = shell_common(runs_dev, states_dev, R, runs_host, bufs, nullptr, &maxD, &maxN, &maxS))) return rc;
  if (!term) return fail(HIPNS_EINVAL, "null termination condition");
  // 256 threads: at 95 VGPRs a 1024-thread block fills a CU, so 512 runs took two rounds (214 us);
  // four-wavefront blocks keep every run resident (145 us per shell for 512 runs of N = 2048)
  hipLaunchKernelGGL(hipns::k_register, dim3(R), dim3(256), 0, as_stream(stream), runs_dev, states_dev, *bufs, *term);
  HIPNS_CHECK_LAUNCH("k_register");
  return HIPNS_OK;
}

int hipns_collect_shell(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                        const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, const hipns_launch_t* launch,
                        const hipns_term_cond_t* term, void* workspace, size_t workspace_bytes, void* stream) {
  int rc;
  if (!term) return fail(HIPNS_EINVAL, "null termination condition");
  if ((rc = hipns_shell_discard(runs_dev, states_dev, R, runs_host, bufs, stream))) return rc;
  if ((rc = hipns_shell_sample(runs_dev, states_dev, R, runs_host, bufs, launch, stream))) return rc;
  if ((rc = hipns_shell_replace(runs_dev, states_dev, R, runs_host, bufs, stream))) return rc;
  if ((rc = hipns_shell_register(runs_dev, states_dev, R, runs_host, bufs, term, stream))) return rc;
  (void)workspace;
  (void)workspace_bytes;
  return HIPNS_OK;
}

int hipns_consume_live(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                       const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, void* stream) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  int maxD, maxN, maxS;
  if ((rc = check_runs(runs_host, R, bufs, &maxD, &maxN, &maxS))) return rc;
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(hipns::k_consume, dim3(rows_blocks(maxN, bufs->dps / 4), R), dim3(256), 0, s, runs_dev, states_dev,
                     *bufs);
  HIPNS_CHECK_LAUNCH("k_consume");
  hipLaunchKernelGGL(hipns::k_consume_advance, dim3((R + 63) / 64), dim3(64), 0, s, runs_dev, states_dev, R);
  HIPNS_CHECK_LAUNCH("k_consume_advance");
  return HIPNS_OK;
}

int hipns_check_termination(const hipns_run_desc_t* runs_dev, hipns_run_state_t* states_dev, int32_t R,
                                 const hipns_buffers_t* bufs, const hipns_term_cond_t* term, int32_t set_no_seed,
                                 void* stream) {
  if (!runs_dev || !states_dev || !term || !bufs || R <= 0) return fail(HIPNS_EINVAL, "bad arguments");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(hipns::k_check_term, dim3((R + 63) / 64), dim3(64), 0, s, runs_dev, states_dev, R, *term,
                     set_no_seed, *bufs);
  HIPNS_CHECK_LAUNCH("k_check_term");
  return HIPNS_OK;
}

size_t hipns_evidence_scan_workspace_size(int64_t M) { return hipns::evidence_scan_ws(M); }

int hipns_evidence_scan(const double* log_L_dev, const double* num_live_points_dev, int64_t M, int64_t stop_idx,
                        double* per_sample_dev, double* final_dev, void* workspace, size_t workspace_bytes,
                        void* stream) {
  if (M < 0) return fail(HIPNS_EINVAL, "M < 0");
  if (M > 0 && (!log_L_dev || !num_live_points_dev)) return fail(HIPNS_EINVAL, "null inputs");
  if (!per_sample_dev && !final_dev) return fail(HIPNS_EINVAL, "no outputs");
  if (M >= ((int64_t)1 << 32) * 2048) return fail(HIPNS_ERANGE, "M too large");
  if (workspace_bytes < hipns::evidence_scan_ws(M) || !workspace)
    return fail(HIPNS_ENOMEM, "workspace %zu < %zu", workspace_bytes, hipns::evidence_scan_ws(M));
  hipError_t e = hipns::evidence_scan(log_L_dev, num_live_points_dev, M, stop_idx, per_sample_dev, final_dev,
                                      workspace, workspace_bytes, as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "evidence_scan");
  return HIPNS_OK;
}

size_t hipns_count_crossed_edges_workspace_size(int64_t N) { return hipns::count_crossed_edges_ws(N); }

size_t hipns_batch_evidence_workspace_size(int32_t R, const hipns_run_desc_t* runs_host) {
  if (R <= 0 || !runs_host) return 0;
  int64_t mx = 0;
  for (int r = 0; r < R; ++r) mx = runs_host[r].max_samples > mx ? runs_host[r].max_samples : mx;
  return hipns::batch_evidence_ws(R, mx);
}

int hipns_batch_evidence(const hipns_run_desc_t* runs_dev, const hipns_run_state_t* states_dev, int32_t R,
                         const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, double* final_dev,
                         int64_t* phantom_dev, void* workspace, size_t workspace_bytes, void* stream) {
  int rc, maxD, maxN, maxS;
  if ((rc = check_buffers(bufs))) return rc;
  if ((rc = check_runs(runs_host, R, bufs, &maxD, &maxN, &maxS))) return rc;
  if (R > 65535) return fail(HIPNS_ERANGE, "at most 65535 runs per call");
  if (!runs_dev || !states_dev || !final_dev || !phantom_dev) return fail(HIPNS_EINVAL, "null pointer");
  const size_t need = hipns_batch_evidence_workspace_size(R, runs_host);
  if (!workspace || workspace_bytes < need)
    return fail(HIPNS_ENOMEM, "workspace of %zu B needed, %zu given", need, workspace_bytes);
  int64_t mx = 0;
  for (int r = 0; r < R; ++r) mx = runs_host[r].max_samples > mx ? runs_host[r].max_samples : mx;
  hipError_t e = hipns::batch_evidence(runs_dev, states_dev, R, mx, *bufs, final_dev, phantom_dev, workspace,
                                       as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "batch evidence");
  return HIPNS_OK;
}

int hipns_count_crossed_edges(const int32_t* sender_node_idx_dev, const double* log_L_dev, int64_t N,
                              int64_t num_samples, int64_t* samples_indices_dev, int64_t* num_live_points_dev,
                              void* workspace, size_t workspace_bytes, void* stream) {
  if (N < 0 || N >= 0x7ffffffe) return fail(HIPNS_ERANGE, "N out of range");
  if (num_samples > N) return fail(HIPNS_EINVAL, "num_samples > N");
  if (N > 0 && (!sender_node_idx_dev || !log_L_dev || !samples_indices_dev || !num_live_points_dev))
    return fail(HIPNS_EINVAL, "null pointer");
  if (!workspace || workspace_bytes < hipns::count_crossed_edges_ws(N))
    return fail(HIPNS_ENOMEM, "workspace %zu < %zu", workspace_bytes, hipns::count_crossed_edges_ws(N));
  hipError_t e = hipns::count_crossed_edges(sender_node_idx_dev, log_L_dev, N, num_samples, samples_indices_dev,
                                            num_live_points_dev, workspace, workspace_bytes, as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "count_crossed_edges");
  return HIPNS_OK;
}

size_t hipns_argsort_workspace_size(int64_t n) { return hipns::argsort_ws(n) + (size_t)n * 4 + 256; }

int hipns_argsort_stable(const double* keys_dev, int64_t n, int64_t* perm_dev, void* workspace,
                         size_t workspace_bytes, void* stream) {
  if (n < 0 || n >= 0x7fffffff) return fail(HIPNS_ERANGE, "n out of range");
  if (n == 0) return HIPNS_OK;
  if (!keys_dev || !perm_dev) return fail(HIPNS_EINVAL, "null pointer");
  if (!workspace || workspace_bytes < hipns_argsort_workspace_size(n))
    return fail(HIPNS_ENOMEM, "workspace too small");
  int32_t* p32 = (int32_t*)workspace;
  void* ws = (void*)(((uintptr_t)(p32 + n) + 255) & ~(uintptr_t)255);
  hipStream_t s = as_stream(stream);
  hipError_t e = hipns::argsort_stable(keys_dev, n, p32, ws, hipns::argsort_ws(n), s);
  if (e != hipSuccess) return hip_fail(e, "argsort");
  hipLaunchKernelGGL(k_widen_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p32, n, perm_dev);
  HIPNS_CHECK_LAUNCH("k_widen_i32");
  return HIPNS_OK;
}

int hipns_log_likelihood(const float* U_dev, int64_t B, double* logL_dev, const hipns_run_desc_t* runs_dev,
                         const hipns_run_desc_t* runs_host, const hipns_buffers_t* bufs, const hipns_launch_t* launch,
                         void* stream) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  if ((rc = check_launch(launch))) return rc;
  if (!runs_host || !runs_dev) return fail(HIPNS_EINVAL, "null runs");
  if (B < 0) return fail(HIPNS_EINVAL, "B < 0");
  if (B == 0) return HIPNS_OK;
  if (!U_dev || !logL_dev) return fail(HIPNS_EINVAL, "null pointer");
  const int D = runs_host[0].D;
  if (D < 1 || D > bufs->dps) return fail(HIPNS_EINVAL, "D outside [1, dps]");
  int lpc = hipns::sampler_lpc(launch->family, launch->lanes_per_chain, bufs->dps, D);
  size_t lds = hipns::sampler_lds_floats(launch->family, bufs->dps, D, launch->gmm_K, lpc, launch->waves_per_block) *
               sizeof(float);
  if ((int)lds > max_lds_bytes()) return fail(HIPNS_ERANGE, "parameters exceed LDS");
  hipns_launch_t L2 = *launch;
  L2.lanes_per_chain = lpc;
  hipns::LinRegLds lr{};
  if (launch->family == HIPNS_FAM_LINREG) {
    if ((rc = check_panels(launch, D, lpc))) return rc;
    lr = linreg_view(bufs, launch);
  }
  hipError_t e = hipns::launch_loglik(U_dev, B, logL_dev, runs_dev, *bufs, L2, D, lds, as_stream(stream), lr);
  if (e != hipSuccess) return hip_fail(e, "loglik launch");
  return HIPNS_OK;
}

int hipns_log_likelihood_grad(const float* U_dev, int64_t B, double* logL_dev, float* grad_dev,
                              const hipns_run_desc_t* runs_dev, const hipns_run_desc_t* runs_host,
                              const hipns_buffers_t* bufs, const hipns_launch_t* launch, void* stream) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  if ((rc = check_launch(launch))) return rc;
  if (launch->family == HIPNS_FAM_LINREG)
    return fail(HIPNS_EINVAL, "hipns_log_likelihood_grad covers the MVN and mixture families");
  if (!runs_host || !runs_dev) return fail(HIPNS_EINVAL, "null runs");
  if (B < 0) return fail(HIPNS_EINVAL, "B < 0");
  if (B == 0) return HIPNS_OK;
  if (!U_dev || !logL_dev || !grad_dev) return fail(HIPNS_EINVAL, "null pointer");
  const int D = runs_host[0].D;
  if (D < 1 || D > bufs->dps) return fail(HIPNS_EINVAL, "D outside [1, dps]");
  const int lpc = hipns::sampler_lpc(launch->family, launch->lanes_per_chain, bufs->dps, D);
  const size_t base_f =
      hipns::sampler_lds_floats(launch->family, bufs->dps, D, launch->gmm_K, lpc, launch->waves_per_block);
  const int gscr_off = (int)((base_f + 3) / 4 * 4);
  const int nj = hipns::sampler_nj(D, lpc);
  const size_t lds =
      ((size_t)gscr_off + (size_t)hipns::grad_rows(nj, lpc) * (HIPNS_WAVE / lpc + 1) * launch->waves_per_block) *
      sizeof(float);
  if ((int)lds > max_lds_bytes()) return fail(HIPNS_ERANGE, "parameters and gradient scratch exceed LDS");
  hipns_launch_t L2 = *launch;
  L2.lanes_per_chain = lpc;
  hipError_t e =
      hipns::launch_loglik_grad(U_dev, B, logL_dev, grad_dev, runs_dev, *bufs, L2, D, gscr_off, lds, as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "loglik grad launch");
  return HIPNS_OK;
}

int hipns_linreg_stats(const float* U_dev, int64_t B, double* logL_dev, double* rss_dev, double* yr_dev,
                       const hipns_run_desc_t* runs_dev, const hipns_run_desc_t* runs_host,
                       const hipns_buffers_t* bufs, const hipns_launch_t* launch, void* stream) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  if ((rc = check_launch(launch))) return rc;
  if (launch->family != HIPNS_FAM_LINREG) return fail(HIPNS_EINVAL, "hipns_linreg_stats needs the regression family");
  if (launch->lr_terms == 5) return fail(HIPNS_EINVAL, "hipns_linreg_stats takes lr_terms 1, 3 or 4");
  if (!runs_host || !runs_dev) return fail(HIPNS_EINVAL, "null runs");
  if (B < 0) return fail(HIPNS_EINVAL, "B < 0");
  if (B == 0) return HIPNS_OK;
  if (!U_dev || !logL_dev || !rss_dev || !yr_dev) return fail(HIPNS_EINVAL, "null pointer");
  const int D = runs_host[0].D;
  if (D < 1 || D > bufs->dps) return fail(HIPNS_EINVAL, "D outside [1, dps]");
  const int lpc = bufs->dps > 64 ? 4 : launch->lanes_per_chain;
  if ((rc = check_panels(launch, D, lpc))) return rc;
  const size_t lds = hipns::sampler_lds_floats(HIPNS_FAM_LINREG, bufs->dps, D, 1, lpc, 1) * sizeof(float);
  hipError_t e = hipns::launch_linreg_stats(U_dev, B, logL_dev, rss_dev, yr_dev, runs_dev, *bufs, lpc, D, lds,
                                            as_stream(stream), linreg_view(bufs, launch));
  if (e != hipSuccess) return hip_fail(e, "linreg stats launch");
  return HIPNS_OK;
}

int hipns_prior_transform(const float* U_dev, int64_t B, int32_t D, int32_t ldu, const float* loc_dev,
                          const float* scale_dev, const float* tril_dev, const int32_t* qkind_dev, float* X_dev,
                          int32_t ldx, void* stream) {
  if (B < 0 || D < 1 || ldu < D || ldx < D) return fail(HIPNS_EINVAL, "bad sizes");
  if (B == 0) return HIPNS_OK;
  if (!U_dev || !loc_dev || !qkind_dev || !X_dev || (!tril_dev && !scale_dev)) return fail(HIPNS_EINVAL, "null");
  dim3 block(64, 4);
  hipLaunchKernelGGL(k_prior_transform, dim3((unsigned)((B + 3) / 4)), block, 0, as_stream(stream), U_dev, B, D, ldu,
                     loc_dev, scale_dev, tril_dev, qkind_dev, X_dev, ldx);
  HIPNS_CHECK_LAUNCH("k_prior_transform");
  return HIPNS_OK;
}

int hipns_slice_chains_injected(const hipns_run_desc_t* runs_dev, const hipns_run_desc_t* runs_host,
                                const hipns_buffers_t* bufs, const hipns_launch_t* launch, int32_t buf, int32_t C,
                                double contour, const int32_t* seed_idx_dev, const float* dirs_dev,
                                const float* tu_dev, float* trace_t_dev, double* trace_logL_dev, int32_t* steps_dev,
                                float* ph_U_dev, double* ph_logL_dev, void* stream) {
  int rc;
  if ((rc = check_buffers(bufs))) return rc;
  int maxD, maxN, maxS;
  if ((rc = check_runs(runs_host, 1, bufs, &maxD, &maxN, &maxS))) return rc;
  if ((rc = check_launch(launch))) return rc;
  if (C < 1 || C > runs_host[0].N) return fail(HIPNS_EINVAL, "C outside [1, N] (new-sample capacity)");
  if (buf != 0 && buf != 1) return fail(HIPNS_EINVAL, "buf must be 0 or 1");
  if (!seed_idx_dev || !dirs_dev || !tu_dev || !trace_t_dev || !trace_logL_dev || !steps_dev)
    return fail(HIPNS_EINVAL, "null pointer");
  if (runs_host[0].k_phantom > 0 && (!ph_U_dev || !ph_logL_dev)) return fail(HIPNS_EINVAL, "null phantom outputs");
  hipns::SampleArgs A = {};
  A.runs = runs_dev;
  A.states = nullptr;
  A.B = *bufs;
  A.max_steps = launch->max_steps;
  A.gmm_K = launch->gmm_K;
  A.inj_buf = buf;
  A.inj_C = C;
  A.inj_contour = contour;
  A.inj_seed = seed_idx_dev;
  A.inj_dirs = dirs_dev;
  A.inj_tu = tu_dev;
  A.trace_t = trace_t_dev;
  A.trace_logL = trace_logL_dev;
  A.trace_steps = steps_dev;
  A.ph_U = ph_U_dev;
  A.ph_logL = ph_logL_dev;
  return sampler_launch(hipns::MODE_INJ, launch, A, 1, C, maxD, bufs->dps, runs_host[0].num_slices,
                        as_stream(stream));
}

int hipns_philox4x32(const uint32_t* ctr_dev, int64_t n, uint32_t key0, uint32_t key1, uint32_t* out_dev,
                     void* stream) {
  if (n < 0) return fail(HIPNS_EINVAL, "n < 0");
  if (n == 0) return HIPNS_OK;
  if (!ctr_dev || !out_dev) return fail(HIPNS_EINVAL, "null pointer");
  hipLaunchKernelGGL(k_philox_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), ctr_dev, n,
                     key0, key1, out_dev);
  HIPNS_CHECK_LAUNCH("k_philox_fill");
  return HIPNS_OK;
}

int hipns_ndtri_f32(const float* u_dev, int64_t n, float* z_dev, void* stream) {
  if (n < 0) return fail(HIPNS_EINVAL, "n < 0");
  if (n == 0) return HIPNS_OK;
  if (!u_dev || !z_dev) return fail(HIPNS_EINVAL, "null pointer");
  hipLaunchKernelGGL(k_ndtri, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), u_dev, n, z_dev);
  HIPNS_CHECK_LAUNCH("k_ndtri");
  return HIPNS_OK;
}

int hipns_slice_quotients(const float* a_dev, const float* n_dev, int64_t count, float* q_dev, void* stream) {
  if (count < 0) return fail(HIPNS_EINVAL, "count < 0");
  if (count == 0) return HIPNS_OK;
  if (!n_dev || !q_dev) return fail(HIPNS_EINVAL, "null pointer");
  hipLaunchKernelGGL(k_slice_quotients, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, as_stream(stream), a_dev,
                     n_dev, count, q_dev);
  HIPNS_CHECK_LAUNCH("k_slice_quotients");
  return HIPNS_OK;
}

int hipns_cholesky_f64(const double* A_dev, int32_t D, double* L_dev, int32_t* info_dev, void* stream) {
  if (D < 1 || D > HIPNS_MAX_DIM) return fail(HIPNS_EINVAL, "D outside [1, %d]", HIPNS_MAX_DIM);
  if (!A_dev || !L_dev) return fail(HIPNS_EINVAL, "null pointer");
  hipError_t e = hipns::cholesky_f64(A_dev, D, L_dev, info_dev, as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "k_cholesky");
  return HIPNS_OK;
}

size_t hipns_mvn_prepare_workspace_size(int32_t D) { return D < 1 ? 0 : hipns::mvn_prepare_ws(D); }

int hipns_mvn_prepare_f64(const double* cov_dev, const double* mean_dev, const double* loc_dev, const double* S_dev,
                          const int32_t* qkind_dev, int32_t D, int32_t dps, double* L_dev, float* params_dev,
                          double* log_const_dev, int32_t* info_dev, void* workspace, size_t workspace_bytes,
                          void* stream) {
  if (D < 1 || D > HIPNS_MAX_DIM) return fail(HIPNS_EINVAL, "D outside [1, %d]", HIPNS_MAX_DIM);
  if (dps < D || dps % 32 || dps > HIPNS_MAX_DIM) return fail(HIPNS_EINVAL, "dps must be a multiple of 32 in [D, %d]",
                                                              HIPNS_MAX_DIM);
  if (!mean_dev || !loc_dev || !S_dev || !qkind_dev || !L_dev || !params_dev || !log_const_dev)
    return fail(HIPNS_EINVAL, "null pointer");
  if (!workspace || workspace_bytes < hipns::mvn_prepare_ws(D))
    return fail(HIPNS_ENOMEM, "workspace of %zu bytes needed", hipns::mvn_prepare_ws(D));
  hipError_t e = hipns::mvn_prepare_f64(cov_dev, mean_dev, loc_dev, S_dev, qkind_dev, D, dps, L_dev, workspace,
                                        params_dev, log_const_dev, info_dev, as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "k_mvn_prepare");
  return HIPNS_OK;
}

int hipns_evidence_scan_noalloc(const double* log_L_dev, const double* num_live_points_dev, int64_t M,
                                int64_t stop_idx, double* per_sample_dev, double* final_dev, void* stream) {
  if (M < 0) return fail(HIPNS_EINVAL, "M < 0");
  if (M > 0 && (!log_L_dev || !num_live_points_dev)) return fail(HIPNS_EINVAL, "null input");
  if (!final_dev && !per_sample_dev) return fail(HIPNS_EINVAL, "no output");
  if ((size_t)M > hipns::evidence_scan_static_max())
    return fail(HIPNS_ERANGE, "M=%lld beyond the %zu samples the library's fixed scan workspace holds; use "
                "hipns_evidence_scan with a caller workspace", (long long)M, hipns::evidence_scan_static_max());
  hipError_t e = hipns::evidence_scan_static(log_L_dev, num_live_points_dev, M, stop_idx, per_sample_dev, final_dev,
                                             as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "evidence scan (fixed workspace)");
  return HIPNS_OK;
}

size_t hipns_evidence_scan_noalloc_max(void) { return hipns::evidence_scan_static_max(); }

int hipns_results_finals(const double* log_dZ_dev, const double* log_L_dev, const double* log_X_dev, int64_t M,
                         double log_Z_mean, double* log_dp_dev, double* H_dev, void* workspace, size_t workspace_bytes,
                         void* stream) {
  if (M < 1) return fail(HIPNS_EINVAL, "M < 1");
  if (!log_dZ_dev || !log_L_dev || !log_X_dev || !log_dp_dev || !H_dev) return fail(HIPNS_EINVAL, "null pointer");
  if (!workspace || workspace_bytes < hipns::results_finals_ws(M))
    return fail(HIPNS_ENOMEM, "workspace of %zu bytes needed", hipns::results_finals_ws(M));
  hipError_t e = hipns::results_finals(log_dZ_dev, log_L_dev, log_X_dev, M, log_Z_mean, log_dp_dev, H_dev, workspace,
                                       as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "results finals");
  return HIPNS_OK;
}

size_t hipns_results_finals_workspace_size(int64_t M) { return M < 1 ? 0 : hipns::results_finals_ws(M); }

int hipns_prior_finals(const float* samp_U_dev, int32_t ldu, const int64_t* perm_dev, int64_t valid, int64_t M, int32_t D,
                       const double* loc_dev, const double* S_dev, const int32_t* qkind_dev, int32_t diag,
                       const double* log_L_dev, double* U_dev, double* X_dev, double* log_post_dev, void* stream) {
  if (M < 0 || D < 1 || ldu < D) return fail(HIPNS_EINVAL, "need M >= 0, D >= 1, ldu >= D");
  if (M == 0) return HIPNS_OK;
  if (!samp_U_dev || !loc_dev || !S_dev || !qkind_dev) return fail(HIPNS_EINVAL, "null pointer");
  if (log_post_dev && !log_L_dev) return fail(HIPNS_EINVAL, "log_post needs log_L");
  hipError_t e = hipns::prior_finals(samp_U_dev, ldu, perm_dev, valid, M, D, loc_dev, S_dev, qkind_dev, diag, log_L_dev,
                                     U_dev, X_dev, log_post_dev, as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "k_prior_finals");
  return HIPNS_OK;
}

int hipns_seed_indices(uint32_t key0, uint32_t key1, uint32_t iteration, int32_t first, int32_t N, int32_t C,
                       int32_t* idx_dev, double* u_dev, void* stream) {
  if (N < 1 || first < 0 || first > N) return fail(HIPNS_EINVAL, "need N >= 1 and 0 <= first <= N");
  if (C < 0) return fail(HIPNS_EINVAL, "C < 0");
  if (C == 0) return HIPNS_OK;
  if (!idx_dev || !u_dev) return fail(HIPNS_EINVAL, "null pointer");
  hipError_t e = hipns::launch_seed_indices(key0, key1, iteration, first, N, C, idx_dev, u_dev, as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "k_seed_indices");
  return HIPNS_OK;
}

int hipns_slice_directions(uint32_t key0, uint32_t key1, uint32_t iteration, int32_t C, int32_t num_slices, int32_t D,
                           int32_t lanes_per_chain, float* dirs_dev, void* stream) {
  if (D < 1 || D > HIPNS_MAX_DIM) return fail(HIPNS_EINVAL, "D outside [1, %d]", HIPNS_MAX_DIM);
  if (num_slices < 1 || num_slices > 65535) return fail(HIPNS_EINVAL, "num_slices outside [1, 65535]");
  if (lanes_per_chain != 2 && lanes_per_chain != 4) return fail(HIPNS_EINVAL, "lanes_per_chain must be 2 or 4");
  if (C < 0) return fail(HIPNS_EINVAL, "C < 0");
  if (C == 0) return HIPNS_OK;
  if (!dirs_dev) return fail(HIPNS_EINVAL, "null pointer");
  hipError_t e = hipns::launch_slice_dirs(key0, key1, iteration, C, num_slices, D, lanes_per_chain, dirs_dev,
                                          as_stream(stream));
  if (e != hipSuccess) return hip_fail(e, "k_slice_dirs");
  return HIPNS_OK;
}

}  // extern "C"
