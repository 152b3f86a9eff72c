// mcr_api.cpp -- C ABI of the HIP model (include/mcr_hip.h).
//
// One query (mcr_run) replaces the body of cuda_simulate
// (src/05-GPU/monte_carlos_cuda.cu:75-163): inputs go to the device once, the
// factor and every layout are built there (k_factor, k_prep), the fused path
// kernel reduces on device and only a few hundred bytes come back, plus the
// handful of losses that bracket each requested quantile.  Statistics follow
// src/05-GPU/monte_carlos_cuda.cu:146-150 (P_hat, SE, clamped 95% CI).
//
// Device memory is owned by the library: one grow-only arena per process and a
// pinned staging buffer, reused across queries, released by mcr_shutdown().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/mcr_hip.h"
#include "mcr_host.h"
#include "mcr_internal.h"

using namespace mcr;

namespace mcr {
namespace host {

thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

constexpr uint64_t kLaunchPaths = 1ull << 32;   // paths per launch (keeps lane tallies 32-bit)
constexpr uint64_t kPerPathChunk = 1ull << 24;  // per-path outputs staged per launch

// Device scratch of the allocation-free entry point (mcr_run_noalloc): every
// device buffer of one query (covariance, factor, layouts, accumulators, VaR
// gather buffers, per-path output chunks) is carved from this array instead of
// the arena.  16 MiB holds the largest query (15.3 MiB: n = 608, four VaR levels,
// per-path outputs staged 2^18 paths at a time); every device the library
// runs on holds it until the process ends.  Host code reaches it only through
// g_ctx.scratch, its device address from hipGetSymbolAddress.
constexpr size_t kScratchBytes = 16u << 20;
constexpr uint64_t kNoAllocChunk = 1ull << 18;  // per-path outputs staged per launch
static __device__ __align__(256) char scratch[kScratchBytes];

// Host buffers of the allocation-free entry point (static storage, no heap):
// equal weights, the gathered losses of one VaR window, the refinement
// histogram and the S histogram.  Guarded by g_mu like the entry point.
static double g_na_weights[kMaxN];
static float g_na_gather[kGatherCap];
static unsigned g_na_hist[kHistBins];
static unsigned long long g_na_shist[kMaxN + 1];

Context g_ctx;
std::mutex g_mu;

int ensure_init(int device) {
  if (g_ctx.ready) return MCR_OK;
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count));
  if (count <= 0) {
    set_error("no HIP device");
    return MCR_E_DEVICE;
  }
  if (device < 0 || device >= count) {
    set_error("device %d out of range (%d devices)", device, count);
    return MCR_E_ARG;
  }
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  g_ctx.device = device;
  g_ctx.num_cus = prop.multiProcessorCount;
  HIP_TRY(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&g_ctx.ev0));
  HIP_TRY(hipEventCreate(&g_ctx.ev1));
  HIP_TRY(hipEventCreate(&g_ctx.evs));
  MCR_TRY(create_forks(&g_ctx.forks, kMaxForks));
  // the scratch array's address on this device (the device is current)
  void* sp = nullptr;
  HIP_TRY(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  if (!sp) {
    set_error("hipGetSymbolAddress returned no address for the scratch array");
    return MCR_E_DEVICE;
  }
  g_ctx.scratch = sp;
  g_ctx.ready = true;
  return MCR_OK;
}

int reserve_arena(size_t bytes) {
  if (bytes <= g_ctx.arena_bytes) return MCR_OK;
  if (g_ctx.arena) {
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    HIP_TRY(hipFree(g_ctx.arena));
    g_ctx.arena = nullptr;
    g_ctx.arena_bytes = 0;
  }
  size_t want = std::max(bytes, (size_t)(g_ctx.arena_bytes * 3 / 2));
  if (hipMalloc(&g_ctx.arena, want) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device allocation of %zu bytes failed", want);
    return MCR_E_NOMEM;
  }
  g_ctx.arena_bytes = want;
  ++g_ctx.arena_gen;
  return MCR_OK;
}

int reserve_host(size_t bytes) {
  if (bytes <= g_ctx.host_bytes) return MCR_OK;
  if (g_ctx.host) {
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    HIP_TRY(hipHostFree(g_ctx.host));
    g_ctx.host = nullptr;
    g_ctx.host_bytes = 0;
  }
  size_t want = std::max(bytes, (size_t)(g_ctx.host_bytes * 3 / 2));
  if (hipHostMalloc(&g_ctx.host, want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    set_error("pinned host allocation of %zu bytes failed", want);
    return MCR_E_NOMEM;
  }
  g_ctx.host_bytes = want;
  return MCR_OK;
}

// Inverse standard normal cdf by Newton on erfc from a bisection bracket.
double inv_norm_cdf(double p) {
  double lo = -40.0, hi = 40.0;
  for (int i = 0; i < 200; ++i) {
    const double mid = 0.5 * (lo + hi);
    if (0.5 * std::erfc(-mid / std::sqrt(2.0)) < p) lo = mid; else hi = mid;
    if (hi - lo < 1e-13) break;
  }
  double x = 0.5 * (lo + hi);
  for (int i = 0; i < 3; ++i) {
    const double f = 0.5 * std::erfc(-x / std::sqrt(2.0)) - p;
    const double d = std::exp(-0.5 * x * x) / std::sqrt(2.0 * M_PI);
    if (d <= 0.0) break;
    x -= f / d;
  }
  return x;
}

// Exact ceil(alpha * M) using the binary value of alpha.
int64_t rank_of(double alpha, int64_t M) {
  int e = 0;
  const double m = std::frexp(alpha, &e);           // alpha = m 2^e, m in [0.5, 1)
  const unsigned __int128 mant = (unsigned __int128)std::ldexp(m, 53);  // exact 53-bit
  const int sh = 53 - e;                             // alpha = mant / 2^sh, sh >= 53
  const unsigned __int128 prod = mant * (unsigned __int128)M;
  unsigned __int128 q = sh >= 128 ? 0 : (prod >> sh);
  const unsigned __int128 rem = sh >= 128 ? prod : (prod - (q << sh));
  if (rem != 0) q += 1;
  int64_t r = (int64_t)q;
  if (r < 1) r = 1;
  if (r > M) r = M;
  return r;
}

struct DevBuffers {
  double *A, *mu, *w, *L, *v64;
  DevInfo* info;
  DevAccum* acc;
  float *colpack, *L32, *mu32, *v32, *rowinit;
  uint4* lfrag;
  float* gather;
  unsigned* hist;
  unsigned long long* s_hist;
  int32_t* S_out;
  float* loss_out;
};

struct Plan {
  int n, T, S, r8;
  int kernel;
  int n_alpha;
  bool per_path;
  size_t chunk;  // per-path chunk
};

size_t layout(const Plan& p, char* base, DevBuffers* b) {
  Carver c{base};
  const size_t n = (size_t)p.n;
  b->A = c.take<double>(n * n);
  b->mu = c.take<double>(n);
  b->w = c.take<double>(n);
  b->L = c.take<double>(n * n);
  b->v64 = c.take<double>(n);
  b->info = c.take<DevInfo>(1);
  b->acc = c.take<DevAccum>(1);
  b->colpack = c.take<float>(colpack_size(p.n));
  b->L32 = c.take<float>((size_t)p.r8 * p.r8);
  b->mu32 = c.take<float>(n);
  b->v32 = c.take<float>(n);
  b->lfrag = c.take<uint4>((size_t)p.T * p.S * 64 * 2);
  b->rowinit = c.take<float>((size_t)32 * p.T);
  b->gather = c.take<float>((size_t)std::max(p.n_alpha, 1) * kGatherCap);
  b->hist = c.take<unsigned>(kHistBins);
  b->s_hist = c.take<unsigned long long>(n + 1);
  b->S_out = p.per_path ? c.take<int32_t>(p.chunk) : nullptr;
  b->loss_out = p.per_path ? c.take<float>(p.chunk) : nullptr;
  return c.off;
}

static_assert(MCR_MAX_N == kChunkMaxN && MCR_MAX_N_BATCH == kBatchMaxN &&
                  MCR_MAX_N_NOALLOC == kMaxN,
              "public size limits follow the kernels'");

int choose_kernel(const mcr_query* q, Plan* p) {
  const int n = q->n;
  int k = q->kernel;
  // The near-threshold count (paths with some r_i within 1 ulp of the fp32
  // threshold) is kept by the fp32 kernels only: the bf16x3 products carry an
  // error of several ulp, so a 1-ulp window there would name no definite set.
  const bool near = (q->flags & MCR_F_NEAR) != 0;
  if (k == MCR_KERNEL_AUTO) {
    if (q->precision != MCR_PREC_FP32 && q->precision != MCR_PREC_BF16X3 &&
        q->precision != MCR_PREC_AUTO) {
      set_error("unknown precision %d", q->precision);
      return MCR_E_ARG;
    }
    const bool fp32 = q->precision == MCR_PREC_FP32 || (q->precision == MCR_PREC_AUTO && near);
    if (fp32) {
      k = n <= kRegMaxN ? MCR_KERNEL_VALU_REG
                        : (n <= kMaxN ? MCR_KERNEL_VALU_TILE : MCR_KERNEL_VALU_CHUNK);
    } else if (n <= kMfmaLdsMaxN) {
      k = n <= 32 * kMfmaMaxT - 1 ? MCR_KERNEL_MFMA : MCR_KERNEL_MFMA_LDS;
    } else {
      // past the LDS MFMA kernel's 24 row tiles the column-chunk fp32 kernel
      // runs every precision (fp32 error is no larger than the bf16x3 products')
      k = MCR_KERNEL_VALU_CHUNK;
    }
  }
  if (k == MCR_KERNEL_VALU_REG && n > kRegMaxN) {
    set_error("register kernel supports n <= %d (n = %d)", kRegMaxN, n);
    return MCR_E_UNSUPPORTED;
  }
  if (k == MCR_KERNEL_MFMA && n > 32 * kMfmaMaxT - 1) {
    set_error("MFMA kernel supports n <= %d (n = %d)", 32 * kMfmaMaxT - 1, n);
    return MCR_E_UNSUPPORTED;
  }
  if (k < MCR_KERNEL_VALU_REG || k > MCR_KERNEL_VALU_CHUNK) {
    set_error("unknown kernel %d", k);
    return MCR_E_ARG;
  }
  if ((k == MCR_KERNEL_VALU_TILE && n > kMaxN) || (k == MCR_KERNEL_MFMA_LDS && n > kMfmaLdsMaxN)) {
    set_error("%s kernel supports n <= %d (n = %d); MCR_KERNEL_VALU_CHUNK takes n <= %d",
              k == MCR_KERNEL_VALU_TILE ? "tile" : "LDS MFMA",
              k == MCR_KERNEL_VALU_TILE ? kMaxN : kMfmaLdsMaxN, n, kChunkMaxN);
    return MCR_E_UNSUPPORTED;
  }
  if (near && (k == MCR_KERNEL_MFMA || k == MCR_KERNEL_MFMA_LDS)) {
    set_error("MCR_F_NEAR needs an fp32 kernel (precision MCR_PREC_FP32 or MCR_PREC_AUTO); "
              "the bf16x3 MFMA kernels keep no near-threshold count");
    return MCR_E_UNSUPPORTED;
  }
  p->kernel = k;
  if (k == MCR_KERNEL_MFMA || k == MCR_KERNEL_MFMA_LDS) {
    p->T = (n + 1 + 31) / 32;
    p->S = (n + 15) / 16;
  } else {
    p->T = 0;
    p->S = 0;
  }
  return MCR_OK;
}

int validate(const mcr_query* q) {
  if (!q) {
    set_error("query is NULL");
    return MCR_E_ARG;
  }
  if (q->n < 1 || q->n > MCR_MAX_N) {
    set_error("n = %d outside [1, %d]", q->n, MCR_MAX_N);
    return q->n > MCR_MAX_N ? MCR_E_UNSUPPORTED : MCR_E_ARG;
  }
  if (!q->mu || !q->sigma) {
    set_error("mu and sigma are required");
    return MCR_E_ARG;
  }
  if (q->ld_sigma < q->n) {
    set_error("ld_sigma = %d < n = %d", q->ld_sigma, q->n);
    return MCR_E_ARG;
  }
  if (q->paths < 1) {
    set_error("paths = %lld must be >= 1", (long long)q->paths);
    return MCR_E_ARG;
  }
  if (!std::isfinite(q->x)) {
    set_error("x is not finite");
    return MCR_E_ARG;
  }
  if ((q->flags & MCR_F_VAR) && (q->n_alpha < 1 || q->n_alpha > MCR_MAX_ALPHA)) {
    set_error("n_alpha = %d outside [1, %d]", q->n_alpha, MCR_MAX_ALPHA);
    return MCR_E_ARG;
  }
  if (q->flags & MCR_F_VAR) {
    for (int i = 0; i < q->n_alpha; ++i)
      if (!(q->alpha[i] > 0.0 && q->alpha[i] < 1.0)) {
        set_error("alpha[%d] = %g outside (0, 1)", i, q->alpha[i]);
        return MCR_E_ARG;
      }
  }
  if ((q->flags & MCR_F_SHIST) && !q->s_hist) {
    set_error("MCR_F_SHIST needs s_hist");
    return MCR_E_ARG;
  }
  if (q->path_offset + (uint64_t)q->paths < q->path_offset) {
    set_error("path ids overflow 64 bits");
    return MCR_E_ARG;
  }
  return MCR_OK;
}

struct Staged {
  Plan plan;
  DevBuffers d;
  char* host;  // pinned staging; nullptr in the allocation-free mode
};

// Uploads inputs, factors and prepares on the device.  Leaves the context
// locked state untouched; caller holds g_mu.  With noalloc the device buffers
// are carved from the scratch array and the inputs are copied straight from
// the caller's arrays (no arena growth, no pinned staging).
int stage_query(const mcr_query* q, Staged* st, bool want_var_windows, bool noalloc = false) {
  Plan& p = st->plan;
  p.n = q->n;
  MCR_TRY(choose_kernel(q, &p));
  p.r8 = ((q->n + kTileRows - 1) / kTileRows) * kTileRows;
  p.n_alpha = (q->flags & MCR_F_VAR) ? q->n_alpha : 0;
  p.per_path = q->S_out || q->loss_out;
  p.chunk = p.per_path ? (size_t)std::min<uint64_t>((uint64_t)q->paths,
                                                    noalloc ? kNoAllocChunk : kPerPathChunk)
                       : 0;
  const size_t dev_bytes = layout(p, nullptr, &st->d);
  const size_t n = (size_t)q->n;
  if (noalloc) {
    if (q->n > kMaxN) {
      set_error("the allocation-free entry point takes n <= %d (n = %d)", kMaxN, q->n);
      return MCR_E_UNSUPPORTED;
    }
    if (dev_bytes > kScratchBytes) {
      set_error("query needs %zu bytes of device scratch, more than the %zu the "
                "allocation-free entry point holds", dev_bytes, kScratchBytes);
      return MCR_E_UNSUPPORTED;
    }
    layout(p, (char*)g_ctx.scratch, &st->d);
    st->host = nullptr;
    for (size_t i = 0; i < n; ++i) {
      const double wi = q->weights ? q->weights[i] : 1.0 / (double)n;
      if (!std::isfinite(q->mu[i]) || !std::isfinite(wi)) {
        set_error("mu or weights not finite at %zu", i);
        return MCR_E_ARG;
      }
      g_na_weights[i] = wi;
    }
    HIP_TRY(hipMemcpy2DAsync(st->d.A, n * sizeof(double), q->sigma,
                             (size_t)q->ld_sigma * sizeof(double), n * sizeof(double), n,
                             hipMemcpyHostToDevice, g_ctx.stream));
    HIP_TRY(hipMemcpyAsync(st->d.mu, q->mu, n * sizeof(double), hipMemcpyHostToDevice,
                           g_ctx.stream));
    HIP_TRY(hipMemcpyAsync(st->d.w, g_na_weights, n * sizeof(double), hipMemcpyHostToDevice,
                           g_ctx.stream));
  } else {
    MCR_TRY(reserve_arena(dev_bytes));
    layout(p, (char*)g_ctx.arena, &st->d);
    const size_t in_bytes = align_up((n * n + 2 * n) * sizeof(double));
    const size_t out_bytes = align_up(sizeof(DevInfo)) + align_up(sizeof(DevAccum)) +
                             align_up(kHistBins * sizeof(unsigned)) +
                             align_up((n + 1) * sizeof(unsigned long long));
    MCR_TRY(reserve_host(in_bytes + out_bytes));
    st->host = (char*)g_ctx.host;
    double* hA = (double*)st->host;
    double* hmu = hA + n * n;
    double* hw = hmu + n;
    for (size_t i = 0; i < n; ++i) {
      std::memcpy(hA + i * n, q->sigma + i * (size_t)q->ld_sigma, n * sizeof(double));
      hmu[i] = q->mu[i];
      hw[i] = q->weights ? q->weights[i] : 1.0 / (double)n;
    }
    for (size_t i = 0; i < n; ++i)
      if (!std::isfinite(hmu[i]) || !std::isfinite(hw[i])) {
        set_error("mu or weights not finite at %zu", i);
        return MCR_E_ARG;
      }
    HIP_TRY(hipMemcpyAsync(st->d.A, hA, n * n * sizeof(double), hipMemcpyHostToDevice,
                           g_ctx.stream));
    HIP_TRY(hipMemcpyAsync(st->d.mu, hmu, n * sizeof(double), hipMemcpyHostToDevice,
                           g_ctx.stream));
    HIP_TRY(hipMemcpyAsync(st->d.w, hw, n * sizeof(double), hipMemcpyHostToDevice,
                           g_ctx.stream));
  }

  FactorArgs fa = {};
  fa.A = st->d.A;
  fa.mu = st->d.mu;
  fa.w = st->d.w;
  fa.L = st->d.L;
  fa.v64 = st->d.v64;
  fa.n = q->n;
  set_windows(q, want_var_windows ? p.n_alpha : 0, &fa);
  fa.info = st->d.info;
  fa.acc = st->d.acc;
  fa.hist = st->d.hist;
  HIP_TRY(hipEventRecord(g_ctx.evs, g_ctx.stream));
  HIP_TRY(launch_factor(fa, g_ctx.stream));
  HIP_TRY(hipMemsetAsync(st->d.s_hist, 0, (n + 1) * sizeof(unsigned long long), g_ctx.stream));

  PrepArgs pa = {};
  pa.L = st->d.L;
  pa.mu = st->d.mu;
  pa.n = q->n;
  pa.T = p.T;
  pa.S = p.S;
  pa.info = st->d.info;
  pa.v64 = st->d.v64;
  pa.colpack = st->d.colpack;
  pa.L32 = st->d.L32;
  pa.mu32 = st->d.mu32;
  pa.v32 = st->d.v32;
  pa.lfrag = st->d.lfrag;
  pa.rowinit = st->d.rowinit;
  HIP_TRY(launch_prep(pa, g_ctx.stream));
  return MCR_OK;
}

int read_info(const Staged& st, DevInfo* info, DevAccum* acc) {
  if (!st.host) {  // allocation-free mode: straight into the caller's structs
    HIP_TRY(hipMemcpyAsync(info, st.d.info, sizeof(DevInfo), hipMemcpyDeviceToHost, g_ctx.stream));
    HIP_TRY(hipMemcpyAsync(acc, st.d.acc, sizeof(DevAccum), hipMemcpyDeviceToHost, g_ctx.stream));
    HIP_TRY(hipStreamSynchronize(g_ctx.stream));
    return MCR_OK;
  }
  char* out = st.host + align_up((size_t)st.plan.n * st.plan.n * sizeof(double) +
                                 2 * (size_t)st.plan.n * sizeof(double));
  DevInfo* hi = (DevInfo*)out;
  DevAccum* ha = (DevAccum*)(out + align_up(sizeof(DevInfo)));
  HIP_TRY(hipMemcpyAsync(hi, st.d.info, sizeof(DevInfo), hipMemcpyDeviceToHost, g_ctx.stream));
  HIP_TRY(hipMemcpyAsync(ha, st.d.acc, sizeof(DevAccum), hipMemcpyDeviceToHost, g_ctx.stream));
  HIP_TRY(hipStreamSynchronize(g_ctx.stream));
  *info = *hi;
  *acc = *ha;
  return MCR_OK;
}

int check_factor(const DevInfo& info) {
  if (info.attempts < 0) {
    set_error("covariance or mean has non-finite entries");
    return MCR_E_FACTOR;
  }
  if (info.fail_col != 0) {
    set_error("factorisation failed at column %d even with the clamp fallback", info.fail_col - 1);
    return MCR_E_FACTOR;
  }
  return MCR_OK;
}

PathArgs base_args(const mcr_query* q, const Staged& st) {
  PathArgs a = {};
  a.seed = q->seed;
  a.n = q->n;
  a.k = q->k;
  a.thr = (float)(-q->x);
  if (q->flags & MCR_F_NEAR) {
    a.thr_lo = std::nextafter(a.thr, -INFINITY);
    a.thr_hi = std::nextafter(a.thr, INFINITY);
  } else {
    a.thr_lo = 1.0f;
    a.thr_hi = 0.0f;  // lo > hi: near accounting off
  }
  a.colpack = st.d.colpack;
  a.L32 = st.d.L32;
  a.mu32 = st.d.mu32;
  a.v32 = st.d.v32;
  a.lfrag = st.d.lfrag;
  a.rowinit = st.d.rowinit;
  a.T = st.plan.T;
  a.S = st.plan.S;
  a.info = st.d.info;
  a.acc = st.d.acc;
  a.gather = st.d.gather;
  a.gather_cap = kGatherCap;
  a.hist = st.d.hist;
  return a;
}

// Runs every path of the query once with the given loss mode.  Per-path
// outputs (if any) are staged per chunk and copied to the caller's arrays.
int run_paths(const mcr_query* q, const Staged& st, PathArgs a, bool per_path, bool timed,
              float* ms) {
  const uint64_t M = (uint64_t)q->paths;
  const uint64_t step = per_path ? (uint64_t)st.plan.chunk : kLaunchPaths;
  if (timed) HIP_TRY(hipEventRecord(g_ctx.ev0, g_ctx.stream));
  for (uint64_t s = 0; s < M; s += step) {
    const uint64_t cnt = std::min<uint64_t>(step, M - s);
    a.path0 = q->path_offset + s;
    a.npaths = cnt;
    a.S_out = (per_path && q->S_out) ? st.d.S_out : nullptr;
    a.loss_out = (per_path && q->loss_out) ? st.d.loss_out : nullptr;
    HIP_TRY(launch_paths(st.plan.kernel, a, g_ctx.num_cus, g_ctx.stream));
    if (a.S_out)
      HIP_TRY(hipMemcpyAsync(q->S_out + s, st.d.S_out, cnt * sizeof(int32_t), hipMemcpyDeviceToHost,
                             g_ctx.stream));
    if (a.loss_out)
      HIP_TRY(hipMemcpyAsync(q->loss_out + s, st.d.loss_out, cnt * sizeof(float),
                             hipMemcpyDeviceToHost, g_ctx.stream));
    if (per_path) HIP_TRY(hipStreamSynchronize(g_ctx.stream));
  }
  if (timed) {
    HIP_TRY(hipEventRecord(g_ctx.ev1, g_ctx.stream));
    HIP_TRY(hipEventSynchronize(g_ctx.ev1));
    HIP_TRY(hipEventElapsedTime(ms, g_ctx.ev0, g_ctx.ev1));
  }
  return MCR_OK;
}

int64_t fixed32_host(float v) { return (int64_t)((double)v * 4294967296.0); }

// Whether a bracket pass's window (counts in acc slot `slot`) holds rank R
// with every in-window loss gathered (at most cap of them).
bool window_holds(int64_t M, const DevAccum& acc, int slot, unsigned cap, int64_t R) {
  const int64_t G = (int64_t)acc.gathered[slot];
  const int64_t below = M - (int64_t)acc.above[slot] - G;
  return G <= (int64_t)cap && below < R && R <= below + G;
}

// The quantile, tail sum and ES from a window that holds rank R, its G
// gathered losses in host memory g (reordered in place).  The (idx+1)-th
// smallest in-window loss is the quantile; the values after it in the
// partition are exactly the larger ones, and their fixed-point sum does not
// depend on order, so a partition gives what a sort would.
void resolve_window(int64_t M, const DevAccum& acc, int slot, float* g, int64_t R, double* var,
                    double* es, int64_t* tail) {
  const int64_t G = (int64_t)acc.gathered[slot];
  const int64_t below = M - (int64_t)acc.above[slot] - G;
  const int64_t idx = R - below - 1;
  std::nth_element(g, g + idx, g + G);
  int64_t t = acc.sum_above[slot];
  for (int64_t i = idx; i < G; ++i) t += fixed32_host(g[(size_t)i]);
  *var = (double)g[(size_t)idx];
  *tail = t;
  *es = (double)t * std::ldexp(1.0, -32) / (double)(M - R + 1);
}

// Resolves a quantile from a bracket pass (window counts in acc slot `slot`,
// in-window losses in gather[0 .. min(G, cap))); returns 1 when resolved, 0
// when the rank missed the window or the buffer overflowed, -1 on copy error.
int resolve_quantile(int64_t M, const DevAccum& acc, int slot, const float* gather, unsigned cap,
                     int64_t R, double* var, double* es, int64_t* tail, float* hostbuf) {
  if (!window_holds(M, acc, slot, cap, R)) return 0;
  const int64_t G = (int64_t)acc.gathered[slot];  // >= 1: the window holds rank R
  // hostbuf (>= cap floats) serves the allocation-free mode; otherwise a heap copy
  std::vector<float> heap(hostbuf ? 0 : (size_t)G);
  float* g = hostbuf ? hostbuf : heap.data();
  if (hipMemcpy(g, gather, (size_t)G * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return -1;
  resolve_window(M, acc, slot, g, R, var, es, tail);
  return 1;
}

int fetch_windows(const std::vector<WindowFetch>& w, float** hbuf, size_t* hcap,
                  hipStream_t stream) {
  size_t need = 0;
  for (const WindowFetch& f : w) need = std::max(need, f.off + (size_t)f.count);
  if (need == 0) return MCR_OK;
  if (need > *hcap) {
    if (*hbuf) HIP_TRY(hipHostFree(*hbuf));
    *hbuf = nullptr;
    *hcap = 0;
    const size_t cap = std::max<size_t>(need + need / 2, 1 << 16);
    HIP_TRY(hipHostMalloc((void**)hbuf, cap * sizeof(float), hipHostMallocDefault));
    *hcap = cap;
  }
  for (const WindowFetch& f : w)
    HIP_TRY(hipMemcpyAsync(*hbuf + f.off, f.dev, (size_t)f.count * sizeof(float),
                           hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  return MCR_OK;
}

int try_resolve(const mcr_query* q, const Staged& st, const DevAccum& acc, int slot, int64_t R,
                double* var, double* es, int64_t* tail) {
  const float* gd = st.d.gather + (size_t)slot * kGatherCap;
  // allocation-free mode: a synchronous copy into the static host buffer
  if (!st.host) return resolve_quantile(q->paths, acc, slot, gd, kGatherCap, R, var, es, tail,
                                        g_na_gather);
  // otherwise on the query's own stream into the pinned window buffer (a copy
  // on the null stream could wait behind another stream's kernels, such as a
  // serving round's, when the two share a hardware queue)
  if (!window_holds(q->paths, acc, slot, kGatherCap, R)) return 0;
  const std::vector<WindowFetch> w{{gd, (int64_t)acc.gathered[slot], 0}};
  if (fetch_windows(w, &g_ctx.hwin, &g_ctx.hwin_cap, g_ctx.stream) != MCR_OK) return -1;
  resolve_window(q->paths, acc, slot, g_ctx.hwin, R, var, es, tail);
  return 1;
}

int upload_windows(const Staged& st, const DevInfo& info) {
  HIP_TRY(hipStreamSynchronize(g_ctx.stream));
  HIP_TRY(hipMemcpy(st.d.info, &info, sizeof(DevInfo), hipMemcpyHostToDevice));
  return MCR_OK;
}

// Quantile refinement when the pass-1 bracket missed the rank (or overflowed
// the gather buffer): histogram passes narrow the interval, then a bracket
// pass on the chosen bins.  Each pass regenerates the same paths.
int refine_quantile(const mcr_query* q, const Staged& st, DevInfo info0, int al, int64_t R,
                    const DevAccum& acc1, double* var, double* es, int64_t* tail, int* passes) {
  const int64_t M = q->paths;
  const double sd = info0.loss_sd > 0 ? info0.loss_sd : 1e-30;
  const double center = -info0.mp;
  // interval known to contain rank R
  double a_lo, a_hi;
  const int64_t G = (int64_t)acc1.gathered[al], above = (int64_t)acc1.above[al];
  const int64_t below = M - above - G;
  if (R <= below) {
    a_lo = center - 64.0 * sd;
    a_hi = info0.lo[al];
  } else if (R > below + G) {
    a_lo = info0.hi[al];
    a_hi = center + 64.0 * sd;
  } else {
    a_lo = info0.lo[al];
    a_hi = info0.hi[al];
  }
  std::vector<unsigned> heap_hist(st.host ? kHistBins : 0);
  unsigned* hist = st.host ? heap_hist.data() : g_na_hist;
  for (int iter = 0; iter < 12; ++iter) {
    if (!(a_hi > a_lo)) a_hi = a_lo + 1e-6 * sd;
    DevInfo info = info0;
    info.hist_lo = (float)a_lo;
    info.hist_scale = (float)(kHistBins / (a_hi - a_lo));
    MCR_TRY(upload_windows(st, info));
    HIP_TRY(hipMemsetAsync(st.d.hist, 0, kHistBins * sizeof(unsigned), g_ctx.stream));
    HIP_TRY(hipMemsetAsync(st.d.acc, 0, sizeof(DevAccum), g_ctx.stream));
    PathArgs a = base_args(q, st);
    a.loss_mode = kLossHist;
    a.thr_lo = 1.0f;
    a.thr_hi = 0.0f;
    float ms = 0;
    MCR_TRY(run_paths(q, st, a, false, false, &ms));
    ++*passes;
    DevAccum acc;
    DevInfo tmp;
    MCR_TRY(read_info(st, &tmp, &acc));
    HIP_TRY(hipMemcpy(hist, st.d.hist, kHistBins * sizeof(unsigned), hipMemcpyDeviceToHost));
    const int64_t hb = (int64_t)acc.hist_below;
    if (R <= hb) {
      a_hi = a_lo;
      a_lo = a_lo - 64.0 * (a_hi - a_lo + sd);
      continue;
    }
    int64_t cum = hb;
    int bin = -1;
    for (int b = 0; b < kHistBins; ++b) {
      if (R <= cum + (int64_t)hist[b]) {
        bin = b;
        break;
      }
      cum += hist[b];
    }
    if (bin < 0) {
      a_lo = a_hi;
      a_hi = a_hi + 64.0 * (a_hi - a_lo + sd);
      continue;
    }
    // bracket pass on bins [bin-1, bin+1]
    const double w = (a_hi - a_lo) / kHistBins;
    const double lo = a_lo + (bin - 1) * w, hi = a_lo + (bin + 2) * w;
    int64_t in_bins = 0;
    for (int b = std::max(0, bin - 1); b <= std::min(kHistBins - 1, bin + 1); ++b) in_bins += hist[b];
    if (in_bins > (int64_t)kGatherCap) {
      a_lo = a_lo + bin * w;
      a_hi = a_lo + w;
      continue;
    }
    info = info0;
    for (int s = 0; s < kMaxAlpha; ++s) {
      info.lo[s] = (float)lo;
      info.hi[s] = (float)hi;
    }
    MCR_TRY(upload_windows(st, info));
    HIP_TRY(hipMemsetAsync(st.d.acc, 0, sizeof(DevAccum), g_ctx.stream));
    a = base_args(q, st);
    a.loss_mode = kLossBracket;
    a.n_alpha = 1;
    a.thr_lo = 1.0f;
    a.thr_hi = 0.0f;
    MCR_TRY(run_paths(q, st, a, false, false, &ms));
    ++*passes;
    MCR_TRY(read_info(st, &tmp, &acc));
    const int r = try_resolve(q, st, acc, 0, R, var, es, tail);
    if (r < 0) {
      set_error("copy of gathered losses failed");
      return MCR_E_DEVICE;
    }
    if (r == 1) return MCR_OK;
    a_lo = lo - w;
    a_hi = hi + w;
  }
  set_error("quantile refinement did not converge for alpha[%d]", al);
  return MCR_E_DEVICE;
}

int run_locked(const mcr_query* q, mcr_result* r, bool noalloc) {
  const auto t0 = std::chrono::steady_clock::now();
  MCR_TRY(validate(q));
  if (noalloc) {
    // initialisation creates the stream, events and the scratch address:
    // the allocation-free entry point only runs after mcr_init
    if (!g_ctx.ready) {
      set_error("mcr_run_noalloc needs mcr_init first (it creates nothing itself)");
      return MCR_E_ARG;
    }
  } else {
    MCR_TRY(ensure_init(g_ctx.device));
  }
  HIP_TRY(hipSetDevice(g_ctx.device));
  std::memset(r, 0, sizeof(*r));
  Staged st;
  MCR_TRY(stage_query(q, &st, true, noalloc));

  PathArgs a = base_args(q, st);
  a.loss_mode = (q->flags & MCR_F_VAR) ? kLossBracket : kLossNone;
  a.n_alpha = st.plan.n_alpha;
  a.s_hist = (q->flags & MCR_F_SHIST) ? st.d.s_hist : nullptr;
  float ms = 0.0f;
  MCR_TRY(run_paths(q, st, a, st.plan.per_path, true, &ms));
  DevInfo info;
  DevAccum acc;
  MCR_TRY(read_info(st, &info, &acc));
  MCR_TRY(check_factor(info));

  const int64_t M = q->paths;
  fill_statistics(M, info, acc, r);
  r->near_threshold = (int64_t)acc.near;
  r->kernel = st.plan.kernel;
  r->ms_kernel = ms;
  float ms_dev = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms_dev, g_ctx.evs, g_ctx.ev1));
  r->ms_device = ms_dev;
  r->var_passes = (q->flags & MCR_F_VAR) ? 1 : 0;

  if (q->flags & MCR_F_SHIST) {
    const size_t nb = (size_t)q->n + 1;
    std::vector<unsigned long long> heap(noalloc ? 0 : nb);
    unsigned long long* h = noalloc ? g_na_shist : heap.data();
    HIP_TRY(hipMemcpy(h, st.d.s_hist, nb * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nb; ++i) q->s_hist[i] = (int64_t)h[i];
  }

  if (q->flags & MCR_F_VAR) {
    for (int al = 0; al < st.plan.n_alpha; ++al) {
      const int64_t R = rank_of(q->alpha[al], M);
      int res = try_resolve(q, st, acc, al, R, &r->var[al], &r->es[al], &r->tail_sum_fixed[al]);
      if (res < 0) {
        set_error("copy of gathered losses failed");
        return MCR_E_DEVICE;
      }
      if (res == 0)
        MCR_TRY(refine_quantile(q, st, info, al, R, acc, &r->var[al], &r->es[al],
                                &r->tail_sum_fixed[al], &r->var_passes));
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  r->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return MCR_OK;
}


// VaR windows of a query: centre Phi^-1(alpha), half-width six standard errors
// of the empirical quantile (in loss sd units) plus 1e-4 sd for the fp32
// stream's departure from the exact Gaussian.
void set_windows(const mcr_query* q, int n_alpha, FactorArgs* f) {
  f->n_alpha = n_alpha;
  for (int al = 0; al < n_alpha; ++al) {
    const double av = q->alpha[al];
    const double za = inv_norm_cdf(av);
    const double dens = std::exp(-0.5 * za * za) / std::sqrt(2.0 * M_PI);
    f->za[al] = za;
    f->ha[al] = (q->flags & MCR_F_TEST_NARROW)
                    ? 0.0
                    : (q->flags & MCR_F_TEST_WIDE)
                          ? 1e6
                          : 6.0 * std::sqrt(av * (1.0 - av) / (double)q->paths) / dens + 1e-4;
  }
}

// Gather capacity per VaR level for a query whose results are copied back
// with its record: the expected number of in-window losses (the portfolio
// loss is Gaussian, the window is za +- ha standard deviations) plus eight
// standard deviations and 64, at most cap.  A window that still overflows is
// resolved by the refinement passes, as any missed window.
unsigned window_capacity(const mcr_query* q, unsigned cap) {
  const int na = (q->flags & MCR_F_VAR) ? q->n_alpha : 0;
  if (na <= 0) return 0;
  FactorArgs f = {};
  set_windows(q, na, &f);
  double need = 0.0;
  for (int al = 0; al < na; ++al) {
    const double frac = 0.5 * (std::erfc(-(f.za[al] + f.ha[al]) / std::sqrt(2.0)) -
                               std::erfc(-(f.za[al] - f.ha[al]) / std::sqrt(2.0)));
    const double eg = (double)q->paths * std::max(0.0, frac);
    need = std::max(need, eg + 8.0 * std::sqrt(eg) + 64.0);
  }
  return (unsigned)std::min<double>((double)cap, std::ceil(need));
}

// P-hat, its standard error and 95% interval (src/05-GPU/monte_carlos_cuda.cu:146-150),
// plus the factor outcome and loss moments.
void fill_statistics(int64_t M, const DevInfo& info, const DevAccum& acc, mcr_result* r) {
  r->count = (int64_t)acc.count;
  r->p_hat = (double)r->count / (double)M;
  r->std_error = std::sqrt(r->p_hat * (1.0 - r->p_hat) / (double)M);
  r->ci_lower = std::fmax(0.0, r->p_hat - 1.96 * r->std_error);
  r->ci_upper = std::fmin(1.0, r->p_hat + 1.96 * r->std_error);
  r->loss_mean = -info.mp;
  r->loss_sd = info.loss_sd;
  r->jitter = info.jitter;
  r->factor_attempts = info.attempts;
  r->factor_clamped = info.clamped;
}

// ---------------------------------------------------------------------------
// Batches: batched factor (one workgroup per query), batched prepare, then one
// persistent launch of k_batch_paths over (query, 64-path group run) work
// items, optionally captured once in a hipGraph and replayed.
// ---------------------------------------------------------------------------
// Tuned workgroup shape per row-tile range of the LDS MFMA batch kernel
// (index into the shapes of batch_big_kernel; 0 = the measured default).  The
// batch autotuner writes them; batches and serving rounds read them.
std::atomic<int> g_shape[kBigRanges] = {};

void tuned_shapes(int out[kBigRanges]) {
  for (int r = 0; r < kBigRanges; ++r) out[r] = g_shape[r].load(std::memory_order_relaxed);
}

// Class of an n-asset query: the register-operand MFMA kernel of its (NB, T)
// for n <= 127, else the LDS MFMA kernel of shape shapes[range of T] (the
// tuned shapes when shapes is null).
BatchKernel class_kernel(int n, const int* shapes) {
  const int T = (n + 1 + 31) / 32;
  if (n <= kSmallMaxN) return batch_small_kernel(n, T);
  const int r = T <= 4 ? 0 : big_range(T);
  return batch_big_kernel(T, shapes ? shapes[r] : g_shape[r].load(std::memory_order_relaxed));
}

int class_block(int n, const int* shapes) {
  const int T = (n + 1 + 31) / 32;
  if (n <= kSmallMaxN) return 256;
  const int r = T <= 4 ? 0 : big_range(T);
  return 64 * batch_big_waves(T, shapes ? shapes[r] : g_shape[r].load(std::memory_order_relaxed));
}

int class_grid(BatchKernel k, int block) {
  int cus = g_ctx.num_cus;
  if (cus <= 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 1;
  }
  return resident_blocks(reinterpret_cast<const void*>(k), block) * cus;
}

void order_by_class(std::vector<WorkItem>* items, const std::vector<BatchKernel>& qk,
                    const std::vector<int>& qb, const std::vector<double>& qcost, bool lpt,
                    std::vector<ClassRun>* runs) {
  // classes by descending total cost, then items longest first (or as given);
  // the keys are computed once per item, so a round of many thousand items
  // sorts in well under a millisecond
  std::map<BatchKernel, double> ccost;
  for (const WorkItem& w : *items) ccost[qk[w.query]] += qcost[w.query] * w.ngroups;
  std::vector<std::pair<double, BatchKernel>> order;
  for (auto& kv : ccost) order.push_back({kv.second, kv.first});
  std::sort(order.begin(), order.end(),
            [](const std::pair<double, BatchKernel>& a, const std::pair<double, BatchKernel>& b) {
              return a.first > b.first;
            });
  std::map<BatchKernel, int> rank;
  for (size_t i = 0; i < order.size(); ++i) rank[order[i].second] = (int)i;
  // per query: its class's rank (queries index qk / qb / qcost)
  std::vector<int> qrank(qk.size(), 0);
  for (size_t q = 0; q < qk.size(); ++q)
    if (qk[q]) qrank[q] = rank.count(qk[q]) ? rank[qk[q]] : 0;
  const size_t N = items->size();
  std::vector<std::pair<int, double>> key(N);
  for (size_t i = 0; i < N; ++i) {
    const WorkItem& w = (*items)[i];
    key[i] = {qrank[w.query], lpt ? -(qcost[w.query] * w.ngroups) : 0.0};
  }
  std::vector<uint32_t> idx(N);
  for (size_t i = 0; i < N; ++i) idx[i] = (uint32_t)i;
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
  std::vector<WorkItem> sorted(N);
  for (size_t i = 0; i < N; ++i) sorted[i] = (*items)[idx[i]];
  items->swap(sorted);
  runs->clear();
  for (size_t i = 0; i < N; ++i) {
    const uint32_t qi = (*items)[i].query;
    const BatchKernel k = qk[qi];
    if (runs->empty() || runs->back().k != k)
      runs->push_back({k, (uint32_t)i, 0u, class_grid(k, qb[qi]), qb[qi], ccost[k]});
    runs->back().count++;
  }
}

int create_forks(ForkSet* f, int n) {
  n = std::min(n, kMaxForks);
  HIP_TRY(hipEventCreateWithFlags(&f->fork, hipEventDisableTiming));
  for (int j = 0; j < n; ++j) {
    HIP_TRY(hipStreamCreateWithFlags(&f->s[j], hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&f->join[j], hipEventDisableTiming));
    f->n = j + 1;
  }
  return MCR_OK;
}

void destroy_forks(ForkSet* f) {
  for (int j = 0; j < kMaxForks; ++j) {
    if (f->s[j]) (void)hipStreamDestroy(f->s[j]);
    if (f->join[j]) (void)hipEventDestroy(f->join[j]);
  }
  if (f->fork) (void)hipEventDestroy(f->fork);
  *f = ForkSet();
}

int launch_runs(const std::vector<ClassRun>& runs, const PathArgs* qs, const WorkItem* items,
                unsigned* next, int grid_override, hipStream_t stream, const ForkSet* forks,
                int nfork) {
  if (runs.empty()) return MCR_OK;
  HIP_TRY(launch_batch_clear(next, (int)runs.size(), stream));
  const int k = forks ? std::min(nfork, forks->n) : 1;
  double load[kMaxForks] = {};
  bool used[kMaxForks] = {};
  if (k > 1) HIP_TRY(hipEventRecord(forks->fork, stream));
  // runs come in descending class cost: each goes to the least-loaded stream
  for (size_t c = 0; c < runs.size(); ++c) {
    BatchArgs b;
    b.qs = qs;
    b.items = items + runs[c].off;
    b.nitems = runs[c].count;
    b.next = next + c;
    hipStream_t st = stream;
    if (k > 1) {
      int j = 0;
      for (int i = 1; i < k; ++i)
        if (load[i] < load[j]) j = i;
      if (!used[j]) HIP_TRY(hipStreamWaitEvent(forks->s[j], forks->fork, 0));
      used[j] = true;
      load[j] += runs[c].cost;
      st = forks->s[j];
    }
    HIP_TRY(launch_batch_kernel(runs[c].k, b, grid_override > 0 ? grid_override : runs[c].grid,
                                runs[c].block, st));
  }
  for (int j = 0; j < k; ++j) {
    if (!used[j]) continue;
    HIP_TRY(hipEventRecord(forks->join[j], forks->s[j]));
    HIP_TRY(hipStreamWaitEvent(stream, forks->join[j], 0));
  }
  return MCR_OK;
}

struct Tuned {
  int chunk_paths = 2048;
  int order = MCR_ORDER_LPT;
  int streams = 3;  // measured: 256 uneven portfolios 125.4 (1) -> 121.5 ms (3)
  bool tuned = false;
};
Tuned g_tune;

struct GraphCache {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t signature = 0;
};
GraphCache g_graph;

void drop_graph() {
  if (g_graph.exec) (void)hipGraphExecDestroy(g_graph.exec);
  if (g_graph.graph) (void)hipGraphDestroy(g_graph.graph);
  g_graph = GraphCache();
}

uint64_t mix64(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h;
}

struct BatchPlan {
  int nq = 0, max_n = 0;
  std::vector<int> T, S, na;
  std::vector<WorkItem> items;
  std::vector<ClassRun> runs;  // contiguous item ranges, one persistent launch each
  // device regions
  size_t in_bytes = 0, out_bytes = 0, total = 0;
  // offsets within the input region
  std::vector<size_t> offA, offMu, offW;
  size_t offFa = 0, offPa = 0, offQa = 0, offItems = 0;
  // offsets of the output region (infos, accs), then scratch
  size_t offInfo = 0, offAcc = 0;
  std::vector<size_t> offL, offV, offFrag, offInit, offGather;
  size_t offNext = 0, offHist = 0;
};

// The last persistent batch's plan and the shape it was made for.
struct PlanCache {
  bool valid = false;
  std::vector<int64_t> key;
  BatchPlan bp;
};
PlanCache g_plan;

// Cost of one 64-path group of an n-asset query on the LDS MFMA kernel:
// S steps of RNG plus sum_t steps(t) tile-steps of MFMA.
double group_cost(int T, int S) {
  double tiles = 0.0;
  for (int t = 0; t < T; ++t) tiles += (t == T - 1) ? S : std::min(2 * t + 2, S);
  return 4.0 * S + 1.5 * tiles;
}

// Whether query i can run in a persistent batch (the checks every call makes,
// whether or not its plan is reused).
int check_batch_query(const mcr_query* q, int i) {
  int r = validate(q);
  if (r != MCR_OK) {
    char msg[512];
    snprintf(msg, sizeof(msg), "query %d: %s", i, g_err);
    set_error("%s", msg);
    return r;
  }
  if (q->precision != MCR_PREC_BF16X3 && q->precision != MCR_PREC_AUTO) {
    set_error("query %d: persistent batches run the bf16x3 MFMA kernels (precision %d)", i,
              q->precision);
    return MCR_E_UNSUPPORTED;
  }
  if (q->n > kBatchMaxN) {
    set_error("query %d: batches take n <= %d (n = %d)", i, kBatchMaxN, q->n);
    return MCR_E_UNSUPPORTED;
  }
  if (q->S_out || q->loss_out || (q->flags & (MCR_F_SHIST | MCR_F_NEAR))) {
    set_error("query %d: per-path outputs, S histograms and near counts are single-query options",
              i);
    return MCR_E_UNSUPPORTED;
  }
  return MCR_OK;
}

// What a batch plan depends on: item size, order, workgroup shapes and each
// query's n, path count and number of VaR levels.
std::vector<int64_t> plan_key(const mcr_query* qs, int nq, int chunk_paths, int order,
                              const int* shapes) {
  std::vector<int64_t> k;
  k.reserve(3 + kBigRanges + 3 * (size_t)nq);
  k.push_back(nq);
  k.push_back(chunk_paths);
  k.push_back(order);
  for (int r = 0; r < kBigRanges; ++r) k.push_back(shapes[r]);
  for (int i = 0; i < nq; ++i) {
    k.push_back(qs[i].n);
    k.push_back(qs[i].paths);
    k.push_back((qs[i].flags & MCR_F_VAR) ? qs[i].n_alpha : 0);
  }
  return k;
}

int plan_batch(const mcr_query* qs, int nq, int chunk_paths, int order, const int* shapes,
               BatchPlan* bp) {
  bp->nq = nq;
  bp->T.resize(nq);
  bp->S.resize(nq);
  bp->na.resize(nq);
  for (int i = 0; i < nq; ++i) {
    const mcr_query* q = &qs[i];
    MCR_TRY(check_batch_query(q, i));
    bp->T[i] = (q->n + 1 + 31) / 32;
    bp->S[i] = (q->n + 15) / 16;
    bp->na[i] = (q->flags & MCR_F_VAR) ? q->n_alpha : 0;
    bp->max_n = std::max(bp->max_n, q->n);
  }
  // work items: per class, runs of cg groups, with at least four items per
  // resident workgroup (small classes: per wave) so the class's launch balances
  std::vector<BatchKernel> qk(nq);
  std::vector<int> qb(nq);
  std::vector<double> qcost(nq);
  std::map<BatchKernel, uint64_t> class_groups;
  for (int i = 0; i < nq; ++i) {
    qk[i] = class_kernel(qs[i].n, shapes);
    qb[i] = class_block(qs[i].n, shapes);
    qcost[i] = group_cost(bp->T[i], bp->S[i]);
    class_groups[qk[i]] += ((uint64_t)qs[i].paths + 63) / 64;
  }
  const uint64_t cg = (uint64_t)std::max(1, chunk_paths / 64);
  bp->items.clear();
  for (int i = 0; i < nq; ++i) {
    const uint64_t slots = (uint64_t)class_grid(qk[i], qb[i]) * (qs[i].n <= kSmallMaxN ? 4 : 1);
    const uint64_t cgi = std::max<uint64_t>(1, std::min(cg, class_groups[qk[i]] / (4 * slots)));
    const uint64_t G = ((uint64_t)qs[i].paths + 63) / 64;
    for (uint64_t f = 0; f < G; f += cgi) {
      WorkItem w;
      w.query = (uint32_t)i;
      w.first_group = f;
      w.ngroups = (uint32_t)std::min<uint64_t>(cgi, G - f);
      bp->items.push_back(w);
    }
  }
  order_by_class(&bp->items, qk, qb, qcost, order == MCR_ORDER_LPT, &bp->runs);
  // layout: [inputs | outputs | scratch]
  Carver c{nullptr};
  bp->offA.resize(nq);
  bp->offMu.resize(nq);
  bp->offW.resize(nq);
  for (int i = 0; i < nq; ++i) {
    const size_t n = (size_t)qs[i].n;
    bp->offA[i] = c.off;
    c.take<double>(n * n);
    bp->offMu[i] = c.off;
    c.take<double>(n);
    bp->offW[i] = c.off;
    c.take<double>(n);
  }
  bp->offFa = c.off;
  c.take<FactorArgs>(nq);
  bp->offPa = c.off;
  c.take<PrepArgs>(nq);
  bp->offQa = c.off;
  c.take<PathArgs>(nq);
  bp->offItems = c.off;
  c.take<WorkItem>(bp->items.size());
  bp->in_bytes = c.off;
  bp->offInfo = c.off;
  c.take<DevInfo>(nq);
  bp->offAcc = c.off;
  c.take<DevAccum>(nq);
  bp->out_bytes = c.off - bp->in_bytes;
  bp->offL.resize(nq);
  bp->offV.resize(nq);
  bp->offFrag.resize(nq);
  bp->offInit.resize(nq);
  bp->offGather.resize(nq);
  for (int i = 0; i < nq; ++i) {
    const size_t n = (size_t)qs[i].n;
    bp->offL[i] = c.off;
    c.take<double>(n * n);
    bp->offV[i] = c.off;
    c.take<double>(n);
    bp->offFrag[i] = c.off;
    c.take<uint4>((size_t)bp->T[i] * bp->S[i] * 128);
    bp->offInit[i] = c.off;
    c.take<float>((size_t)32 * bp->T[i]);
    bp->offGather[i] = c.off;
    c.take<float>((size_t)std::max(bp->na[i], 1) * kBatchGatherCap);
  }
  bp->offNext = c.off;
  c.take<unsigned>(bp->runs.size());
  bp->offHist = c.off;
  c.take<unsigned>(kHistBins);
  bp->total = c.off;
  return MCR_OK;
}

// Fills the pinned image of the input region (host) with inputs and args.
void fill_batch_inputs(const mcr_query* qs, const BatchPlan& bp, char* h, char* dev) {
  const int nq = bp.nq;
  FactorArgs* fa = (FactorArgs*)(h + bp.offFa);
  PrepArgs* pa = (PrepArgs*)(h + bp.offPa);
  PathArgs* qa = (PathArgs*)(h + bp.offQa);
  for (int i = 0; i < nq; ++i) {
    const mcr_query* q = &qs[i];
    const size_t n = (size_t)q->n;
    double* A = (double*)(h + bp.offA[i]);
    double* mu = (double*)(h + bp.offMu[i]);
    double* w = (double*)(h + bp.offW[i]);
    for (size_t r = 0; r < n; ++r) {
      std::memcpy(A + r * n, q->sigma + r * (size_t)q->ld_sigma, n * sizeof(double));
      mu[r] = q->mu[r];
      w[r] = q->weights ? q->weights[r] : 1.0 / (double)n;
    }
    DevInfo* info = (DevInfo*)(dev + bp.offInfo) + i;
    DevAccum* acc = (DevAccum*)(dev + bp.offAcc) + i;
    FactorArgs f = {};
    f.A = (const double*)(dev + bp.offA[i]);
    f.mu = (const double*)(dev + bp.offMu[i]);
    f.w = (const double*)(dev + bp.offW[i]);
    f.L = (double*)(dev + bp.offL[i]);
    f.v64 = (double*)(dev + bp.offV[i]);
    f.n = q->n;
    set_windows(q, bp.na[i], &f);
    f.info = info;
    f.acc = acc;
    f.hist = (unsigned*)(dev + bp.offHist);
    fa[i] = f;
    PrepArgs pr = {};
    pr.L = f.L;
    pr.mu = f.mu;
    pr.n = q->n;
    pr.T = bp.T[i];
    pr.S = bp.S[i];
    pr.info = info;
    pr.v64 = f.v64;
    pr.lfrag = (uint4*)(dev + bp.offFrag[i]);
    pr.rowinit = (float*)(dev + bp.offInit[i]);
    pa[i] = pr;
    PathArgs pp = {};
    pp.seed = q->seed;
    pp.path0 = q->path_offset;
    pp.npaths = (uint64_t)q->paths;
    pp.n = q->n;
    pp.k = q->k;
    pp.thr = (float)(-q->x);
    pp.thr_lo = 1.0f;
    pp.thr_hi = 0.0f;
    pp.lfrag = pr.lfrag;
    pp.rowinit = pr.rowinit;
    pp.T = pr.T;
    pp.S = pr.S;
    pp.info = info;
    pp.acc = acc;
    pp.loss_mode = bp.na[i] ? kLossBracket : kLossNone;
    pp.n_alpha = bp.na[i];
    pp.gather = (float*)(dev + bp.offGather[i]);
    pp.gather_cap = kBatchGatherCap;
    pp.hist = f.hist;
    qa[i] = pp;
  }
  std::memcpy(h + bp.offItems, bp.items.data(), bp.items.size() * sizeof(WorkItem));
}

uint64_t batch_signature(const mcr_query* qs, const BatchPlan& bp, int grid, int chunk, int order,
                         int streams, const int* shapes) {
  uint64_t h = mix64(0, (uint64_t)(uintptr_t)g_ctx.arena);
  h = mix64(h, (uint64_t)streams);
  for (int r = 0; r < kBigRanges; ++r) h = mix64(h, (uint64_t)shapes[r]);
  h = mix64(h, g_ctx.arena_gen);
  h = mix64(h, (uint64_t)(uintptr_t)g_ctx.host);  // the captured copy targets it
  h = mix64(h, (uint64_t)bp.nq);
  h = mix64(h, (uint64_t)grid);
  h = mix64(h, (uint64_t)chunk);
  h = mix64(h, (uint64_t)order);
  h = mix64(h, bp.total);
  for (int i = 0; i < bp.nq; ++i) {
    h = mix64(h, (uint64_t)qs[i].n);
    h = mix64(h, (uint64_t)qs[i].paths);
    h = mix64(h, (uint64_t)bp.na[i]);
  }
  return h;
}

// Device work of a batch on g_ctx.stream (the part a graph captures).
int enqueue_batch(const BatchPlan& bp, char* dev, char* hout, int grid, int streams) {
  const FactorArgs* fa = (const FactorArgs*)(dev + bp.offFa);
  const PrepArgs* pa = (const PrepArgs*)(dev + bp.offPa);
  HIP_TRY(launch_factor_batch(fa, bp.nq, bp.max_n, g_ctx.stream));
  HIP_TRY(launch_prep_batch(pa, bp.nq, bp.max_n, g_ctx.stream));
  MCR_TRY(launch_runs(bp.runs, (const PathArgs*)(dev + bp.offQa),
                      (const WorkItem*)(dev + bp.offItems), (unsigned*)(dev + bp.offNext), grid,
                      g_ctx.stream, &g_ctx.forks, streams));
  HIP_TRY(hipMemcpyAsync(hout, dev + bp.offInfo, bp.out_bytes, hipMemcpyDeviceToHost, g_ctx.stream));
  return MCR_OK;
}

int run_batch_locked(const mcr_query* qs, int nq, mcr_result* rs, mcr_batch_opts* o) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!qs || !rs || nq < 1) {
    set_error("batch needs queries, results and nq >= 1");
    return MCR_E_ARG;
  }
  MCR_TRY(ensure_init(g_ctx.device));
  HIP_TRY(hipSetDevice(g_ctx.device));
  if (o->mode == MCR_BATCH_PER_QUERY) {
    double dev_ms = 0.0;
    for (int i = 0; i < nq; ++i) {
      int r = run_locked(&qs[i], &rs[i]);
      if (r != MCR_OK) return r;
      // factor, layouts and path kernels of the query (HIP events), the
      // same device work the persistent mode times
      dev_ms += rs[i].ms_device;
    }
    o->ms_device = dev_ms;
    o->ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return MCR_OK;
  }
  if (o->mode != MCR_BATCH_PERSISTENT) {
    set_error("unknown batch mode %d", o->mode);
    return MCR_E_ARG;
  }
  const int chunk = o->chunk_paths > 0 ? o->chunk_paths : g_tune.chunk_paths;
  const int order = o->chunk_paths > 0 ? o->order : g_tune.order;
  const int streams = std::min(o->streams > 0 ? o->streams : g_tune.streams, kMaxForks);
  int shapes[kBigRanges];
  tuned_shapes(shapes);
  for (int r = 0; r < kBigRanges; ++r)
    if (o->shapes[r] > 0) shapes[r] = std::min(o->shapes[r] - 1, batch_big_shapes(r) - 1);
  if (chunk % 64 != 0) {
    set_error("chunk_paths = %d is not a multiple of 64", chunk);
    return MCR_E_ARG;
  }
  using SClock = std::chrono::steady_clock;
  auto ms_between = [](SClock::time_point a, SClock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  const SClock::time_point tp0 = SClock::now();
  // The plan (work items, class launches, layout) depends only on the batch's
  // shape: a call that repeats the previous call's shape reuses it, as it
  // replays its graph, and only checks the queries (a 256-portfolio batch
  // plans 581,340 items in about 10 ms).
  std::vector<int64_t> key = plan_key(qs, nq, chunk, order, shapes);
  if (g_plan.valid && g_plan.key == key) {
    for (int i = 0; i < nq; ++i) MCR_TRY(check_batch_query(&qs[i], i));
  } else {
    g_plan.valid = false;
    g_plan.bp = BatchPlan();
    MCR_TRY(plan_batch(qs, nq, chunk, order, shapes, &g_plan.bp));
    g_plan.key.swap(key);
    g_plan.valid = true;
  }
  const BatchPlan& bp = g_plan.bp;
  const SClock::time_point tp1 = SClock::now();
  const int grid = o->grid;  // 0: each class kernel at its own occupancy
  MCR_TRY(reserve_arena(bp.total));
  MCR_TRY(reserve_host(bp.in_bytes + bp.out_bytes));
  char* dev = (char*)g_ctx.arena;
  char* hin = (char*)g_ctx.host;
  char* hout = hin + bp.in_bytes;
  fill_batch_inputs(qs, bp, hin, dev);
  HIP_TRY(hipMemcpyAsync(dev, hin, bp.in_bytes, hipMemcpyHostToDevice, g_ctx.stream));
  HIP_TRY(hipEventRecord(g_ctx.ev0, g_ctx.stream));
  o->graph_replayed = 0;
  if (o->use_graph) {
    const uint64_t sig = batch_signature(qs, bp, grid, chunk, order, streams, shapes);
    if (!g_graph.exec || g_graph.signature != sig) {
      drop_graph();
      HIP_TRY(hipStreamBeginCapture(g_ctx.stream, hipStreamCaptureModeThreadLocal));
      const int r = enqueue_batch(bp, dev, hout, grid, streams);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(g_ctx.stream, &graph);
      if (r != MCR_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
      }
      HIP_TRY(e);
      g_graph.graph = graph;
      HIP_TRY(hipGraphInstantiate(&g_graph.exec, graph, nullptr, nullptr, 0));
      g_graph.signature = sig;
    } else {
      o->graph_replayed = 1;
    }
    HIP_TRY(hipGraphLaunch(g_graph.exec, g_ctx.stream));
  } else {
    MCR_TRY(enqueue_batch(bp, dev, hout, grid, streams));
  }
  HIP_TRY(hipEventRecord(g_ctx.ev1, g_ctx.stream));
  const SClock::time_point tp2 = SClock::now();
  HIP_TRY(hipEventSynchronize(g_ctx.ev1));
  const SClock::time_point tp3 = SClock::now();
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1));
  o->ms_device = ms;
  o->chunk_used = chunk;
  o->order_used = order;
  o->streams_used = streams;
  for (int r = 0; r < kBigRanges; ++r) o->big_shapes[r] = shapes[r];
  o->work_items = (int64_t)bp.items.size();

  const DevInfo* infos = (const DevInfo*)hout;
  const DevAccum* accs = (const DevAccum*)(hout + (bp.offAcc - bp.offInfo));
  std::vector<int> refine;
  // every window that holds its rank is copied to host memory together (one
  // wait), then resolved; a query with a missed window is recomputed below
  struct Pending {
    int query, level;
    int64_t R;
  };
  std::vector<WindowFetch> fetch;
  std::vector<Pending> pend;
  size_t off = 0;
  for (int i = 0; i < nq; ++i) {
    const mcr_query* q = &qs[i];
    mcr_result* r = &rs[i];
    std::memset(r, 0, sizeof(*r));
    const int fr = check_factor(infos[i]);
    if (fr != MCR_OK) {
      char msg[512];
      snprintf(msg, sizeof(msg), "query %d: %s", i, g_err);
      set_error("%s", msg);
      return fr;
    }
    const int64_t M = q->paths;
    fill_statistics(M, infos[i], accs[i], r);
    r->kernel = q->n <= kSmallMaxN ? MCR_KERNEL_MFMA : MCR_KERNEL_MFMA_LDS;
    r->ms_kernel = ms;
    r->var_passes = bp.na[i] ? 1 : 0;
    bool ok = true;
    for (int al = 0; al < bp.na[i]; ++al)
      ok = ok && window_holds(M, accs[i], al, kBatchGatherCap, rank_of(q->alpha[al], M));
    if (!ok) {
      refine.push_back(i);
      continue;
    }
    for (int al = 0; al < bp.na[i]; ++al) {
      const int64_t G = (int64_t)accs[i].gathered[al];
      fetch.push_back({(const float*)(dev + bp.offGather[i]) + (size_t)al * kBatchGatherCap, G, off});
      pend.push_back({i, al, rank_of(q->alpha[al], M)});
      off += (size_t)G;
    }
  }
  MCR_TRY(fetch_windows(fetch, &g_ctx.hwin, &g_ctx.hwin_cap, g_ctx.stream));
  for (size_t k = 0; k < pend.size(); ++k) {
    const Pending& pw = pend[k];
    mcr_result* r = &rs[pw.query];
    resolve_window(qs[pw.query].paths, accs[pw.query], pw.level, g_ctx.hwin + fetch[k].off, pw.R,
                   &r->var[pw.level], &r->es[pw.level], &r->tail_sum_fixed[pw.level]);
  }
  // Rare: a quantile missed its window.  Recompute that query alone on the
  // same kernel (identical per-path losses) with the refinement passes.
  for (int i : refine) {
    mcr_query q = qs[i];
    q.kernel = MCR_KERNEL_AUTO;
    q.precision = MCR_PREC_BF16X3;
    mcr_result one;
    MCR_TRY(run_locked(&q, &one));
    for (int al = 0; al < bp.na[i]; ++al) {
      rs[i].var[al] = one.var[al];
      rs[i].es[al] = one.es[al];
      rs[i].tail_sum_fixed[al] = one.tail_sum_fixed[al];
    }
    rs[i].var_passes = 1 + one.var_passes;
  }
  const SClock::time_point tp4 = SClock::now();
  o->host_plan_ms = ms_between(tp0, tp1);
  o->host_stage_ms = ms_between(tp1, tp2);
  o->host_wait_ms = ms_between(tp2, tp3);
  o->host_finalize_ms = ms_between(tp3, tp4);
  o->ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < nq; ++i) rs[i].ms_total = o->ms_total;
  return MCR_OK;
}

}  // namespace host
}  // namespace mcr

using namespace mcr::host;

extern "C" {

int mcr_version(void) { return 1; }

const char* mcr_last_error(void) { return g_err; }

int mcr_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ctx.ready && g_ctx.device != device) {
    set_error("library already initialised on device %d", g_ctx.device);
    return MCR_E_ARG;
  }
  return ensure_init(device);
}

int mcr_shutdown(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx.ready) return MCR_OK;
  HIP_TRY(hipStreamSynchronize(g_ctx.stream));
  drop_graph();
  g_plan = PlanCache();
  if (g_ctx.arena) HIP_TRY(hipFree(g_ctx.arena));
  if (g_ctx.host) HIP_TRY(hipHostFree(g_ctx.host));
  if (g_ctx.hwin) HIP_TRY(hipHostFree(g_ctx.hwin));
  HIP_TRY(hipEventDestroy(g_ctx.ev0));
  HIP_TRY(hipEventDestroy(g_ctx.ev1));
  HIP_TRY(hipEventDestroy(g_ctx.evs));
  destroy_forks(&g_ctx.forks);
  HIP_TRY(hipStreamDestroy(g_ctx.stream));
  g_ctx = Context();
  return MCR_OK;
}

int mcr_run(const mcr_query* q, mcr_result* r) {
  if (!r) {
    set_error("result is NULL");
    return MCR_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  return run_locked(q, r);
}

int mcr_run_noalloc(const mcr_query* q, mcr_result* r) {
  if (!r) {
    set_error("result is NULL");
    return MCR_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  return run_locked(q, r, true);
}

int mcr_prepare(const mcr_query* q, mcr_model_export* out) {
  if (!out) {
    set_error("export is NULL");
    return MCR_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  MCR_TRY(validate(q));
  MCR_TRY(ensure_init(g_ctx.device));
  HIP_TRY(hipSetDevice(g_ctx.device));
  mcr_query qq = *q;
  qq.S_out = nullptr;
  qq.loss_out = nullptr;
  Staged st;
  MCR_TRY(stage_query(&qq, &st, false));
  DevInfo info;
  DevAccum acc;
  MCR_TRY(read_info(st, &info, &acc));
  MCR_TRY(check_factor(info));
  const int n = q->n;
  if (out->L64)
    HIP_TRY(hipMemcpy(out->L64, st.d.L, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost));
  if (out->L32) {
    std::vector<float> l((size_t)st.plan.r8 * st.plan.r8);
    HIP_TRY(hipMemcpy(l.data(), st.d.L32, l.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) out->L32[(size_t)i * n + j] = l[l32_offset(st.plan.r8, i, j)];
  }
  if (out->mu32) HIP_TRY(hipMemcpy(out->mu32, st.d.mu32, n * sizeof(float), hipMemcpyDeviceToHost));
  if (out->v32) HIP_TRY(hipMemcpy(out->v32, st.d.v32, n * sizeof(float), hipMemcpyDeviceToHost));
  out->mp32 = (float)info.mp;
  out->thr32 = (float)(-q->x);
  out->loss_sd = info.loss_sd;
  return MCR_OK;
}

int mcr_factor(int n, const double* sigma, int ld_sigma, double* L_out, mcr_factor_info* info) {
  if (!L_out || !sigma) {
    set_error("sigma and L_out are required");
    return MCR_E_ARG;
  }
  std::vector<double> mu((size_t)std::max(n, 1), 0.0);
  mcr_query q = {};
  q.n = n;
  q.mu = mu.data();
  q.sigma = sigma;
  q.ld_sigma = ld_sigma;
  q.paths = 1;
  q.precision = MCR_PREC_FP32;
  q.kernel = MCR_KERNEL_AUTO;
  mcr_model_export ex = {};
  ex.L64 = L_out;
  std::lock_guard<std::mutex> lk(g_mu);
  MCR_TRY(validate(&q));
  MCR_TRY(ensure_init(g_ctx.device));
  HIP_TRY(hipSetDevice(g_ctx.device));
  Staged st;
  MCR_TRY(stage_query(&q, &st, false));
  DevInfo di;
  DevAccum acc;
  MCR_TRY(read_info(st, &di, &acc));
  MCR_TRY(check_factor(di));
  HIP_TRY(hipMemcpy(L_out, st.d.L, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost));
  if (info) {
    info->attempts = di.attempts;
    info->jitter = di.jitter;
    info->clamped = di.clamped;
  }
  return MCR_OK;
}

int mcr_normals(uint64_t seed, uint64_t path0, int64_t npaths, int n, float* out) {
  if (!out || npaths < 0 || n < 1) {
    set_error("bad arguments to mcr_normals");
    return MCR_E_ARG;
  }
  if (npaths == 0) return MCR_OK;
  std::lock_guard<std::mutex> lk(g_mu);
  MCR_TRY(ensure_init(g_ctx.device));
  HIP_TRY(hipSetDevice(g_ctx.device));
  const size_t bytes = (size_t)npaths * n * sizeof(float);
  MCR_TRY(reserve_arena(bytes));
  HIP_TRY(launch_normals(seed, path0, (uint64_t)npaths, n, (float*)g_ctx.arena, g_ctx.stream));
  HIP_TRY(hipMemcpyAsync(out, g_ctx.arena, bytes, hipMemcpyDeviceToHost, g_ctx.stream));
  HIP_TRY(hipStreamSynchronize(g_ctx.stream));
  return MCR_OK;
}


int mcr_run_batch(const mcr_query* qs, int nq, mcr_result* rs, mcr_batch_opts* opts) {
  mcr_batch_opts local = {};
  mcr_batch_opts* o = opts ? opts : &local;
  std::lock_guard<std::mutex> lk(g_mu);
  return run_batch_locked(qs, nq, rs, o);
}

int mcr_autotune_batch(const mcr_query* qs, int nq, mcr_batch_opts* best, int* n_candidates) {
  // Staged search on the caller's batch, device time by HIP events: work-item
  // size x item order with the class launches on one stream, then the number
  // of concurrent class-launch streams at the best of those, then the
  // workgroup shape of each row-tile range the batch uses.
  static const int kChunks[] = {512, 2048, 8192, 32768};
  static const int kOrders[] = {MCR_ORDER_LPT, MCR_ORDER_NATURAL};
  static const int kStreams[] = {2, 3, 4};
  std::lock_guard<std::mutex> lk(g_mu);
  std::vector<mcr_result> rs((size_t)std::max(nq, 1));
  double best_ms = 1e300;
  mcr_batch_opts pick = {};
  int tried = 0;
  int cur[kBigRanges];
  tuned_shapes(cur);
  auto trial = [&](int chunk, int order, int streams) -> int {
    mcr_batch_opts o = {};
    o.mode = MCR_BATCH_PERSISTENT;
    o.chunk_paths = chunk;
    o.order = order;
    o.streams = streams;
    for (int r = 0; r < kBigRanges; ++r) o.shapes[r] = cur[r] + 1;
    MCR_TRY(run_batch_locked(qs, nq, rs.data(), &o));
    ++tried;
    if (o.ms_device < best_ms) {
      best_ms = o.ms_device;
      pick = o;
    }
    return MCR_OK;
  };
  for (int c : kChunks)
    for (int od : kOrders) MCR_TRY(trial(c, od, 1));
  const int c0 = pick.chunk_used, o0 = pick.order_used;
  for (int st : kStreams) MCR_TRY(trial(c0, o0, st));
  // then, for each row-tile range the batch uses, its alternative workgroup
  // shapes (one range at a time, the others at their best so far)
  bool in_range[kBigRanges] = {};
  for (int i = 0; i < nq; ++i) {
    const int T = (qs[i].n + 1 + 31) / 32;
    if (qs[i].n > kSmallMaxN && T > 4) in_range[big_range(T)] = true;
  }
  for (int r = 0; r < kBigRanges; ++r) {
    if (!in_range[r]) continue;
    const int keep = pick.big_shapes[r];
    for (int sh = 0; sh < batch_big_shapes(r); ++sh) {
      if (sh == keep) continue;
      for (int q = 0; q < kBigRanges; ++q) cur[q] = pick.big_shapes[q];
      cur[r] = sh;
      MCR_TRY(trial(pick.chunk_used, pick.order_used, pick.streams_used));
    }
  }
  for (int r = 0; r < kBigRanges; ++r) g_shape[r].store(pick.big_shapes[r]);
  g_tune.chunk_paths = pick.chunk_used;
  g_tune.order = pick.order_used;
  g_tune.streams = pick.streams_used;
  g_tune.tuned = true;
  if (best) *best = pick;
  if (n_candidates) *n_candidates = tried;
  return MCR_OK;
}

}  // extern "C"
