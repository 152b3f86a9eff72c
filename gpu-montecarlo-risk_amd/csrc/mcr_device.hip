// mcr_device.hip -- gfx950 kernels of the Monte Carlo crash-risk model.
//
// Replaces monte_carlo_kernel (src/05-GPU/monte_carlos_cuda.cu:23-72) and the host
// Cholesky it relies on (:84-93).  Per path p the reference computes
//   Z ~ N(0, I_N), R = mu + L Z (lower-triangular, :51-57), S = #{R_i < -x} (:60-65),
//   extreme iff S >= k (:69-71), and stores S (:67).
// Here:
//   k_factor     FP64 blocked Cholesky on one workgroup with the jitter policy, then
//                the portfolio loading v = L^T w, w^T mu and ||v|| (one launch).
//   k_prep       fp32 / bf16 layouts of the factor for the path kernels.
//   k_valu_reg   fused path kernel, one path per lane, n <= 64 rows in registers,
//                factor read through the scalar cache: r_i = mu_i then fmaf(L_ij, z_j)
//                for j = 0..i, exactly the order the C mirror uses.
//   k_valu_tile  fused path kernel for n <= 608: 64 paths per workgroup, their
//                normals staged once in LDS, rows in 8-row register blocks.
//   k_mfma       32 paths per wave on v_mfma_f32_32x32x16_bf16: A = factor rows
//                (bf16 hi/lo), B = normals (bf16 hi/lo), three products per step
//                (hi*hi + hi*lo + lo*hi), fp32 accumulation, n <= 127.
// Every path kernel reduces the exceedance count, tail counts and fixed-point
// tail sums per lane, then per wave, with one atomic per wave per counter; no
// normal ever leaves the chip.
#include <map>
#include <mutex>

#include "mcr_kern_common.h"

namespace mcr {

// ---------------------------------------------------------------------------
// Factor: FP64 Cholesky of A + delta I on one workgroup of 16 waves (blocked,
// right-looking; see cholesky_body), then the portfolio moments.
// Jitter policy: delta_0 = 0, delta_t = mean(diag) 10^(t-12) for t = 1..8, then
// the modified factorisation with failing pivots clamped to delta_8.  A pivot
// fails when it is not above n eps (A_jj + delta).
// ---------------------------------------------------------------------------
__device__ double block_sum_f64(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_sum_f64(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

// FP64 Cholesky of A + delta I into L (row-major, zero upper triangle) on one
// workgroup of 1024 threads, right-looking by panels of kPanel columns:
//   panel:    thread r holds row k0 + r of the panel in registers; column c is
//             finished with one barrier: rows c..kb-1 publish their column-c
//             entry (double-buffered LDS slot), every row tests the same
//             pivot, scales its entry and updates its later panel columns;
//   trailing: A22 -= L21 L21^T on the lower triangle, 4 x 4 register tiles
//             read from a column-major LDS copy of the panel.
// Jitter policy: attempt t adds delta_t = mean(diag A) * 10^(min(t,8)-12)
// (t >= 1); attempt 9 clamps failing pivots to delta.  A pivot fails when it
// is not above n eps (A_jj + delta).  Returns false (and records attempts =
// -1) for non-finite A.
constexpr int kPanel = 16;
// MAXN: the largest n the kernel is built for (kMaxN: 78 KiB of LDS; the
// column-chunk queries' kChunkMaxN: 129 KiB); a panel row per thread.
template <int MAXN>
struct FactorShared {
  static constexpr int kLd = (MAXN + 3) / 4 * 4 + 4;  // rows + tile overrun, multiple of 4
  __attribute__((aligned(16))) double panel[kPanel * kLd];  // [c * kLd + r]
  double pivot[2][kPanel];
  double adiag[kPanel];  // A_jj + delta of the panel's columns (pivot tolerance)
  double red[16];
};

template <int MAXN>
__device__ bool cholesky_body(const FactorArgs& a, FactorShared<MAXN>& sh, int* attempts_out,
                              int* clamped_out, int* fail_out, double* delta_out) {
  const int n = a.n;
  const int tid = threadIdx.x, nthr = blockDim.x;  // nthr >= MAXN: every panel row has an owner
  constexpr int kPanelLd = FactorShared<MAXN>::kLd;
  const int lane = tid & 63, wave = tid >> 6, nwave = nthr >> 6;
  const double* __restrict__ A = a.A;
  double* __restrict__ L = a.L;

  double bad = 0.0, dsum = 0.0;
  for (int e = tid; e < n * n; e += nthr)
    if (!isfinite(A[e])) bad = 1.0;
  if (tid < n) dsum = A[tid * n + tid];
  bad = block_sum_f64(bad, sh.red);
  dsum = block_sum_f64(dsum, sh.red);
  if (bad > 0.0) {
    if (tid == 0) {
      a.info->attempts = -1;
      a.info->clamped = 0;
      a.info->fail_col = 0;
      a.info->jitter = 0.0;
      if (a.finfo) *a.finfo = *a.info;
    }
    return false;
  }
  const double mean = dsum / (double)n;

  int attempts = 0, clamped = 0, fail = 0, last_fail = 0;
  double delta = 0.0;
  for (int t = 0; t <= 9; ++t) {
    const bool clamp = (t == 9);
    delta = (t == 0) ? 0.0 : mean * pow(10.0, (double)((t < 8 ? t : 8) - 12));
    fail = 0;
    clamped = 0;
    for (int i = wave; i < n; i += nwave)
      for (int j = lane; j < n; j += 64) {
        const size_t e = (size_t)i * n + j;
        L[e] = (j < i) ? A[e] : (j == i ? A[e] + delta : 0.0);
      }
    __syncthreads();

    for (int k0 = 0; k0 < n && !fail; k0 += kPanel) {
      const int kb = min(kPanel, n - k0), rows = n - k0;
      const int r = tid;
      const bool own = r < rows;
      double x[kPanel];
#pragma unroll
      for (int c = 0; c < kPanel; ++c)
        x[c] = (own && c < kb) ? L[(size_t)(k0 + r) * n + k0 + c] : 0.0;
      if (tid < kb) sh.adiag[tid] = A[(size_t)(k0 + tid) * n + k0 + tid] + delta;

#pragma unroll
      for (int c = 0; c < kPanel; ++c) {
        const bool live = c < kb && !fail;  // uniform over the workgroup
        if (live && own && r >= c && r < kb) sh.pivot[c & 1][r] = x[c];
        __syncthreads();
        if (live) {
          const int j = k0 + c;
          double d = sh.pivot[c & 1][c];
          const double tol = (double)n * DBL_EPSILON * sh.adiag[c];
          if (!(d > tol) || !isfinite(d)) {
            if (clamp && isfinite(d)) {
              d = delta;
              ++clamped;
            } else {
              fail = j + 1;
            }
          }
          if (!fail) {
            const double ljj = sqrt(d);
            const double rl = 1.0 / ljj;
            if (own && r == c) x[c] = ljj;
            if (own && r > c) {
              const double lr = x[c] * rl;
              x[c] = lr;
#pragma unroll
              for (int c2 = c + 1; c2 < kPanel; ++c2)
                if (c2 < kb && r >= c2) x[c2] = fma(-lr, sh.pivot[c & 1][c2] * rl, x[c2]);
            }
          }
        }
      }
      if (fail) break;

      if (own) {
#pragma unroll
        for (int c = 0; c < kPanel; ++c) {
          if (c < kb) L[(size_t)(k0 + r) * n + k0 + c] = x[c];
          sh.panel[c * kPanelLd + r] = x[c];
        }
      }
      for (int e = tid; e < kPanel * (kPanelLd - rows); e += nthr) {
        const int c = e / (kPanelLd - rows), q = e % (kPanelLd - rows);
        sh.panel[c * kPanelLd + rows + q] = 0.0;
      }
      __syncthreads();

      // trailing update of columns k1..n-1 (only when the panel was full)
      const int k1 = k0 + kb, m = n - k1;
      if (m > 0) {
        const int TB = (m + 3) >> 2;
        const int ntiles = TB * (TB + 1) / 2;
        for (int tt = tid; tt < ntiles; tt += nthr) {
          int bi = (int)((sqrt(8.0 * tt + 1.0) - 1.0) * 0.5);
          while (bi * (bi + 1) / 2 > tt) --bi;
          while ((bi + 1) * (bi + 2) / 2 <= tt) ++bi;
          const int bj = tt - bi * (bi + 1) / 2;
          const int ri = kPanel + 4 * bi, rj = kPanel + 4 * bj;
          double acc[4][4];
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
#pragma unroll 1
          for (int p = 0; p < kPanel; ++p) {
            const double2* pa = (const double2*)&sh.panel[p * kPanelLd + ri];
            const double2* pb = (const double2*)&sh.panel[p * kPanelLd + rj];
            const double2 a01 = pa[0], a23 = pa[1], b01 = pb[0], b23 = pb[1];
            const double av[4] = {a01.x, a01.y, a23.x, a23.y};
            const double bv[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
              for (int v = 0; v < 4; ++v) acc[u][v] = fma(av[u], bv[v], acc[u][v]);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = k1 + 4 * bi + u;
            if (i >= n) break;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int j = k1 + 4 * bj + v;
              if (j <= i) L[(size_t)i * n + j] -= acc[u][v];
            }
          }
        }
      }
      __syncthreads();
    }
    attempts = t + 1;
    last_fail = fail;
    if (!fail) break;
  }
  *attempts_out = attempts;
  *clamped_out = clamped;
  *fail_out = last_fail;
  *delta_out = delta;
  return true;
}

// Portfolio loading v = L^T w, mean w^T mu and sd ||v|| in FP64, and the
// query's record (factor outcome, moments, VaR windows).
__device__ void moments_body(const FactorArgs& a, double* red, int attempts, int clamped,
                             int last_fail, double delta) {
  const int n = a.n;
  const int tid = threadIdx.x;
  const double* __restrict__ L = a.L;
  __syncthreads();
  for (int j = tid; j < n; j += blockDim.x) {
    double s = 0.0;
    for (int i = j; i < n; ++i) s += L[(size_t)i * n + j] * a.w[i];
    a.v64[j] = s;
  }
  __syncthreads();
  double mp = 0.0, ss = 0.0;
  for (int i = tid; i < n; i += blockDim.x) {
    mp += a.w[i] * a.mu[i];
    ss += a.v64[i] * a.v64[i];
  }
  mp = block_sum_f64(mp, red);
  ss = block_sum_f64(ss, red);
  if (tid == 0) {
    DevInfo* info = a.info;
    info->attempts = attempts;
    info->clamped = clamped;
    info->fail_col = last_fail;
    info->jitter = delta;
    info->mp = mp;
    const double sd = sqrt(ss);
    info->loss_sd = sd;
    for (int al = 0; al < kMaxAlpha; ++al) {
      if (al < a.n_alpha) {
        info->lo[al] = (float)(-mp + (a.za[al] - a.ha[al]) * sd);
        info->hi[al] = (float)(-mp + (a.za[al] + a.ha[al]) * sd);
      } else {
        info->lo[al] = 0.0f;
        info->hi[al] = 0.0f;
      }
    }
    info->hist_lo = 0.0f;
    info->hist_scale = 0.0f;
  }
}


template <int MAXN>
__device__ void factor_body(const FactorArgs& a, FactorShared<MAXN>& sh) {
  const int tid = threadIdx.x;

  // Clear the query's accumulators.
  if (tid == 0) {
    DevAccum z = {};
    *a.acc = z;
  }
  for (int i = tid; i < kHistBins; i += blockDim.x) a.hist[i] = 0u;

  int attempts = 0, clamped = 0, last_fail = 0;
  double delta = 0.0;
  if (a.reuse) {
    // L holds this covariance's factor from an earlier launch.
    const DevInfo fi = *a.finfo;
    attempts = fi.attempts;
    clamped = fi.clamped;
    last_fail = fi.fail_col;
    delta = fi.jitter;
    if (attempts < 0) {
      if (tid == 0) {
        a.info->attempts = -1;
        a.info->clamped = 0;
        a.info->fail_col = 0;
        a.info->jitter = 0.0;
      }
      return;
    }
  } else if (!cholesky_body(a, sh, &attempts, &clamped, &last_fail, &delta)) {
    return;
  }
  if (a.finfo && !a.reuse && tid == 0) {
    a.finfo->attempts = attempts;
    a.finfo->clamped = clamped;
    a.finfo->fail_col = last_fail;
    a.finfo->jitter = delta;
  }
  moments_body(a, sh.red, attempts, clamped, last_fail, delta);
}

template <int MAXN>
__global__ __launch_bounds__(1024) void k_factor(FactorArgs a) {
  __shared__ FactorShared<MAXN> sh;
  factor_body(a, sh);
}

// One workgroup per query of a batch (n <= MAXN).
template <int MAXN>
__global__ __launch_bounds__(1024) void k_factor_batch(const FactorArgs* fas) {
  __shared__ FactorShared<MAXN> sh;
  const FactorArgs a = fas[blockIdx.x];
  factor_body(a, sh);
}

hipError_t launch_factor(const FactorArgs& a, hipStream_t stream) {
  if (a.n < 1 || a.n > kChunkMaxN) return hipErrorInvalidValue;
  if (a.n <= kMaxN)
    hipLaunchKernelGGL(k_factor<kMaxN>, dim3(1), dim3(1024), 0, stream, a);
  else
    hipLaunchKernelGGL(k_factor<kChunkMaxN>, dim3(1), dim3(1024), 0, stream, a);
  return hipGetLastError();
}

// max_n: the batch's largest query (a panel of 78 KiB up to kMaxN, 99 KiB
// up to kBatchMaxN)
hipError_t launch_factor_batch(const FactorArgs* fas, int nq, int max_n, hipStream_t stream) {
  if (max_n < 1 || max_n > kBatchMaxN) return hipErrorInvalidValue;
  if (max_n <= kMaxN)
    hipLaunchKernelGGL(k_factor_batch<kMaxN>, dim3(nq), dim3(1024), 0, stream, fas);
  else
    hipLaunchKernelGGL(k_factor_batch<kBatchMaxN>, dim3(nq), dim3(1024), 0, stream, fas);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Prep: layouts of the factor for the three path kernels.
// ---------------------------------------------------------------------------
__device__ void prep_body(const PrepArgs& a, size_t gid, size_t gsz) {
  const int n = a.n;
  const double* __restrict__ L = a.L;
  const int r8 = ((n + kTileRows - 1) / kTileRows) * kTileRows;
  if (a.colpack) {
    // fp32 loading and mean
    for (size_t e = gid; e < (size_t)n; e += gsz) {
      a.v32[e] = (float)a.v64[e];
      a.mu32[e] = (float)a.mu[e];
      a.colpack[colpack_offset(n, (int)e) + n - (int)e] = (float)a.v64[e];
    }
    // register kernel: column-packed lower triangle
    for (size_t e = gid; e < (size_t)n * n; e += gsz) {
      const int i = (int)(e / n), j = (int)(e % n);
      if (j <= i) a.colpack[colpack_offset(n, j) + i - j] = (float)L[e];
    }
    // tile kernel: zero-padded r8 x r8 in 8 x 4 blocks (l32_offset)
    for (size_t e = gid; e < (size_t)r8 * r8; e += gsz) {
      const int i = (int)(e / r8), j = (int)(e % r8);
      a.L32[l32_offset(r8, i, j)] = (i < n && j <= i) ? (float)L[(size_t)i * n + j] : 0.0f;
    }
  }
  // MFMA: augmented rows (assets, zero pads, loss row 32T-1 = -v), hi/lo bf16
  if (a.T > 0) {
    const int T = a.T, S = a.S;
    const float mp32 = (float)a.info->mp;
    for (size_t e = gid; e < (size_t)T * S * 64; e += gsz) {
      const int lane = (int)(e % 64);
      const int s = (int)((e / 64) % S);
      const int t = (int)(e / (64 * (size_t)S));
      const int row = 32 * t + (lane & 31);
      const int k0 = 16 * s + 8 * (lane >> 5);
      uint16_t hb[8], lb[8];
      for (int q = 0; q < 8; ++q) {
        const int k = k0 + q;
        float m = 0.0f;
        if (k < n) {
          if (row < n) {
            m = (k <= row) ? (float)L[(size_t)row * n + k] : 0.0f;
          } else if (row == 32 * T - 1) {
            m = -(float)a.v64[k];
          }
        }
        hb[q] = bf16_bits(m);
        lb[q] = bf16_bits(m - bf16_val(hb[q]));
      }
      uint4 h, l;
      h.x = hb[0] | ((uint32_t)hb[1] << 16);
      h.y = hb[2] | ((uint32_t)hb[3] << 16);
      h.z = hb[4] | ((uint32_t)hb[5] << 16);
      h.w = hb[6] | ((uint32_t)hb[7] << 16);
      l.x = lb[0] | ((uint32_t)lb[1] << 16);
      l.y = lb[2] | ((uint32_t)lb[3] << 16);
      l.z = lb[4] | ((uint32_t)lb[5] << 16);
      l.w = lb[6] | ((uint32_t)lb[7] << 16);
      a.lfrag[2 * e] = h;
      a.lfrag[2 * e + 1] = l;
    }
    for (size_t e = gid; e < (size_t)32 * T; e += gsz) {
      const int row = (int)e;
      a.rowinit[e] = row < n ? (float)a.mu[row] : (row == 32 * T - 1 ? -mp32 : INFINITY);
    }
  }
}

__global__ __launch_bounds__(256) void k_prep(PrepArgs a) {
  prep_body(a, (size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x);
}

// blockIdx.y = query of a batch
__global__ __launch_bounds__(256) void k_prep_batch(const PrepArgs* pas) {
  const PrepArgs a = pas[blockIdx.y];
  prep_body(a, (size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x);
}

hipError_t launch_prep_batch(const PrepArgs* pas, int nq, int max_n, hipStream_t stream) {
  int blocks = (max_n * max_n + 255) / 256;
  if (blocks > 256) blocks = 256;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_prep_batch, dim3(blocks, nq), dim3(256), 0, stream, pas);
  return hipGetLastError();
}

hipError_t launch_prep(const PrepArgs& a, hipStream_t stream) {
  size_t work = (size_t)a.n * a.n;
  const int r8 = ((a.n + kTileRows - 1) / kTileRows) * kTileRows;
  if ((size_t)r8 * r8 > work) work = (size_t)r8 * r8;
  int blocks = (int)((work + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_prep, dim3(blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tile kernel: 64 paths per workgroup (lane = path), NW waves.  Phase 1 writes
// the paths' normals to LDS as float4 [block][lane]; phase 2 splits the rows
// into 8-row register blocks dealt to the waves in snake order (block cost grows
// with the row index), the loss row being one more block.  Row i still sees
// mu_i, then z_0 .. z_i in order, so results equal the register kernel's.
// The normals of 64 paths take up to 152 KiB, so at large n one workgroup
// fills a CU's LDS: NW grows as fewer workgroups fit (4, 8, 16 waves), keeping
// several waves per SIMD to cover the scalar factor loads and LDS reads.
// ---------------------------------------------------------------------------
constexpr int kTileStaticLds = 16 * 64 * 4 + 64 * 4 + (kMaxN + 1) * 4;

template <int NW>
__global__ __launch_bounds__(64 * NW) void k_valu_tile(PathArgs a) {
  extern __shared__ float4 zs[];  // [nb][64]
  __shared__ int sS[NW][64];
  __shared__ float sLoss[64];
  __shared__ unsigned sh_hist[kMaxN + 1];
  const int n = a.n;
  const int nb = (n + 3) >> 2;
  const int nrb = (n + kTileRows - 1) / kTileRows;
  const int ldl = nrb * kTileRows;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool want_hist = a.s_hist != nullptr;
  if (want_hist) {
    for (int i = threadIdx.x; i <= n; i += blockDim.x) sh_hist[i] = 0u;
  }
  MCR_CONSTP(float) L = cview(a.L32);
  MCR_CONSTP(float) mu = cview(a.mu32);
  MCR_CONSTP(float) v = cview(a.v32);
  const float mp32 = (float)cview(a.info)->mp;
  const Key2 key = key_of(a.seed);
  const bool want_near = a.thr_lo < a.thr_hi;
  LaneTally t;
  zero_tally(t);

  for (uint64_t tile = blockIdx.x; tile * kTilePaths < a.npaths; tile += gridDim.x) {
    const uint64_t q = tile * kTilePaths + lane;
    const uint64_t path = a.path0 + q;
    for (int b = wave; b < nb; b += NW) {
      float z[4];
      normals4(key, path, (uint32_t)b, z);
      zs[b * 64 + lane] = make_float4(z[0], z[1], z[2], z[3]);
    }
    __syncthreads();

    int S = 0;
    bool near = false;
    for (int rb = 0; rb <= nrb; ++rb) {
      const int p = rb / NW, rw = rb % NW;
      const int owner = (p & 1) ? NW - 1 - rw : rw;
      if (owner != wave) continue;
      if (rb == nrb) {
        // loss row: g = w^T mu + sum_j v_j z_j, j ascending
        float g = mp32;
        for (int jb = 0; jb < nb - 1; ++jb) {
          const float4 z4 = zs[jb * 64 + lane];
          g = __builtin_fmaf(v[4 * jb + 0], z4.x, g);
          g = __builtin_fmaf(v[4 * jb + 1], z4.y, g);
          g = __builtin_fmaf(v[4 * jb + 2], z4.z, g);
          g = __builtin_fmaf(v[4 * jb + 3], z4.w, g);
        }
        {
          const int jb = nb - 1;
          const float4 z4 = zs[jb * 64 + lane];
          const int rem = n - 4 * jb;
          g = __builtin_fmaf(v[4 * jb + 0], z4.x, g);
          if (rem > 1) g = __builtin_fmaf(v[4 * jb + 1], z4.y, g);
          if (rem > 2) g = __builtin_fmaf(v[4 * jb + 2], z4.z, g);
          if (rem > 3) g = __builtin_fmaf(v[4 * jb + 3], z4.w, g);
        }
        sLoss[lane] = -g;
        continue;
      }
      const int i0 = rb * kTileRows;
      float acc[kTileRows];
#pragma unroll
      for (int r = 0; r < kTileRows; ++r) acc[r] = (i0 + r < n) ? mu[i0 + r] : 0.0f;
      // Row block rb's 8 x 4 factor blocks (32 floats each, column block jb at
      // Lb + 32 jb): one K step is two 64-byte scalar loads.
      MCR_CONSTP(float) Lb = L + l32_offset(ldl, i0, 0);
      const int J = i0 / 4;
      // Two K steps per wait where the row blocks are long (NW >= 8: n >~ 150;
      // n = 400 -8 %); with four waves the unrolled loop measured slower.
      constexpr int kStepUnroll = NW >= 8 ? 2 : 1;
#pragma unroll kStepUnroll
      for (int jb = 0; jb < J; ++jb) {
        const float4 z4 = zs[jb * 64 + lane];
        MCR_CONSTP(float) lv = Lb + 32 * jb;
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) {
          acc[r] = __builtin_fmaf(lv[4 * r + 0], z4.x, acc[r]);
          acc[r] = __builtin_fmaf(lv[4 * r + 1], z4.y, acc[r]);
          acc[r] = __builtin_fmaf(lv[4 * r + 2], z4.z, acc[r]);
          acc[r] = __builtin_fmaf(lv[4 * r + 3], z4.w, acc[r]);
        }
      }
      // diagonal 8x8 block: row i0+r uses columns i0 .. i0+r
      float zd[8];
      {
        const float4 z4 = zs[(i0 / 4) * 64 + lane];
        zd[0] = z4.x; zd[1] = z4.y; zd[2] = z4.z; zd[3] = z4.w;
        if (i0 / 4 + 1 < nb) {
          const float4 z5 = zs[(i0 / 4 + 1) * 64 + lane];
          zd[4] = z5.x; zd[5] = z5.y; zd[6] = z5.z; zd[7] = z5.w;
        } else {
          zd[4] = zd[5] = zd[6] = zd[7] = 0.0f;
        }
      }
#pragma unroll
      for (int r = 0; r < kTileRows; ++r) {
#pragma unroll
        for (int c = 0; c <= r; ++c)
          acc[r] = __builtin_fmaf(Lb[32 * (J + (c >> 2)) + 4 * r + (c & 3)], zd[c], acc[r]);
      }
#pragma unroll
      for (int r = 0; r < kTileRows; ++r) {
        if (i0 + r < n) {
          S += (acc[r] < a.thr) ? 1 : 0;
          if (want_near) near |= (acc[r] >= a.thr_lo) & (acc[r] <= a.thr_hi);
        }
      }
    }
    sS[wave][lane] = S | (near ? (1 << 30) : 0);
    __syncthreads();
    if (wave == 0 && q < a.npaths) {
      int tot = 0;
      bool nr = false;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const int e = sS[w][lane];
        tot += e & 0xffff;
        nr |= (e >> 30) & 1;
      }
      tally_path(a, t, q, tot, sLoss[lane], nr, want_hist ? sh_hist : nullptr);
    }
    __syncthreads();
  }
  flush_tally(a, t);
  if (want_hist) {
    __syncthreads();
    for (int i = threadIdx.x; i <= n; i += blockDim.x)
      if (sh_hist[i]) atomicAdd(&a.s_hist[i], (unsigned long long)sh_hist[i]);
  }
}

// ---------------------------------------------------------------------------
// Column-chunk kernel for n up to kChunkMaxN (and any smaller n): 64 paths per
// workgroup, 16 waves.  The paths' normals no longer fit the LDS whole, so the
// columns are taken in chunks of kChunkBlocks 4-column blocks (512 columns,
// 128 KiB): the waves generate a chunk's normals into LDS, then every wave
// advances all of its 8-row register blocks (snake order, the loss row one
// more block) over the chunk's columns.  Each wave's blocks stay in registers
// across the chunks, so row i still sees mu_i, then z_0 .. z_i in order and
// the results equal the tile and register kernels' bit for bit.
// ---------------------------------------------------------------------------
constexpr int kChunkBlocks = 128;  // 4-column blocks per chunk (even: a diagonal
                                   // block's two column blocks share a chunk)
constexpr int kChunkNW = 16;
constexpr int kChunkMB = (kChunkMaxN / kTileRows + kChunkNW - 1) / kChunkNW;  // row blocks per wave

__global__ __launch_bounds__(64 * kChunkNW) void k_valu_chunk(PathArgs a) {
  constexpr int NW = kChunkNW;
  __shared__ float4 zs[kChunkBlocks * 64];  // [block - chunk start][lane]
  __shared__ int sS[NW][64];
  __shared__ float sLoss[64];
  __shared__ unsigned sh_hist[kChunkMaxN + 1];
  const int n = a.n;
  const int nb = (n + 3) >> 2;
  const int nrb = (n + kTileRows - 1) / kTileRows;  // row blocks
  const int ldl = nrb * kTileRows;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // the loss row goes to the wave the snake deal would give block nrb
  const bool loss_wave = ((nrb / NW) & 1 ? NW - 1 - nrb % NW : nrb % NW) == wave;
  const bool want_hist = a.s_hist != nullptr;
  if (want_hist) {
    for (int i = threadIdx.x; i <= n; i += blockDim.x) sh_hist[i] = 0u;
  }
  MCR_CONSTP(float) L = cview(a.L32);
  MCR_CONSTP(float) mu = cview(a.mu32);
  MCR_CONSTP(float) v = cview(a.v32);
  const float mp32 = (float)cview(a.info)->mp;
  const Key2 key = key_of(a.seed);
  const bool want_near = a.thr_lo < a.thr_hi;
  LaneTally t;
  zero_tally(t);

  for (uint64_t tile = blockIdx.x; tile * kTilePaths < a.npaths; tile += gridDim.x) {
    const uint64_t q = tile * kTilePaths + lane;
    const uint64_t path = a.path0 + q;
    // this wave's k-th row block: rb = k NW + (k odd ? NW - 1 - wave : wave)
    float acc[kChunkMB][kTileRows];
#pragma unroll
    for (int k = 0; k < kChunkMB; ++k) {
      const int rb = k * NW + ((k & 1) ? NW - 1 - wave : wave);
#pragma unroll
      for (int r = 0; r < kTileRows; ++r) {
        const int i = rb * kTileRows + r;
        acc[k][r] = (rb < nrb && i < n) ? mu[i] : 0.0f;
      }
    }
    float g = mp32;  // loss row: g = w^T mu + sum_j v_j z_j, j ascending
    for (int jb0 = 0; jb0 < nb; jb0 += kChunkBlocks) {
      const int jb1 = min(nb, jb0 + kChunkBlocks);
      // a fresh copy of the wave index per chunk keeps the compiler from
      // holding every block's addresses across the chunk loop (SGPR spills)
      int wv = wave;
      asm volatile("" : "+s"(wv));
      for (int b = jb0 + wv; b < jb1; b += NW) {
        float z[4];
        normals4(key, path, (uint32_t)b, z);
        zs[(b - jb0) * 64 + lane] = make_float4(z[0], z[1], z[2], z[3]);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kChunkMB; ++k) {
        const int rb = k * NW + ((k & 1) ? NW - 1 - wv : wv);
        if (rb >= nrb) continue;
        const int i0 = rb * kTileRows;
        const int J = i0 / 4;  // column blocks left of the diagonal block
        MCR_CONSTP(float) Lb = L + l32_offset(ldl, i0, 0);
        const int je = min(jb1, J);
        for (int jb = jb0; jb < je; ++jb) {
          const float4 z4 = zs[(jb - jb0) * 64 + lane];
          MCR_CONSTP(float) lv = Lb + 32 * jb;
#pragma unroll
          for (int r = 0; r < kTileRows; ++r) {
            acc[k][r] = __builtin_fmaf(lv[4 * r + 0], z4.x, acc[k][r]);
            acc[k][r] = __builtin_fmaf(lv[4 * r + 1], z4.y, acc[k][r]);
            acc[k][r] = __builtin_fmaf(lv[4 * r + 2], z4.z, acc[k][r]);
            acc[k][r] = __builtin_fmaf(lv[4 * r + 3], z4.w, acc[k][r]);
          }
        }
        if (J >= jb0 && J < jb1) {
          // diagonal 8x8 block (column blocks J and J+1, both in this chunk)
          float zd[8];
          const float4 z4 = zs[(J - jb0) * 64 + lane];
          zd[0] = z4.x; zd[1] = z4.y; zd[2] = z4.z; zd[3] = z4.w;
          if (J + 1 < nb) {
            const float4 z5 = zs[(J + 1 - jb0) * 64 + lane];
            zd[4] = z5.x; zd[5] = z5.y; zd[6] = z5.z; zd[7] = z5.w;
          } else {
            zd[4] = zd[5] = zd[6] = zd[7] = 0.0f;
          }
#pragma unroll
          for (int r = 0; r < kTileRows; ++r) {
#pragma unroll
            for (int c = 0; c <= r; ++c)
              acc[k][r] = __builtin_fmaf(Lb[32 * (J + (c >> 2)) + 4 * r + (c & 3)], zd[c], acc[k][r]);
          }
        }
      }
      if (loss_wave) {
        for (int jb = jb0; jb < jb1; ++jb) {
          const float4 z4 = zs[(jb - jb0) * 64 + lane];
          const int rem = n - 4 * jb;  // >= 4 except in the last block
          g = __builtin_fmaf(v[4 * jb + 0], z4.x, g);
          if (rem > 1) g = __builtin_fmaf(v[4 * jb + 1], z4.y, g);
          if (rem > 2) g = __builtin_fmaf(v[4 * jb + 2], z4.z, g);
          if (rem > 3) g = __builtin_fmaf(v[4 * jb + 3], z4.w, g);
        }
      }
      __syncthreads();  // the chunk's normals are read; the next chunk may overwrite them
    }

    if (loss_wave) sLoss[lane] = -g;
    int S = 0;
    bool near = false;
#pragma unroll
    for (int k = 0; k < kChunkMB; ++k) {
      const int rb = k * NW + ((k & 1) ? NW - 1 - wave : wave);
      if (rb >= nrb) continue;
#pragma unroll
      for (int r = 0; r < kTileRows; ++r) {
        if (rb * kTileRows + r < n) {
          S += (acc[k][r] < a.thr) ? 1 : 0;
          if (want_near) near |= (acc[k][r] >= a.thr_lo) & (acc[k][r] <= a.thr_hi);
        }
      }
    }
    sS[wave][lane] = S | (near ? (1 << 30) : 0);
    __syncthreads();
    if (wave == 0 && q < a.npaths) {
      int tot = 0;
      bool nr = false;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const int e = sS[w][lane];
        tot += e & 0xffff;
        nr |= (e >> 30) & 1;
      }
      tally_path(a, t, q, tot, sLoss[lane], nr, want_hist ? sh_hist : nullptr);
    }
    __syncthreads();
  }
  flush_tally(a, t);
  if (want_hist) {
    __syncthreads();
    for (int i = threadIdx.x; i <= n; i += blockDim.x)
      if (sh_hist[i]) atomicAdd(&a.s_hist[i], (unsigned long long)sh_hist[i]);
  }
}

// ---------------------------------------------------------------------------
// Debug: the normal stream itself, one path per lane.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_normals(uint64_t seed, uint64_t path0, uint64_t npaths,
                                                 int n, float* out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npaths) return;
  const Key2 key = key_of(seed);
  for (int b = 0; 4 * b < n; ++b) {
    float z[4];
    normals4(key, path0 + q, (uint32_t)b, z);
    for (int u = 0; u < 4; ++u)
      if (4 * b + u < n) out[q * n + 4 * b + u] = z[u];
  }
}

hipError_t launch_normals(uint64_t seed, uint64_t path0, uint64_t npaths, int n, float* out,
                          hipStream_t stream) {
  const uint64_t blocks = (npaths + 255) / 256;
  hipLaunchKernelGGL(k_normals, dim3((unsigned)blocks), dim3(256), 0, stream, seed, path0, npaths,
                     n, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Dispatch.
// ---------------------------------------------------------------------------
static uint64_t min_u64(uint64_t x, uint64_t y) { return x < y ? x : y; }

int resident_blocks(const void* kernel, int block) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find({kernel, block});
  if (it != cache.end()) return it->second;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  cache[{kernel, block}] = per_cu;
  return per_cu;
}

hipError_t launch_paths(int kernel, const PathArgs& a, int num_cus, hipStream_t stream) {
  if (a.npaths == 0) return hipSuccess;
  if (kernel == 1) {
    PathKernel k = reg_kernel(a.n);
    if (!k) return hipErrorInvalidValue;
    const uint64_t blocks = min_u64((a.npaths + 255) / 256, (uint64_t)num_cus * 8);
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    return hipGetLastError();
  }
  if (kernel == 2) {
    if (a.n < 1 || a.n > kMaxN) return hipErrorInvalidValue;
    const int nb = (a.n + 3) >> 2;
    const size_t lds = (size_t)nb * 64 * sizeof(float4);
    static bool attr = false;
    if (!attr) {
      // the largest dynamic request (n = kMaxN); static LDS comes on top
      const int max_dyn = ((kMaxN + 3) / 4) * 64 * (int)sizeof(float4);
      for (const void* f : {(const void*)k_valu_tile<4>, (const void*)k_valu_tile<8>,
                            (const void*)k_valu_tile<16>}) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_dyn);
        if (e != hipSuccess) return e;
      }
      attr = true;
    }
    uint64_t per_cu = (160 * 1024) / (lds + kTileStaticLds);
    if (per_cu < 1) per_cu = 1;
    if (per_cu > 8) per_cu = 8;
    const uint64_t blocks = min_u64((a.npaths + kTilePaths - 1) / kTilePaths, (uint64_t)num_cus * per_cu);
    // waves per workgroup: about 16 per CU where the LDS limits the
    // workgroups, at most 8 per SIMD
    if (per_cu == 1)
      hipLaunchKernelGGL(k_valu_tile<16>, dim3((unsigned)blocks), dim3(1024), lds, stream, a);
    else if (per_cu <= 3)
      hipLaunchKernelGGL(k_valu_tile<8>, dim3((unsigned)blocks), dim3(512), lds, stream, a);
    else
      hipLaunchKernelGGL(k_valu_tile<4>, dim3((unsigned)blocks), dim3(256), lds, stream, a);
    return hipGetLastError();
  }
  if (kernel == 5) {
    if (a.n < 1 || a.n > kChunkMaxN) return hipErrorInvalidValue;
    // one 1024-thread workgroup per CU (139.5 KiB of LDS)
    const uint64_t blocks = min_u64((a.npaths + kTilePaths - 1) / kTilePaths, (uint64_t)num_cus);
    hipLaunchKernelGGL(k_valu_chunk, dim3((unsigned)blocks), dim3(64 * kChunkNW), 0, stream, a);
    return hipGetLastError();
  }
  if (kernel == 4) {
    PathKernel k = mfma_big_kernel(a.T);
    if (!k || a.n > kMfmaLdsMaxN) return hipErrorInvalidValue;
    const int block = 64 * big_waves(a.T);
    const uint64_t per_cu = (uint64_t)resident_blocks((const void*)k, block);
    const uint64_t blocks = min_u64((a.npaths + 63) / 64, (uint64_t)num_cus * per_cu);
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(block), 0, stream, a);
    return hipGetLastError();
  }
  if (kernel == 3) {
    PathKernel k = mfma_kernel(a.n, a.T);
    if (!k) return hipErrorInvalidValue;
    const uint64_t blocks = min_u64((a.npaths + 255) / 256, (uint64_t)num_cus * 8);
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

}  // namespace mcr
