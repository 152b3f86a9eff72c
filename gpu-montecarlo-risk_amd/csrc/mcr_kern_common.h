// mcr_kern_common.h -- device helpers shared by the path kernels: scalar
// views of read-only inputs, wave reductions, per-path tallies (crash count,
// VaR window bookkeeping, refinement histogram) and their flush.
#pragma once
#include <float.h>
#include <math.h>

#include "mcr_internal.h"
#include "mcr_rng.h"

namespace mcr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
#define MCR_CONSTP(T) const __attribute__((address_space(4))) T*

// Read-only inputs viewed through the constant address space: uniform reads
// become scalar loads whose results feed VALU ops as SGPR operands.
template <typename T>
__device__ __forceinline__ MCR_CONSTP(T) cview(const T* p) {
  return (MCR_CONSTP(T))p;
}

// A read-only struct loaded word by word through the constant address space
// (scalar loads when the address is uniform).
template <typename T>
__device__ __forceinline__ T load_const(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized struct expected");
  uint32_t w[sizeof(T) / 4];
  MCR_CONSTP(uint32_t) src = cview(reinterpret_cast<const uint32_t*>(p));
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = src[i];
  return __builtin_bit_cast(T, w);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Loss -> fixed point in units of 2^-32, truncating (exact for |loss| >= 2^-9).
__device__ __forceinline__ long long fixed32(float v) {
  return (long long)((double)v * 4294967296.0);
}

struct LaneTally {
  uint32_t count;
  uint32_t near;
  uint32_t above[kMaxAlpha];
  long long sum_above[kMaxAlpha];
  uint32_t hist_below, hist_above;
};

// Wave-aggregated append of losses that fall inside window `al`.
__device__ __forceinline__ void append_loss(const PathArgs& a, int al, bool in, float v) {
  const unsigned long long m = __ballot(in);
  if (m == 0ull) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)m) - 1;
  const unsigned rank = __popcll(m & ((1ull << lane) - 1ull));
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(&a.acc->gathered[al], (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  if (in) {
    const unsigned long long idx = base + rank;
    if (idx < a.gather_cap) a.gather[(size_t)al * a.gather_cap + idx] = v;
  }
}

// Per-path bookkeeping shared by every path kernel.  Lanes that reach it are
// active paths; ballots inside see exactly those lanes.
// Window bounds are re-read from the (scalar-cached) info block each path
// rather than held in SGPRs across the path loop.
__device__ __forceinline__ void tally_path(const PathArgs& a, LaneTally& t, uint64_t q, int S,
                                           float loss, bool near, unsigned* sh_hist) {
  const DevInfo* ip = a.info;
  asm volatile("" : "+s"(ip));
  MCR_CONSTP(DevInfo) w = cview(ip);
  t.count += (S >= a.k) ? 1u : 0u;
  t.near += near ? 1u : 0u;
  if (a.S_out) a.S_out[q] = S;
  if (a.loss_out) a.loss_out[q] = loss;
  if (sh_hist) atomicAdd(&sh_hist[S], 1u);
  if (a.loss_mode == kLossBracket) {
#pragma unroll
    for (int al = 0; al < kMaxAlpha; ++al) {
      if (al < a.n_alpha) {
        if (loss > w->hi[al]) {
          t.above[al] += 1u;
          t.sum_above[al] += fixed32(loss);
        }
        append_loss(a, al, loss >= w->lo[al] && loss <= w->hi[al], loss);
      }
    }
  } else if (a.loss_mode == kLossHist) {
    const float pos = (loss - w->hist_lo) * w->hist_scale;
    if (!(pos >= 0.0f)) {  // also catches NaN
      t.hist_below += 1u;
    } else if (!(pos < (float)kHistBins)) {
      t.hist_above += 1u;
    } else {
      atomicAdd(&a.hist[(int)pos], 1u);
    }
  }
}

__device__ __forceinline__ void flush_tally(const PathArgs& a, const LaneTally& t) {
  const int lane = __lane_id();
  const uint32_t c = wave_sum_u32(t.count);
  const uint32_t nr = wave_sum_u32(t.near);
  if (lane == 0) {
    if (c) atomicAdd(&a.acc->count, (unsigned long long)c);
    if (nr) atomicAdd(&a.acc->near, (unsigned long long)nr);
  }
  if (a.loss_mode == kLossBracket) {
#pragma unroll
    for (int al = 0; al < kMaxAlpha; ++al) {
      if (al < a.n_alpha) {
        const uint32_t ab = wave_sum_u32(t.above[al]);
        const long long sm = wave_sum_i64(t.sum_above[al]);
        if (lane == 0 && ab) {
          atomicAdd(&a.acc->above[al], (unsigned long long)ab);
          atomicAdd((unsigned long long*)&a.acc->sum_above[al], (unsigned long long)sm);
        }
      }
    }
  } else if (a.loss_mode == kLossHist) {
    const uint32_t b = wave_sum_u32(t.hist_below);
    const uint32_t ab = wave_sum_u32(t.hist_above);
    if (lane == 0) {
      if (b) atomicAdd(&a.acc->hist_below, (unsigned long long)b);
      if (ab) atomicAdd(&a.acc->hist_above, (unsigned long long)ab);
    }
  }
}

// Workgroup form of flush_tally for kernels whose waves all tally paths: each
// wave sums its lanes, the waves' sums meet in LDS, and wave 0 issues one
// atomic per counter for the whole workgroup (a per-wave flush costs every
// wave its own L2 atomics: at 8 waves per CU they were most of the launch's
// HBM write traffic).  Every thread of the workgroup must call it (barrier).
__device__ __forceinline__ void flush_tally_block(const PathArgs& a, const LaneTally& t) {
  constexpr int kWaves = 16;              // workgroups of up to 1024 threads
  constexpr int kU = 4 + kMaxAlpha;       // count, near, hist below/above, above[]
  __shared__ uint32_t s_u[kWaves][kU];
  __shared__ long long s_l[kWaves][kMaxAlpha];
  const int lane = __lane_id(), wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  const bool bracket = a.loss_mode == kLossBracket, hist = a.loss_mode == kLossHist;
  uint32_t u[kU];
  long long l[kMaxAlpha];
  u[0] = wave_sum_u32(t.count);
  u[1] = wave_sum_u32(t.near);
  u[2] = hist ? wave_sum_u32(t.hist_below) : 0u;
  u[3] = hist ? wave_sum_u32(t.hist_above) : 0u;
#pragma unroll
  for (int al = 0; al < kMaxAlpha; ++al) {
    const bool on = bracket && al < a.n_alpha;
    u[4 + al] = on ? wave_sum_u32(t.above[al]) : 0u;
    l[al] = on ? wave_sum_i64(t.sum_above[al]) : 0ll;
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kU; ++k) s_u[wave][k] = u[k];
#pragma unroll
    for (int al = 0; al < kMaxAlpha; ++al) s_l[wave][al] = l[al];
  }
  __syncthreads();
  if (wave == 0) {
    unsigned long long* const dst[kU] = {&a.acc->count, &a.acc->near, &a.acc->hist_below,
                                         &a.acc->hist_above, &a.acc->above[0], &a.acc->above[1],
                                         &a.acc->above[2], &a.acc->above[3]};
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const uint32_t v = wave_sum_u32(lane < nw ? s_u[lane][k] : 0u);
      if (lane == 0 && v) atomicAdd(dst[k], (unsigned long long)v);
    }
#pragma unroll
    for (int al = 0; al < kMaxAlpha; ++al) {
      if (bracket && al < a.n_alpha) {
        const long long v = wave_sum_i64(lane < nw ? s_l[lane][al] : 0ll);
        if (lane == 0 && v) atomicAdd((unsigned long long*)&a.acc->sum_above[al], (unsigned long long)v);
      }
    }
  }
}

__device__ __forceinline__ void zero_tally(LaneTally& t) {
  t.count = 0;
  t.near = 0;
#pragma unroll
  for (int i = 0; i < kMaxAlpha; ++i) {
    t.above[i] = 0;
    t.sum_above[i] = 0;
  }
  t.hist_below = 0;
  t.hist_above = 0;
}

__device__ __forceinline__ uint16_t bf16_bits(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ __forceinline__ float bf16_val(uint16_t b) { return as_f32((uint32_t)b << 16); }


__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  return (uint32_t)bf16_bits(lo) | ((uint32_t)bf16_bits(hi) << 16);
}

// z[0..7] -> bf16 hi and lo parts, two per 32-bit word
__device__ __forceinline__ void split8(const float* z, uint32_t hi[4], uint32_t lo[4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    hi[q] = pack_bf16x2(z[2 * q], z[2 * q + 1]);
    const float r0 = z[2 * q] - as_f32(hi[q] << 16);
    const float r1 = z[2 * q + 1] - as_f32(hi[q] & 0xffff0000u);
    lo[q] = pack_bf16x2(r0, r1);
  }
}

// lanes 32-63 of x <-> lanes 0-31 of y
__device__ __forceinline__ void swap32(uint32_t& x, uint32_t& y) {
  const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
  x = r[0];
  y = r[1];
}

__device__ __forceinline__ bf16x8 as_bf16x8(const uint32_t w[4]) {
  return __builtin_bit_cast(bf16x8, make_uint4(w[0], w[1], w[2], w[3]));
}

typedef void (*PathKernel)(PathArgs);
PathKernel reg_kernel(int n);          // mcr_reg.hip
PathKernel mfma_kernel(int n, int T);  // mcr_mfma.hip
PathKernel mfma_big_kernel(int T);     // mcr_mfma_big.hip

}  // namespace mcr
