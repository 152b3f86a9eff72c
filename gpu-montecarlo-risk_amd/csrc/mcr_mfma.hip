// mcr_mfma.hip -- the bf16x3 MFMA path kernel (n <= 127) and its persistent
// batch form.
#include "mcr_kern_common.h"

namespace mcr {

// ---------------------------------------------------------------------------
// MFMA kernel: T tiles of 32 rows (assets, zero pads, loss row 32T-1), S steps
// of 16 normals, v_mfma_f32_32x32x16_bf16 with A = factor rows (bf16 hi/lo,
// resident in registers for the whole kernel) and B = normals (bf16 hi/lo).
// A wave handles 64 paths as two 32-path tiles; for a tile, the two lanes l
// and l + 32 share path l % 32, generating its normals 16s..16s+7 and
// 16s+8..16s+15 of each step s: their B-operand slots.  D[row][path]: lane l
// holds rows 8(r/4) + 4(l/32) + r%4 of its path; one v_permlane32_swap sums
// the two halves' crash counts, and the upper lane, which holds the loss row,
// tallies the path.
// ---------------------------------------------------------------------------
// NB = Philox blocks per path (n <= 4 NB), T = row tiles; both compile-time
// so every block, step and tile bound below is a constant.
// Groups g0, g0 + gstep, ... < g1 of 64 paths of query a, accumulated into
// the lane tally t (no workgroup barriers).  A wave runs a group as two
// 32-path tiles.  For tile path p = lane & 31, lane l supplies normals
// 16s + 8h .. 16s + 8h + 7 of step s (h = l >> 5, the half-wave): exactly its
// slot of the B operand (lane l: column l % 32, k = 8 (l / 32) + e), so the
// normals need no lane exchange and the half-wave pair shares one path.
template <int NB, int T>
__device__ __forceinline__ void mfma_wave_groups(const PathArgs& a, uint64_t g0, uint64_t g1,
                                                 uint64_t gstep, LaneTally& t, unsigned* sh_hist) {
  constexpr int S = (NB + 3) / 4;  // 16-wide K steps
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5;
  const Key2 key = key_of(a.seed);
  // Factor fragments are streamed per step from L1/L2 (16 B per lane per
  // tile and part) instead of being held for the whole kernel.
  const uint4* __restrict__ lf = a.lfrag + lane * 2;
  // Row initial values (pads +inf so they never count; the loss row, last
  // tile r = 15 in the upper half-wave, is excluded explicitly).
  const float4* __restrict__ rinit = (const float4*)a.rowinit;

  for (uint64_t grp = g0; grp < g1; grp += gstep) {
#pragma unroll 1
    for (int sub = 0; sub < 2; ++sub) {
      const uint64_t q = grp * 64 + 32 * sub + (lane & 31);
      const uint64_t path = a.path0 + q;
      f32x16 acc[T];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        bf16x8 Lhi[T], Llo[T];
#pragma unroll
        for (int tt = 0; tt < T; ++tt) {
          const int steps = (tt == T - 1) ? S : (2 * tt + 2 < S ? 2 * tt + 2 : S);
          if (s < steps) {
            const size_t e = (size_t)(tt * S + s) * 128;
            Lhi[tt] = __builtin_bit_cast(bf16x8, lf[e]);
            Llo[tt] = __builtin_bit_cast(bf16x8, lf[e + 1]);
          }
        }
        float z[8];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int b = 4 * s + 2 * h + u;
          if (b < NB - 1 || (b == NB - 1 && 4 * b + 2 < a.n)) {
            normals4(key, path, (uint32_t)b, &z[4 * u]);
          } else if (b == NB - 1) {
            // the last block's second pair (normals 4b+2, 4b+3 >= n) only
            // meets zero factor columns: skipped, as zeros
            normals2(key, path, (uint32_t)b, &z[4 * u]);
            z[4 * u + 2] = z[4 * u + 3] = 0.0f;
          } else {
            z[4 * u] = z[4 * u + 1] = z[4 * u + 2] = z[4 * u + 3] = 0.0f;
          }
        }
        uint32_t zh[4], zl[4];
        split8(z, zh, zl);
        const bf16x8 bh = as_bf16x8(zh), bl = as_bf16x8(zl);
#pragma unroll
        for (int tt = 0; tt < T; ++tt) {
          const int steps = (tt == T - 1) ? S : (2 * tt + 2 < S ? 2 * tt + 2 : S);
          if (s < steps) {
            if (s == 0) {
              f32x16 init;
#pragma unroll
              for (int r4 = 0; r4 < 4; ++r4) {
                const float4 v = rinit[(32 * tt + 8 * r4 + 4 * h) >> 2];
                init[4 * r4 + 0] = v.x;
                init[4 * r4 + 1] = v.y;
                init[4 * r4 + 2] = v.z;
                init[4 * r4 + 3] = v.w;
              }
              acc[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo[tt], bh, init, 0, 0, 0);
            } else {
              acc[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo[tt], bh, acc[tt], 0, 0, 0);
            }
            acc[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi[tt], bl, acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi[tt], bh, acc[tt], 0, 0, 0);
          }
        }
      }
      // D: lane l holds column l % 32 (its path), rows 8(r/4) + 4h + r%4
      uint32_t c = 0;
#pragma unroll
      for (int tt = 0; tt < T; ++tt)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const bool excl = tt == T - 1 && r == 15 && h == 1;  // the loss row
          c += (acc[tt][r] < a.thr && !excl) ? 1u : 0u;
        }
      uint32_t c2 = c;
      swap32(c, c2);  // c = the lower half's counts, c2 = the upper half's, in every lane
      const int Sp = (int)(c + c2);
      const float loss = acc[T - 1][15];  // row 32T - 1 in the upper half-wave
      if (h == 1 && q < a.npaths) tally_path(a, t, q, Sp, loss, false, sh_hist);
    }
  }
}

template <int NB, int T>
__global__ __launch_bounds__(256) void k_mfma(PathArgs a) {
  __shared__ unsigned sh_hist[32 * kMfmaMaxT];
  const int n = a.n;
  const bool want_hist = a.s_hist != nullptr;
  if (want_hist) {
    for (int i = threadIdx.x; i <= n; i += blockDim.x) sh_hist[i] = 0u;
    __syncthreads();
  }
  LaneTally t;
  zero_tally(t);
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  mfma_wave_groups<NB, T>(a, (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6),
                          (a.npaths + 63) / 64, waves, t, want_hist ? sh_hist : nullptr);
  flush_tally_block(a, t);
  if (want_hist) {
    __syncthreads();
    for (int i = threadIdx.x; i <= n; i += blockDim.x)
      if (sh_hist[i]) atomicAdd(&a.s_hist[i], (unsigned long long)sh_hist[i]);
  }
}

// Persistent batch kernel for one (NB, T) class: every wave takes work items
// (a run of 64-path groups of one query) from the class's atomic counter until
// the list is exhausted, and flushes its tally into the item's query after
// each item.  Waves never wait for each other.
template <int NB, int T>
__global__ __launch_bounds__(256) void k_batch_small(BatchArgs b) {
  for (;;) {
    unsigned it = 0;
    if ((threadIdx.x & 63) == 0) it = atomicAdd(b.next, 1u);
    it = __builtin_amdgcn_readfirstlane(it);  // lane 0 took the item for the wave
    if (it >= b.nitems) break;
    const WorkItem w = load_const(b.items + it);
    const PathArgs a = load_const(b.qs + w.query);
    LaneTally t;
    zero_tally(t);
    mfma_wave_groups<NB, T>(a, w.first_group, w.first_group + w.ngroups, 1, t, nullptr);
    flush_tally(a, t);
  }
}

// MFMA instantiations: NB = ceil(n/4) Philox blocks, T = ceil((n+1)/32) tiles.
// This file is compiled once per part (MCR_MFMA_PART = 0..3), each part
// instantiating eight consecutive NB, so the instantiations build in parallel.
#ifndef MCR_MFMA_PART
#define MCR_MFMA_PART 0
#endif
constexpr int kNbLo = 8 * MCR_MFMA_PART + 1;
constexpr int kNbHi = 8 * MCR_MFMA_PART + 8;

struct MfmaTables {
  PathKernel path[kMfmaNB + 1][kMfmaMaxT + 1];
  BatchKernel batch[kMfmaNB + 1][kMfmaMaxT + 1];
};

template <int NB>
struct MfmaFill {
  static void fill(MfmaTables* t) {
    constexpr int Tlo = (4 * NB - 2 + 31) / 32;  // n = 4 NB - 3
    constexpr int Thi = (4 * NB + 1 + 31) / 32;  // n = 4 NB
    t->path[NB][Tlo] = k_mfma<NB, Tlo>;
    t->batch[NB][Tlo] = k_batch_small<NB, Tlo>;
    if constexpr (Thi != Tlo && Thi <= kMfmaMaxT) {
      t->path[NB][Thi] = k_mfma<NB, Thi>;
      t->batch[NB][Thi] = k_batch_small<NB, Thi>;
    }
    if constexpr (NB < kNbHi) MfmaFill<NB + 1>::fill(t);
  }
};

#define MCR_CAT2(a, b) a##b
#define MCR_CAT(a, b) MCR_CAT2(a, b)
void MCR_CAT(mfma_fill_part, MCR_MFMA_PART)(MfmaTables* t) { MfmaFill<kNbLo>::fill(t); }

#if MCR_MFMA_PART == 0
void mfma_fill_part1(MfmaTables* t);
void mfma_fill_part2(MfmaTables* t);
void mfma_fill_part3(MfmaTables* t);

static const MfmaTables& mfma_tables() {
  static const MfmaTables tables = [] {
    MfmaTables t = {};
    mfma_fill_part0(&t);
    mfma_fill_part1(&t);
    mfma_fill_part2(&t);
    mfma_fill_part3(&t);
    return t;
  }();
  return tables;
}

PathKernel mfma_kernel(int n, int T) {
  const int nb = (n + 3) / 4;
  if (nb < 1 || nb > kMfmaNB || T < 1 || T > kMfmaMaxT) return nullptr;
  return mfma_tables().path[nb][T];
}

BatchKernel batch_small_kernel(int n, int T) {
  const int nb = (n + 3) / 4;
  if (nb < 1 || nb > kMfmaNB || T < 1 || T > kMfmaMaxT) return nullptr;
  return mfma_tables().batch[nb][T];
}
#endif

}  // namespace mcr
