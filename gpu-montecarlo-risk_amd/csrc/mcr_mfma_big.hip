// mcr_mfma_big.hip -- bf16x3 MFMA path kernel for large n (up to 608).
//
// The large-n form of the reference kernel's correlate step
// (src/05-GPU/monte_carlos_cuda.cu:51-57): R = mu + L Z as a GEMM
// D[rows x paths] = Lfrag[rows x k] . Z[k x paths] on v_mfma_f32_32x32x16_bf16,
// with the same augmented rows as the small MFMA kernel (assets, +inf pads,
// loss row 32T-1 = -v, initial -w^T mu) and the same bf16 hi/lo split.
//
// A workgroup of four waves owns 64 paths.  The 64 x 32T accumulator block is
// spread over the waves' VGPRs: row tile t belongs to wave (t/4 even ? t%4 :
// 3-t%4) (snake order, since tile t needs min(2t+2, S) steps), at most TW
// tiles per wave.  Per 16-wide K step every lane of the workgroup generates one
// Philox block (4 normals of one path) and writes its bf16 hi/lo parts straight
// into a double-buffered LDS tile already in B-operand order, so step s+1 is
// generated while step s is multiplied; factor fragments stream from L2.
#include "mcr_kern_common.h"

namespace mcr {

// LDS operand tiles: [buffer][path tile][hi, lo][operand lane][8 bf16]
struct ZTile {
  uint4 v[2][2][2][64];
};

// Step s, this lane's block: path rp of the group, normals 16 s + 4 ru .. +3.
__device__ __forceinline__ void gen_step(ZTile* zt, int buf, int s, int nb, Key2 key,
                                         uint64_t path, int rp, int ru) {
  const int b = 4 * s + ru;
  float z[4];
  if (b < nb) {
    normals4(key, path, (uint32_t)b, z);
  } else {
    z[0] = z[1] = z[2] = z[3] = 0.0f;
  }
  uint32_t h0 = pack_bf16x2(z[0], z[1]), h1 = pack_bf16x2(z[2], z[3]);
  const uint32_t l0 = pack_bf16x2(z[0] - as_f32(h0 << 16), z[1] - as_f32(h0 & 0xffff0000u));
  const uint32_t l1 = pack_bf16x2(z[2] - as_f32(h1 << 16), z[3] - as_f32(h1 & 0xffff0000u));
  const int pt = rp >> 5;
  const int oplane = (rp & 31) + 32 * (ru >> 1);
  uint2* hi = reinterpret_cast<uint2*>(&zt->v[buf][pt][0][oplane]) + (ru & 1);
  uint2* lo = reinterpret_cast<uint2*>(&zt->v[buf][pt][1][oplane]) + (ru & 1);
  *hi = make_uint2(h0, h1);
  *lo = make_uint2(l0, l1);
}

struct BigShared {
  ZTile zt;
  uint32_t sS[4][64];
  float sLoss[64];
};

// Groups g0, g0 + gstep, ... < g1 of 64 paths of query a, accumulated into t.
template <int TW>
__device__ __forceinline__ void mfma_lds_groups(const PathArgs& a, uint64_t g0, uint64_t g1,
                                                uint64_t gstep, BigShared& sh, LaneTally& t,
                                                unsigned* sh_hist) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 5;
  const int n = a.n, T = a.T, S = a.S;
  const int nb = (n + 3) >> 2;
  const Key2 key = key_of(a.seed);
  const uint4* __restrict__ lf = a.lfrag + lane * 2;
  const float4* __restrict__ rinit = (const float4*)a.rowinit;

  int tile[TW], steps[TW];
#pragma unroll
  for (int i = 0; i < TW; ++i) {
    tile[i] = 4 * i + ((i & 1) ? 3 - wave : wave);
    steps[i] = tile[i] >= T ? 0 : (tile[i] == T - 1 ? S : (2 * tile[i] + 2 < S ? 2 * tile[i] + 2 : S));
  }
  const int loss_slot = (T - 1) / 4;  // i of tile T-1 (its owner wave)
  const int loss_wave = (loss_slot & 1) ? 3 - (T - 1) % 4 : (T - 1) % 4;

  const int rp = threadIdx.x & 63, ru = threadIdx.x >> 6;
  for (uint64_t grp = g0; grp < g1; grp += gstep) {
    const uint64_t rpath = a.path0 + grp * 64 + rp;
    gen_step(&sh.zt, 0, 0, nb, key, rpath, rp, ru);
    __syncthreads();
    f32x16 accA[TW], accB[TW];
    for (int s = 0; s < S; ++s) {
      if (s + 1 < S) gen_step(&sh.zt, (s + 1) & 1, s + 1, nb, key, rpath, rp, ru);
      const int cb = s & 1;
      const bf16x8 ah = __builtin_bit_cast(bf16x8, sh.zt.v[cb][0][0][lane]);
      const bf16x8 al = __builtin_bit_cast(bf16x8, sh.zt.v[cb][0][1][lane]);
      const bf16x8 bh = __builtin_bit_cast(bf16x8, sh.zt.v[cb][1][0][lane]);
      const bf16x8 bl = __builtin_bit_cast(bf16x8, sh.zt.v[cb][1][1][lane]);
#pragma unroll
      for (int i = 0; i < TW; ++i) {
        if (s < steps[i]) {
          const size_t e = (size_t)(tile[i] * S + s) * 128;
          const bf16x8 Lhi = __builtin_bit_cast(bf16x8, lf[e]);
          const bf16x8 Llo = __builtin_bit_cast(bf16x8, lf[e + 1]);
          if (s == 0) {
            f32x16 init;
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
              const float4 v = rinit[(32 * tile[i] + 8 * r4 + 4 * g) >> 2];
              init[4 * r4 + 0] = v.x;
              init[4 * r4 + 1] = v.y;
              init[4 * r4 + 2] = v.z;
              init[4 * r4 + 3] = v.w;
            }
            accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo, ah, init, 0, 0, 0);
            accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo, bh, init, 0, 0, 0);
          } else {
            accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo, ah, accA[i], 0, 0, 0);
            accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo, bh, accB[i], 0, 0, 0);
          }
          accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, al, accA[i], 0, 0, 0);
          accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, bl, accB[i], 0, 0, 0);
          accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, ah, accA[i], 0, 0, 0);
          accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, bh, accB[i], 0, 0, 0);
        }
      }
      __syncthreads();
    }
    // partial crash counts over this wave's tiles
    uint32_t SA = 0, SB = 0;
#pragma unroll
    for (int i = 0; i < TW; ++i) {
      if (steps[i] > 0) {
        const bool last = tile[i] == T - 1;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const bool excl = last && r == 15 && g == 1;  // the loss row
          SA += (accA[i][r] < a.thr && !excl) ? 1u : 0u;
          SB += (accB[i][r] < a.thr && !excl) ? 1u : 0u;
        }
      }
    }
    swap32(SA, SB);
    sh.sS[wave][lane] = SA + SB;
    if (wave == loss_wave) {
      float la = 0.0f, lb = 0.0f;
#pragma unroll
      for (int i = 0; i < TW; ++i)
        if (i == loss_slot) {
          la = accA[i][15];
          lb = accB[i][15];
        }
      uint32_t ua = as_u32(la), ub = as_u32(lb);
      swap32(ua, ub);
      sh.sLoss[lane] = as_f32(ub);
    }
    __syncthreads();
    const uint64_t q = grp * 64 + lane;
    if (wave == 0 && q < a.npaths) {
      const int Sp = (int)(sh.sS[0][lane] + sh.sS[1][lane] + sh.sS[2][lane] + sh.sS[3][lane]);
      tally_path(a, t, q, Sp, sh.sLoss[lane], false, sh_hist);
    }
    __syncthreads();
  }
}

template <int TW>
__global__ __launch_bounds__(256) void k_mfma_big(PathArgs a) {
  __shared__ BigShared sh;
  __shared__ unsigned sh_hist[kMaxN + 1];
  const bool want_hist = a.s_hist != nullptr;
  if (want_hist) {
    for (int i = threadIdx.x; i <= a.n; i += blockDim.x) sh_hist[i] = 0u;
  }
  LaneTally t;
  zero_tally(t);
  mfma_lds_groups<TW>(a, blockIdx.x, (a.npaths + 63) / 64, gridDim.x, sh, t,
                      want_hist ? sh_hist : nullptr);
  flush_tally(a, t);
  if (want_hist) {
    __syncthreads();
    for (int i = threadIdx.x; i <= a.n; i += blockDim.x)
      if (sh_hist[i]) atomicAdd(&a.s_hist[i], (unsigned long long)sh_hist[i]);
  }
}

// Persistent batch kernel for one TW class: every workgroup takes work items
// (a run of 64-path groups of one query) from the class's atomic counter until
// the list is exhausted; the exit test is uniform, so every wave of every
// workgroup leaves the loop.  Tallies are flushed into the item's query after
// each item.
template <int TW>
__global__ __launch_bounds__(256) void k_batch_paths(BatchArgs b) {
  __shared__ BigShared sh;
  __shared__ unsigned s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(b.next, 1u);
    __syncthreads();
    // uniform by construction; readfirstlane lets the descriptor live in SGPRs
    const unsigned it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (it >= b.nitems) break;
    // item and descriptor are read-only for the launch: scalar loads
    const WorkItem w = load_const(b.items + it);
    const PathArgs a = load_const(b.qs + w.query);
    LaneTally t;
    zero_tally(t);
    mfma_lds_groups<TW>(a, w.first_group, w.first_group + w.ngroups, 1, sh, t, nullptr);
    flush_tally(a, t);
  }
}

__global__ void k_batch_clear(unsigned* next, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) next[i] = 0u;
}

hipError_t launch_batch_clear(unsigned* next, int count, hipStream_t stream) {
  hipLaunchKernelGGL(k_batch_clear, dim3(1), dim3(64), 0, stream, next, count);
  return hipGetLastError();
}

hipError_t launch_batch_kernel(BatchKernel k, const BatchArgs& b, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, stream, b);
  return hipGetLastError();
}

int batch_kernel_grid(BatchKernel k, int num_cus) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k), 256,
                                                   0) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  return per_cu * num_cus;
}

BatchKernel batch_big_kernel(int T) {
  switch ((T + 3) / 4) {
    case 1: return k_batch_paths<1>;
    case 2: return k_batch_paths<2>;
    case 3: return k_batch_paths<3>;
    case 4: return k_batch_paths<4>;
    case 5: return k_batch_paths<5>;
    default: return nullptr;
  }
}

PathKernel mfma_big_kernel(int T) {
  switch ((T + 3) / 4) {
    case 1: return k_mfma_big<1>;
    case 2: return k_mfma_big<2>;
    case 3: return k_mfma_big<3>;
    case 4: return k_mfma_big<4>;
    case 5: return k_mfma_big<5>;
    default: return nullptr;
  }
}

}  // namespace mcr
