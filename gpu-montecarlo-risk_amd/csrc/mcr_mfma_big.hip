// mcr_mfma_big.hip -- bf16x3 MFMA path kernels for large n (up to 767: 24 row
// tiles of 32).
//
// The large-n form of the reference kernel's correlate step
// (src/05-GPU/monte_carlos_cuda.cu:51-57): R = mu + L Z as a GEMM
// D[rows x paths] = Lfrag[rows x k] . Z[k x paths] on v_mfma_f32_32x32x16_bf16,
// with the same augmented rows as the small MFMA kernel (assets, +inf pads,
// loss row 32T-1 = -v, initial -w^T mu) and the same bf16 hi/lo split.
//
// A workgroup of NW waves (4; 8 for 13-16 row tiles in batches) owns 64 paths.  The 64 x 32T
// accumulator block is spread over the waves' VGPRs: row tile t belongs to
// wave (t/NW even ? t%NW : NW-1-t%NW) (snake order, since tile t needs
// min(2t+2, S) steps), at most TW tiles per wave; the shape keeps at least two
// waves per SIMD so one's MFMAs overlap another's normal generation.  Per
// phase of NW/4 16-wide K steps every lane of the workgroup generates one
// Philox block (4 normals of one path) and writes its bf16 hi/lo parts
// straight into a double-buffered LDS tile already in B-operand order, so the
// next phase is generated while this one is multiplied; factor fragments
// stream from L2.  For n >= 384 (13-24 row tiles; batches from 17 tiles) the
// roles are split instead: 8 waves only multiply, 4 only generate
// (k_mfma_spec below).
#include "mcr_kern_common.h"

namespace mcr {

// LDS operand tiles: [buffer][step of the phase][path tile][hi, lo][operand
// lane][8 bf16].  A phase is NW/4 consecutive 16-wide K steps, generated
// by the whole workgroup (one Philox block per lane) between two barriers.
template <int NW>
struct ZTile {
  uint4 v[2][NW / 4][2][2][64];
};

// Step s of the group (its normals 16 s .. 16 s + 15 of path rp), this lane's
// Philox block ru (normals 16 s + 4 ru .. + 3), into phase slot sp of buffer buf.
template <int NW>
__device__ __forceinline__ void gen_step(ZTile<NW>* zt, int buf, int sp, int s, int nb, Key2 key,
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
  uint2* hi = reinterpret_cast<uint2*>(&zt->v[buf][sp][pt][0][oplane]) + (ru & 1);
  uint2* lo = reinterpret_cast<uint2*>(&zt->v[buf][sp][pt][1][oplane]) + (ru & 1);
  *hi = make_uint2(h0, h1);
  *lo = make_uint2(l0, l1);
}

// Factor fragments are read through a buffer resource: fragment (tile, s) of
// a lane is at byte (tile S + s) 2048 + 32 lane (+16 for the lo part).  The
// per-step part of the address is then a scalar offset, so no VGPR is
// rewritten while a load is pending; with a per-lane 64-bit address the
// compiler reuses the registers of a pending load for the next address, and
// each step waits for the previous tile's loads before issuing the next.
constexpr int kBufferRsrcWord3 = 0x00020000;  // raw buffer, gfx9 family

// The base and size go through readfirstlane: where the query record came
// from memory the compiler cannot prove them uniform, and a descriptor it
// takes for divergent is used through a loop over lanes at every load.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t frag_rsrc(const PathArgs& a) {
  const uint64_t p = (uint64_t)a.lfrag;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)p);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(p >> 32));
  const int bytes = __builtin_amdgcn_readfirstlane(a.T * a.S * 2048);
  return __builtin_amdgcn_make_buffer_rsrc((void*)(((uint64_t)hi << 32) | lo), (short)0, bytes,
                                           kBufferRsrcWord3);
}

// part 0: hi, 1: lo
__device__ __forceinline__ uint4 load_frag(__amdgpu_buffer_rsrc_t rs, int voff, int tile, int S,
                                           int s, int part) {
  return __builtin_bit_cast(
      uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff + 16 * part, (tile * S + s) * 2048, 0));
}

template <int NW>
struct BigShared {
  ZTile<NW> zt;
  uint32_t sS[NW][64];
  float sLoss[64];
};

// Groups g0, g0 + gstep, ... < g1 of 64 paths of query a, accumulated into t.
// NW waves share each group: row tile t belongs to the wave dealt by a snake
// over the waves (tile t needs min(2t+2, S) steps, so the deal balances the
// triangle), at most TW tiles per wave.  A phase is P = NW/4 K steps.  With
// PF the next K step's factor fragments are loaded before this step's MFMAs
// (also across the phase barrier); that costs 8 TW registers, so it is used
// only where it does not lower the occupancy.
template <int TW, int NW, bool PF>
__device__ __forceinline__ void mfma_lds_groups(const PathArgs& a, uint64_t g0, uint64_t g1,
                                                uint64_t gstep, BigShared<NW>& sh, LaneTally& t,
                                                unsigned* sh_hist) {
  constexpr int P = NW / 4;
  // the wave index is uniform; readfirstlane lets the compiler see it, so
  // the tile deal and every per-tile step test are scalar (no exec masks)
  const int lane = threadIdx.x & 63, g = lane >> 5;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n = a.n, T = a.T, S = a.S;
  const int nb = (n + 3) >> 2;
  const int nph = (S + P - 1) / P;
  const Key2 key = key_of(a.seed);
  const __amdgpu_buffer_rsrc_t lrs = frag_rsrc(a);
  const int voff = lane * 32;
  const float4* __restrict__ rinit = (const float4*)a.rowinit;

  int tile[TW], steps[TW];
#pragma unroll
  for (int i = 0; i < TW; ++i) {
    tile[i] = NW * i + ((i & 1) ? NW - 1 - wave : wave);
    steps[i] = tile[i] >= T ? 0 : (tile[i] == T - 1 ? S : (2 * tile[i] + 2 < S ? 2 * tile[i] + 2 : S));
  }
  const int loss_slot = (T - 1) / NW;  // i of tile T-1 (its owner wave)
  const int loss_wave = (loss_slot & 1) ? NW - 1 - (T - 1) % NW : (T - 1) % NW;

  // generator role of this thread: path rp, block ru of step sp of a phase
  const int rp = threadIdx.x & 63, ru = wave & 3, sp = wave >> 2;
  for (uint64_t grp = g0; grp < g1; grp += gstep) {
    const uint64_t rpath = a.path0 + grp * 64 + rp;
    if (sp < S) gen_step<NW>(&sh.zt, 0, sp, sp, nb, key, rpath, rp, ru);
    __syncthreads();
    // accumulators start at the rows' initial values (+inf pads, -w^T mu):
    // loaded here without the fragment prefetch, at the tile's first step
    // with it (measured: each placement is the faster one for its shapes)
    f32x16 accA[TW], accB[TW];
    auto init_acc = [&](int i) {
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) {
        const float4 v = rinit[(32 * tile[i] + 8 * r4 + 4 * g) >> 2];
        accA[i][4 * r4 + 0] = v.x;
        accA[i][4 * r4 + 1] = v.y;
        accA[i][4 * r4 + 2] = v.z;
        accA[i][4 * r4 + 3] = v.w;
      }
      accB[i] = accA[i];
    };
    if (!PF) {
#pragma unroll
      for (int i = 0; i < TW; ++i)
        if (steps[i] > 0) init_acc(i);
    }
    // One K step s (slot q of the phase in LDS buffer cb) with this step's
    // factor fragments ch/cl; with PF the next step's are loaded into nh/nl
    // before the MFMAs are issued, without it every active tile's fragments
    // of this step are (eh/el), so the tiles' loads are in flight together
    // (the registers this takes spill a few values outside the step loop).
    auto k_step = [&](int s, int q, int cb, const uint4 (&ch)[TW], const uint4 (&cl)[TW],
                      uint4 (&nh)[TW], uint4 (&nl)[TW]) {
#pragma unroll
      for (int i = 0; i < TW; ++i) {
        if (PF && s + 1 < steps[i]) {
          nh[i] = load_frag(lrs, voff, tile[i], S, s + 1, 0);
          nl[i] = load_frag(lrs, voff, tile[i], S, s + 1, 1);
        }
      }
      uint4 eh[TW], el[TW];
      if (!PF) {
#pragma unroll
        for (int i = 0; i < TW; ++i) {
          if (s < steps[i]) {
            eh[i] = load_frag(lrs, voff, tile[i], S, s, 0);
            el[i] = load_frag(lrs, voff, tile[i], S, s, 1);
          }
        }
      }
      const bf16x8 ah = __builtin_bit_cast(bf16x8, sh.zt.v[cb][q][0][0][lane]);
      const bf16x8 al = __builtin_bit_cast(bf16x8, sh.zt.v[cb][q][0][1][lane]);
      const bf16x8 bh = __builtin_bit_cast(bf16x8, sh.zt.v[cb][q][1][0][lane]);
      const bf16x8 bl = __builtin_bit_cast(bf16x8, sh.zt.v[cb][q][1][1][lane]);
#pragma unroll
      for (int i = 0; i < TW; ++i) {
        if (s < steps[i]) {
          uint4 h = ch[i], l = cl[i];
          if (!PF) {
            h = eh[i];
            l = el[i];
          }
          const bf16x8 Lhi = __builtin_bit_cast(bf16x8, h);
          const bf16x8 Llo = __builtin_bit_cast(bf16x8, l);
          if (PF && s == 0) init_acc(i);
          accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo, ah, accA[i], 0, 0, 0);
          accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo, bh, accB[i], 0, 0, 0);
          accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, al, accA[i], 0, 0, 0);
          accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, bl, accB[i], 0, 0, 0);
          accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, ah, accA[i], 0, 0, 0);
          accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, bh, accB[i], 0, 0, 0);
        }
      }
    };
    // factor fragments of the current K step (A) and, with PF, of the next
    // (B, loaded before this step's MFMAs, also across the phase barrier)
    uint4 LAh[TW], LAl[TW], LBh[TW], LBl[TW];
#pragma unroll
    for (int i = 0; i < TW; ++i) {
      if (PF && steps[i] > 0) {
        LAh[i] = load_frag(lrs, voff, tile[i], S, 0, 0);
        LAl[i] = load_frag(lrs, voff, tile[i], S, 0, 1);
      }
    }
    for (int ph = 0; ph < nph; ++ph) {
      const int sn = (ph + 1) * P + sp;
      if (ph + 1 < nph && sn < S) gen_step<NW>(&sh.zt, (ph + 1) & 1, sp, sn, nb, key, rpath, rp, ru);
      const int cb = ph & 1;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int s = ph * P + q;
        if (s >= S) break;
        k_step(s, q, cb, LAh, LAl, LBh, LBl);
        // unconditional, so that the copies are renamings of the unrolled
        // steps (a conditional copy is a move that waits for the load)
        if (PF) {
#pragma unroll
          for (int i = 0; i < TW; ++i) {
            LAh[i] = LBh[i];
            LAl[i] = LBl[i];
          }
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
      int Sp = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) Sp += (int)sh.sS[w][lane];
      tally_path(a, t, q, Sp, sh.sLoss[lane], false, sh_hist);
    }
    // no barrier here: sS and sLoss are next written at the end of the next
    // group, after its phase barriers, which wave 0 reaches only once its
    // tally is done; the next group's first writes go to the operand tiles,
    // whose last reads preceded the final phase barrier
  }
}

// Waves per SIMD the register budget must allow (measured best): three for
// the 3-tile 4-wave shape (it lands within a few registers of 168, and the
// allocation granularity of 8 decides between two and three), four for the
// 2-tile shapes (128 VGPRs, a few spilled; two 8-wave workgroups per CU).
constexpr int big_min_waves(int TW, int NW) {
  return NW == 4 && TW == 3 ? 3 : (TW == 2 ? 4 : 1);
}

template <int TW, int NW, bool PF>
__global__ __launch_bounds__(64 * NW, big_min_waves(TW, NW)) void k_mfma_big(PathArgs a) {
  __shared__ BigShared<NW> sh;
  __shared__ unsigned sh_hist[kMaxN + 1];
  const bool want_hist = a.s_hist != nullptr;
  if (want_hist) {
    for (int i = threadIdx.x; i <= a.n; i += blockDim.x) sh_hist[i] = 0u;
  }
  LaneTally t;
  zero_tally(t);
  mfma_lds_groups<TW, NW, PF>(a, blockIdx.x, (a.npaths + 63) / 64, gridDim.x, sh, t,
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
template <int TW, int NW, bool PF>
__global__ __launch_bounds__(64 * NW, big_min_waves(TW, NW)) void k_batch_paths(BatchArgs b) {
  __shared__ BigShared<NW> sh;
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
    mfma_lds_groups<TW, NW, PF>(a, w.first_group, w.first_group + w.ngroups, 1, sh, t, nullptr);
    flush_tally(a, t);
  }
}

// ---------------------------------------------------------------------------
// Specialised form: NC consumer waves hold the accumulators (64 paths, up to
// TW row tiles each) and only multiply; 4 producer waves only generate the
// next phase's normals (a phase is P K steps: P Philox blocks per producer
// lane).  A SIMD runs one producer beside NC/4 consumers, so generation and
// MFMAs overlap within every phase instead of alternating in every wave.
// ---------------------------------------------------------------------------
constexpr int kProducers = 4;

template <int NC, int P>
struct SpecShared {
  ZTile<4 * P> zt;  // P steps per phase, double-buffered
  uint32_t sS[NC][64];
  float sLoss[64];
};

template <int NC, int TW, int P>
__device__ __forceinline__ void mfma_spec_groups(const PathArgs& a, uint64_t g0, uint64_t g1,
                                                 uint64_t gstep, SpecShared<NC, P>& sh,
                                                 LaneTally& t, unsigned* sh_hist) {
  const int lane = threadIdx.x & 63, g = lane >> 5;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n = a.n, T = a.T, S = a.S;
  const int nb = (n + 3) >> 2;
  const int nph = (S + P - 1) / P;
  const Key2 key = key_of(a.seed);

  // The two roles run separate loops with the same barriers per group
  // (1 + nph + 1), so neither keeps the other's registers live.
  if (wave >= NC) {
    // producer lane: path rp, blocks kP .. kP+P-1 of the phase's 4P per path
    // (block j: step j / 4 of the phase, Philox block j % 4 of that step)
    const int rp = lane, k0 = P * (wave - NC);
    for (uint64_t grp = g0; grp < g1; grp += gstep) {
      const uint64_t rpath = a.path0 + grp * 64 + rp;
#pragma unroll
      for (int u = 0; u < P; ++u) {
        const int j = k0 + u, sp = j >> 2;
        if (sp < S) gen_step<4 * P>(&sh.zt, 0, sp, sp, nb, key, rpath, rp, j & 3);
      }
      __syncthreads();
      for (int ph = 0; ph < nph; ++ph) {
        if (ph + 1 < nph) {
#pragma unroll
          for (int u = 0; u < P; ++u) {
            const int j = k0 + u, sp = j >> 2, sn = (ph + 1) * P + sp;
            if (sn < S) gen_step<4 * P>(&sh.zt, (ph + 1) & 1, sp, sn, nb, key, rpath, rp, j & 3);
          }
        }
        __syncthreads();
      }
      __syncthreads();  // the consumers' count barrier
    }
    return;
  }

  const __amdgpu_buffer_rsrc_t lrs = frag_rsrc(a);
  const int voff = lane * 32;
  const float4* __restrict__ rinit = (const float4*)a.rowinit;
  int tile[TW], steps[TW];
#pragma unroll
  for (int i = 0; i < TW; ++i) {
    tile[i] = NC * i + ((i & 1) ? NC - 1 - wave : wave);
    steps[i] = tile[i] >= T ? 0 : (tile[i] == T - 1 ? S : (2 * tile[i] + 2 < S ? 2 * tile[i] + 2 : S));
  }
  const int loss_slot = (T - 1) / NC;
  const int loss_wave = (loss_slot & 1) ? NC - 1 - (T - 1) % NC : (T - 1) % NC;
  for (uint64_t grp = g0; grp < g1; grp += gstep) {
    __syncthreads();  // the first phase's normals are in
    f32x16 accA[TW], accB[TW];
#pragma unroll
    for (int i = 0; i < TW; ++i) {
      if (steps[i] > 0) {
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
          const float4 v = rinit[(32 * tile[i] + 8 * r4 + 4 * g) >> 2];
          accA[i][4 * r4 + 0] = v.x;
          accA[i][4 * r4 + 1] = v.y;
          accA[i][4 * r4 + 2] = v.z;
          accA[i][4 * r4 + 3] = v.w;
        }
        accB[i] = accA[i];
      }
    }
    for (int ph = 0; ph < nph; ++ph) {
      const int cb = ph & 1;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int s = ph * P + q;
        if (s >= S) break;
        // every active tile's factor fragments are requested before the
        // first MFMA of the step, so the loads of all tiles are in flight
        // together instead of one tile's at a time
        uint4 fh[TW], fl[TW];
#pragma unroll
        for (int i = 0; i < TW; ++i) {
          if (s < steps[i]) {
            fh[i] = load_frag(lrs, voff, tile[i], S, s, 0);
            fl[i] = load_frag(lrs, voff, tile[i], S, s, 1);
          }
        }
        const bf16x8 ah = __builtin_bit_cast(bf16x8, sh.zt.v[cb][q][0][0][lane]);
        const bf16x8 al = __builtin_bit_cast(bf16x8, sh.zt.v[cb][q][0][1][lane]);
        const bf16x8 bh = __builtin_bit_cast(bf16x8, sh.zt.v[cb][q][1][0][lane]);
        const bf16x8 bl = __builtin_bit_cast(bf16x8, sh.zt.v[cb][q][1][1][lane]);
#pragma unroll
        for (int i = 0; i < TW; ++i) {
          if (s < steps[i]) {
            const bf16x8 Lhi = __builtin_bit_cast(bf16x8, fh[i]);
            const bf16x8 Llo = __builtin_bit_cast(bf16x8, fl[i]);
            accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo, ah, accA[i], 0, 0, 0);
            accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Llo, bh, accB[i], 0, 0, 0);
            accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, al, accA[i], 0, 0, 0);
            accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, bl, accB[i], 0, 0, 0);
            accA[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, ah, accA[i], 0, 0, 0);
            accB[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Lhi, bh, accB[i], 0, 0, 0);
          }
        }
      }
      __syncthreads();
    }
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
      int Sp = 0;
#pragma unroll
      for (int v = 0; v < NC; ++v) Sp += (int)sh.sS[v][lane];
      tally_path(a, t, q, Sp, sh.sLoss[lane], false, sh_hist);
    }
  }
}

template <int NC, int TW, int P>
__global__ __launch_bounds__(64 * (NC + kProducers)) void k_mfma_spec(PathArgs a) {
  __shared__ SpecShared<NC, P> sh;
  __shared__ unsigned sh_hist[kMfmaLdsMaxN + 1];  // the 3-tile form runs up to 24 row tiles
  const bool want_hist = a.s_hist != nullptr;
  if (want_hist) {
    for (int i = threadIdx.x; i <= a.n; i += blockDim.x) sh_hist[i] = 0u;
  }
  LaneTally t;
  zero_tally(t);
  mfma_spec_groups<NC, TW, P>(a, blockIdx.x, (a.npaths + 63) / 64, gridDim.x, sh, t,
                              want_hist ? sh_hist : nullptr);
  flush_tally(a, t);
  if (want_hist) {
    __syncthreads();
    for (int i = threadIdx.x; i <= a.n; i += blockDim.x)
      if (sh_hist[i]) atomicAdd(&a.s_hist[i], (unsigned long long)sh_hist[i]);
  }
}

template <int NC, int TW, int P>
__global__ __launch_bounds__(64 * (NC + kProducers)) void k_batch_spec(BatchArgs b) {
  __shared__ SpecShared<NC, P> sh;
  __shared__ unsigned s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(b.next, 1u);
    __syncthreads();
    const unsigned it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (it >= b.nitems) break;
    const WorkItem w = load_const(b.items + it);
    const PathArgs a = load_const(b.qs + w.query);
    LaneTally t;
    zero_tally(t);
    mfma_spec_groups<NC, TW, P>(a, w.first_group, w.first_group + w.ngroups, 1, sh, t, nullptr);
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

hipError_t launch_batch_kernel(BatchKernel k, const BatchArgs& b, int grid, int block,
                               hipStream_t stream) {
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, stream, b);
  return hipGetLastError();
}

// Shapes of the LDS MFMA kernels for T row tiles (measured best per range).
// Four waves with up to 3 tiles each for T <= 12 (several workgroups per CU,
// so one's barrier waits are covered by another's work); beyond, single
// queries run the split form: 8 multiplying waves of 2 (T <= 16) or 3 tiles
// beside 4 generating waves (n = 450: 6.0 ms against 6.4 ms for eight waves of
// 2 tiles; n = 600: 14.85 ms against 15.0 ms for 12 multiplying waves of 2
// tiles and 16.8 ms for eight waves of 3 tiles that all generate).  The 3-tile
// split form runs 4 K steps per phase: half the workgroup barriers of 2-step
// phases, 0.5-1.7 % faster for n = 400-600 in interleaved timings (6- and
// 8-step phases no faster; the 2-tile form gained at n = 450-500 and lost at
// n = 400, so it keeps 2-step phases).  Batches
// keep eight waves of 2 tiles at four waves per SIMD for 13-16 tiles (a
// persistent workgroup there runs many short work items, and two per CU cover
// each other's item changes; the 8 + 4 split form, one workgroup per CU, made
// the uneven batch 3 % slower, and 12 + 4 waves of 1 tile for 9-12 tiles as
// well 10 %), and the single-query shape beyond.  The split form
// loses below 13 tiles (12 + 4 waves of 1 tile: +3 % at n = 300; 8 + 4 of 2:
// +5 %; 4 + 4 of 3: +9 %).
int big_waves(int T) { return T <= 12 ? 4 : 8 + kProducers; }

// Batch shapes per row-tile range, the measured default first, then the
// alternatives the batch autotuner may choose (every shape gives the same
// bits: each accumulator receives the same MFMA sequence whichever wave holds
// it).  Range 0: T = 5-8 (n = 128-255), 1: T = 9-12, 2: T = 13-16, 3: T = 17-24.
struct BigShape {
  BatchKernel k;
  int waves;
};
static const BigShape kBigShapes[kBigRanges][kMaxBigShapes] = {
    {{k_batch_paths<2, 4, false>, 4}, {k_batch_paths<1, 8, true>, 8}},
    {{k_batch_paths<3, 4, false>, 4}, {k_batch_paths<2, 8, false>, 8}},
    {{k_batch_paths<2, 8, false>, 8}, {k_batch_spec<8, 2, 2>, 8 + kProducers}},
    {{k_batch_spec<8, 3, 4>, 8 + kProducers}, {k_batch_spec<12, 2, 2>, 12 + kProducers}},
};

int big_range(int T) { return T <= 8 ? 0 : (T <= 12 ? 1 : (T <= 16 ? 2 : 3)); }
int batch_big_shapes(int range) { return range >= 0 && range < kBigRanges ? kMaxBigShapes : 0; }

static int clamp_shape(int shape) { return shape >= 0 && shape < kMaxBigShapes ? shape : 0; }

int batch_big_waves(int T, int shape) {
  return T <= 4 ? 4 : kBigShapes[big_range(T)][clamp_shape(shape)].waves;
}

BatchKernel batch_big_kernel(int T, int shape) {
  if (T <= 4) return k_batch_paths<1, 4, true>;
  return T <= 24 ? kBigShapes[big_range(T)][clamp_shape(shape)].k : nullptr;
}

PathKernel mfma_big_kernel(int T) {
  if (T <= 4) return k_mfma_big<1, 4, true>;
  if (T <= 8) return k_mfma_big<2, 4, false>;
  if (T <= 12) return k_mfma_big<3, 4, false>;
  if (T <= 16) return k_mfma_spec<8, 2, 2>;
  return T <= 24 ? k_mfma_spec<8, 3, 4> : nullptr;
}

}  // namespace mcr
