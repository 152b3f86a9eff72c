// mcr_reg.hip -- the register path kernel (n <= 64), one instantiation per n.
#include "mcr_kern_common.h"

namespace mcr {

// ---------------------------------------------------------------------------
// Register kernel: one path per lane, N rows in registers (exact N template).
// Column-oriented: normals arrive four at a time; column j updates rows
// j..N-1 and the loss row, so row i sees z_0 .. z_i in order.
// ---------------------------------------------------------------------------
// Column J of the sweep: rows J..N-1 and the loss row take z_J.
// The column's address is re-derived through an empty asm so its scalar loads
// cannot be hoisted above earlier columns: at most ~one column of the factor
// occupies SGPRs at a time (hoisting them all spills the scalar file).
template <int N, int J>
__device__ __forceinline__ void sweep_col(float (&acc)[N], float& g, const float* cp_raw, float zj) {
  if constexpr (J < N) {
    constexpr int off = J * (N + 1) - (J * (J - 1)) / 2;
    // chained on the previous column's last result: its loads are issued
    // once that column is done, so one column at most occupies SGPRs
    asm volatile("" : "+s"(cp_raw) : "v"(g));
    MCR_CONSTP(float) cp = cview(cp_raw);
#pragma unroll
    for (int i = J; i < N; ++i) acc[i] = __builtin_fmaf(cp[off + i - J], zj, acc[i]);
    g = __builtin_fmaf(cp[off + N - J], zj, g);
  }
}

// All normals of the path first (z stays in VGPRs), then the pure column sweep:
// scalar loads of the factor never compete with the RNG for SGPRs.
template <int N, int B>
__device__ __forceinline__ void gen_blocks(float (&z)[4 * ((N + 3) / 4)], Key2 key, uint64_t path) {
  if constexpr (4 * B < N) {
    normals4(key, path, (uint32_t)B, &z[4 * B]);
    gen_blocks<N, B + 1>(z, key, path);
  }
}

template <int N, int J>
__device__ __forceinline__ void sweep_cols(float (&acc)[N], float& g, const float* cp,
                                           const float (&z)[4 * ((N + 3) / 4)]) {
  if constexpr (J < N) {
    sweep_col<N, J>(acc, g, cp, z[J]);
    sweep_cols<N, J + 1>(acc, g, cp, z);
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_valu_reg(PathArgs a) {
  __shared__ unsigned sh_hist[kRegMaxN + 1];
  const bool want_hist = a.s_hist != nullptr;
  if (want_hist) {
    for (int i = threadIdx.x; i <= N; i += blockDim.x) sh_hist[i] = 0u;
    __syncthreads();
  }
  const float mp32 = (float)cview(a.info)->mp;
  const Key2 key = key_of(a.seed);
  const bool want_near = a.thr_lo < a.thr_hi;
  LaneTally t;
  zero_tally(t);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < a.npaths; q += stride) {
    const uint64_t path = a.path0 + q;
    // The factor is loop-invariant; re-deriving its address each path keeps
    // its ~n^2/2 scalar loads inside the loop instead of hoisted into (and
    // spilled out of) the scalar register file.
    // The Philox key schedule is re-derived per path for the same reason.
    // mu is read with vector loads (a lane-varying zero offset) straight into
    // the row registers, keeping its N values out of the scalar file.
    Key2 pkey = key;
    asm volatile("" : "+s"(pkey.k0), "+s"(pkey.k1));
    int vzero = 0;
    asm volatile("" : "+v"(vzero));
    const float* __restrict__ mu = a.mu32 + vzero;
    float acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = mu[i];
    float g = mp32;
    float z[4 * ((N + 3) / 4)];
    gen_blocks<N, 0>(z, pkey, path);
    // the sweep's first scalar loads wait for the last normal
    const float* cp0 = a.colpack;
    asm volatile("" : "+s"(cp0) : "v"(z[N - 1]));
    sweep_cols<N, 0>(acc, g, cp0, z);
    int S = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) S += (acc[i] < a.thr) ? 1 : 0;
    bool near = false;
    if (want_near) {
#pragma unroll
      for (int i = 0; i < N; ++i) near |= (acc[i] >= a.thr_lo) & (acc[i] <= a.thr_hi);
    }
    tally_path(a, t, q, S, -g, near, want_hist ? sh_hist : nullptr);
  }
  flush_tally_block(a, t);
  if (want_hist) {
    __syncthreads();
    for (int i = threadIdx.x; i <= N; i += blockDim.x)
      if (sh_hist[i]) atomicAdd(&a.s_hist[i], (unsigned long long)sh_hist[i]);
  }
}

// Instantiations: this file is compiled once per part (MCR_REG_PART = 0..7);
// compile cost grows with N^2, so the parts cover ranges of similar cost.
#ifndef MCR_REG_PART
#define MCR_REG_PART 0
#endif
constexpr int kRegLo[8] = {1, 33, 41, 47, 52, 56, 60, 63};
constexpr int kRegHi[8] = {32, 40, 46, 51, 55, 59, 62, 64};

template <int N>
struct RegFill {
  static void fill(PathKernel* t) {
    t[N] = k_valu_reg<N>;
    if constexpr (N < kRegHi[MCR_REG_PART]) RegFill<N + 1>::fill(t);
  }
};

#define MCR_CAT2(a, b) a##b
#define MCR_CAT(a, b) MCR_CAT2(a, b)
void MCR_CAT(reg_fill_part, MCR_REG_PART)(PathKernel* t) { RegFill<kRegLo[MCR_REG_PART]>::fill(t); }

#if MCR_REG_PART == 0
void reg_fill_part1(PathKernel* t);
void reg_fill_part2(PathKernel* t);
void reg_fill_part3(PathKernel* t);
void reg_fill_part4(PathKernel* t);
void reg_fill_part5(PathKernel* t);
void reg_fill_part6(PathKernel* t);
void reg_fill_part7(PathKernel* t);

PathKernel reg_kernel(int n) {
  static PathKernel table[kRegMaxN + 1];
  static bool init = false;
  if (!init) {
    for (auto& k : table) k = nullptr;
    reg_fill_part0(table);
    reg_fill_part1(table);
    reg_fill_part2(table);
    reg_fill_part3(table);
    reg_fill_part4(table);
    reg_fill_part5(table);
    reg_fill_part6(table);
    reg_fill_part7(table);
    init = true;
  }
  return (n >= 1 && n <= kRegMaxN) ? table[n] : nullptr;
}
#endif

}  // namespace mcr
