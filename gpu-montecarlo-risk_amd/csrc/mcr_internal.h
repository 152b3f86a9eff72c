// mcr_internal.h -- structures shared by the kernels (mcr_device.hip) and the
// C-ABI implementation (mcr_api.cpp).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcr {

constexpr int kMaxAlpha = 4;
constexpr int kMaxN = 608;      // tile kernel: 64 paths x n normals in LDS (152 KiB)
constexpr int kChunkMaxN = 1024;  // column-chunk kernel and the factor: one thread per row
constexpr int kMfmaLdsMaxN = 32 * 24 - 1;  // LDS MFMA kernels: 24 row tiles of 32
                                           // (assets + loss row)
constexpr int kBatchMaxN = kMfmaLdsMaxN;   // batches and the serving loop
constexpr int kRegMaxN = 64;    // register kernel: one row accumulator per asset
constexpr int kMfmaMaxT = 4;    // MFMA kernel: 32-row tiles (assets + loss row)
constexpr int kMfmaNB = 32;     // MFMA kernel: Philox blocks per path (n <= 127)

constexpr int kTilePaths = 64;  // paths per workgroup in the tile kernel
constexpr int kTileRows = 8;    // rows per register block in the tile kernel
constexpr unsigned kGatherCap = 1u << 18;  // loss values kept per VaR window
constexpr unsigned kBatchGatherCap = 1u << 16;  // per window, per query of a batch
constexpr int kHistBins = 4096;            // refinement histogram

enum LossMode : int { kLossNone = 0, kLossBracket = 1, kLossHist = 2 };

// Written on the device by the factor kernel, read by the path kernels and
// copied back to the host with the accumulators.
struct DevInfo {
  int attempts;      // 1..9, 10 = clamp fallback, -1 = non-finite input
  int clamped;       // pivots clamped
  int fail_col;      // 0, or first failing column + 1 of the final attempt
  int pad0;
  double jitter;     // delta added to the diagonal
  double mp;         // w^T mu (FP64)
  double loss_sd;    // ||L^T w|| (FP64)
  float lo[kMaxAlpha];  // loss windows [lo, hi] per VaR level
  float hi[kMaxAlpha];
  float hist_lo;        // refinement histogram origin and inverse bin width
  float hist_scale;
};

// Device accumulators, cleared by the factor kernel before each query.
struct DevAccum {
  unsigned long long count;                 // paths with S >= k
  unsigned long long near;                  // paths with a row within 1 ulp of thr
  unsigned long long above[kMaxAlpha];      // losses > hi
  long long sum_above[kMaxAlpha];           // fixed-point (2^-32) sum of those
  unsigned long long gathered[kMaxAlpha];   // losses in [lo, hi] (may exceed cap)
  unsigned long long hist_below;            // refinement: losses < hist_lo
  unsigned long long hist_above;            // refinement: losses past the last bin
};

// Kernel arguments (passed by value).
struct PathArgs {
  uint64_t seed;
  uint64_t path0;   // id of path 0 of this launch
  uint64_t npaths;  // paths in this launch
  int n, k;
  float thr;        // (float)(-x)
  float thr_lo, thr_hi;  // thr -/+ 1 ulp (near-threshold accounting)
  const float* colpack;  // register kernel: column j = L[j..n-1][j], v[j]
  const float* L32;      // tile kernel: fp32 factor in 8 x 4 blocks (l32_offset)
  const float* mu32;     // [n]
  const float* v32;      // [n] L^T w
  const uint4* lfrag;    // MFMA: [T][S][64 lanes][hi, lo] 8 x bf16 each
  const float* rowinit;  // MFMA: [32 T] row initial values
  int T, S;              // MFMA tiles and 16-wide K steps
  const DevInfo* info;
  DevAccum* acc;
  int32_t* S_out;        // optional [npaths]
  float* loss_out;       // optional [npaths]
  unsigned long long* s_hist;  // optional [n+1]
  int loss_mode;
  int n_alpha;
  float* gather;         // [n_alpha][gather_cap]
  unsigned gather_cap;   // losses kept per window
  unsigned* hist;        // [kHistBins] refinement
};

// Factor kernel arguments.
struct FactorArgs {
  const double* A;   // [n*n] covariance (row-major, ld = n)
  const double* mu;  // [n]
  const double* w;   // [n]
  double* L;         // [n*n] out
  double* v64;       // [n] out: L^T w
  int n;
  int n_alpha;
  double za[kMaxAlpha];  // inverse normal cdf at alpha
  double ha[kMaxAlpha];  // window half-widths in units of loss_sd
  DevInfo* info;
  DevAccum* acc;
  unsigned* hist;        // cleared too
  // Factor record of the covariance (attempts, clamped, fail_col, jitter):
  // written after factoring when reuse == 0, read when reuse == 1, in which
  // case L already holds the factor of this covariance and A is not read.
  DevInfo* finfo;
  int reuse;
};

struct PrepArgs {
  const double* L;   // [n*n]
  const double* mu;  // [n]
  int n, T, S;
  const DevInfo* info;
  double* v64;       // [n] scratch: L^T w
  float* colpack;
  float* L32;
  float* mu32;
  float* v32;
  uint4* lfrag;
  float* rowinit;
};

// Launchers (mcr_device.hip).  All asynchronous on `stream`.
hipError_t launch_factor(const FactorArgs& a, hipStream_t stream);
hipError_t launch_prep(const PrepArgs& a, hipStream_t stream);
hipError_t launch_factor_batch(const FactorArgs* fas, int nq, int max_n, hipStream_t stream);
hipError_t launch_prep_batch(const PrepArgs* pas, int nq, int max_n, hipStream_t stream);

// Batched persistent execution: work item = a run of 64-path groups of one query.
struct WorkItem {
  uint32_t query;
  uint32_t ngroups;
  uint64_t first_group;
};

struct BatchArgs {
  const PathArgs* qs;       // [nq] per-query kernel arguments (device)
  const WorkItem* items;    // [nitems] in dispatch order
  uint32_t nitems;
  unsigned* next;           // the class's work counter, cleared by k_batch_clear
};

// Persistent batch kernels, one per class: n <= 127 runs the register-operand
// MFMA code per wave (class = Philox blocks NB, tiles T); larger n runs the
// LDS MFMA code per workgroup (class = the workgroup shape chosen for T: waves
// per workgroup and row-tile slots per wave).
typedef void (*BatchKernel)(BatchArgs);
BatchKernel batch_small_kernel(int n, int T);  // n <= 127, 256 threads
int big_waves(int T);                          // waves per workgroup, single-query LDS MFMA kernel
// Batch LDS MFMA kernel shapes: row-tile ranges T = 5-8, 9-12, 13-16, 17-24,
// each with kMaxBigShapes interchangeable shapes (shape 0 the default).
constexpr int kBigRanges = 4;
constexpr int kMaxBigShapes = 2;
int big_range(int T);                          // range of T (T >= 5)
int batch_big_shapes(int range);               // shapes of a range
BatchKernel batch_big_kernel(int T, int shape);  // T <= 24, 64 * batch_big_waves threads
int batch_big_waves(int T, int shape);         // waves per workgroup of that kernel
int resident_blocks(const void* kernel, int block);  // workgroups per CU (occupancy)
hipError_t launch_batch_clear(unsigned* next, int count, hipStream_t stream);
hipError_t launch_batch_kernel(BatchKernel k, const BatchArgs& b, int grid, int block,
                               hipStream_t stream);
hipError_t launch_paths(int kernel, const PathArgs& a, int num_cus, hipStream_t stream);
hipError_t launch_normals(uint64_t seed, uint64_t path0, uint64_t npaths, int n, float* out,
                          hipStream_t stream);

__host__ __device__ inline int colpack_offset(int n, int j) { return j * (n + 1) - (j * (j - 1)) / 2; }
// Tile kernel factor: zero-padded r8 x r8 (r8 = n rounded up to kTileRows),
// stored as 8-row x 4-column blocks, row block major: block (i / 8, j / 4) is
// 32 contiguous floats (row i % 8, column j % 4), so one K step of a row
// block is two 64-byte scalar loads.
__host__ __device__ inline size_t l32_offset(int r8, int i, int j) {
  return ((((size_t)(i >> 3) * (size_t)(r8 >> 2) + (size_t)(j >> 2)) * 8 + (i & 7)) << 2) + (j & 3);
}
__host__ __device__ inline size_t colpack_size(int n) { return (size_t)colpack_offset(n, n); }

}  // namespace mcr
