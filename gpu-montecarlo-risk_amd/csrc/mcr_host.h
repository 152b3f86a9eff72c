// mcr_host.h -- host-side helpers of the C ABI shared by mcr_api.cpp and
// mcr_server.cpp.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/mcr_hip.h"
#include "mcr_internal.h"

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::mcr::host::set_error("%s failed: %s (line %d)", #expr, hipGetErrorString(e_), \
                             __LINE__);                                            \
      return MCR_E_DEVICE;                                                         \
    }                                                                              \
  } while (0)

#define MCR_TRY(expr)              \
  do {                             \
    int r_ = (expr);               \
    if (r_ != MCR_OK) return r_;   \
  } while (0)

namespace mcr {
namespace host {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Bump allocator over a region: base == nullptr only measures.
struct Carver {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t count) {
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += align_up(count * sizeof(T));
    return p;
  }
};

// Streams the class launches of a batch (or a serving round) are forked over:
// after the counters' clear kernel the forks wait for `fork`, each runs the
// class kernels dealt to it, and the launching stream waits for every used
// fork's `join` -- inside a captured graph these become the graph's edges.
constexpr int kMaxForks = 4;
struct ForkSet {
  int n = 0;
  hipStream_t s[kMaxForks] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kMaxForks] = {};
};
int create_forks(ForkSet* f, int n);  // default-priority non-blocking streams
void destroy_forks(ForkSet* f);

struct Context {
  bool ready = false;
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evs = nullptr;  // start of a query's device work (after its uploads)
  void* arena = nullptr;
  size_t arena_bytes = 0;
  void* host = nullptr;  // pinned staging
  size_t host_bytes = 0;
  float* hwin = nullptr;  // pinned copies of batches' gathered in-window losses
  size_t hwin_cap = 0;    // floats
  uint64_t arena_gen = 0;  // bumped whenever the arena moves (invalidates graphs)
  ForkSet forks;           // class-launch streams of batches
  // Device address of the file-scope scratch array of mcr_api.cpp on this
  // device (hipGetSymbolAddress at init); the array's own name is only a
  // host-side handle and must never be passed to a kernel or a copy.
  void* scratch = nullptr;
};

extern Context g_ctx;
extern std::mutex g_mu;  // serialises the single-query / batch entry points
extern thread_local char g_err[512];

void set_error(const char* fmt, ...);
int ensure_init(int device);
int validate(const mcr_query* q);
double inv_norm_cdf(double p);
int64_t rank_of(double alpha, int64_t M);
int check_factor(const DevInfo& info);
double group_cost(int T, int S);
// hostbuf: nullptr, or a host buffer of >= cap floats used instead of a heap copy
int resolve_quantile(int64_t M, const DevAccum& acc, int slot, const float* gather, unsigned cap,
                     int64_t R, double* var, double* es, int64_t* tail, float* hostbuf = nullptr);
bool window_holds(int64_t M, const DevAccum& acc, int slot, unsigned cap, int64_t R);
// gather capacity per VaR level that holds a query's window losses (<= cap)
unsigned window_capacity(const mcr_query* q, unsigned cap);
void resolve_window(int64_t M, const DevAccum& acc, int slot, float* g, int64_t R, double* var,
                    double* es, int64_t* tail);
// One device window of gathered losses to copy: count floats from dev to
// offset off of a pinned host buffer.
struct WindowFetch {
  const float* dev;
  int64_t count;
  size_t off;
};
// Copies every window on `stream` (asynchronously, one wait at the end) into
// *hbuf, a pinned buffer of *hcap floats grown as needed.
int fetch_windows(const std::vector<WindowFetch>& w, float** hbuf, size_t* hcap,
                  hipStream_t stream);
// noalloc: device buffers from the scratch array, host buffers static (no
// hipMalloc, pinned or heap allocation); the library must be initialised
int run_locked(const mcr_query* q, mcr_result* r, bool noalloc = false);
void set_windows(const mcr_query* q, int n_alpha, FactorArgs* f);

// Persistent batch classes: queries with n <= kSmallMaxN run the per-wave
// register-operand MFMA kernel of their (NB, T), larger ones the LDS MFMA
// kernel of their shape.  Items of one class are contiguous and launched
// together.
constexpr int kSmallMaxN = 127;
struct ClassRun {
  BatchKernel k;
  uint32_t off, count;  // item range
  int grid;             // resident workgroups of k
  int block;            // threads per workgroup of k
  double cost;          // summed cost of the class's items (group_cost units)
};
// The tuned LDS MFMA workgroup shape per row-tile range (0 = the default).
void tuned_shapes(int out[kBigRanges]);
// shapes: per row-tile range, the LDS MFMA kernel shape; nullptr = the tuned
// ones (read anew at each call: a caller that needs a kernel and its block
// size to agree passes one snapshot to both)
BatchKernel class_kernel(int n, const int* shapes = nullptr);
int class_block(int n, const int* shapes = nullptr);
int class_grid(BatchKernel k, int block);
void order_by_class(std::vector<WorkItem>* items, const std::vector<BatchKernel>& qk,
                    const std::vector<int>& qb, const std::vector<double>& qcost, bool lpt,
                    std::vector<ClassRun>* runs);
// nfork > 1 (and forks given): the class kernels run on min(nfork, forks->n)
// concurrent streams, dealt longest-class-first to the least-loaded stream.
int launch_runs(const std::vector<ClassRun>& runs, const PathArgs* qs, const WorkItem* items,
                unsigned* next, int grid_override, hipStream_t stream,
                const ForkSet* forks = nullptr, int nfork = 1);
void fill_statistics(int64_t M, const DevInfo& info, const DevAccum& acc, mcr_result* r);

}  // namespace host
}  // namespace mcr
