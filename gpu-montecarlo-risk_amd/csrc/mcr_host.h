// mcr_host.h -- host-side helpers of the C ABI shared by mcr_api.cpp and
// mcr_server.cpp.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/mcr_hip.h"
#include "mcr_internal.h"

namespace mcr {
namespace host {

struct Context {
  bool ready = false;
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  void* host = nullptr;  // pinned staging
  size_t host_bytes = 0;
  uint64_t arena_gen = 0;  // bumped whenever the arena moves (invalidates graphs)
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
int resolve_quantile(int64_t M, const DevAccum& acc, int slot, const float* gather, unsigned cap,
                     int64_t R, double* var, double* es, int64_t* tail);
int run_locked(const mcr_query* q, mcr_result* r);
size_t align_up(size_t x);

}  // namespace host
}  // namespace mcr
