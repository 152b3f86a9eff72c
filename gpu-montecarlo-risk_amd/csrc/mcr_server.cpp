// mcr_server.cpp -- serving loop: risk queries submitted from any thread are
// admitted by device-memory footprint and run in rounds of the persistent
// path kernels (one launch per kernel class), new queries joining between
// rounds.
//
// The reference runs one query per process (src/main_runner.c:198-296: read
// inputs, call the model, write a CSV row); nothing there batches or serves.
// What carries over is the per-query computation, which is the batch engine's:
// FP64 factor (cached here by covariance content), bf16 operand layouts, and
// the (seed, path id)-keyed paths reduced into integer / fixed-point tallies,
// so a query's result does not depend on which rounds or which other queries
// its paths shared the device with.
//
// Device work per round r (parity p = r & 1), on two default-priority streams
// (plus the class-launch forks):
//   setup:  H2D of new queries' inputs and descriptors, k_factor_batch
//           (factorisation skipped on a cache hit), k_prep_batch, H2D of the round's work
//           items; event setup[p].  Overlaps the previous round's path kernel.
//   paths:  wait setup[p]; k_batch_clear; one persistent launch per kernel class
//           over the round's items (k_batch_small / k_batch_paths / k_batch_spec),
//           the classes forked over nfork streams and joined back; events
//           start[p], end[p].
//   out:    on the path stream after end[p]: D2H of the records of queries
//           whose last paths ran in round r; event out[p].
// The worker enqueues round r, then waits for round r-1's outputs and
// finalises it, so the path stream always holds the next round.  Memory,
// slots and cache references of a query are released only after its last
// round's outputs have arrived.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mcr_hip.h"
#include "mcr_host.h"
#include "mcr_internal.h"

using namespace mcr;
using namespace mcr::host;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

constexpr int64_t kDefaultBudget = 4ll << 30;
constexpr int kDefaultCache = 256;
constexpr int kDefaultActive = 1024;
constexpr double kDefaultRoundMs = 4.0;
// Class-launch streams per round: two by default, at most two.  A round uses
// the setup stream, the path stream (which also carries the output copies)
// and the forks, and the process has four hardware queues: with three forks
// and a separate output stream the loop was slower (3000-query stream at 800
// QPS: p99 33 -> 63 ms; saturation 1090 -> 880 queries/s) as setup work
// waited behind path kernels.  With the output copies on the path stream and
// finalising free of copies, two forks against one: 1539-1543 vs 1530-1531
// queries/s completed at 1600 QPS offered, p99 116-118 vs 180-231 ms; at 800
// QPS p99 24-26 ms either way.
constexpr int kDefaultForks = 2;
constexpr int kMaxServerForks = 2;
constexpr uint32_t kMaxItems = 1u << 16;  // work items per round
constexpr uint64_t kItemGroups = 32;      // 64-path groups per work item, at most
constexpr int kMaxRuns = 64;              // kernel classes per round (at most 41 exist)
constexpr int kMaxBypass = 16;            // rounds a blocked head query may be overtaken
constexpr double kInitMsPerUnit = 7e-5;   // initial device ms per cost unit per workgroup

// First-fit offset allocator with coalescing over [0, total).
class OffsetAllocator {
 public:
  void reset(size_t total) {
    free_.clear();
    free_[0] = total;
    used_ = peak_ = 0;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, start = it->first + bytes;
      free_.erase(it);
      if (rest) free_[start] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(bytes);
    used_ -= bytes;
    size_t start = off, len = bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && next->first == off + bytes) {
      len += next->second;
      free_.erase(next);
    }
    free_[start] = len;
  }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  std::map<size_t, size_t> free_;
  size_t used_ = 0, peak_ = 0;
};

// Device bytes of a cached factor: [factor record | L n x n].
size_t entry_bytes(int n) {
  return align_up(sizeof(DevInfo)) + align_up((size_t)n * n * sizeof(double));
}

// Region of an admitted query.  Inputs first (one H2D span), then its record,
// accumulators and gathered window losses (one D2H span), then the rest.
struct JobLayout {
  size_t mu = 0, w = 0, A = 0, in_span = 0;
  size_t info = 0, acc = 0, gather = 0, out_span = 0;
  size_t v64 = 0, rowinit = 0, lfrag = 0, total = 0;
};

// wcap: gathered losses per VaR level (window_capacity)
JobLayout job_layout(int n, int T, int S, int na, bool with_A, unsigned wcap) {
  JobLayout l;
  Carver c{nullptr};
  l.mu = c.off;
  c.take<double>(n);
  l.w = c.off;
  c.take<double>(n);
  l.A = c.off;
  if (with_A) c.take<double>((size_t)n * n);
  l.in_span = c.off;
  l.info = c.off;
  c.take<DevInfo>(1);
  l.acc = c.off;
  c.take<DevAccum>(1);
  l.gather = c.off;
  if (na > 0) c.take<float>((size_t)na * wcap);
  l.out_span = c.off - l.info;
  l.v64 = c.off;
  c.take<double>(n);
  l.rowinit = c.off;
  c.take<float>((size_t)32 * T);
  l.lfrag = c.off;
  c.take<uint4>((size_t)T * S * 128);
  l.total = c.off;
  return l;
}

int tiles_T(int n) { return (n + 1 + 31) / 32; }
int steps_S(int n) { return (n + 15) / 16; }

struct CacheEntry {
  uint64_t hash = 0;
  int n = 0;
  std::vector<double> sigma;  // compact copy: the exact key
  size_t off = 0, bytes = 0;
  int refs = 0;
  uint64_t ready_round = 0;   // setups of rounds >= this may reuse the factor
  uint64_t last_use = 0;
  bool failed = false;
};

enum class JobState { kQueued, kActive, kDone };

struct Job {
  uint64_t ticket = 0;
  mcr_query q{};
  std::vector<double> mu, w, sigma;  // compact copies; q points at them
  uint64_t hash = 0;
  int T = 0, S = 0, na = 0;
  unsigned wcap = 0;  // gathered losses per VaR level (window_capacity)
  JobState state = JobState::kQueued;
  int status = MCR_OK;
  std::string err;
  mcr_result res{};
  Clock::time_point t_submit, t_admit, t_done;
  // admitted
  int slot = -1;
  size_t off = 0;
  JobLayout lay;
  CacheEntry* ce = nullptr;
  bool hit = false;
  uint64_t groups = 0, next_group = 0;
  int rounds = 0;
  int bypassed = 0;
  double dev_ms = 0.0;
};

struct RoundRec {
  uint64_t id = 0;
  int parity = 0;
  std::vector<Job*> members, finishing;
  double units = 0.0;
  size_t nitems = 0;
  std::vector<ClassRun> runs;
  std::vector<size_t> out_off;  // per finishing query: its output span in h_out
};

uint64_t hash_sigma(int n, const double* s) {
  uint64_t h = 0x6A09E667F3BCC909ull ^ (uint64_t)n;
  for (size_t i = 0; i < (size_t)n * n; ++i) {
    uint64_t v;
    std::memcpy(&v, &s[i], sizeof(v));
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h *= 0xBF58476D1CE4E5B9ull;
  }
  return h ^ (h >> 31);
}

// Checks the options a served query may use; sets the error text.
int check_servable(const mcr_query* q) {
  MCR_TRY(validate(q));
  if (q->precision != MCR_PREC_BF16X3 && q->precision != MCR_PREC_AUTO) {
    set_error("served queries run the bf16x3 MFMA kernels (precision %d)", q->precision);
    return MCR_E_UNSUPPORTED;
  }
  if (q->n > kBatchMaxN) {
    set_error("served queries take n <= %d (n = %d)", kBatchMaxN, q->n);
    return MCR_E_UNSUPPORTED;
  }
  if (q->S_out || q->loss_out || (q->flags & (MCR_F_SHIST | MCR_F_NEAR))) {
    set_error("per-path outputs, S histograms and near counts are single-query options");
    return MCR_E_UNSUPPORTED;
  }
  for (int i = 0; i < q->n; ++i) {
    const double w = q->weights ? q->weights[i] : 1.0;
    if (!std::isfinite(q->mu[i]) || !std::isfinite(w)) {
      set_error("mu or weights not finite at %d", i);
      return MCR_E_ARG;
    }
  }
  return MCR_OK;
}

size_t query_footprint(const mcr_query* q, bool with_factor) {
  const int na = (q->flags & MCR_F_VAR) ? q->n_alpha : 0;
  const JobLayout l = job_layout(q->n, tiles_T(q->n), steps_S(q->n), na, with_factor,
                                 window_capacity(q, kBatchGatherCap));
  return l.total + (with_factor ? entry_bytes(q->n) : 0);
}

}  // namespace

struct mcr_server {
  // configuration
  int device = 0;
  size_t budget = 0;
  int cache_cap = kDefaultCache;  // < 0: no cache
  int max_active = kDefaultActive;
  double round_ms = kDefaultRoundMs;
  int grid = 0;           // CUs: the round budget is grid x round_ms of device time
  int grid_override = 0;  // > 0: every class kernel launched with this many workgroups
  int nfork = kDefaultForks;  // streams a round's class launches are forked over
  double age_limit_ms = 0.0;  // > 0: queries older than this are planned first

  // device resources
  hipStream_t s_setup = nullptr, s_paths = nullptr, s_out = nullptr;
  ForkSet forks;  // class-launch streams, joined back into s_paths
  hipEvent_t ev_setup[2] = {}, ev_start[2] = {}, ev_end[2] = {}, ev_out[2] = {};
  char* arena = nullptr;     // budget bytes: query regions and cached factors
  char* fixed = nullptr;     // slot table, round areas, histogram scratch
  PathArgs* d_slots = nullptr;
  char* d_round[2] = {};
  size_t off_fa = 0, off_pa = 0, off_next = 0;
  unsigned* d_hist = nullptr;
  char* h_stage[2] = {};
  size_t h_stage_bytes[2] = {};
  char* h_out[2] = {};
  size_t h_out_bytes[2] = {};

  // worker-owned state
  OffsetAllocator alloc;
  std::unordered_multimap<uint64_t, CacheEntry*> cache;
  std::vector<std::unique_ptr<CacheEntry>> entries;
  std::vector<int> free_slots;
  std::vector<Job*> active;
  double ms_per_unit = kInitMsPerUnit;
  uint64_t use_clock = 0;
  Clock::time_point t_ready{};  // end of finalize_round's wait for the results

  // shared state (mu)
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> queued;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  uint64_t next_ticket = 1;
  bool stop = false;
  bool broken = false;
  std::string broken_msg;
  mcr_server_stats stats{};
  std::thread worker;

  int init(int dev, const mcr_server_opts* o);
  void teardown();
  void run();
  // worker steps
  int admit(uint64_t round, std::vector<Job*>* fresh);
  bool try_admit(Job* j, uint64_t round);
  void evict_unused(size_t need_hint);
  void drop_entry(CacheEntry* e);
  int plan_round(uint64_t round, RoundRec* rr, std::vector<WorkItem>* items);
  int enqueue_round(RoundRec& rr, const std::vector<Job*>& fresh, const std::vector<WorkItem>& items);
  int finalize_round(RoundRec& rr);
  void finish_job(Job* j, int status, const char* err);
  void fail_everything(const char* msg);
  int grow_pinned(char** p, size_t* cap, size_t need);
};

int mcr_server::grow_pinned(char** p, size_t* cap, size_t need) {
  if (need <= *cap) return MCR_OK;
  if (*p) HIP_TRY(hipHostFree(*p));
  *p = nullptr;
  *cap = 0;
  const size_t bytes = std::max<size_t>(need + need / 2, 1 << 20);
  HIP_TRY(hipHostMalloc((void**)p, bytes, hipHostMallocDefault));
  *cap = bytes;
  return MCR_OK;
}

int mcr_server::init(int dev, const mcr_server_opts* o) {
  device = dev;
  budget = (size_t)((o && o->hbm_budget > 0) ? o->hbm_budget : kDefaultBudget);
  if (o && o->cache_entries != 0) cache_cap = o->cache_entries;
  if (o && o->max_active > 0) max_active = o->max_active;
  if (o && o->round_ms > 0) round_ms = o->round_ms;
  if (o && !(o->age_limit_ms >= 0.0)) {
    set_error("age_limit_ms must be >= 0 (got %g)", o->age_limit_ms);
    return MCR_E_ARG;
  }
  if (o) age_limit_ms = o->age_limit_ms;
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count));
  if (dev < 0 || dev >= count) {
    set_error("device %d not present (%d devices)", dev, count);
    return MCR_E_ARG;
  }
  HIP_TRY(hipSetDevice(dev));
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  grid = cus;
  grid_override = (o && o->grid > 0) ? o->grid : 0;
  if (o && o->streams > 0) nfork = std::min(o->streams, kMaxServerForks);
  HIP_TRY(hipStreamCreateWithFlags(&s_setup, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&s_paths, hipStreamNonBlocking));
  s_out = s_paths;  // the output copies follow the path kernels
  if (nfork > 1) MCR_TRY(create_forks(&forks, nfork));
  for (int p = 0; p < 2; ++p) {
    HIP_TRY(hipEventCreateWithFlags(&ev_setup[p], hipEventDisableTiming));
    HIP_TRY(hipEventCreate(&ev_start[p]));
    HIP_TRY(hipEventCreate(&ev_end[p]));
    HIP_TRY(hipEventCreateWithFlags(&ev_out[p], hipEventDisableTiming));
  }
  if (hipMalloc((void**)&arena, budget) != hipSuccess) {
    (void)hipGetLastError();
    set_error("cannot allocate the %zu-byte admission budget on device %d", budget, dev);
    return MCR_E_NOMEM;
  }
  Carver c{nullptr};
  c.take<PathArgs>(max_active);
  const size_t round_base = c.off;
  Carver r{nullptr};
  r.take<WorkItem>(kMaxItems);
  off_fa = r.off;
  r.take<FactorArgs>(max_active);
  off_pa = r.off;
  r.take<PrepArgs>(max_active);
  off_next = r.off;
  r.take<unsigned>(kMaxRuns);
  const size_t round_bytes = r.off;
  const size_t hist_off = round_base + 2 * round_bytes;
  const size_t fixed_bytes = hist_off + align_up(kHistBins * sizeof(unsigned));
  if (hipMalloc((void**)&fixed, fixed_bytes) != hipSuccess) {
    (void)hipGetLastError();
    set_error("cannot allocate %zu bytes of server tables", fixed_bytes);
    return MCR_E_NOMEM;
  }
  d_slots = (PathArgs*)fixed;
  d_round[0] = fixed + round_base;
  d_round[1] = fixed + round_base + round_bytes;
  d_hist = (unsigned*)(fixed + hist_off);
  for (int p = 0; p < 2; ++p) {
    MCR_TRY(grow_pinned(&h_stage[p], &h_stage_bytes[p], 1 << 20));
    MCR_TRY(grow_pinned(&h_out[p], &h_out_bytes[p], 1 << 16));
  }
  alloc.reset(budget);
  free_slots.reserve(max_active);
  for (int i = max_active - 1; i >= 0; --i) free_slots.push_back(i);
  stats.budget_bytes = (int64_t)budget;
  worker = std::thread([this] { run(); });
  return MCR_OK;
}

void mcr_server::teardown() {
  if (worker.joinable()) {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    worker.join();
  }
  (void)hipSetDevice(device);
  for (hipStream_t s : {s_setup, s_paths, s_out})
    if (s) (void)hipStreamSynchronize(s);
  for (int j = 0; j < forks.n; ++j) (void)hipStreamSynchronize(forks.s[j]);
  destroy_forks(&forks);
  for (int p = 0; p < 2; ++p) {
    for (hipEvent_t e : {ev_setup[p], ev_start[p], ev_end[p], ev_out[p]})
      if (e) (void)hipEventDestroy(e);
    if (h_stage[p]) (void)hipHostFree(h_stage[p]);
    if (h_out[p]) (void)hipHostFree(h_out[p]);
  }
  for (hipStream_t s : {s_setup, s_paths})
    if (s) (void)hipStreamDestroy(s);
  if (arena) (void)hipFree(arena);
  if (fixed) (void)hipFree(fixed);
}

void mcr_server::drop_entry(CacheEntry* e) {
  auto range = cache.equal_range(e->hash);
  for (auto it = range.first; it != range.second; ++it)
    if (it->second == e) {
      cache.erase(it);
      break;
    }
  alloc.release(e->off, e->bytes);
  for (size_t i = 0; i < entries.size(); ++i)
    if (entries[i].get() == e) {
      entries[i].swap(entries.back());
      entries.pop_back();
      break;
    }
}

// Evicts unreferenced factors, least recently used first, until the count is
// within the cache capacity and (if need_hint > 0) one eviction was made for
// memory.
void mcr_server::evict_unused(size_t need_hint) {
  for (;;) {
    CacheEntry* victim = nullptr;
    for (auto& e : entries)
      if (e->refs == 0 && (!victim || e->last_use < victim->last_use)) victim = e.get();
    if (!victim) return;
    const int cap = cache_cap < 0 ? 0 : cache_cap;
    const bool over = (int)entries.size() > cap;
    if (!over && need_hint == 0) return;
    drop_entry(victim);
    if (!over) return;  // one eviction for memory; caller retries
  }
}

bool mcr_server::try_admit(Job* j, uint64_t round) {
  if (free_slots.empty()) return false;
  CacheEntry* ce = nullptr;
  auto range = cache.equal_range(j->hash);
  for (auto it = range.first; it != range.second; ++it) {
    CacheEntry* e = it->second;
    if (!e->failed && e->n == j->q.n && e->sigma == j->sigma) {
      ce = e;
      break;
    }
  }
  if (ce && ce->ready_round > round) return false;  // its factor is being made this round
  const bool hit = ce != nullptr;
  const JobLayout lay = job_layout(j->q.n, j->T, j->S, j->na, !hit, j->wcap);
  for (;;) {
    size_t joff = 0, eoff = 0;
    bool ok = alloc.alloc(lay.total, &joff);
    if (ok && !hit) {
      ok = alloc.alloc(entry_bytes(j->q.n), &eoff);
      if (!ok) alloc.release(joff, lay.total);
    }
    if (ok) {
      if (!hit) {
        auto e = std::make_unique<CacheEntry>();
        e->hash = j->hash;
        e->n = j->q.n;
        e->sigma = j->sigma;
        e->off = eoff;
        e->bytes = entry_bytes(j->q.n);
        e->ready_round = round + 1;
        ce = e.get();
        entries.push_back(std::move(e));
        cache.emplace(j->hash, ce);
      }
      ce->refs++;
      ce->last_use = ++use_clock;
      j->ce = ce;
      j->hit = hit;
      j->off = joff;
      j->lay = lay;
      j->slot = free_slots.back();
      free_slots.pop_back();
      j->groups = ((uint64_t)j->q.paths + 63) / 64;
      j->next_group = 0;
      return true;
    }
    // make room: drop one unreferenced factor (never the one this job would hit)
    CacheEntry* victim = nullptr;
    for (auto& e : entries)
      if (e->refs == 0 && e.get() != ce && (!victim || e->last_use < victim->last_use))
        victim = e.get();
    if (!victim) return false;
    drop_entry(victim);
  }
}

// Moves queued jobs that fit into the active set.  FIFO, except that a job
// which does not fit may be overtaken for at most kMaxBypass rounds.
int mcr_server::admit(uint64_t round, std::vector<Job*>* fresh) {
  std::vector<Job*> cand;
  {
    std::lock_guard<std::mutex> lk(mu);
    cand.assign(queued.begin(), queued.end());
  }
  bool blocked = false;
  std::vector<Job*> taken;
  for (Job* j : cand) {
    if ((int)active.size() >= max_active) break;
    if (blocked) {
      // the head that did not fit has been overtaken long enough
      break;
    }
    if (try_admit(j, round)) {
      taken.push_back(j);
      active.push_back(j);
      fresh->push_back(j);
    } else if (j->bypassed++ >= kMaxBypass) {
      blocked = true;
    }
  }
  if (!taken.empty()) {
    const auto now = Clock::now();
    std::lock_guard<std::mutex> lk(mu);
    for (Job* j : taken) {
      j->state = JobState::kActive;
      j->t_admit = now;
      queued.erase(std::find(queued.begin(), queued.end(), j));
      if (j->hit)
        stats.cache_hits++;
      else
        stats.cache_misses++;
    }
  }
  evict_unused(0);
  return MCR_OK;
}

// The round's budget (in cost units) goes to the active jobs with paths left
// smallest remaining work first (with an age limit, jobs older than it before
// them, oldest first), each taking all it can of what is left and every job
// at least one group (so the largest still advance every round); work items
// sized for load balance, longest first.
int mcr_server::plan_round(uint64_t round, RoundRec* rr, std::vector<WorkItem>* items) {
  rr->id = round;
  rr->parity = (int)(round & 1);
  items->clear();
  struct Open {
    double key;  // remaining cost units, or minus the age (ms) of a job past the age limit
    double units;
    Job* j;
  };
  std::vector<Open> open;
  const Clock::time_point now = Clock::now();
  for (Job* j : active)
    if (j->next_group < j->groups) {
      const double units = (double)(j->groups - j->next_group) * group_cost(j->T, j->S);
      const double age = ms_since(j->t_submit, now);
      open.push_back({age_limit_ms > 0 && age > age_limit_ms ? -age : units, units, j});
    }
  if (open.empty()) return MCR_OK;
  std::sort(open.begin(), open.end(), [](const Open& a, const Open& b) { return a.key < b.key; });
  // Round length: round_ms while the admitted work is short; under a backlog
  // it grows with it (an eighth of the backlog, at most 8 round_ms), so each
  // class launch of the round carries more work and its tail costs less.  A
  // 3000-query stream offered at 1600 QPS completed 1061 queries/s with 4 ms
  // rounds and 1418 / 1447 with 8 / 16 ms ones (p99 1.9 s -> 0.5 / 0.4 s),
  // while at 800 QPS, below saturation, the length changed nothing but p50
  // (5.0 -> 5.9 ms at 16 ms).  Past saturation (3200 QPS offered) 32 ms
  // rounds complete 1747-1763 queries/s against 1637-1679 with 16 ms ones,
  // and 64 ms ones 1633-1679 (the worker's staging of such rounds leaves the
  // device idle); below it nothing changes, as the backlog never asks for them.
  double backlog_units = 0.0;
  for (auto& o : open) backlog_units += o.units;
  const double backlog_ms = backlog_units * ms_per_unit / grid;
  const double len_ms = std::min(8.0 * round_ms, std::max(round_ms, backlog_ms / 8.0));
  // Smallest remaining first rather than a max-min fair split: on 5000-query
  // Poisson streams the fair split gave p99 12.0 / 25.0-25.2 / 128-130 ms at
  // 200 / 800 / 1600 QPS offered and this grant 11.4-11.5 / 19.6-20.2 /
  // 78-89 ms, at the same completed rate (p50 at 1600 QPS 12 -> 15-17 ms: the
  // largest jobs wait behind the small ones, and the worst latency at 1600
  // QPS grew from 0.28 s to 0.6-2.5 s).  The age limit trades back: at 256 ms
  // the worst latency at 1600 QPS was 0.30 s, the p99 128-136 ms against
  // 81 ms without it, and past saturation (3200 QPS offered) the completed
  // rate fell 1800-1807 -> 1719-1751 queries/s and the p50 rose 81 -> 700 ms
  // (every job there ages past the limit: the order becomes first come, first
  // served); at 64 ms, 1600 QPS gave a worst latency of 117 ms and p50 / p99
  // 26-28 / 105-108 ms against 13 / 72-73 ms.
  double left = len_ms * grid / ms_per_unit;
  std::vector<std::pair<Job*, uint64_t>> grant;
  double total_units = 0.0;
  for (auto& o : open) {
    Job* j = o.j;
    const double c = group_cost(j->T, j->S);
    const double share = std::max(left, 0.0);
    const uint64_t rem = j->groups - j->next_group;
    uint64_t g = o.units <= share ? rem : (uint64_t)std::ceil(share / c);
    g = std::max<uint64_t>(1, std::min(g, rem));
    left -= (double)g * c;
    total_units += (double)g * c;
    grant.push_back({j, g});
  }
  // Item size per kernel class, as batches size them: at least four items
  // per resident slot of the class's launch (a workgroup; a wave for the
  // small-n classes, whose waves take items on their own) and at most
  // kItemGroups groups.  Sizing from the round's total instead gave a class
  // with little work a few huge items, one wave or workgroup grinding through
  // each while the rest of the device idled inside that launch.
  // One snapshot of the tuned workgroup shapes for the whole round: the batch
  // autotuner may change them from another thread, and a class's kernel and
  // its block size must come from the same shape.
  int shapes[kBigRanges];
  tuned_shapes(shapes);
  std::map<BatchKernel, uint64_t> class_groups;
  for (auto& gr : grant) class_groups[class_kernel(gr.first->q.n, shapes)] += gr.second;
  auto item_groups = [&](const Job* j) {
    const int n = j->q.n;
    const BatchKernel k = class_kernel(n, shapes);
    const uint64_t slots =
        (uint64_t)class_grid(k, class_block(n, shapes)) * (n <= kSmallMaxN ? 4 : 1);
    return std::max<uint64_t>(1, std::min<uint64_t>(kItemGroups, class_groups[k] / (4 * slots)));
  };
  size_t est = 0;
  for (auto& gr : grant) {
    const uint64_t ig = item_groups(gr.first);
    est += (size_t)((gr.second + ig - 1) / ig);
  }
  const double scale = est > kMaxItems ? (double)est / kMaxItems + 1.0 : 1.0;
  std::vector<std::pair<double, WorkItem>> it;  // (cost, item)
  for (auto& gr : grant) {
    Job* j = gr.first;
    const double c = group_cost(j->T, j->S);
    const uint64_t ig = std::max<uint64_t>(1, (uint64_t)std::ceil(scale * (double)item_groups(j)));
    const uint64_t g0 = j->next_group, g1 = g0 + gr.second;
    for (uint64_t g = g0; g < g1; g += ig) {
      WorkItem w;
      w.query = (uint32_t)j->slot;
      w.first_group = g;
      w.ngroups = (uint32_t)std::min<uint64_t>(ig, g1 - g);
      it.push_back({c * w.ngroups, w});
    }
    j->next_group = g1;
    j->rounds++;
    rr->members.push_back(j);
    if (g1 == j->groups) rr->finishing.push_back(j);
  }
  if (it.size() > kMaxItems) {
    set_error("round planned %zu work items (limit %u)", it.size(), kMaxItems);
    return MCR_E_DEVICE;
  }
  for (auto& x : it) items->push_back(x.second);
  std::vector<BatchKernel> qk(max_active, nullptr);
  std::vector<int> qb(max_active, 0);
  std::vector<double> qcost(max_active, 0.0);
  for (auto& gr : grant) {
    qk[gr.first->slot] = class_kernel(gr.first->q.n, shapes);
    qb[gr.first->slot] = class_block(gr.first->q.n, shapes);
    qcost[gr.first->slot] = group_cost(gr.first->T, gr.first->S);
  }
  order_by_class(items, qk, qb, qcost, true, &rr->runs);
  if (rr->runs.size() > (size_t)kMaxRuns) {
    set_error("round needs %zu kernel classes (limit %d)", rr->runs.size(), kMaxRuns);
    return MCR_E_DEVICE;
  }
  rr->units = total_units;
  rr->nitems = items->size();
  return MCR_OK;
}

int mcr_server::enqueue_round(RoundRec& rr, const std::vector<Job*>& fresh,
                              const std::vector<WorkItem>& items) {
  const int p = rr.parity;
  const size_t nf = fresh.size();
  // pinned staging: [items | FactorArgs | PrepArgs | PathArgs | inputs...]
  Carver c{nullptr};
  const size_t o_items = c.off;
  c.take<WorkItem>(items.size());
  const size_t o_fa = c.off;
  c.take<FactorArgs>(nf);
  const size_t o_pa = c.off;
  c.take<PrepArgs>(nf);
  const size_t o_qa = c.off;
  c.take<PathArgs>(nf);
  std::vector<size_t> o_in(nf);
  for (size_t i = 0; i < nf; ++i) {
    o_in[i] = c.off;
    c.take<char>(fresh[i]->lay.in_span);
  }
  MCR_TRY(grow_pinned(&h_stage[p], &h_stage_bytes[p], c.off));
  char* h = h_stage[p];
  std::memcpy(h + o_items, items.data(), items.size() * sizeof(WorkItem));
  FactorArgs* fa = (FactorArgs*)(h + o_fa);
  PrepArgs* pa = (PrepArgs*)(h + o_pa);
  PathArgs* qa = (PathArgs*)(h + o_qa);
  int max_n = 1;
  for (size_t i = 0; i < nf; ++i) {
    Job* j = fresh[i];
    const mcr_query* q = &j->q;
    const size_t n = (size_t)q->n;
    char* in = h + o_in[i];
    std::memcpy(in + j->lay.mu, j->mu.data(), n * sizeof(double));
    std::memcpy(in + j->lay.w, j->w.data(), n * sizeof(double));
    if (!j->hit) std::memcpy(in + j->lay.A, j->sigma.data(), n * n * sizeof(double));
    char* dj = arena + j->off;
    char* de = arena + j->ce->off;
    DevInfo* info = (DevInfo*)(dj + j->lay.info);
    FactorArgs f = {};
    f.A = j->hit ? nullptr : (const double*)(dj + j->lay.A);
    f.mu = (const double*)(dj + j->lay.mu);
    f.w = (const double*)(dj + j->lay.w);
    f.L = (double*)(de + align_up(sizeof(DevInfo)));
    f.v64 = (double*)(dj + j->lay.v64);
    f.n = q->n;
    set_windows(q, j->na, &f);
    f.info = info;
    f.acc = (DevAccum*)(dj + j->lay.acc);
    f.hist = d_hist;
    f.finfo = (DevInfo*)de;
    f.reuse = j->hit ? 1 : 0;
    fa[i] = f;
    PrepArgs pr = {};
    pr.L = f.L;
    pr.mu = f.mu;
    pr.n = q->n;
    pr.T = j->T;
    pr.S = j->S;
    pr.info = info;
    pr.v64 = f.v64;
    pr.lfrag = (uint4*)(dj + j->lay.lfrag);
    pr.rowinit = (float*)(dj + j->lay.rowinit);
    pa[i] = pr;
    PathArgs a = {};
    a.seed = q->seed;
    a.path0 = q->path_offset;
    a.npaths = (uint64_t)q->paths;
    a.n = q->n;
    a.k = q->k;
    a.thr = (float)(-q->x);
    a.thr_lo = 1.0f;
    a.thr_hi = 0.0f;
    a.lfrag = pr.lfrag;
    a.rowinit = pr.rowinit;
    a.T = j->T;
    a.S = j->S;
    a.info = info;
    a.acc = f.acc;
    a.loss_mode = j->na ? kLossBracket : kLossNone;
    a.n_alpha = j->na;
    a.gather = j->na ? (float*)(dj + j->lay.gather) : nullptr;
    a.gather_cap = j->wcap;
    a.hist = d_hist;
    qa[i] = a;
    max_n = std::max(max_n, q->n);
  }
  char* dr = d_round[p];
  for (size_t i = 0; i < nf; ++i) {
    Job* j = fresh[i];
    HIP_TRY(hipMemcpyAsync(arena + j->off, h + o_in[i], j->lay.in_span, hipMemcpyHostToDevice,
                           s_setup));
    HIP_TRY(hipMemcpyAsync(d_slots + j->slot, qa + i, sizeof(PathArgs), hipMemcpyHostToDevice,
                           s_setup));
  }
  if (nf) {
    HIP_TRY(hipMemcpyAsync(dr + off_fa, fa, nf * sizeof(FactorArgs), hipMemcpyHostToDevice, s_setup));
    HIP_TRY(hipMemcpyAsync(dr + off_pa, pa, nf * sizeof(PrepArgs), hipMemcpyHostToDevice, s_setup));
    HIP_TRY(launch_factor_batch((const FactorArgs*)(dr + off_fa), (int)nf, max_n, s_setup));
    HIP_TRY(launch_prep_batch((const PrepArgs*)(dr + off_pa), (int)nf, max_n, s_setup));
  }
  if (!items.empty())
    HIP_TRY(hipMemcpyAsync(dr, h + o_items, items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                           s_setup));
  HIP_TRY(hipEventRecord(ev_setup[p], s_setup));

  HIP_TRY(hipStreamWaitEvent(s_paths, ev_setup[p], 0));
  HIP_TRY(hipEventRecord(ev_start[p], s_paths));
  MCR_TRY(launch_runs(rr.runs, d_slots, (const WorkItem*)dr, (unsigned*)(dr + off_next),
                      grid_override, s_paths, &forks, nfork));
  HIP_TRY(hipEventRecord(ev_end[p], s_paths));

  // each finishing query's record, accumulators and gathered window losses
  // in one copy, in order behind the path kernels: finalising the round then
  // needs no copy of its own (a synchronous copy on another stream can land
  // on the hardware queue of the path stream and wait for the next round)
  rr.out_off.resize(rr.finishing.size());
  size_t out_bytes = 0;
  for (size_t i = 0; i < rr.finishing.size(); ++i) {
    rr.out_off[i] = out_bytes;
    out_bytes += align_up(rr.finishing[i]->lay.out_span);
  }
  MCR_TRY(grow_pinned(&h_out[p], &h_out_bytes[p], out_bytes));
  HIP_TRY(hipStreamWaitEvent(s_out, ev_end[p], 0));
  for (size_t i = 0; i < rr.finishing.size(); ++i) {
    Job* j = rr.finishing[i];
    HIP_TRY(hipMemcpyAsync(h_out[p] + rr.out_off[i], arena + j->off + j->lay.info, j->lay.out_span,
                           hipMemcpyDeviceToHost, s_out));
  }
  HIP_TRY(hipEventRecord(ev_out[p], s_out));
  return MCR_OK;
}

void mcr_server::finish_job(Job* j, int status, const char* err) {
  j->t_done = Clock::now();
  j->status = status;
  if (err) j->err = err;
  j->res.ms_total = ms_since(j->t_submit, j->t_done);
  j->res.ms_kernel = j->dev_ms;
  if (j->slot >= 0) {
    free_slots.push_back(j->slot);
    j->slot = -1;
    alloc.release(j->off, j->lay.total);
  }
  if (j->ce) {
    if (status == MCR_E_FACTOR) j->ce->failed = true;
    j->ce->refs--;
    if (j->ce->failed && j->ce->refs == 0) drop_entry(j->ce);
    j->ce = nullptr;
  }
  active.erase(std::remove(active.begin(), active.end(), j), active.end());
  {
    std::lock_guard<std::mutex> lk(mu);
    j->state = JobState::kDone;
    if (status == MCR_OK) {
      stats.completed++;
      stats.paths_done += j->q.paths;
    } else {
      stats.failed++;
    }
  }
  cv_done.notify_all();
}

int mcr_server::finalize_round(RoundRec& rr) {
  const int p = rr.parity;
  HIP_TRY(hipEventSynchronize(ev_out[p]));
  t_ready = Clock::now();
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, ev_start[p], ev_end[p]));
  {
    std::lock_guard<std::mutex> lk(mu);
    stats.rounds++;
    stats.device_ms += ms;
    stats.peak_bytes = (int64_t)alloc.peak();
  }
  for (Job* j : rr.members) j->dev_ms += ms;
  // calibrate the cost model on rounds that filled the device
  if (rr.nitems >= (size_t)grid && rr.units > 0.0 && ms > 0.0f) {
    const double obs = (double)ms * grid / rr.units;
    ms_per_unit = 0.8 * ms_per_unit + 0.2 * obs;
  }
  for (size_t i = 0; i < rr.finishing.size(); ++i) {
    Job* j = rr.finishing[i];
    const char* o = h_out[p] + rr.out_off[i];
    const DevInfo& info = *(const DevInfo*)o;
    const DevAccum& acc = *(const DevAccum*)(o + (j->lay.acc - j->lay.info));
    float* win = (float*)(h_out[p] + rr.out_off[i] + (j->lay.gather - j->lay.info));
    mcr_result* r = &j->res;
    std::memset(r, 0, sizeof(*r));
    const int fr = check_factor(info);
    if (fr != MCR_OK) {
      finish_job(j, fr, g_err);
      continue;
    }
    const int64_t M = j->q.paths;
    fill_statistics(M, info, acc, r);
    r->kernel = j->q.n <= kSmallMaxN ? MCR_KERNEL_MFMA : MCR_KERNEL_MFMA_LDS;
    r->var_passes = j->na ? 1 : 0;
    bool ok = true;
    for (int al = 0; al < j->na; ++al)
      ok = ok && window_holds(M, acc, al, j->wcap, rank_of(j->q.alpha[al], M));
    int status = MCR_OK;
    if (ok) {
      for (int al = 0; al < j->na; ++al)
        resolve_window(M, acc, al, win + (size_t)al * j->wcap, rank_of(j->q.alpha[al], M),
                       &r->var[al], &r->es[al], &r->tail_sum_fixed[al]);
    } else {
      // Rare: a quantile missed (or overflowed) its window.  Same kernel
      // family alone, with the refinement passes (identical per-path losses).
      mcr_query q1 = j->q;
      q1.kernel = MCR_KERNEL_AUTO;
      q1.precision = MCR_PREC_BF16X3;
      mcr_result one;
      {
        std::lock_guard<std::mutex> lk(g_mu);
        status = run_locked(&q1, &one);
      }
      if (status == MCR_OK) {
        for (int al = 0; al < j->na; ++al) {
          r->var[al] = one.var[al];
          r->es[al] = one.es[al];
          r->tail_sum_fixed[al] = one.tail_sum_fixed[al];
        }
        r->var_passes = 1 + one.var_passes;
      }
    }
    finish_job(j, status, status == MCR_OK ? nullptr : g_err);
  }
  return MCR_OK;
}

void mcr_server::fail_everything(const char* msg) {
  std::vector<Job*> victims(active.begin(), active.end());
  for (Job* j : victims) finish_job(j, MCR_E_DEVICE, msg);
  std::vector<Job*> q;
  {
    std::lock_guard<std::mutex> lk(mu);
    broken = true;
    broken_msg = msg;
    q.assign(queued.begin(), queued.end());
    queued.clear();
  }
  for (Job* j : q) finish_job(j, MCR_E_DEVICE, msg);
}

void mcr_server::run() {
  if (hipSetDevice(device) != hipSuccess) {
    fail_everything("worker could not select the device");
    return;
  }
  uint64_t round = 1;
  std::unique_ptr<RoundRec> inflight;
  std::vector<WorkItem> items;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !queued.empty() || !active.empty() || inflight; });
      if (stop && !queued.empty()) {
        std::vector<Job*> q(queued.begin(), queued.end());
        queued.clear();
        lk.unlock();
        for (Job* j : q) finish_job(j, MCR_E_ARG, "server shut down before the query was admitted");
        lk.lock();
      }
      if (stop && active.empty() && !inflight) return;
    }
    std::vector<Job*> fresh;
    const Clock::time_point t0 = Clock::now();
    int rc = admit(round, &fresh);
    const Clock::time_point t1 = Clock::now();
    auto rr = std::make_unique<RoundRec>();
    if (rc == MCR_OK) rc = plan_round(round, rr.get(), &items);
    const Clock::time_point t2 = Clock::now();
    const bool have = !items.empty() || !fresh.empty();
    if (rc == MCR_OK && have) rc = enqueue_round(*rr, fresh, items);
    const Clock::time_point t3 = Clock::now();
    if (rc == MCR_OK && inflight) {
      rc = finalize_round(*inflight);
      inflight.reset();
    }
    const Clock::time_point t4 = Clock::now();
    {
      // t_ready: when finalize_round's wait for the results ended (t3 if
      // there was no round to finalise)
      const Clock::time_point tw = t_ready > t3 ? t_ready : t3;
      std::lock_guard<std::mutex> lk(mu);
      stats.host_admit_ms += ms_since(t0, t1);
      stats.host_plan_ms += ms_since(t1, t2);
      stats.host_enqueue_ms += ms_since(t2, t3);
      stats.host_wait_ms += ms_since(t3, tw);
      stats.host_finalize_ms += ms_since(tw, t4);
    }
    if (rc != MCR_OK) {
      fail_everything(g_err);
      return;
    }
    if (have) {
      inflight = std::move(rr);
      ++round;
    } else if (!inflight) {
      // nothing runnable: queued jobs wait for memory held by nobody, which
      // the submit check rules out; sleep until something changes
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait_for(lk, std::chrono::milliseconds(1));
    }
  }
}

extern "C" {

int mcr_server_footprint(const mcr_query* q, int64_t* bytes) {
  if (!bytes) {
    set_error("bytes is NULL");
    return MCR_E_ARG;
  }
  MCR_TRY(check_servable(q));
  *bytes = (int64_t)query_footprint(q, true);
  return MCR_OK;
}

int mcr_server_create(int device, const mcr_server_opts* opts, mcr_server** out) {
  if (!out) {
    set_error("out is NULL");
    return MCR_E_ARG;
  }
  *out = nullptr;
  auto s = std::make_unique<mcr_server>();
  const int r = s->init(device, opts);
  if (r != MCR_OK) {
    std::string msg = g_err;
    s->teardown();
    set_error("%s", msg.c_str());
    return r;
  }
  *out = s.release();
  return MCR_OK;
}

int mcr_server_submit(mcr_server* s, const mcr_query* q, uint64_t* ticket) {
  if (!s || !ticket) {
    set_error("server and ticket are required");
    return MCR_E_ARG;
  }
  MCR_TRY(check_servable(q));
  const size_t need = query_footprint(q, true);
  if (need > s->budget) {
    set_error("query needs %zu device bytes, the admission budget is %zu", need, s->budget);
    return MCR_E_NOMEM;
  }
  auto j = std::make_unique<Job>();
  const int n = q->n;
  j->q = *q;
  j->mu.assign(q->mu, q->mu + n);
  j->w.resize(n);
  for (int i = 0; i < n; ++i) j->w[i] = q->weights ? q->weights[i] : 1.0 / (double)n;
  j->sigma.resize((size_t)n * n);
  for (int r = 0; r < n; ++r)
    std::memcpy(&j->sigma[(size_t)r * n], q->sigma + (size_t)r * q->ld_sigma, n * sizeof(double));
  j->q.mu = j->mu.data();
  j->q.weights = j->w.data();
  j->q.sigma = j->sigma.data();
  j->q.ld_sigma = n;
  j->hash = hash_sigma(n, j->sigma.data());
  j->T = tiles_T(n);
  j->S = steps_S(n);
  j->na = (q->flags & MCR_F_VAR) ? q->n_alpha : 0;
  j->wcap = window_capacity(&j->q, kBatchGatherCap);
  j->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop || s->broken) {
      set_error("server is %s", s->broken ? s->broken_msg.c_str() : "shutting down");
      return MCR_E_DEVICE;
    }
    j->ticket = s->next_ticket++;
    *ticket = j->ticket;
    s->queued.push_back(j.get());
    s->stats.submitted++;
    s->jobs.emplace(j->ticket, std::move(j));
  }
  s->cv_work.notify_one();
  return MCR_OK;
}

static int take_result(mcr_server* s, std::unique_lock<std::mutex>& lk, uint64_t ticket,
                       mcr_result* r, mcr_ticket_info* info) {
  auto it = s->jobs.find(ticket);
  Job* j = it->second.get();
  if (r) *r = j->res;
  if (info) {
    info->latency_ms = ms_since(j->t_submit, j->t_done);
    info->queue_ms = j->t_admit.time_since_epoch().count() ? ms_since(j->t_submit, j->t_admit) : 0.0;
    info->rounds = j->rounds;
    info->cache_hit = j->hit ? 1 : 0;
    info->status = j->status;
  }
  const int status = j->status;
  if (status != MCR_OK) set_error("%s", j->err.c_str());
  s->jobs.erase(it);
  (void)lk;
  return status;
}

int mcr_server_wait(mcr_server* s, uint64_t ticket, mcr_result* r, mcr_ticket_info* info) {
  if (!s) {
    set_error("server is NULL");
    return MCR_E_ARG;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->jobs.find(ticket) == s->jobs.end()) {
    set_error("unknown ticket %llu", (unsigned long long)ticket);
    return MCR_E_ARG;
  }
  s->cv_done.wait(lk, [&] {
    auto it = s->jobs.find(ticket);
    return it == s->jobs.end() || it->second->state == JobState::kDone;
  });
  if (s->jobs.find(ticket) == s->jobs.end()) {
    set_error("ticket %llu was taken by another wait", (unsigned long long)ticket);
    return MCR_E_ARG;
  }
  return take_result(s, lk, ticket, r, info);
}

int mcr_server_poll(mcr_server* s, uint64_t ticket, mcr_result* r, mcr_ticket_info* info) {
  if (!s) {
    set_error("server is NULL");
    return MCR_E_ARG;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end()) {
    set_error("unknown ticket %llu", (unsigned long long)ticket);
    return MCR_E_ARG;
  }
  if (it->second->state != JobState::kDone) return 0;
  const int status = take_result(s, lk, ticket, r, info);
  return status == MCR_OK ? 1 : status;
}

int mcr_server_get_stats(mcr_server* s, mcr_server_stats* st) {
  if (!s || !st) {
    set_error("server and stats are required");
    return MCR_E_ARG;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->stats;
  st->queued = (int)s->queued.size();
  int act = 0;
  for (auto& kv : s->jobs)
    if (kv.second->state == JobState::kActive) ++act;
  st->active = act;
  return MCR_OK;
}

int mcr_server_destroy(mcr_server* s) {
  if (!s) return MCR_OK;
  s->teardown();
  delete s;
  return MCR_OK;
}

}  // extern "C"
