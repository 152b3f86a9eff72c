/*
 * mcr_hip.h -- C ABI of the MI355X (gfx950) Monte Carlo crash-risk library.
 *
 * Plain C types and host pointers only.  Every entry point returns MCR_OK (0)
 * or a negative MCR_E_* code; mcr_last_error() describes the last failure of
 * the calling thread.  Nothing here aborts or exits.
 *
 * Reference interfaces each entry point replaces:
 *   get_hip_model / hip_simulate  -> get_cuda_model / cuda_simulate
 *       (src/05-GPU/monte_carlos_cuda.cu:165-171, :75-163), dispatched by
 *       src/main_runner.c:198-241 through ModelFunctions (src/model_interface.h:33-36)
 *   mcr_run                      -> the body of cuda_simulate: Cholesky (:84-93),
 *       flatten + H2D (:96-122), monte_carlo_kernel launch (:125-129, kernel :23-72),
 *       D2H (:138-143) and statistics (:146-150), with the count reduced on device
 *   mcr_factor                   -> gsl_linalg_cholesky_decomp1 as called at
 *       src/05-GPU/monte_carlos_cuda.cu:87-93 (FP64, here on the GPU with jitter)
 *   mcr_normals                  -> curand_init + curand_normal_double
 *       (src/05-GPU/monte_carlos_cuda.cu:30-31, :45-47): the per-path normal stream
 *   mcr_synth_factor_model       -> the mu/Sigma producer preprocessing/preprocessing_stats.py:5-28
 *       (sample mean, ddof=1 covariance), fed by a synthetic factor model
 */
#ifndef MCR_HIP_H
#define MCR_HIP_H

#include <stdint.h>

#include "model_interface.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MCR_OK 0
#define MCR_E_ARG -1         /* invalid argument */
#define MCR_E_UNSUPPORTED -2 /* size or option outside what a kernel supports */
#define MCR_E_DEVICE -3      /* HIP runtime error or no device */
#define MCR_E_NOMEM -4       /* device or host allocation failed */
#define MCR_E_FACTOR -5      /* covariance not factorable (non-finite input) */

#define MCR_MAX_ALPHA 4
#define MCR_MAX_N 1024        /* largest asset count of a single query (mcr_run) */
#define MCR_MAX_N_BATCH 767   /* largest in batches and the serving loop */
#define MCR_MAX_N_NOALLOC 608 /* largest through mcr_run_noalloc */

/* precision / kernel family */
#define MCR_PREC_FP32 0   /* fused fp32 VALU kernels: bit-exact with the C mirror */
#define MCR_PREC_BF16X3 1 /* MFMA bf16 hi/lo split products, fp32 accumulation */
#define MCR_PREC_AUTO 2   /* fastest supported for the size */

/* explicit kernel override (0 = choose from precision) */
#define MCR_KERNEL_AUTO 0
#define MCR_KERNEL_VALU_REG 1  /* one path per lane, rows in registers, n <= 64 */
#define MCR_KERNEL_VALU_TILE 2 /* 64 paths per workgroup, normals in LDS, n <= 608 */
#define MCR_KERNEL_MFMA 3      /* 64 paths per wave on 32x32x16 bf16 MFMA, n <= 127 */
#define MCR_KERNEL_MFMA_LDS 4  /* 64 paths per workgroup, normals staged in LDS, n <= 767
                                  (24 row tiles) */
#define MCR_KERNEL_VALU_CHUNK 5 /* fp32, 64 paths per workgroup, normals in 512-column LDS
                                   chunks, any n <= 1024: fp32 queries past 608, and every
                                   precision past 767 */

/* flags */
#define MCR_F_VAR 0x1   /* portfolio-loss quantiles (VaR) and tail sums at alpha[] */
#define MCR_F_NEAR 0x2  /* count paths with some r_i within 1 ulp of the threshold; fp32
                           kernels only (MCR_PREC_AUTO then picks fp32, MCR_PREC_BF16X3
                           or an MFMA kernel fails with MCR_E_UNSUPPORTED) */
#define MCR_F_SHIST 0x4 /* histogram of S (n+1 bins) into s_hist */
/* test hook: zero-width VaR windows, forcing the quantile refinement passes */
#define MCR_F_TEST_NARROW 0x100
/* test hook: VaR windows that hold every loss (every path appends its loss);
   needs paths <= 2^18 to stay one pass */
#define MCR_F_TEST_WIDE 0x200

typedef struct mcr_query {
  int n;                /* assets, 1 <= n <= MCR_MAX_N (MCR_MAX_N_BATCH in batches
                           and the serving loop, MCR_MAX_N_NOALLOC in
                           mcr_run_noalloc) */
  const double *mu;     /* [n] mean daily returns */
  const double *sigma;  /* [n * ld_sigma] covariance, row-major */
  int ld_sigma;         /* row stride of sigma (>= n) */
  const double *weights; /* [n] portfolio weights, NULL = equal weights 1/n */
  double x;             /* crash threshold: asset i crashes iff r_i < -x */
  int k;                /* extreme event iff #crashes >= k */
  int64_t paths;        /* Monte Carlo paths M (>= 1) */
  uint64_t seed;        /* Philox key */
  uint64_t path_offset; /* id of the first path (paths are keyed by id) */
  int precision;        /* MCR_PREC_* */
  int kernel;           /* MCR_KERNEL_* override, 0 = auto */
  int flags;            /* MCR_F_* */
  int n_alpha;          /* number of VaR levels (MCR_F_VAR), <= MCR_MAX_ALPHA */
  double alpha[MCR_MAX_ALPHA];
  int32_t *S_out;       /* optional host [paths]: crash count per path */
  float *loss_out;      /* optional host [paths]: portfolio loss per path */
  int64_t *s_hist;      /* optional host [n+1] (MCR_F_SHIST) */
} mcr_query;

typedef struct mcr_result {
  int64_t count;          /* paths with S >= k */
  int64_t near_threshold; /* MCR_F_NEAR: paths with some |r_i - thr| <= 1 ulp */
  double p_hat, std_error, ci_lower, ci_upper;
  /* MCR_F_VAR: VaR_a = ceil(a M)-th smallest loss; es = mean of the
   * M - ceil(a M) + 1 largest losses; tail_sum in units of 2^-32 */
  double var[MCR_MAX_ALPHA];
  double es[MCR_MAX_ALPHA];
  int64_t tail_sum_fixed[MCR_MAX_ALPHA];
  int var_passes;         /* path passes used for the quantiles (1 normally) */
  double loss_mean;       /* w^T mu negated: mean of the simulated loss */
  double loss_sd;         /* ||L^T w||: sd of the simulated loss */
  double jitter;          /* delta added to the diagonal before factoring */
  int factor_attempts;    /* 1 = plain Cholesky succeeded */
  int factor_clamped;     /* pivots clamped by the last-resort fallback */
  int kernel;             /* MCR_KERNEL_* that ran */
  double ms_kernel;       /* path kernel time (HIP events) */
  double ms_total;        /* host wall time of the call */
  double ms_device;       /* device time from the factor to the end of the path
                             kernels (HIP events; inputs already uploaded, no
                             host<->device copies); first VaR pass only */
} mcr_result;

typedef struct mcr_factor_info {
  int attempts;  /* 1..9 jitter attempts; 10 = clamp fallback ran */
  double jitter; /* delta finally added to the diagonal */
  int clamped;   /* pivots replaced by delta in the fallback */
} mcr_factor_info;

/* fp32 model the path kernels read, for parity checks (host copies) */
typedef struct mcr_model_export {
  double *L64;  /* [n*n] FP64 factor, row-major, upper zero (or NULL) */
  float *L32;   /* [n*n] fp32 factor (or NULL) */
  float *mu32;  /* [n] (or NULL) */
  float *v32;   /* [n] L^T w in fp32 (or NULL) */
  float mp32;   /* w^T mu in fp32 */
  float thr32;  /* (float)(-x) */
  double loss_sd;
} mcr_model_export;

/* Library lifetime.  mcr_init is optional (first use initialises device 0). */
int mcr_init(int device);
int mcr_shutdown(void);
const char *mcr_last_error(void);
int mcr_version(void);

/* One risk query: factor, prepare, run all paths, reduce, statistics. */
int mcr_run(const mcr_query *q, mcr_result *r);

/* mcr_run without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate (tight serving loops, harnesses that forbid
 * hipMalloc or host allocation while they run; the call synchronises the
 * library's stream before it returns, so it is not itself recorded into a
 * caller's HIP graph).  Same query, same kernels, same
 * result as mcr_run, but every device buffer of the query comes from a
 * 16 MiB array of static device memory that the library holds on the device
 * until the process ends, and the host side uses static buffers: no
 * hipMalloc, hipFree, pinned or heap allocation.  Needs mcr_init first;
 * fails with MCR_E_UNSUPPORTED when n > MCR_MAX_N_NOALLOC or the query's
 * buffers exceed the array (they do not for n <= MCR_MAX_N_NOALLOC; per-path
 * outputs are staged 2^18 paths at a time).  Calls share the one array, so calls that overlap on one
 * device race: they are serialised with mcr_run by the library's lock. */
int mcr_run_noalloc(const mcr_query *q, mcr_result *r);

/* GPU FP64 Cholesky of sigma with the jitter policy; L_out [n*n] row-major. */
int mcr_factor(int n, const double *sigma, int ld_sigma, double *L_out, mcr_factor_info *info);

/* Factor + prepare only; copies the device model to the caller's arrays. */
int mcr_prepare(const mcr_query *q, mcr_model_export *out);

/* The device normal stream: out[p*n + j] = z_j of path path0 + p. */
int mcr_normals(uint64_t seed, uint64_t path0, int64_t npaths, int n, float *out);

/* ---- batches of independent queries (portfolios) ---- */
#define MCR_BATCH_PERSISTENT 0 /* batched factor + prepare, then one persistent path launch
                                  per kernel class, all in one hipGraph */
#define MCR_BATCH_PER_QUERY 1  /* one mcr_run per query (the comparison baseline) */
#define MCR_ORDER_LPT 0        /* work items by descending cost (longest first) */
#define MCR_ORDER_NATURAL 1    /* work items in query order */

typedef struct mcr_batch_opts {
  int mode;          /* MCR_BATCH_* */
  int use_graph;     /* capture the device work in a hipGraph (replayed while the batch
                        layout is unchanged) */
  int chunk_paths;   /* paths per work item, multiple of 64; 0 = autotuned value */
  int order;         /* MCR_ORDER_*; used when chunk_paths != 0 */
  int grid;          /* persistent workgroups per launch; 0 = each class kernel's
                        resident maximum (occupancy x CUs) */
  double ms_device;  /* out: device time of the batch (HIP events): factor, layouts and
                        path kernels; per-query mode sums each query's ms_device */
  double ms_total;   /* out: host wall time of the call */
  int chunk_used;    /* out: chunk_paths actually used */
  int order_used;    /* out */
  int graph_replayed; /* out: 1 if an existing graph was replayed */
  int64_t work_items; /* out */
  int streams;       /* concurrent streams the class launches are forked over (1..4);
                        0 = the tuned default */
  int shapes[4];     /* workgroup shape of the LDS MFMA kernel per row-tile range
                        (n 128-255, 256-383, 384-511, 512-767): 0 = the tuned shape,
                        s > 0 = shape s - 1 (all shapes give identical results) */
  int streams_used;  /* out */
  int big_shapes[4]; /* out: shape used per range (0 = the measured default) */
  /* out, persistent mode: host time of the call by phase (ms): planning the
   * work items, staging the inputs (pinned copies, upload, graph lookup or
   * capture and launch), waiting for the device, finalising the results
   * (window copies and quantiles, any refinement) */
  double host_plan_ms, host_stage_ms, host_wait_ms, host_finalize_ms;
} mcr_batch_opts;

/* Runs nq independent queries.  Persistent mode: every query uses the bf16x3
 * MFMA kernel family (precision MCR_PREC_BF16X3 or MCR_PREC_AUTO; no per-path
 * outputs); results equal those of mcr_run on the same query with precision
 * MCR_PREC_BF16X3 and kernel MCR_KERNEL_AUTO (MCR_KERNEL_MFMA for n <= 127,
 * MCR_KERNEL_MFMA_LDS above).  rs[i] receives query i's result. */
int mcr_run_batch(const mcr_query *qs, int nq, mcr_result *rs, mcr_batch_opts *opts);

/* Times candidate settings on this batch -- work-item size x item order with
 * the class launches on one stream, then 2-4 concurrent class-launch streams
 * at the best of those, then the alternative workgroup shapes of each
 * row-tile range the batch uses -- keeps the fastest as the library's tuned
 * default (chunk and order used when chunk_paths = 0, streams when
 * streams = 0, shapes always, also by the serving loop) and returns it in
 * *best (chunk_used, order_used, streams_used, big_shapes, ms_device).
 * n_candidates (optional) receives how many settings were timed. */
int mcr_autotune_batch(const mcr_query *qs, int nq, mcr_batch_opts *best, int *n_candidates);

/* ---- serving loop (continuous batching) ----
 * No counterpart in the reference (it runs one query per process,
 * src/main_runner.c:198-296).  A server owns a device-memory arena of
 * hbm_budget bytes, a cache of FP64 Cholesky factors keyed by covariance
 * content and a worker thread.  Submitted queries are admitted while their
 * device footprint fits the free budget; the worker runs rounds of the
 * persistent path kernels over every admitted query, admitting new queries
 * between rounds, so queries that arrive while others run share the device
 * with them.  Each query's paths are keyed by (seed, path id) and reduced
 * exactly, so its result equals mcr_run on the same query with precision
 * MCR_PREC_BF16X3 and kernel MCR_KERNEL_AUTO.  Queries: precision
 * MCR_PREC_BF16X3 or MCR_PREC_AUTO, flags MCR_F_VAR only, no per-path outputs.
 * All calls are thread-safe. */
typedef struct mcr_server mcr_server;

typedef struct mcr_server_opts {
  int64_t hbm_budget;  /* device bytes for admitted queries and cached factors; 0 = 4 GiB */
  int cache_entries;   /* Cholesky factors kept on the device; 0 = 256, -1 = no cache */
  int max_active;      /* queries admitted at once; 0 = 1024 */
  double round_ms;     /* target device time of one round; 0 = 4 ms */
  int grid;            /* persistent workgroups; 0 = one per CU */
  int streams;         /* concurrent streams a round's class launches are forked
                          over (1..2; 0 = 2, larger values are taken as 2) */
  double age_limit_ms; /* rounds go to the queries with the least work left
                          first; a query waiting longer than this since submit
                          goes before them, oldest first (bounds the worst
                          latency below saturation at some cost in p50/p99);
                          0 = no limit */
} mcr_server_opts;

typedef struct mcr_ticket_info {
  double latency_ms;  /* submit -> result ready */
  double queue_ms;    /* submit -> admitted */
  int rounds;         /* rounds the query's paths were spread over */
  int cache_hit;      /* 1 if its Cholesky factor came from the device cache */
  int status;         /* MCR_OK or the query's error code */
} mcr_ticket_info;

typedef struct mcr_server_stats {
  int64_t submitted, completed, failed;
  int64_t rounds, cache_hits, cache_misses;
  int64_t paths_done;
  double device_ms;      /* summed device time of all rounds */
  int64_t peak_bytes;    /* high-water mark of the budget in use */
  int64_t budget_bytes;
  int active, queued;    /* at the time of the call */
  /* Worker-thread time summed over rounds (ms): admission with the
   * factor-cache lookups, round planning, staging and launching, waiting for
   * the previous round's results, finalising it (quantiles, releases).  A
   * round's host work overlaps the previous round's kernels; once it is
   * longer, the device waits. */
  double host_admit_ms, host_plan_ms, host_enqueue_ms, host_wait_ms, host_finalize_ms;
} mcr_server_stats;

/* Device bytes a query needs while admitted (inputs, operands, accumulators,
 * VaR buffers, plus its factor when not cached); host-only. */
int mcr_server_footprint(const mcr_query *q, int64_t *bytes);

int mcr_server_create(int device, const mcr_server_opts *opts, mcr_server **out);
/* Copies the query's inputs and queues it; *ticket identifies it.  Fails with
 * MCR_E_NOMEM when the query could never fit the budget. */
int mcr_server_submit(mcr_server *s, const mcr_query *q, uint64_t *ticket);
/* Blocks until the ticket's query is done; returns its status.  info may be
 * NULL.  A ticket is consumed by the wait (or poll) that returns it. */
int mcr_server_wait(mcr_server *s, uint64_t ticket, mcr_result *r, mcr_ticket_info *info);
/* 1 (and the result) if done, 0 if not yet, < 0 for an unknown ticket. */
int mcr_server_poll(mcr_server *s, uint64_t ticket, mcr_result *r, mcr_ticket_info *info);
int mcr_server_get_stats(mcr_server *s, mcr_server_stats *st);
/* Finishes the admitted queries, fails the queued ones, frees everything. */
int mcr_server_destroy(mcr_server *s);

/* Synthetic factor-model data (host): r_t = drift + B f_t + e_t, t < T, with
 * B ~ U(0.2, 0.8) [n x K], f_t ~ N(0, fvol^2 I_K), e_t ~ N(0, ivol^2 I_n);
 * mu = sample mean, sigma = ddof-1 sample covariance [n*n].  returns (optional)
 * receives the [T*n] return matrix. */
int mcr_synth_factor_model(int n, int K, int T, double drift, double fvol, double ivol,
                           uint64_t seed, double *mu, double *sigma, double *returns);

/* Plug-in model for the reference driver (src/main_runner.c). */
int hip_simulate(MonteCarloParams *params, MonteCarloResult *result);
ModelFunctions get_hip_model(void);

#ifdef __cplusplus
}
#endif
#endif /* MCR_HIP_H */
