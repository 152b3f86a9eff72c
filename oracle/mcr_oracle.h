/*
 * mcr_oracle.h -- CPU oracle for the Monte Carlo crash-risk path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP model:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product never links, calls or falls back to it.
 *
 * It restates, in plain C99 (gcc, -ffp-contract=off), the reference path:
 *   - per-trial algorithm  R = mu + L Z, S = #{R_i < -x}, extreme iff S >= k
 *     (src/05-GPU/monte_carlos_cuda.cu:23-72, src/02-C-serial/monte_carlo_serial.c:62-85,
 *      src/01-python-serial/python_serial.py:38-75), with the lower factor tril(L)
 *     as the CUDA and Python paths use it;
 *   - the trial loop and statistics of the serial model
 *     (src/02-C-serial/monte_carlo_serial.c:98-164, stats :152-158);
 *   - the OpenMP model's static split with a count reduction
 *     (src/03-openMP/monte_carlo_omp.c:152-163), here on a counter-based stream so
 *     its counts equal the serial ones by construction;
 *   - the Cholesky factorisation the models call (gsl_linalg_cholesky_decomp1,
 *     src/05-GPU/monte_carlos_cuda.cu:87-93), restated as textbook Cholesky-Crout in
 *     FP64 plus the HIP model's jitter policy.
 * The reference's own random streams (GSL MT19937, cuRAND XORWOW) cannot be
 * reproduced here (GSL is not installed, cuRAND is closed); the oracle follows the
 * HIP model's documented Philox4x32-10 + Box-Muller stream instead, restated from
 * Salmon et al. (SC'11) and the stream definition.
 *
 * Pinning (tests/test_oracle_*.py): Philox against the Random123 known-answer
 * vectors; the fp32 transcendental kernels against libm error bounds and the
 * Box-Muller identity; the normals against N(0,1) moments and a KS test; Cholesky
 * against LAPACK (numpy.linalg.cholesky); the path statistic against closed
 * forms (one-factor equicorrelated integral, marginal expectations, Genz orthant
 * probabilities via scipy) and brute force on n = 1, 2; VaR against the Gaussian
 * quantile of the portfolio loss.
 */
#ifndef MCR_ORACLE_H
#define MCR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 and the normal stream ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_neg_log_u24(uint32_t a);
float orc_sqrt_twice(float h);
void orc_sincos_turn(uint32_t b, float *c, float *s);
void orc_box_muller_f32(uint32_t a, uint32_t b, float *z0, float *z1);
/* z[0..n-1] of one path: fp32 mirror of the device stream */
void orc_normals_f32(uint64_t seed, uint64_t path, int n, float *z);
/* same integers, fp64 Box-Muller through libm (the precision oracle) */
void orc_normals_f64(uint64_t seed, uint64_t path, int n, double *z);
/* z[p*n + j] for paths path0 .. path0+npaths-1 */
void orc_normals_batch_f32(uint64_t seed, uint64_t path0, int64_t npaths, int n, float *z);

/* ---- FP64 Cholesky with the jitter policy ---- */
/* Plain Cholesky-Crout of A + delta*I (lower, row-major out L, upper zeroed).
 * Returns 0 on success or j+1 for the first column whose pivot fails
 * (pivot <= n*eps*(A_jj+delta) or not finite).  clamp != 0: failing pivots are
 * replaced by clamp_value instead (modified Cholesky); *clamped counts them. */
int orc_cholesky(int n, const double *A, int lda, double delta, int clamp, double clamp_value,
                 double *L, int *clamped);
/* Jitter policy: delta_0 = 0, delta_t = mean(diag) * 10^(t-12) for t = 1..8;
 * if all fail, modified Cholesky with delta_8 and pivots clamped to delta_8.
 * Returns attempts used (1..9), or 10 when the clamp fallback ran; -1 on
 * non-finite input. */
int orc_factor(int n, const double *A, int lda, double *L, double *delta_out, int *clamped);

/* ---- path simulation ---- */
typedef struct {
  int n;
  const float *L;  /* n*n row-major; only j <= i read */
  const float *mu; /* n */
  const float *v;  /* n: L^T w (portfolio loading on z) */
  float mp;        /* w^T mu */
  float thr;       /* (float)(-x) */
} orc_model_f32;

typedef struct {
  int n;
  const double *L, *mu, *v;
  double mp;
  double x;
} orc_model_f64;

/* Mirrored fp32 path: r_i = mu_i; r_i = fmaf(L_ij, z_j, r_i) for j = 0..i;
 * S = #{r_i < thr}; g = mp; g = fmaf(v_j, z_j, g) for j = 0..n-1; loss = -g.
 * Paths path0 .. path0+npaths-1.  Optional per-path outputs S_out, loss_out.
 * near_out: paths with some |r_i - thr| <= 1 ulp(thr).  s_hist: n+1 bins.
 * Returns #{S >= k}.  nthreads = 1 is the serial model, > 1 the OpenMP model. */
int64_t orc_simulate_f32(const orc_model_f32 *m, uint64_t seed, uint64_t path0, int64_t npaths, int k,
                         int32_t *S_out, float *loss_out, int64_t *near_out, int64_t *s_hist,
                         int nthreads);
/* FP64 oracle on the same integer stream (libm Box-Muller, fp64 mat-vec). */
int64_t orc_simulate_f64(const orc_model_f64 *m, uint64_t seed, uint64_t path0, int64_t npaths, int k,
                         int32_t *S_out, double *loss_out, int nthreads);

/* P_hat, standard error and clamped 95% CI (monte_carlo_serial.c:152-158). */
void orc_stats(int64_t count, int64_t M, double out[4]);

/* fp32 -> fixed point 2^-32, truncating: the tail-sum unit. */
int64_t orc_fixed32(float v);

#ifdef __cplusplus
}
#endif
#endif
