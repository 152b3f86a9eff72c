/*
 * mcr_model.c -- the "hip" model behind the reference plug-in boundary.
 *
 * get_hip_model() / hip_simulate() take the place of get_cuda_model() /
 * cuda_simulate() (src/05-GPU/monte_carlos_cuda.cu:165-171, :75-163) and follow the
 * contract src/main_runner.c:198-298 relies on:
 *   - params and its gsl objects belong to the caller and are not modified
 *     (strides and tda are honoured);
 *   - result is filled on success; S_values is NULL unless MCR_HIP_S_VALUES=1,
 *     in which case it is malloc'd here and freed by the caller's make_clean
 *     (src/main_runner.c:47-52);
 *   - 0 on success, -1 on failure with a message on stderr; never exits.
 * MCR_HIP_PRECISION selects fp32 (default: bit-exact with the C mirror),
 * bf16x3 (MFMA) or auto.
 */
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/mcr_hip.h"

static int env_precision(void) {
  const char *p = getenv("MCR_HIP_PRECISION");
  if (!p || !*p || strcmp(p, "fp32") == 0) return MCR_PREC_FP32;
  if (strcmp(p, "bf16x3") == 0) return MCR_PREC_BF16X3;
  if (strcmp(p, "auto") == 0) return MCR_PREC_AUTO;
  fprintf(stderr, "Warning: unknown MCR_HIP_PRECISION '%s', using fp32\n", p);
  return MCR_PREC_FP32;
}

int hip_simulate(MonteCarloParams *params, MonteCarloResult *result) {
  if (!params || !result || !params->mu || !params->Sigma) {
    fprintf(stderr, "Error: hip model called without parameters\n");
    return -1;
  }
  const int n = params->N;
  printf("Starting HIP Monte Carlo simulation with M = %d trials...\n", params->M);
  if (n < 1 || (size_t)n != params->mu->size || (size_t)n != params->Sigma->size1 ||
      (size_t)n != params->Sigma->size2) {
    fprintf(stderr, "Error: N = %d does not match mu (%zu) / Sigma (%zux%zu)\n", n,
            params->mu->size, params->Sigma->size1, params->Sigma->size2);
    return -1;
  }
  if (params->M < 1) {
    fprintf(stderr, "Error: M = %d must be positive\n", params->M);
    return -1;
  }
  double *mu = (double *)malloc(sizeof(double) * (size_t)n);
  if (!mu) {
    fprintf(stderr, "Error: Failed to allocate mu copy\n");
    return -1;
  }
  for (int i = 0; i < n; ++i) mu[i] = params->mu->data[(size_t)i * params->mu->stride];

  const char *sv = getenv("MCR_HIP_S_VALUES");
  int *S_values = NULL;
  if (sv && strcmp(sv, "1") == 0) {
    S_values = (int *)malloc(sizeof(int) * (size_t)params->M);
    if (!S_values) {
      fprintf(stderr, "Error: Failed to allocate S_values array\n");
      free(mu);
      return -1;
    }
  }

  mcr_query q;
  memset(&q, 0, sizeof(q));
  q.n = n;
  q.mu = mu;
  q.sigma = params->Sigma->data;
  q.ld_sigma = (int)params->Sigma->tda;
  q.weights = NULL;
  q.x = params->x;
  q.k = params->k;
  q.paths = params->M;
  q.seed = (uint64_t)params->random_seed;
  q.precision = env_precision();
  q.S_out = (int32_t *)S_values;

  mcr_result r;
  const int st = mcr_run(&q, &r);
  free(mu);
  if (st != MCR_OK) {
    fprintf(stderr, "Error: HIP model: %s\n", mcr_last_error());
    free(S_values);
    return -1;
  }
  if (r.count > INT_MAX) {
    fprintf(stderr, "Error: count %lld does not fit MonteCarloResult.count\n", (long long)r.count);
    free(S_values);
    return -1;
  }
  if (r.factor_attempts > 1)
    fprintf(stderr, "Warning: covariance factored with jitter %.3e (attempts %d, clamped %d)\n",
            r.jitter, r.factor_attempts, r.factor_clamped);
  result->count = (int)r.count;
  result->P_hat = r.p_hat;
  result->std_error = r.std_error;
  result->ci_lower = r.ci_lower;
  result->ci_upper = r.ci_upper;
  result->S_values = S_values;
  printf("HIP Simulation complete\n");
  return 0;
}

ModelFunctions get_hip_model(void) {
  ModelFunctions model = {"hip", hip_simulate};
  return model;
}
