/*
 * mcr_runner.c -- the reference driver's flow (src/main_runner.c:105-298) for
 * this build: load data/mu.bin, data/sigma.bin, data/params.bin, read
 * config.yaml, ask for a comment, run every listed model through its
 * ModelFunctions entry, time it with CLOCK_MONOTONIC around run_model, print
 * the result and append one row to results/simulation_results.csv.
 *
 * The only model in this build is "hip" (get_hip_model); "serial", "openmp",
 * "mpi", "cuda"/"gpu" are reported as not available, unknown names as
 * unknown, and the run continues with the next model (main_runner.c:206-240).
 *
 * Usage: mcr_runner [config.yaml [data_dir [results.csv]]]
 *        (defaults: config.yaml, data, results/simulation_results.csv)
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mcr_hip.h"
#include "mcr_io.h"

static void print_result(const char *name, const MonteCarloResult *r, int M) {
  printf("\n%s RESULTS\n", name);
  printf("Estimated Probability (P_MIN): %.6f\n", r->P_hat);
  printf("Extreme Events Count: %d out of %d trials\n", r->count, M);
  printf("Standard Error: %.6f\n", r->std_error);
  printf("95%% Confidence Interval: [%.6f, %.6f]\n", r->ci_lower, r->ci_upper);
}

static int env_int(const char *name, int fallback) {
  const char *v = getenv(name);
  return v ? atoi(v) : fallback;
}

static char *read_comment(void) {
  char line[256];
  printf("\nEnter a comment for this simulation (max 255 chars, or press Enter to skip): ");
  fflush(stdout);
  if (!fgets(line, sizeof(line), stdin)) return NULL;
  size_t len = strlen(line);
  if (len > 0 && line[len - 1] == '\n') line[--len] = '\0';
  char *c = (char *)malloc(len + 1);
  if (c) memcpy(c, line, len + 1);
  return c;
}

static int run_models(const mcr_config *cfg, MonteCarloParams *params, double actual_freq,
                      const char *comment, const char *csv_path) {
  const int nodes = env_int("SLURM_NNODES", 1);
  const int threads = getenv("OMP_NUM_THREADS") ? env_int("OMP_NUM_THREADS", 1)
                                                 : env_int("SLURM_CPUS_PER_TASK", 1);
  const int processes = env_int("SLURM_NTASKS", 1);
  for (int i = 0; i < cfg->num_models; ++i) {
    const char *type = cfg->models[i];
    printf("\n=== Running %s model ===\n", type);
    ModelFunctions model;
    if (strcmp(type, "hip") == 0) {
      model = get_hip_model();
    } else if (strcmp(type, "serial") == 0 || strcmp(type, "openmp") == 0 ||
               strcmp(type, "mpi") == 0 || strcmp(type, "cuda") == 0 ||
               strcmp(type, "gpu") == 0) {
      fprintf(stderr, "Error: %s model not available in this build\n", type);
      continue;
    } else {
      fprintf(stderr, "Error: Unknown model type '%s'\n", type);
      continue;
    }
    MonteCarloResult result;
    memset(&result, 0, sizeof(result));
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    const int status = model.run_model(params, &result);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    const long ms = (long)((t1.tv_sec - t0.tv_sec) * 1000L + (t1.tv_nsec - t0.tv_nsec) / 1000000L);
    if (status == 0) {
      print_result(model.name, &result, params->M);
      mcr_csv_row row;
      memset(&row, 0, sizeof(row));
      row.iteration_id = 0; /* one run per process, as the reference's counter gives */
      row.timestamp = (long)time(NULL);
      row.execution_time_ms = ms;
      row.comment = comment;
      row.start_date = cfg->start;
      row.end_date = cfg->end;
      row.train_ratio = cfg->train_ratio;
      row.M = params->M;
      row.k = params->k;
      row.x = params->x;
      row.model_name = model.name;
      row.seed = params->random_seed;
      row.nodes = nodes;
      row.threads = threads;
      row.processes = processes;
      row.indices = cfg->indices;
      row.num_indices = cfg->num_indices;
      row.actual_freq = actual_freq;
      row.P_hat = result.P_hat;
      row.count = result.count;
      row.std_error = result.std_error;
      row.ci_lower = result.ci_lower;
      row.ci_upper = result.ci_upper;
      if (mcr_append_csv(csv_path, &row) != MCR_OK)
        fprintf(stderr, "Warning: Failed to write results to CSV: %s\n", mcr_last_error());
      else
        printf("Results written to %s\n", csv_path);
    } else {
      fprintf(stderr, "Error: %s model failed\n", model.name);
    }
    free(result.S_values); /* the caller owns S_values (main_runner.c:47-52) */
    printf("\n");
  }
  return 0;
}

int main(int argc, char **argv) {
  const char *cfg_path = argc > 1 ? argv[1] : "config.yaml";
  const char *data_dir = argc > 2 ? argv[2] : "data";
  const char *csv_path = argc > 3 ? argv[3] : "results/simulation_results.csv";
  char path[4096];

  printf("Monte Carlo Financial Risk Simulation\n\n");
  printf("Loading parameters from binary files...\n");
  double *mu = NULL, *sigma = NULL;
  int n = 0, rows = 0, cols = 0;
  snprintf(path, sizeof(path), "%s/mu.bin", data_dir);
  int ok_mu = mcr_read_mu_bin(path, &mu, &n) == MCR_OK;
  if (ok_mu)
    printf("Loaded mu (N=%d) from %s\n", n, path);
  else
    fprintf(stderr, "Error: %s\n", mcr_last_error());
  snprintf(path, sizeof(path), "%s/sigma.bin", data_dir);
  int ok_sigma = mcr_read_sigma_bin(path, &sigma, &rows, &cols) == MCR_OK;
  if (ok_sigma)
    printf("Loaded Sigma (%dx%d) from %s\n", rows, cols, path);
  else
    fprintf(stderr, "Error: %s\n", mcr_last_error());
  double actual_freq = NAN;
  snprintf(path, sizeof(path), "%s/params.bin", data_dir);
  if (mcr_read_params_bin(path, &actual_freq) == MCR_OK) {
    printf("Loaded actual_freq (%.6f) from %s\n", actual_freq, path);
  } else {
    fprintf(stderr, "Error: %s\n", mcr_last_error());
    fprintf(stderr, "Warning: Failed to load actual_freq from params.bin, using NaN\n");
    actual_freq = NAN;
  }
  if (!ok_mu || !ok_sigma) {
    fprintf(stderr, "Error: Failed to load binary files. Run the preprocessing first.\n");
    mcr_free(mu);
    mcr_free(sigma);
    return 1;
  }

  mcr_config cfg;
  if (mcr_load_config(cfg_path, &cfg) != MCR_OK) {
    fprintf(stderr, "Error: %s\nError: Failed to load %s. Exiting.\n", mcr_last_error(), cfg_path);
    mcr_free(mu);
    mcr_free(sigma);
    return 1;
  }
  printf("Loaded configuration from %s\n", cfg_path);
  printf("\nSimulation Configuration:\n");
  printf("  Number of assets (N): %d\n", n);
  printf("  Crash threshold (k): %d\n", cfg.k);
  printf("  Return threshold (x): %.2f%%\n", cfg.x * 100);
  printf("  Number of trials (M): %d\n", cfg.M);
  printf("  Training ratio: %.2f%%\n", cfg.train_ratio * 100);

  char *user_comment = read_comment();
  const char *comment = (user_comment && *user_comment) ? user_comment
                        : (cfg.comment && *cfg.comment)   ? cfg.comment
                                                          : "";

  gsl_vector vmu;
  memset(&vmu, 0, sizeof(vmu));
  vmu.size = (size_t)n;
  vmu.stride = 1;
  vmu.data = mu;
  gsl_matrix msig;
  memset(&msig, 0, sizeof(msig));
  msig.size1 = (size_t)rows;
  msig.size2 = (size_t)cols;
  msig.tda = (size_t)cols;
  msig.data = sigma;
  MonteCarloParams params;
  memset(&params, 0, sizeof(params));
  params.N = n;
  params.k = cfg.k;
  params.x = cfg.x;
  params.M = cfg.M;
  params.random_seed = cfg.random_seed;
  params.mu = &vmu;
  params.Sigma = &msig;

  run_models(&cfg, &params, actual_freq, comment, csv_path);

  free(user_comment);
  mcr_free_config(&cfg);
  mcr_free(mu);
  mcr_free(sigma);
  printf("Simulation complete.\n");
  return 0;
}
