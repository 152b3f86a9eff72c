/*
 * mcr_io.h -- data formats around the Monte Carlo risk path: the binary
 * inputs the reference's preprocessing writes, its YAML configuration and
 * its results CSV.  Plain C types; every function returns MCR_OK (0) or a
 * negative MCR_E_* code with mcr_last_error() set, and never prints or exits.
 *
 * Reference interfaces each function replaces:
 *   mcr_read_mu_bin      -> load_mu_binary        (src/utilities/load_binary.c:5-45)
 *   mcr_read_sigma_bin   -> load_sigma_binary     (src/utilities/load_binary.c:47-98)
 *   mcr_read_params_bin  -> load_actual_freq_binary (src/utilities/load_binary.c:100-124)
 *   mcr_write_bins       -> save_results_binary   (preprocessing/save_binary.py:5-45)
 *   mcr_load_config      -> load_config           (src/utilities/load_config.c:207-343)
 *   mcr_free_config      -> free_config           (src/utilities/load_config.c:27-62)
 *   mcr_append_csv       -> write_results_to_csv  (src/utilities/csv_writer.c:163-378)
 */
#ifndef MCR_IO_H
#define MCR_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mu.bin: int32 N, then N float64.  1 <= N <= 1e6.  *mu_out is malloc'd
 * (release with mcr_free). */
int mcr_read_mu_bin(const char *path, double **mu_out, int *n_out);
/* sigma.bin: int32 rows, int32 cols, then rows*cols float64 row-major.
 * rows, cols >= 1 and rows*cols <= 2e9. */
int mcr_read_sigma_bin(const char *path, double **sigma_out, int *rows_out, int *cols_out);
/* params.bin: one float64 actual crash frequency in [0, 1]. */
int mcr_read_params_bin(const char *path, double *actual_freq_out);
/* Writes dir/mu.bin, dir/sigma.bin (n x n) and dir/params.bin, creating dir. */
int mcr_write_bins(const char *dir, const double *mu, int n, const double *sigma,
                   double actual_freq);
void mcr_free(void *p);

/* config.yaml: flat keys start, end, x, N, M, k, random_seed, train_ratio,
 * comment; `indices:` a mapping of index name -> CSV file; `models:` a
 * sequence of model names (block or flow style, '#' comments).  Defaults and
 * conversions as the reference's loader: x 0.02, N 0, M 100000, k 5,
 * random_seed 42, train_ratio 0.8; integers by atoi/atol (so "1e8" is 1). */
typedef struct mcr_index {
  char *name;
  char *file;
} mcr_index;

typedef struct mcr_config {
  char *start, *end;
  double x;
  int N, M, k;
  unsigned long random_seed;
  double train_ratio;
  mcr_index *indices;
  int num_indices;
  char **models;
  int num_models;
  char *comment;
} mcr_config;

int mcr_load_config(const char *path, mcr_config *cfg);
void mcr_free_config(mcr_config *cfg);

/* One row of the 22-column results CSV.  The header is written when the
 * file does not exist yet; the parent directory is created; text fields
 * holding ',', '"', CR or LF are quoted with '"' doubled.  Appends are
 * serialised within the process. */
typedef struct mcr_csv_row {
  int iteration_id;
  long timestamp;          /* Unix seconds, written as local "%Y-%m-%d %H:%M:%S" */
  long execution_time_ms;
  const char *comment;
  const char *start_date, *end_date;
  double train_ratio;
  int M, k;
  double x;
  const char *model_name;
  unsigned long seed;
  int nodes, threads, processes; /* <= 0 written as 0 */
  const mcr_index *indices;      /* written as name:file;name:file */
  int num_indices;
  double actual_freq;
  double P_hat;
  int count;
  double std_error, ci_lower, ci_upper;
} mcr_csv_row;

int mcr_append_csv(const char *path, const mcr_csv_row *row);
/* The header line (without newline) and the formatted data line for a row
 * (without newline) into buf; returns the length or a negative code. */
int mcr_csv_header(char *buf, int cap);
int mcr_csv_format_row(const mcr_csv_row *row, char *buf, int cap);

#ifdef __cplusplus
}
#endif
#endif /* MCR_IO_H */
