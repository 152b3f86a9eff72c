/*
 * model_interface.h -- the reference's model plug-in boundary, ABI-compatible.
 *
 * Same layout and meaning as src/model_interface.h:9-36 of the reference:
 * MonteCarloResult (:9-16), MonteCarloParams (:19-27), ModelRunFunc (:30) and
 * ModelFunctions (:33-36).  The reference includes GSL for gsl_vector and
 * gsl_matrix; when GSL's headers are present they are used, otherwise the two
 * structs are declared here with GSL's documented field layout
 * (gsl_vector: size, stride, data, block, owner; gsl_matrix: size1, size2, tda,
 * data, block, owner), so a driver built against GSL and this library agree.
 */
#ifndef MODEL_INTERFACE_H
#define MODEL_INTERFACE_H

#include <stddef.h>

#if defined(__has_include)
#if __has_include(<gsl/gsl_vector.h>) && !defined(MCR_NO_GSL)
#define MCR_HAVE_GSL 1
#endif
#endif

#ifdef MCR_HAVE_GSL
#include <gsl/gsl_matrix.h>
#include <gsl/gsl_vector.h>
#else
typedef struct {
  size_t size;
  double *data;
} gsl_block;

typedef struct {
  size_t size;
  size_t stride;
  double *data;
  gsl_block *block;
  int owner;
} gsl_vector;

typedef struct {
  size_t size1;
  size_t size2;
  size_t tda;
  double *data;
  gsl_block *block;
  int owner;
} gsl_matrix;
#endif

/* What a model hands back to the driver after one run_model call. */
typedef struct {
  double P_hat;     /* count / M: share of paths with S >= k */
  int count;        /* paths whose crash tally S reached k */
  int *S_values;    /* optional per-path S array, malloc'd here and free()d by
                       the driver; NULL when the model keeps none */
  double std_error; /* sqrt(P_hat (1 - P_hat) / M) */
  double ci_lower;  /* P_hat - 1.96 std_error, floored at 0 */
  double ci_upper;  /* P_hat + 1.96 std_error, capped at 1 */
} MonteCarloResult;

/* One query as the driver fills it; the model reads it and never frees it. */
typedef struct {
  double x;                  /* loss level: asset i crashes on a path iff r_i < -x */
  int N;                     /* asset count, the length of mu and order of Sigma */
  int k;                     /* a path is extreme once at least k assets crash */
  gsl_vector *mu;            /* daily mean return per asset (may be strided) */
  gsl_matrix *Sigma;         /* N x N return covariance, row stride tda */
  int M;                     /* Monte Carlo path count */
  unsigned long random_seed; /* seeds the path generator (the hip model keys
                                Philox with it) */
} MonteCarloParams;

/* The entry every model exposes: 0 when the run completed, non-zero otherwise */
typedef int (*ModelRunFunc)(MonteCarloParams *params, MonteCarloResult *result);

typedef struct {
  const char *name;       /* the key the driver's config lists ("hip" here) */
  ModelRunFunc run_model; /* runs one query synchronously */
} ModelFunctions;

#endif /* MODEL_INTERFACE_H */
