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

/* Holds the final results of the Monte Carlo simulation */
typedef struct {
  double P_hat;     /* Estimated probability of extreme event */
  int count;        /* Number of trials that exceeded threshold k */
  int *S_values;    /* Crash count S per trial (malloc'd by the model) or NULL */
  double std_error; /* Standard error of the estimate */
  double ci_lower;  /* Lower bound of 95% confidence interval */
  double ci_upper;  /* Upper bound of 95% confidence interval */
} MonteCarloResult;

/* Input parameters for the Monte Carlo simulation */
typedef struct {
  double x;                  /* Crash threshold (0.02 = 2%), compare R_i < -x */
  int N;                     /* Number of assets */
  int k;                     /* Minimum number of crashes for an extreme event */
  gsl_vector *mu;            /* Mean returns of assets */
  gsl_matrix *Sigma;         /* Covariance matrix of asset returns */
  int M;                     /* Number of Monte Carlo trials to run */
  unsigned long random_seed; /* Random seed for reproducibility */
} MonteCarloParams;

/* Common interface that all models implement: 0 on success, non-zero on failure */
typedef int (*ModelRunFunc)(MonteCarloParams *params, MonteCarloResult *result);

typedef struct {
  const char *name;       /* Model name ("serial", "openmp", "cuda", "hip") */
  ModelRunFunc run_model; /* Main simulation function */
} ModelFunctions;

#endif /* MODEL_INTERFACE_H */
