/*
 * mcr_oracle.c -- CPU oracle (test infrastructure only; see mcr_oracle.h).
 * Build: oracle/Makefile (gcc -std=c11 -O2 -ffp-contract=off -fopenmp).
 */
#include "mcr_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------------
 * Philox4x32-10, restated from Salmon et al. SC'11 (sec. 4.2 and Table 2):
 * each round multiplies words 0 and 2 by M0 and M1, the high halves are xored
 * with words 1 and 3 and the round key, the low halves become words 1 and 3;
 * the key is bumped by the Weyl constants between rounds.
 * ------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t a = (uint64_t)0xD2511F53u * (uint64_t)c0;
    uint64_t b = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)b;
    uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)a;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static float bits_f(uint32_t b) {
  float f;
  memcpy(&f, &b, 4);
  return f;
}
static uint32_t f_bits(float f) {
  uint32_t b;
  memcpy(&b, &f, 4);
  return b;
}

/* -ln(((a>>8)+1) 2^-24): u = 2^e m, m in [sqrt(1/2), sqrt(2)), ln(m) = ln(1+f)
 * as f - f^2/2 + f^3 P(f) with the stream's fitted coefficients, then e ln2 in
 * two parts. */
float orc_neg_log_u24(uint32_t a) {
  float u = fmaf((float)(a >> 8), 0x1p-24f, 0x1p-24f);
  int32_t ix = (int32_t)f_bits(u) - 0x3f3504f3;
  int32_t e = ix >> 23; /* arithmetic: gcc defines >> on negative int32 as such */
  float f = bits_f((uint32_t)(ix & 0x007fffff) + 0x3f3504f3u) - 1.0f;
  float p = 0x1.645d88p-4f;
  p = fmaf(p, f, -0x1.2432a4p-3f);
  p = fmaf(p, f, 0x1.31747cp-3f);
  p = fmaf(p, f, -0x1.53824cp-3f);
  p = fmaf(p, f, 0x1.98d7f0p-3f);
  p = fmaf(p, f, -0x1.000382p-2f);
  p = fmaf(p, f, 0x1.5556d8p-2f);
  float q = fmaf(p, f, -0.5f);
  float ff = f * f;
  float l1p = fmaf(ff, q, f);
  float ef = (float)e;
  float lo = fmaf(ef, 0x1.7f7d1cp-20f, l1p);
  return -fmaf(ef, 0x1.62e400p-1f, lo);
}

/* sqrt(2h) via three Newton steps on 1/sqrt(2h) from a bit-pattern seed; the
 * steps use h + 2^-60 (= h unless h = 0, where g would overflow). */
float orc_sqrt_twice(float h) {
  float hn = h + 0x1p-60f;
  float g = bits_f(0x5ef75a86u - (f_bits(hn) >> 1));
  for (int it = 0; it < 3; ++it) {
    float t = g * g;
    t = fmaf(-hn, t, 1.5f);
    g = g * t;
  }
  return (h + h) * g;
}

/* cos, sin of 2 pi b / 2^32: nearest quarter turn from the top bits, reduced
 * argument in [-1/8, 1/8) turns, odd/even polynomials, quadrant rotation. */
void orc_sincos_turn(uint32_t b, float *c_out, float *s_out) {
  int32_t rem = (int32_t)(b << 2) >> 2;
  uint32_t q = (b - (uint32_t)rem) >> 30;
  float a = (float)rem * 0x1p-32f;
  float t = a * a;
  float s = fmaf(-0x1.2d9bf2p+6f, t, 0x1.465ec8p+6f);
  s = fmaf(s, t, -0x1.4abbbap+5f);
  s = fmaf(s, t, 0x1.921fb6p+2f);
  s = s * a;
  float c = fmaf(0x1.d9f7bep+5f, t, -0x1.55c664p+6f);
  c = fmaf(c, t, 0x1.03c1dep+6f);
  c = fmaf(c, t, -0x1.3bd3ccp+4f);
  c = fmaf(c, t, 1.0f);
  float sin_v, cos_v;
  switch (q & 3u) {
    case 0: sin_v = s; cos_v = c; break;
    case 1: sin_v = c; cos_v = -s; break;
    case 2: sin_v = -s; cos_v = -c; break;
    default: sin_v = -c; cos_v = s; break;
  }
  *c_out = cos_v;
  *s_out = sin_v;
}

void orc_box_muller_f32(uint32_t a, uint32_t b, float *z0, float *z1) {
  float r = orc_sqrt_twice(orc_neg_log_u24(a));
  float c, s;
  orc_sincos_turn(b, &c, &s);
  *z0 = r * c;
  *z1 = r * s;
}

static void block4(uint64_t seed, uint64_t path, uint32_t blk, uint32_t out[4]) {
  uint32_t ctr[4] = {blk, 0u, (uint32_t)path, (uint32_t)(path >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox4x32_10(ctr, key, out);
}

void orc_normals_f32(uint64_t seed, uint64_t path, int n, float *z) {
  for (int j = 0; j < n; j += 4) {
    uint32_t x[4];
    float t[4];
    block4(seed, path, (uint32_t)(j / 4), x);
    orc_box_muller_f32(x[0], x[1], &t[0], &t[1]);
    orc_box_muller_f32(x[2], x[3], &t[2], &t[3]);
    for (int q = 0; q < 4 && j + q < n; ++q) z[j + q] = t[q];
  }
}

void orc_normals_batch_f32(uint64_t seed, uint64_t path0, int64_t npaths, int n, float *z) {
  for (int64_t p = 0; p < npaths; ++p) orc_normals_f32(seed, path0 + (uint64_t)p, n, z + p * n);
}

void orc_normals_f64(uint64_t seed, uint64_t path, int n, double *z) {
  const double two_pi = 6.283185307179586476925286766559;
  for (int j = 0; j < n; j += 4) {
    uint32_t x[4];
    double t[4];
    block4(seed, path, (uint32_t)(j / 4), x);
    for (int h = 0; h < 2; ++h) {
      double u = ((double)(x[2 * h] >> 8) + 1.0) * 0x1p-24;
      double r = sqrt(-2.0 * log(u));
      double th = two_pi * ((double)x[2 * h + 1] * 0x1p-32);
      t[2 * h] = r * cos(th);
      t[2 * h + 1] = r * sin(th);
    }
    for (int q = 0; q < 4 && j + q < n; ++q) z[j + q] = t[q];
  }
}

/* ---------------------------------------------------------------------------
 * Cholesky (FP64), Cholesky-Crout order, row-major lower factor.
 * ------------------------------------------------------------------------- */
int orc_cholesky(int n, const double *A, int lda, double delta, int clamp, double clamp_value,
                 double *L, int *clamped) {
  int nclamp = 0;
  memset(L, 0, sizeof(double) * (size_t)n * (size_t)n);
  for (int j = 0; j < n; ++j) {
    double d = A[(size_t)j * lda + j] + delta;
    double tol = (double)n * DBL_EPSILON * d;
    for (int k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    if (!(d > tol) || !isfinite(d)) {
      if (!clamp || !isfinite(d)) {
        if (clamped) *clamped = nclamp;
        return j + 1;
      }
      d = clamp_value;
      ++nclamp;
    }
    double ljj = sqrt(d);
    L[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double s = A[(size_t)i * lda + j];
      for (int k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = s / ljj;
    }
  }
  if (clamped) *clamped = nclamp;
  return 0;
}

int orc_factor(int n, const double *A, int lda, double *L, double *delta_out, int *clamped) {
  double mean = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j)
      if (!isfinite(A[(size_t)i * lda + j])) return -1;
    mean += A[(size_t)i * lda + i];
  }
  mean /= (double)n;
  if (clamped) *clamped = 0;
  double delta = 0.0;
  for (int t = 0; t <= 8; ++t) {
    delta = t == 0 ? 0.0 : mean * pow(10.0, (double)(t - 12));
    if (orc_cholesky(n, A, lda, delta, 0, 0.0, L, NULL) == 0) {
      *delta_out = delta;
      return t + 1;
    }
  }
  int nc = 0;
  orc_cholesky(n, A, lda, delta, 1, delta, L, &nc);
  *delta_out = delta;
  if (clamped) *clamped = nc;
  return 10;
}

/* ---------------------------------------------------------------------------
 * Path simulation.
 * ------------------------------------------------------------------------- */
static float next_up(float x) { return nextafterf(x, INFINITY); }
static float next_down(float x) { return nextafterf(x, -INFINITY); }

int64_t orc_simulate_f32(const orc_model_f32 *m, uint64_t seed, uint64_t path0, int64_t npaths, int k,
                         int32_t *S_out, float *loss_out, int64_t *near_out, int64_t *s_hist,
                         int nthreads) {
  const int n = m->n;
  const float lo = next_down(m->thr), hi = next_up(m->thr);
  int64_t count = 0, near = 0;
  if (nthreads < 1) nthreads = 1;
  int64_t *hist_all = NULL;
  if (s_hist) hist_all = (int64_t *)calloc((size_t)nthreads * (size_t)(n + 1), sizeof(int64_t));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) reduction(+ : count, near)
#endif
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    float *z = (float *)malloc(sizeof(float) * (size_t)(n + 4));
    float *r = (float *)malloc(sizeof(float) * (size_t)(n + 4));
    int64_t *hist = hist_all ? hist_all + (size_t)tid * (size_t)(n + 1) : NULL;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t p = 0; p < npaths; ++p) {
      orc_normals_f32(seed, path0 + (uint64_t)p, n, z);
      int S = 0, is_near = 0;
      for (int i = 0; i < n; ++i) {
        float acc = m->mu[i];
        const float *Li = m->L + (size_t)i * n;
        for (int j = 0; j <= i; ++j) acc = fmaf(Li[j], z[j], acc);
        r[i] = acc;
        if (acc < m->thr) ++S;
        if (acc >= lo && acc <= hi) is_near = 1;
      }
      float g = m->mp;
      for (int j = 0; j < n; ++j) g = fmaf(m->v[j], z[j], g);
      if (S_out) S_out[p] = S;
      if (loss_out) loss_out[p] = -g;
      if (hist) hist[S]++;
      if (S >= k) ++count;
      near += is_near;
    }
    free(z);
    free(r);
  }
  if (s_hist) {
    for (int b = 0; b <= n; ++b) {
      int64_t s = 0;
      for (int t = 0; t < nthreads; ++t) s += hist_all[(size_t)t * (size_t)(n + 1) + b];
      s_hist[b] = s;
    }
    free(hist_all);
  }
  if (near_out) *near_out = near;
  return count;
}

int64_t orc_simulate_f64(const orc_model_f64 *m, uint64_t seed, uint64_t path0, int64_t npaths, int k,
                         int32_t *S_out, double *loss_out, int nthreads) {
  const int n = m->n;
  int64_t count = 0;
  if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) reduction(+ : count)
#endif
  {
    double *z = (double *)malloc(sizeof(double) * (size_t)(n + 4));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t p = 0; p < npaths; ++p) {
      orc_normals_f64(seed, path0 + (uint64_t)p, n, z);
      int S = 0;
      for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        const double *Li = m->L + (size_t)i * n;
        for (int j = 0; j <= i; ++j) acc += Li[j] * z[j];
        if (m->mu[i] + acc < -m->x) ++S;
      }
      double g = 0.0;
      for (int j = 0; j < n; ++j) g += m->v[j] * z[j];
      if (S_out) S_out[p] = S;
      if (loss_out) loss_out[p] = -(m->mp + g);
      if (S >= k) ++count;
    }
    free(z);
  }
  return count;
}

void orc_stats(int64_t count, int64_t M, double out[4]) {
  double p = (double)count / (double)M;
  double se = sqrt(p * (1.0 - p) / (double)M);
  double margin = 1.96 * se;
  out[0] = p;
  out[1] = se;
  out[2] = fmax(0.0, p - margin);
  out[3] = fmin(1.0, p + margin);
}

int64_t orc_fixed32(float v) { return (int64_t)((double)v * 4294967296.0); }
