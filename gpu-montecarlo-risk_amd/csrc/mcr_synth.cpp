// mcr_synth.cpp -- deterministic synthetic factor-model inputs (host only).
//
// Produces mu and Sigma the way preprocessing/preprocessing_stats.py:18,21 does
// from a return matrix (column means; pandas cov, i.e. ddof = 1), but the
// returns come from a seeded factor model instead of index CSVs:
//   r_t = drift + B f_t + e_t,  B_ik ~ U(0.2, 0.8),  f_t ~ N(0, fvol^2 I_K),
//   e_t ~ N(0, ivol^2 I_n),  t = 0 .. T-1.
// Every draw is keyed by (seed, stream, element index) through Philox4x32-10,
// so the output does not depend on threading: stream 1 = loadings, 2 = factors,
// 3 = idiosyncratic terms; normals by FP64 Box-Muller.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/mcr_hip.h"
#include "mcr_rng.h"

namespace {

void draw4(uint64_t seed, uint32_t stream, uint64_t e, uint32_t out[4]) {
  const mcr::Ctr4 c = mcr::philox4x32_10(mcr::Ctr4{(uint32_t)e, stream, (uint32_t)(e >> 32), 0u},
                                         (uint32_t)seed, (uint32_t)(seed >> 32));
  out[0] = c.x;
  out[1] = c.y;
  out[2] = c.z;
  out[3] = c.w;
}

double unit53(uint32_t hi, uint32_t lo) {  // (0, 1)
  const uint64_t m = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
  return ((double)m + 0.5) * 0x1p-53;
}

double uniform(uint64_t seed, uint32_t stream, uint64_t e) {
  uint32_t x[4];
  draw4(seed, stream, e, x);
  return unit53(x[0], x[1]);
}

double normal(uint64_t seed, uint32_t stream, uint64_t e) {
  uint32_t x[4];
  draw4(seed, stream, e, x);
  const double u = unit53(x[0], x[1]);
  const double v = unit53(x[2], x[3]);
  return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586476925286766559 * v);
}

}  // namespace

extern "C" int mcr_synth_factor_model(int n, int K, int T, double drift, double fvol, double ivol,
                                      uint64_t seed, double* mu, double* sigma, double* returns) {
  if (n < 1 || K < 0 || T < 2 || !mu || !sigma || !(fvol >= 0.0) || !(ivol >= 0.0))
    return MCR_E_ARG;
  std::vector<double> B((size_t)n * (K > 0 ? K : 1));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < K; ++k) B[(size_t)i * K + k] = 0.2 + 0.6 * uniform(seed, 1u, (uint64_t)i * K + k);
  std::vector<double> R((size_t)T * n);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < T; ++t) {
    std::vector<double> f((size_t)(K > 0 ? K : 1));
    for (int k = 0; k < K; ++k) f[k] = fvol * normal(seed, 2u, (uint64_t)t * K + k);
    for (int i = 0; i < n; ++i) {
      double r = drift;
      for (int k = 0; k < K; ++k) r += B[(size_t)i * K + k] * f[k];
      r += ivol * normal(seed, 3u, (uint64_t)t * n + i);
      R[(size_t)t * n + i] = r;
    }
  }
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int t = 0; t < T; ++t) s += R[(size_t)t * n + i];
    mu[i] = s / (double)T;
  }
  std::vector<double> C((size_t)T * n);
  for (int t = 0; t < T; ++t)
    for (int i = 0; i < n; ++i) C[(size_t)t * n + i] = R[(size_t)t * n + i] - mu[i];
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j <= i; ++j) {
      double s = 0.0;
      for (int t = 0; t < T; ++t) s += C[(size_t)t * n + i] * C[(size_t)t * n + j];
      s /= (double)(T - 1);
      sigma[(size_t)i * n + j] = s;
      sigma[(size_t)j * n + i] = s;
    }
  }
  if (returns)
    for (size_t e = 0; e < R.size(); ++e) returns[e] = R[e];
  return MCR_OK;
}
