"""bf16x3 MFMA kernel against the FP64 oracle on the identical stream, and
statistical agreement on independent streams.

Tolerances (stated here, checked below):
  * exceedance count: |count_mfma - count_fp64| <= max(3, 1e-4 * count_fp64)
    (flips only for paths with some asset within ~1e-6 of the threshold), and
    on large samples strictly <= 1e-4 relative (the bar on the exceedance
    probability);
  * VaR at 0.99 / 0.999: relative error vs the FP64 order statistic <= 1e-4;
  * independent seeds: |P_a - P_b| <= 3 sqrt(SE_a^2 + SE_b^2)."""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from helpers import equicorrelated_sigma, equicorrelated_tail, factor_like, rank_of

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 5, 20, 31, 32, 50, 63, 64, 100, 127])
def test_mfma_vs_fp64_same_stream(mcr, orc, n):
    M, seed, x = 200_000, 77, 0.02
    k = max(1, n // 5)
    mu, sigma = factor_like(n, min(5, n), 40 + n)
    r = mcr.run(mu, sigma, x, k, M, seed, precision=mcr.PREC_BF16X3, alphas=(0.99, 0.999),
                want_S=True)
    assert r.kernel == "mfma"
    L, *_ = mcr.factor(sigma)
    w = np.full(n, 1.0 / n)
    ref = orc.simulate_f64(L, mu, w, x, seed, 0, M, k, want_S=True, want_loss=True, nthreads=8)
    assert abs(r.count - ref["count"]) <= max(3, 1e-4 * ref["count"]), (r.count, ref["count"])
    assert np.mean(r.S != ref["S"]) <= 2e-3
    srt = np.sort(ref["loss"])
    for i, a in enumerate((0.99, 0.999)):
        v64 = srt[rank_of(a, M) - 1]
        assert abs(r.var[i] - v64) <= 1e-4 * abs(v64), (a, r.var[i], v64)


@pytest.mark.parametrize("n,M", [(5, 100_000), (50, 100_000), (128, 40_000), (200, 20_000),
                                 (333, 10_000), (450, 8_000), (511, 6_000), (512, 6_000),
                                 (520, 6_000), (600, 6_000), (608, 5_000), (609, 5_000),
                                 (700, 4_000), (767, 3_000)])
def test_mfma_lds_vs_fp64_same_stream(mcr, orc, n, M):
    """Large-n MFMA kernel (normals staged in LDS), same tolerances.  The sizes
    cover every single-query workgroup shape (T = 1, 2, 5, 7, 11 row tiles
    with four waves; 15, 16 with eight multiplying waves of 2 tiles beside
    four generating waves; 17, 19, 20, 22, 24 with eight of 3 tiles), the boundary
    between the last two (n = 511: T = 16; n = 512: T = 17), an odd step count with
    two-step phases (n = 520: S = 33, the last phase half full), the batches'
    largest n (608) and the single queries' (767: 24 tiles, every multiplying
    wave holding three).  The batch shapes are covered by the batch tests."""
    seed, x = 91, 0.02
    k = max(1, n // 5)
    mu, sigma = factor_like(n, min(20, n), 70 + n, T=max(4000, 2 * n))
    r = mcr.run(mu, sigma, x, k, M, seed, precision=mcr.PREC_BF16X3,
                kernel=mcr.KERNEL_MFMA_LDS, alphas=(0.99,), want_S=True)
    assert r.kernel == "mfma_lds"
    L, *_ = mcr.factor(sigma)
    w = np.full(n, 1.0 / n)
    ref = orc.simulate_f64(L, mu, w, x, seed, 0, M, k, want_S=True, want_loss=True, nthreads=8)
    assert abs(r.count - ref["count"]) <= max(3, 1e-4 * ref["count"]), (r.count, ref["count"])
    assert np.mean(r.S != ref["S"]) <= 2e-3
    v64 = np.sort(ref["loss"])[rank_of(0.99, M) - 1]
    assert abs(r.var[0] - v64) <= 1e-4 * abs(v64), (r.var[0], v64)


@pytest.mark.parametrize("n,M,kernel", [(50, 300_000, 3), (50, 200_000, 4), (600, 50_000, 4),
                                        (767, 20_000, 4)])
def test_mfma_s_histogram_matches_per_path_counts(mcr, n, M, kernel):
    """The in-kernel (n+1)-bin histogram of S equals the histogram of the
    per-path S the same launch stores, and its tail above k is the count."""
    mu, sigma = factor_like(n, min(20, n), 500 + n, T=max(4000, 2 * n))
    k = 10
    r = mcr.run(mu, sigma, 0.02, k, M, 17, precision=mcr.PREC_BF16X3, kernel=kernel,
                want_S=True, want_hist=True)
    assert r.s_hist.shape == (n + 1,)
    assert np.array_equal(r.s_hist, np.bincount(r.S, minlength=n + 1))
    assert int(r.s_hist[k:].sum()) == r.count
    assert int(r.s_hist.sum()) == M


def test_mfma_kernels_agree(mcr):
    """Small-n MFMA (register operands) and LDS MFMA kernels: same products,
    summation order may differ in the last fp32 bit, so counts agree to a few
    paths and VaR to 1e-6."""
    mu, sigma = factor_like(50, 5, 1002)
    a = mcr.run(mu, sigma, 0.02, 10, 2_000_000, 3, precision=mcr.PREC_BF16X3,
                kernel=mcr.KERNEL_MFMA, alphas=(0.99,))
    b = mcr.run(mu, sigma, 0.02, 10, 2_000_000, 3, precision=mcr.PREC_BF16X3,
                kernel=mcr.KERNEL_MFMA_LDS, alphas=(0.99,))
    assert abs(a.count - b.count) <= 5
    assert abs(a.var[0] - b.var[0]) <= 1e-6 * abs(a.var[0])


def test_mfma_independent_streams_agree_with_fp32(mcr):
    mu, sigma = factor_like(50, 5, 1002)
    a = mcr.run(mu, sigma, 0.02, 10, 20_000_000, 1, precision=mcr.PREC_BF16X3)
    b = mcr.run(mu, sigma, 0.02, 10, 20_000_000, 2, precision=mcr.PREC_FP32)
    assert abs(a.p_hat - b.p_hat) <= 3 * math.sqrt(a.std_error**2 + b.std_error**2)


@pytest.mark.parametrize("precision", [0, 1])
def test_closed_form_on_gpu(mcr, precision):
    """Reference prototype's example (python_serial.py:112-122) at 1e7 paths."""
    n, k, x, var, rho = 10, 5, 0.2, 0.04, 0.3
    exact = equicorrelated_tail(n, k, x, 0.0, var, rho)
    r = mcr.run(np.zeros(n), equicorrelated_sigma(n, var, rho), x, k, 10_000_000, 67,
                precision=precision)
    se = math.sqrt(exact * (1 - exact) / 10_000_000)
    assert abs(r.p_hat - exact) < 4 * se, (r.p_hat, exact)


@pytest.mark.parametrize("n,M,kernel", [(50, 4_000_000, 3), (600, 200_000, 4)])
def test_mfma_relative_error_bar(mcr, orc, n, M, kernel):
    """Exceedance probability of the bf16x3 kernels within 1e-4 relative of
    the FP64 computation on the identical stream, at the bench's sizes of n."""
    seed, x = 20251202, 0.02
    K, T = (5, 2000) if n == 50 else (20, 2500)
    mu, sigma = mcr.synth_factor_model(n, K, T, 1002 if n == 50 else 1003)
    r = mcr.run(mu, sigma, x, 10, M, seed, precision=mcr.PREC_BF16X3, kernel=kernel,
                alphas=(0.99,))
    L, *_ = mcr.factor(sigma)
    ref = orc.simulate_f64(L, mu, np.full(n, 1.0 / n), x, seed, 0, M, 10, want_loss=True,
                           nthreads=16)
    rel = abs(r.count - ref["count"]) / ref["count"]
    print(f"n={n} M={M}: count {r.count} vs FP64 {ref['count']} (rel {rel:.2e})")
    assert rel <= 1e-4, (r.count, ref["count"])
    v64 = np.sort(ref["loss"])[rank_of(0.99, M) - 1]
    assert abs(r.var[0] - v64) <= 1e-4 * abs(v64)


# MCR_MFMA_RANDOM_CASES raises the number of randomized cases for a longer run
@pytest.mark.parametrize("case", range(int(os.environ.get("MCR_MFMA_RANDOM_CASES", "4"))))
def test_randomized_bf16x3_queries_within_bar_of_fp64(mcr, orc, case):
    """Random bf16x3 queries (n = 1-767 on the register-operand or LDS MFMA
    kernel, random path counts and offsets, seeds, thresholds, k, weights,
    VaR levels) against the FP64 oracle on the identical stream, with the
    tolerances above."""
    rng = np.random.default_rng(5151 + case)
    n = int(rng.choice([int(rng.integers(1, 128)), int(rng.integers(128, 768))]))
    M = int(np.exp(rng.uniform(np.log(50), np.log(1e7 / max(n, 64)))))
    seed = int(rng.integers(0, 2**63))
    path0 = int(rng.choice([0, int(rng.integers(0, 2**40))]))
    x = float(rng.uniform(0.0, 0.04))
    k = int(rng.integers(1, n + 1))
    alphas = tuple(sorted(rng.choice([0.9, 0.95, 0.99, 0.999], int(rng.integers(1, 3)),
                                     replace=False)))
    mu, sigma = factor_like(n, int(rng.integers(1, min(20, n) + 1)), 950 + case, T=max(1500, 2 * n))
    w = rng.dirichlet(np.ones(n)) if rng.uniform() < 0.5 else np.full(n, 1.0 / n)
    r = mcr.run(mu, sigma, x, k, M, seed, weights=w, precision=mcr.PREC_BF16X3, alphas=alphas,
                path_offset=path0)
    assert r.kernel == ("mfma" if n <= 127 else "mfma_lds")
    L, *_ = mcr.factor(sigma)
    ref = orc.simulate_f64(L, mu, w, x, seed, path0, M, k, want_loss=True, nthreads=8)
    tag = (n, M, k, path0)
    assert abs(r.count - ref["count"]) <= max(3, 1e-4 * ref["count"]), (tag, r.count, ref["count"])
    srt = np.sort(ref["loss"])
    for i, a in enumerate(alphas):
        v64 = srt[rank_of(a, M) - 1]
        assert abs(r.var[i] - v64) <= 1e-4 * abs(v64), (tag, a, r.var[i], v64)


def test_kernel_choice_between_608_and_767_assets(mcr):
    """609-767 assets: bf16x3 (and the automatic choice) run the LDS MFMA
    kernel, fp32 (and the automatic choice when the near count is asked for)
    the column-chunk kernel; from 768 every precision runs the chunk kernel and
    naming the LDS MFMA kernel is refused."""
    mu, sigma = factor_like(700, 10, 7000, T=2000)
    b = mcr.run(mu, sigma, 0.02, 70, 5000, 3, precision=mcr.PREC_BF16X3)
    a = mcr.run(mu, sigma, 0.02, 70, 5000, 3, precision=mcr.PREC_AUTO)
    f = mcr.run(mu, sigma, 0.02, 70, 5000, 3, precision=mcr.PREC_FP32)
    nr = mcr.run(mu, sigma, 0.02, 70, 5000, 3, precision=mcr.PREC_AUTO, near=True)
    assert (b.kernel, a.kernel, f.kernel, nr.kernel) == ("mfma_lds", "mfma_lds", "valu_chunk",
                                                          "valu_chunk")
    assert b.count == a.count and f.count == nr.count
    mu, sigma = factor_like(768, 10, 7001, T=2000)
    assert mcr.run(mu, sigma, 0.02, 70, 1000, 3, precision=mcr.PREC_BF16X3).kernel == "valu_chunk"
    with pytest.raises(mcr.McrError) as e:
        mcr.run(mu, sigma, 0.02, 70, 1000, 3, kernel=mcr.KERNEL_MFMA_LDS)
    assert e.value.code == mcr.MCR_E_UNSUPPORTED
