"""GPU portfolio-loss quantiles and tail sums against the oracle's losses.

Bar (fp32 kernels): VaR_a is bit-identical to the ceil(a M)-th smallest of the
oracle's per-path losses; the tail sum (fixed point 2^-32, truncating) and ES
are identical integers / the same double.  The refinement passes (forced with
zero-width windows) must give the same answers."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import factor_like, rank_of

pytestmark = pytest.mark.gpu


def oracle_quantiles(orc, mcr, mu, sigma, x, w, seed, M, alphas, path_offset=0):
    m = mcr.prepare(mu, sigma, x, w)
    model = orc.F32Model(m["L32"], m["mu32"], m["v32"], m["mp32"], m["thr32"])
    loss = orc.simulate_f32(model, seed, path_offset, M, 1, want_loss=True, nthreads=8)["loss"]
    srt = np.sort(loss)
    out = []
    for a in alphas:
        R = rank_of(a, M)
        tail = srt[R - 1:]
        fixed = int(np.sum((tail.astype(np.float64) * 2.0**32).astype(np.int64)))
        out.append((float(srt[R - 1]), fixed, fixed * 2.0**-32 / (M - R + 1)))
    return out


@pytest.mark.parametrize("kernel", [1, 2, 5])
@pytest.mark.parametrize("weights", ["equal", "dirichlet"])
def test_var_exact(mcr, orc, kernel, weights):
    n, M, seed = 24, 200_000, 99
    mu, sigma = factor_like(n, 3, 21)
    w = None if weights == "equal" else np.random.default_rng(5).dirichlet(np.ones(n))
    alphas = (0.95, 0.99, 0.999)
    r = mcr.run(mu, sigma, 0.02, 4, M, seed, weights=w, kernel=kernel, alphas=alphas)
    ref = oracle_quantiles(orc, mcr, mu, sigma, 0.02, w, seed, M, alphas)
    for i, (var, fixed, es) in enumerate(ref):
        assert r.var[i] == var
        assert r.tail_sum_fixed[i] == fixed
        assert r.es[i] == es
    assert r.var_passes == 1


@pytest.mark.parametrize("kernel", [1, 2, 5])
def test_var_refinement_path(mcr, orc, kernel):
    n, M, seed = 10, 50_000, 3
    mu, sigma = factor_like(n, 2, 22)
    alphas = (0.9, 0.99)
    r = mcr.run(mu, sigma, 0.02, 2, M, seed, kernel=kernel, alphas=alphas,
                extra_flags=mcr.F_TEST_NARROW)
    ref = oracle_quantiles(orc, mcr, mu, sigma, 0.02, None, seed, M, alphas)
    assert r.var_passes > 1
    for i, (var, fixed, es) in enumerate(ref):
        assert r.var[i] == var and r.tail_sum_fixed[i] == fixed and r.es[i] == es


@pytest.mark.parametrize("kernel", [1, 2])
def test_var_windows_holding_every_loss(mcr, orc, kernel):
    """Windows wide enough to hold every loss (every path appends its loss to
    the gather buffer): the order statistics, tail sums and ES stay exact, in
    one pass."""
    n, M, seed = 24, 200_000, 7
    mu, sigma = factor_like(n, 3, 26)
    alphas = (0.5, 0.99)
    r = mcr.run(mu, sigma, 0.02, 4, M, seed, kernel=kernel, alphas=alphas,
                extra_flags=mcr.F_TEST_WIDE)
    ref = oracle_quantiles(orc, mcr, mu, sigma, 0.02, None, seed, M, alphas)
    assert r.var_passes == 1
    for i, (var, fixed, es) in enumerate(ref):
        assert r.var[i] == var and r.tail_sum_fixed[i] == fixed and r.es[i] == es


@pytest.mark.parametrize("n", [24, 100])
def test_mfma_var_independent_of_window_width(mcr, n):
    """The register-operand MFMA kernel gives the same order statistics, tail
    sums and ES whether a few thousand losses or all of them fall in the
    windows."""
    M, seed = 200_000, 8
    mu, sigma = factor_like(n, 3, 27)
    alphas = (0.5, 0.99, 0.999)
    a = mcr.run(mu, sigma, 0.02, 4, M, seed, precision=mcr.PREC_BF16X3, alphas=alphas)
    b = mcr.run(mu, sigma, 0.02, 4, M, seed, precision=mcr.PREC_BF16X3, alphas=alphas,
                extra_flags=mcr.F_TEST_WIDE)
    assert a.kernel == b.kernel == "mfma" and b.var_passes == 1
    assert list(a.var) == list(b.var)
    assert list(a.tail_sum_fixed) == list(b.tail_sum_fixed)
    assert list(a.es) == list(b.es)


@pytest.mark.parametrize("refine", [False, True])
def test_var_exact_past_608_assets(mcr, orc, refine):
    """n = 1000 (two 512-column chunks of the column-chunk kernel, chosen
    automatically at every precision): VaR, tail sums and ES equal the
    oracle's, in one pass and through the refinement passes."""
    n, M, seed = 1000, 6000, 17
    mu, sigma = factor_like(n, 12, 23, T=2500)
    w = np.random.default_rng(6).dirichlet(np.ones(n))
    alphas = (0.9, 0.99)
    r = mcr.run(mu, sigma, 0.02, 100, M, seed, weights=w, precision=mcr.PREC_BF16X3,
                alphas=alphas, extra_flags=mcr.F_TEST_NARROW if refine else 0)
    assert r.kernel == "valu_chunk" and (r.var_passes > 1) == refine
    ref = oracle_quantiles(orc, mcr, mu, sigma, 0.02, w, seed, M, alphas)
    for i, (var, fixed, es) in enumerate(ref):
        assert r.var[i] == var and r.tail_sum_fixed[i] == fixed and r.es[i] == es


@pytest.mark.parametrize("M", [1, 2, 10, 100])
def test_var_tiny_path_counts(mcr, orc, M):
    mu, sigma = factor_like(6, 2, 23)
    alphas = (0.5, 0.99)
    r = mcr.run(mu, sigma, 0.02, 2, M, 11, alphas=alphas)
    ref = oracle_quantiles(orc, mcr, mu, sigma, 0.02, None, 11, M, alphas)
    for i, (var, fixed, es) in enumerate(ref):
        assert r.var[i] == var and r.tail_sum_fixed[i] == fixed


def test_var_with_path_offset(mcr, orc):
    mu, sigma = factor_like(16, 3, 24)
    r = mcr.run(mu, sigma, 0.02, 3, 30_000, 5, alphas=(0.99,), path_offset=2**40)
    ref = oracle_quantiles(orc, mcr, mu, sigma, 0.02, None, 5, 30_000, (0.99,), path_offset=2**40)
    assert r.var[0] == ref[0][0] and r.tail_sum_fixed[0] == ref[0][1]


def test_loss_moments_reported(mcr):
    mu, sigma = factor_like(16, 3, 25)
    w = np.full(16, 1 / 16)
    r = mcr.run(mu, sigma, 0.02, 3, 1000, 5, alphas=(0.99,))
    assert r.loss_mean == pytest.approx(-(w @ mu), rel=1e-12)
    assert r.loss_sd == pytest.approx(np.sqrt(w @ sigma @ w), rel=1e-9)
