"""Pins the oracle's path statistic, Cholesky and quantiles (CPU) against what
the mathematics fixes: exact fp32 brute force on tiny n, the one-factor
closed form of the reference prototype's example, marginal expectations, Genz
orthant probabilities, LAPACK Cholesky, the Gaussian loss quantile.  Two
planted mistakes (transposed factor, dropped mean) show the checks have power."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest
from scipy import stats

from helpers import (equicorrelated_sigma, equicorrelated_tail, factor_like, fma32, random_spd,
                     rank_of)

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _model_from(orc, mu, sigma, x, w=None):
    L = np.linalg.cholesky(sigma)
    n = len(mu)
    w = np.full(n, 1.0 / n) if w is None else np.asarray(w)
    return orc.F32Model.from_f64(L, mu, w, x), L, w


# ---------------------------------------------------------------- Cholesky ---
@pytest.mark.parametrize("n,cond", [(1, 1.0), (2, 10.0), (5, 100.0), (50, 1e4), (200, 1e6)])
def test_cholesky_matches_lapack(orc, n, cond):
    A = random_spd(n, seed=n, cond=cond)
    st, L, _ = orc.cholesky(A)
    assert st == 0
    ref = np.linalg.cholesky(A)
    assert np.allclose(L, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    assert np.linalg.norm(L @ L.T - A) <= 64 * n * np.finfo(float).eps * np.linalg.norm(A)
    assert np.all(np.triu(L, 1) == 0)


def test_cholesky_reports_failing_column(orc):
    A = np.diag([1.0, 2.0, -1.0, 4.0])
    st, _, _ = orc.cholesky(A)
    assert st == 3


def test_jitter_policy_rank_deficient(orc):
    # B B^T with rank 3 < n: plain factorisation fails, a small jitter succeeds
    rng = np.random.default_rng(4)
    B = rng.uniform(0.2, 0.8, (20, 3)) * 0.01
    A = B @ B.T
    att, L, delta, clamped = orc.factor(A)
    mean = np.trace(A) / 20
    assert 1 < att <= 9 and clamped == 0
    assert delta == pytest.approx(mean * 10.0 ** (att - 1 - 12), rel=1e-12)
    assert np.allclose(L @ L.T, A + delta * np.eye(20), atol=1e-12 * mean)


def test_jitter_policy_indefinite_falls_back_to_clamp(orc):
    A = np.diag([1.0, 1.0, -0.5, 1.0]) * 1e-4
    att, L, delta, clamped = orc.factor(A)
    assert att == 10 and clamped == 1
    assert np.all(np.isfinite(L)) and L[2, 2] == pytest.approx(math.sqrt(delta))


def test_jitter_policy_rejects_nan(orc):
    A = np.eye(3)
    A[1, 2] = np.nan
    att, _, _, _ = orc.factor(A)
    assert att == -1


def test_plain_spd_needs_no_jitter(orc):
    att, L, delta, _ = orc.factor(random_spd(30, 9, 1e3))
    assert att == 1 and delta == 0.0


# -------------------------------------------------------- exact fp32 paths ---
@pytest.mark.parametrize("n", [1, 2, 3])
def test_tiny_n_brute_force_exact(orc, n):
    """S, loss of every path equal an exact-rational fp32 evaluation of
    r_i = fma chain from mu_i over j = 0..i (lower factor only)."""
    rng = np.random.default_rng(n)
    L = np.tril(rng.uniform(-0.02, 0.03, (n, n))).astype(np.float32)
    L[np.diag_indices(n)] = np.abs(L[np.diag_indices(n)]) + 0.005
    U = rng.uniform(-0.05, 0.05, (n, n)).astype(np.float32)  # upper garbage must be ignored
    Lg = np.where(np.tri(n, dtype=bool), L, U).astype(np.float32)
    mu = rng.uniform(-0.01, 0.01, n).astype(np.float32)
    v = rng.uniform(-0.01, 0.01, n).astype(np.float32)
    mp = np.float32(0.0007)
    thr = np.float32(-0.02)
    model = orc.F32Model(Lg, mu, v, mp, thr)
    M = 300
    out = orc.simulate_f32(model, 31, 1000, M, 1, want_S=True, want_loss=True)
    Z = orc.normals_f32(31, 1000, M, n)
    for p in range(M):
        S = 0
        for i in range(n):
            acc = mu[i]
            for j in range(i + 1):
                acc = fma32(L[i, j], Z[p, j], acc)
            S += int(acc < thr)
        g = mp
        for j in range(n):
            g = fma32(v[j], Z[p, j], g)
        assert out["S"][p] == S
        assert np.float32(-g).view(np.uint32) == out["loss"][p].view(np.uint32)


def test_one_factor_closed_form(orc):
    cf = json.load(open(os.path.join(GOLD, "closed_forms.json")))
    for c in cf["equicorrelated"]:
        exact = equicorrelated_tail(c["n"], c["k"], c["x"], c["mu"], c["var"], c["rho"])
        assert exact == pytest.approx(c["p"], abs=2e-10)
        sigma = equicorrelated_sigma(c["n"], c["var"], c["rho"])
        model, _, _ = _model_from(orc, np.zeros(c["n"]), sigma, c["x"])
        M = 400000
        cnt = orc.simulate_f32(model, 67, 0, M, c["k"], nthreads=8)["count"]
        p = cnt / M
        se = math.sqrt(exact * (1 - exact) / M)
        assert abs(p - exact) < 4 * se, (p, exact, se)


def test_marginal_expectation_and_planted_mistakes(orc):
    """E[S] = sum_i Phi((-x - mu_i)/sqrt(Sigma_ii)).  A transposed factor or a
    dropped mean (mistakes planted in the oracle's input here, on the CPU) must
    fail the same check."""
    n, x = 12, 0.015
    mu, sigma = factor_like(n, 3, 5)
    mu = mu + np.linspace(0.0, 0.008, n)
    expect = np.sum(stats.norm.cdf((-x - mu) / np.sqrt(np.diag(sigma))))
    M = 200000
    s_vals = np.arange(n + 1)

    def moments(model):
        h = orc.simulate_f32(model, 3, 0, M, 1, want_hist=True, nthreads=8)["hist"]
        m = float(np.dot(s_vals, h)) / M
        return m, math.sqrt(float(np.dot((s_vals - m) ** 2, h)) / M)

    def mean_S(model):
        return moments(model)[0]

    good, L, w = _model_from(orc, mu, sigma, x)
    m_good, sd_good = moments(good)
    tol = 5 * sd_good / math.sqrt(M)  # five standard errors of the mean of S
    assert abs(m_good - expect) < tol
    transposed = orc.F32Model(np.ascontiguousarray(L.T, dtype=np.float32), good.mu, good.v,
                              good.mp, good.thr)
    assert abs(mean_S(transposed) - expect) > tol
    no_mean = orc.F32Model(good.L, np.zeros(n, np.float32), good.v, good.mp, good.thr)
    assert abs(mean_S(no_mean) - expect) > tol


def test_orthant_k1_against_genz(orc):
    n, x = 5, 0.012
    mu, sigma = factor_like(n, 2, 6)
    # P(S >= 1) = 1 - P(r_i >= -x for all i) = 1 - P(-r <= x)
    mvn = stats.multivariate_normal(mean=-mu, cov=sigma)
    exact = 1.0 - mvn.cdf(np.full(n, x))
    model, _, _ = _model_from(orc, mu, sigma, x)
    M = 300000
    p = orc.simulate_f32(model, 8, 0, M, 1, nthreads=8)["count"] / M
    se = math.sqrt(exact * (1 - exact) / M)
    assert abs(p - exact) < 4 * se + 2e-4, (p, exact)


def test_serial_equals_openmp(orc):
    mu, sigma = factor_like(20, 3, 7)
    model, _, _ = _model_from(orc, mu, sigma, 0.02)
    a = orc.simulate_f32(model, 1, 0, 50000, 5, want_S=True, want_loss=True, nthreads=1)
    b = orc.simulate_f32(model, 1, 0, 50000, 5, want_S=True, want_loss=True, nthreads=8)
    assert a["count"] == b["count"]
    assert np.array_equal(a["S"], b["S"]) and np.array_equal(a["loss"], b["loss"])


def test_path_ranges_compose(orc):
    mu, sigma = factor_like(9, 2, 8)
    model, _, _ = _model_from(orc, mu, sigma, 0.02)
    whole = orc.simulate_f32(model, 2, 0, 30000, 3, want_S=True)
    a = orc.simulate_f32(model, 2, 0, 12345, 3, want_S=True)
    b = orc.simulate_f32(model, 2, 12345, 30000 - 12345, 3, want_S=True)
    assert whole["count"] == a["count"] + b["count"]
    assert np.array_equal(whole["S"], np.concatenate([a["S"], b["S"]]))


def test_fp32_vs_fp64_same_stream(orc):
    n, x, k = 20, 0.02, 5
    mu, sigma = factor_like(n, 3, 10)
    model, L, w = _model_from(orc, mu, sigma, x)
    M = 100000
    a = orc.simulate_f32(model, 4, 0, M, k, want_S=True, nthreads=8)
    b = orc.simulate_f64(L, mu, w, x, 4, 0, M, k, want_S=True, nthreads=8)
    diff = np.nonzero(a["S"] != b["S"])[0]
    assert len(diff) <= 5, len(diff)
    assert abs(a["count"] - b["count"]) <= 5


def test_loss_quantile_against_gaussian(orc):
    n = 15
    mu, sigma = factor_like(n, 3, 11)
    model, L, w = _model_from(orc, mu, sigma, 0.02)
    M = 400000
    loss = orc.simulate_f32(model, 6, 0, M, 1, want_loss=True, nthreads=8)["loss"]
    sd = math.sqrt(w @ sigma @ w)
    for a in (0.99, 0.999):
        q = float(np.sort(loss)[rank_of(a, M) - 1])
        za = stats.norm.ppf(a)
        exact = -w @ mu + za * sd
        se = math.sqrt(a * (1 - a) / M) / stats.norm.pdf(za) * sd
        assert abs(q - exact) < 5 * se, (a, q, exact, se)


def test_stats_formula(orc):
    p, se, lo, hi = orc.stats(400, 10000)
    assert p == 0.04 and se == pytest.approx(math.sqrt(0.04 * 0.96 / 10000))
    assert lo == pytest.approx(0.04 - 1.96 * se) and hi == pytest.approx(0.04 + 1.96 * se)
    assert orc.stats(0, 10)[2] == 0.0 and orc.stats(10, 10)[3] == 1.0


def test_fixed_point_truncates(orc):
    assert orc.fixed32(0.5) == 2**31
    assert orc.fixed32(-0.5) == -(2**31)
    assert orc.fixed32(np.float32(1e-12)) == 0
    v = np.float32(0.0312)
    assert orc.fixed32(v) == int(float(v) * 2**32)
