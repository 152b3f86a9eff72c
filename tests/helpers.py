"""Test helpers: exact fp32 arithmetic, closed forms, model builders."""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
from scipy import integrate, stats


def f32_round(v: Fraction) -> np.float32:
    """Round an exact rational to the nearest float32, ties to even."""
    if v == 0:
        return np.float32(0.0)
    sign = -1 if v < 0 else 1
    a = abs(v)
    e = math.floor(math.log2(a.numerator) - math.log2(a.denominator))
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    e = max(e, -126)
    scale = Fraction(2) ** (23 - e)
    m = a * scale
    q, r = divmod(m.numerator, m.denominator)
    rem = Fraction(r, m.denominator)
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and q % 2 == 1):
        q += 1
    return np.float32(sign * float(Fraction(q) / scale))


def fma32(a, b, c) -> np.float32:
    """Correctly rounded fp32 fused multiply-add, from exact rationals."""
    return f32_round(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def equicorrelated_tail(n: int, k: int, x: float, mu: float, var: float, rho: float) -> float:
    """P(#{r_i < -x} >= k) for r = mu + sqrt(var)(sqrt(rho) f + sqrt(1-rho) e_i):
    one-factor integral of a binomial tail (textbook conditional independence)."""
    s = math.sqrt(var)

    def integrand(f):
        p = stats.norm.cdf((-x - mu - s * math.sqrt(rho) * f) / (s * math.sqrt(1 - rho)))
        return stats.norm.pdf(f) * stats.binom.sf(k - 1, n, p)

    val, _ = integrate.quad(integrand, -12, 12, limit=400, epsabs=1e-13)
    return val


def equicorrelated_sigma(n: int, var: float, rho: float) -> np.ndarray:
    # python_serial.py:119-120
    return np.full((n, n), rho * var) + np.eye(n) * var * (1 - rho)


def random_spd(n: int, seed: int, cond: float = 100.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.geomspace(1.0, 1.0 / cond, n) * 1e-4
    A = (Q * ev) @ Q.T
    return 0.5 * (A + A.T)


def factor_like(n: int, K: int, seed: int, T: int = 4000):
    """A factor-model style covariance from numpy (CPU tests)."""
    rng = np.random.default_rng(seed)
    B = rng.uniform(0.2, 0.8, (n, K))
    f = rng.standard_normal((T, K)) * 0.01
    e = rng.standard_normal((T, n)) * 0.01
    R = 2e-4 + f @ B.T + e
    return R.mean(axis=0), np.cov(R, rowvar=False, ddof=1)


def rank_of(alpha: float, M: int) -> int:
    """ceil(alpha M) on the exact binary value of alpha, clamped to [1, M]."""
    r = math.ceil(Fraction(alpha) * M)
    return min(max(r, 1), M)
