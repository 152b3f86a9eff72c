"""Derives the fp32 polynomial coefficients of the normal stream (mcr_rng.h).

Method: least squares on Chebyshev nodes, reweighted towards the minimax
solution (Lawson's algorithm), then rounded to fp32.  Prints the coefficients
as C hex floats and their maximum error after rounding.

  ln(1+f) = f - f^2/2 + f^3 P(f),  f in [sqrt(1/2)-1, sqrt(2)-1], P degree 6
  sin(2 pi a) = a S(a^2),          a in [-1/8, 1/8] turns,       S degree 3
  cos(2 pi a) = 1 + a^2 C(a^2),                                   C degree 3

Run: python derive_normal_poly.py
"""
import numpy as np


def cheb_nodes(a, b, m):
    k = np.arange(m)
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos((2 * k + 1) * np.pi / (2 * m))


def lawson(V, y, w, iters=50):
    """Weighted least squares with Lawson reweighting towards minimax."""
    u = np.ones_like(y)
    c = None
    for _ in range(iters):
        s = np.sqrt(u) * w
        c = np.linalg.lstsq(V * s[:, None], y * s, rcond=None)[0]
        e = np.abs((V @ c - y) * w)
        u *= e / e.max() + 1e-12
        u /= u.sum()
    return c.astype(np.float32)


def main():
    lo, hi = np.sqrt(0.5) - 1, np.sqrt(2.0) - 1
    f = cheb_nodes(lo, hi, 6000)
    tgt = np.log1p(f)
    log_c = lawson(np.vander(f, 7, increasing=True), (tgt - f + f * f / 2) / f**3,
                   np.abs(f**3 / tgt))
    a = cheb_nodes(0.0, 0.125, 6000)
    t = a * a
    sin_c = lawson(np.vander(t, 4, increasing=True), np.sin(2 * np.pi * a) / a, np.ones_like(t))
    cos_c = lawson(np.vander(t, 5, increasing=True)[:, 1:], np.cos(2 * np.pi * a) - 1,
                   np.ones_like(t))
    for name, cs in (("kLogP", log_c), ("kSin", sin_c), ("kCos", cos_c)):
        base = 1 if name == "kCos" else 0
        for i, v in enumerate(cs):
            print(f"constexpr float {name}{i + base} = {float(v).hex()}f;")
    ff = np.linspace(lo, hi, 400001)
    ff = ff[ff != 0]
    p = ff - ff * ff / 2 + ff**3 * np.polyval(log_c.astype(np.float64)[::-1], ff)
    print("log max rel err", np.max(np.abs(p - np.log1p(ff)) / np.abs(np.log1p(ff))))
    aa = np.linspace(1e-9, 0.125, 400001)
    s = aa * np.polyval(sin_c.astype(np.float64)[::-1], aa * aa)
    print("sin max rel err", np.max(np.abs(s - np.sin(2 * np.pi * aa)) / np.sin(2 * np.pi * aa)))
    c = 1 + aa * aa * np.polyval(cos_c.astype(np.float64)[::-1], aa * aa)
    print("cos max abs err", np.max(np.abs(c - np.cos(2 * np.pi * aa))))


if __name__ == "__main__":
    main()
