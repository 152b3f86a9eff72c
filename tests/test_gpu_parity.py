"""GPU parity (through the C ABI): the device stream, the FP64 factor and the
fused fp32 path kernels against the CPU oracle.

Bar: bit-exact for the normals, per-path crash counts S, per-path losses,
the exceedance count, the near-threshold count and the S histogram (the fp32
kernels mirror the oracle's fixed fma order); FP64 factor within 1e-10
relative of LAPACK."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from helpers import factor_like, random_spd

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def device_model(mcr, orc, mu, sigma, x, w=None):
    m = mcr.prepare(mu, sigma, x, w)
    return orc.F32Model(m["L32"], m["mu32"], m["v32"], m["mp32"], m["thr32"]), m


# ------------------------------------------------------------------ stream ---
@pytest.mark.parametrize("seed,path0,n", [(0, 0, 13), (67, 2**32 - 5, 8), (2**64 - 1, 2**63, 5),
                                          (0x123456789ABCDEF0, 12345, 64), (1, 7, 1)])
def test_device_stream_equals_oracle(mcr, orc, seed, path0, n):
    z = mcr.normals(seed, path0, 8192, n)
    ref = orc.normals_f32(seed, path0, 8192, n)
    assert np.array_equal(z.view(np.uint32), ref.view(np.uint32))


def test_device_stream_golden(mcr):
    g = json.load(open(os.path.join(GOLD, "normals_golden.json")))
    for c in g["cases"]:
        z = mcr.normals(int(c["seed"], 16), int(c["path"], 16), 1, c["n"])[0]
        assert [f"0x{int(b):08x}" for b in z.view(np.uint32)] == c["bits"]


def test_device_stream_one_million_paths(mcr, orc):
    z = mcr.normals(42, 1000, 1 << 20, 4)
    ref = orc.normals_f32(42, 1000, 1 << 20, 4)
    assert np.array_equal(z.view(np.uint32), ref.view(np.uint32))


# ------------------------------------------------------------------ factor ---
@pytest.mark.parametrize("n,cond", [(1, 1.0), (2, 10.0), (7, 1e2), (50, 1e4), (200, 1e6),
                                    (600, 1e4)])
def test_factor_matches_lapack(mcr, n, cond):
    A = random_spd(n, seed=100 + n, cond=cond)
    L, att, jit, cl = mcr.factor(A)
    assert att == 1 and jit == 0.0 and cl == 0
    ref = np.linalg.cholesky(A)
    assert np.all(np.triu(L, 1) == 0)
    assert np.max(np.abs(L - ref)) <= 1e-10 * np.max(np.abs(ref)) * max(1.0, cond**0.5)
    assert np.linalg.norm(L @ L.T - A) <= 64 * n * np.finfo(float).eps * np.linalg.norm(A)


def test_factor_jitter_policy_matches_oracle(mcr, orc):
    rng = np.random.default_rng(4)
    B = rng.uniform(0.2, 0.8, (20, 3)) * 0.01
    A = B @ B.T
    L, att, jit, cl = mcr.factor(A)
    oatt, oL, ojit, ocl = orc.factor(A)
    assert (att, cl) == (oatt, ocl) and att > 1
    # same policy step; the diagonal mean is summed in another order
    assert jit == pytest.approx(ojit, rel=1e-13)
    assert np.allclose(L, oL, rtol=1e-8, atol=1e-12)


def test_factor_clamp_fallback(mcr, orc):
    A = np.diag([1.0, 1.0, -0.5, 1.0]) * 1e-4
    L, att, jit, cl = mcr.factor(A)
    oatt, oL, ojit, ocl = orc.factor(A)
    assert att == oatt == 10 and cl == ocl == 1
    assert jit == pytest.approx(ojit, rel=1e-13)
    assert np.allclose(L, oL)


def test_factor_rejects_nonfinite(mcr):
    A = np.eye(4)
    A[2, 1] = np.inf
    with pytest.raises(mcr.McrError) as e:
        mcr.factor(A)
    assert e.value.code == mcr.MCR_E_FACTOR


def test_device_model_is_rounded_factor(mcr, orc):
    mu, sigma = factor_like(30, 3, 12)
    w = np.random.default_rng(1).dirichlet(np.ones(30))
    model, m = device_model(mcr, orc, mu, sigma, 0.02, w)
    L64 = m["L64"]
    assert np.array_equal(m["L32"], np.tril(L64).astype(np.float32))
    assert np.array_equal(m["mu32"], mu.astype(np.float32))
    v = L64.T @ w
    assert np.all(np.abs(m["v32"].astype(np.float64) - v) <= 1.5 * np.spacing(np.abs(v).astype(np.float32)))
    assert m["thr32"] == np.float32(-0.02)
    assert m["loss_sd"] == pytest.approx(np.sqrt(w @ sigma @ w), rel=1e-9)


# ------------------------------------------------------------ exact paths ---
def _exact_case(mcr, orc, n, M, k, kernel, seed=2024, path_offset=0, x=0.02, K=3, msd=31):
    mu, sigma = factor_like(n, min(K, n), msd + n)
    model, _ = device_model(mcr, orc, mu, sigma, x)
    r = mcr.run(mu, sigma, x, k, M, seed, kernel=kernel, want_S=True, want_loss=True, near=True,
                want_hist=True, path_offset=path_offset)
    ref = orc.simulate_f32(model, seed, path_offset, M, k, want_S=True, want_loss=True,
                           want_hist=True, nthreads=8)
    bad = np.nonzero(r.S != ref["S"])[0]
    assert bad.size == 0, f"{bad.size} paths differ, first {bad[:5]}"
    assert np.array_equal(r.loss.view(np.uint32), ref["loss"].view(np.uint32))
    assert r.count == ref["count"]
    assert r.near_threshold == ref["near"]
    assert np.array_equal(r.s_hist, ref["hist"])
    return r


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 13, 20, 31, 32, 33, 50, 63, 64])
def test_register_kernel_exact(mcr, orc, n):
    r = _exact_case(mcr, orc, n, 20000, max(1, n // 5), mcr.KERNEL_VALU_REG)
    assert r.kernel == "valu_reg"


@pytest.mark.parametrize("n", [1, 5, 8, 9, 50, 64, 65, 100, 129, 257, 600, 608])
def test_tile_kernel_exact(mcr, orc, n):
    M = 20000 if n <= 129 else 4000
    r = _exact_case(mcr, orc, n, M, max(1, n // 5), mcr.KERNEL_VALU_TILE)
    assert r.kernel == "valu_tile"


@pytest.mark.parametrize("n", [1, 5, 9, 50, 129, 257, 600, 608, 609, 700, 1000, 1024])
def test_chunk_kernel_exact(mcr, orc, n):
    """The column-chunk kernel (any n <= 1024; fp32 past 608): per-path
    S and loss bits, count, near-threshold count and S histogram equal the
    oracle's.  n = 609-1024 covers one and two 512-column chunks, a last
    chunk of one 4-column block (n = 1024 - 3 .. 1024 ends on block 255) and
    n not a multiple of 4 or 8."""
    M = 20000 if n <= 129 else (4000 if n <= 608 else 1500)
    T = max(4000, 2 * n)
    mu, sigma = factor_like(n, min(12, n), 31 + n, T=T)
    model, _ = device_model(mcr, orc, mu, sigma, 0.02)
    k = max(1, n // 5)
    r = mcr.run(mu, sigma, 0.02, k, M, 2024, kernel=mcr.KERNEL_VALU_CHUNK, want_S=True,
                want_loss=True, near=True, want_hist=True)
    ref = orc.simulate_f32(model, 2024, 0, M, k, want_S=True, want_loss=True, want_hist=True,
                           nthreads=8)
    assert r.kernel == "valu_chunk"
    bad = np.nonzero(r.S != ref["S"])[0]
    assert bad.size == 0, f"{bad.size} paths differ, first {bad[:5]}"
    assert np.array_equal(r.loss.view(np.uint32), ref["loss"].view(np.uint32))
    assert r.count == ref["count"] and r.near_threshold == ref["near"]
    assert np.array_equal(r.s_hist, ref["hist"])


def test_chunk_and_tile_kernels_agree_at_scale(mcr):
    """Size-independent property at the n = 600 workload's shape (1e6
    paths): the column-chunk and tile kernels compute identical per-path
    results, so their counts, tails and quantiles are identical."""
    mu, sigma = factor_like(600, 20, 6000, T=4000)
    a = mcr.run(mu, sigma, 0.02, 597, 1_000_000, 5, kernel=mcr.KERNEL_VALU_TILE, alphas=(0.99,))
    b = mcr.run(mu, sigma, 0.02, 597, 1_000_000, 5, kernel=mcr.KERNEL_VALU_CHUNK, alphas=(0.99,))
    assert a.count == b.count and a.var == b.var and a.tail_sum_fixed == b.tail_sum_fixed


def test_past_767_assets_every_precision_runs_the_chunk_kernel(mcr):
    """Above 767 assets (the LDS MFMA kernel's 24 row tiles) every precision
    (and the automatic choice) runs the column-chunk kernel and gives the same
    result; naming the tile or an MFMA kernel there is refused, as is
    n = 1025."""
    mu, sigma = factor_like(800, 10, 8000, T=2000)
    rs = [mcr.run(mu, sigma, 0.02, 80, 5000, 3, precision=p, alphas=(0.99,))
          for p in (mcr.PREC_FP32, mcr.PREC_BF16X3, mcr.PREC_AUTO)]
    assert all(r.kernel == "valu_chunk" for r in rs)
    assert len({(r.count, r.var[0]) for r in rs}) == 1
    for kernel in (mcr.KERNEL_VALU_TILE, mcr.KERNEL_MFMA_LDS, mcr.KERNEL_MFMA):
        with pytest.raises(mcr.McrError) as e:
            mcr.run(mu, sigma, 0.02, 80, 100, 3, kernel=kernel)
        assert e.value.code == mcr.MCR_E_UNSUPPORTED
    with pytest.raises(mcr.McrError) as e:
        mcr.run(np.zeros(1025), np.eye(1025) * 1e-4, 0.02, 1, 10, 1)
    assert e.value.code == mcr.MCR_E_UNSUPPORTED


# MCR_PARITY_RANDOM_CASES raises the number of randomized cases for a longer run
@pytest.mark.parametrize("case", range(int(os.environ.get("MCR_PARITY_RANDOM_CASES", "4"))))
def test_randomized_fp32_queries_match_oracle(mcr, orc, case):
    """Random single queries on the fp32 kernels (every kernel valid for the
    drawn n in 1-1024: register, tile, column chunk), path counts, path
    offsets, seeds, thresholds, k and weights: per-path S and loss bits,
    count, near-threshold count and S histogram equal the oracle's."""
    rng = np.random.default_rng(4242 + case)
    n = int(rng.choice([int(rng.integers(1, 65)), int(rng.integers(65, 609)),
                        int(rng.integers(609, 1025))]))
    kernels = ([mcr.KERNEL_VALU_REG] if n <= 64 else []) + \
        ([mcr.KERNEL_VALU_TILE] if n <= 608 else []) + [mcr.KERNEL_VALU_CHUNK]
    kernel = int(rng.choice(kernels))
    M = int(np.exp(rng.uniform(0.0, np.log(4e6 / max(n, 40)))))
    seed = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
    path0 = int(rng.choice([0, int(rng.integers(0, 2**40)), 2**32 - M // 2]))
    x = float(rng.uniform(-0.005, 0.04))
    k = int(rng.integers(0, n + 2))
    mu, sigma = factor_like(n, int(rng.integers(1, min(20, n) + 1)), 900 + case, T=max(1500, 2 * n))
    w = None if rng.uniform() < 0.5 else rng.dirichlet(np.ones(n))
    model, _ = device_model(mcr, orc, mu, sigma, x, w)
    r = mcr.run(mu, sigma, x, k, M, seed, weights=w, kernel=kernel, want_S=True, want_loss=True,
                near=True, want_hist=True, path_offset=path0)
    ref = orc.simulate_f32(model, seed, path0, M, k, want_S=True, want_loss=True, want_hist=True,
                           nthreads=8)
    tag = (n, kernel, M, path0, k)
    assert np.array_equal(r.S, ref["S"]), tag
    assert np.array_equal(r.loss.view(np.uint32), ref["loss"].view(np.uint32)), tag
    assert (r.count, r.near_threshold) == (ref["count"], ref["near"]), tag
    assert np.array_equal(r.s_hist, ref["hist"]), tag


@pytest.mark.parametrize("M", [1, 63, 64, 65, 257, 1000])
@pytest.mark.parametrize("kernel", [1, 2, 5])
def test_ragged_path_counts(mcr, orc, M, kernel):
    _exact_case(mcr, orc, 20, M, 5, kernel)


@pytest.mark.parametrize("kernel", [1, 2, 5])
def test_path_ids_cross_32_bits_and_large_seed(mcr, orc, kernel):
    _exact_case(mcr, orc, 11, 5000, 3, kernel, seed=2**64 - 1, path_offset=2**32 - 2500)


@pytest.mark.parametrize("k", [0, 1, 20, 21])
def test_k_edges(mcr, orc, k):
    r = _exact_case(mcr, orc, 20, 3000, k, mcr.KERNEL_VALU_REG)
    if k == 0:
        assert r.count == 3000
    if k == 21:
        assert r.count == 0


def test_threshold_edges(mcr, orc):
    # x huge: nobody crashes; x negative enough: everybody crashes on every asset
    mu, sigma = factor_like(10, 2, 3)
    r = mcr.run(mu, sigma, 10.0, 1, 1000, 1)
    assert r.count == 0
    r = mcr.run(mu, sigma, -10.0, 10, 1000, 1)
    assert r.count == 1000
    _exact_case(mcr, orc, 10, 5000, 2, mcr.KERNEL_VALU_REG, x=0.0)


def test_register_and_tile_agree_at_scale(mcr):
    """Size-independent property at the fused-kernel workload's size (n = 50,
    1e8 paths): the two fp32 kernels compute identical per-path results, so
    their counts, tails and quantiles are identical."""
    mu, sigma = factor_like(50, 5, 1002)
    a = mcr.run(mu, sigma, 0.02, 10, 100_000_000, 7, kernel=mcr.KERNEL_VALU_REG, alphas=(0.99,))
    b = mcr.run(mu, sigma, 0.02, 10, 100_000_000, 7, kernel=mcr.KERNEL_VALU_TILE, alphas=(0.99,))
    assert a.count == b.count
    assert a.var == b.var and a.tail_sum_fixed == b.tail_sum_fixed


def test_subrange_of_large_run_matches_oracle(mcr, orc):
    """Path ids address the stream directly: a 20000-path window at offset
    99,980,000 of the n = 50 workload equals the oracle's same window."""
    _exact_case(mcr, orc, 50, 20000, 10, mcr.KERNEL_VALU_REG, seed=7, path_offset=99_980_000, K=5,
                msd=952)
