"""The allocation-free entry point (mcr_run_noalloc) against mcr_run.

Both run the same kernels on the same stream; mcr_run_noalloc only carves the
query's device buffers from the library's static scratch array and uses static
host buffers.  Bar: bit-identical results (counts, near counts, VaR, ES, tail
sums, per-path S and losses, S histograms), including on a second call with
a smaller query after a larger one has filled the scratch, with per-path
outputs staged over several chunks, and through the quantile refinement
passes."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import factor_like

pytestmark = pytest.mark.gpu


def same(a, b):
    assert a.kernel == b.kernel
    assert (a.count, a.near_threshold) == (b.count, b.near_threshold)
    assert (a.p_hat, a.std_error, a.ci_lower, a.ci_upper) == (b.p_hat, b.std_error, b.ci_lower,
                                                              b.ci_upper)
    assert a.var == b.var and a.es == b.es and a.tail_sum_fixed == b.tail_sum_fixed
    assert a.var_passes == b.var_passes
    assert (a.jitter, a.factor_attempts, a.loss_mean, a.loss_sd) == (
        b.jitter, b.factor_attempts, b.loss_mean, b.loss_sd)
    for x, y in ((a.S, b.S), (a.loss, b.loss), (a.s_hist, b.s_hist)):
        assert (x is None) == (y is None)
        if x is not None:
            assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


def both(mcr, *args, **kw):
    a = mcr.run_noalloc(*args, **kw)
    b = mcr.run(*args, **kw)
    same(a, b)
    return a


def test_noalloc_equals_run_large_then_smaller(mcr):
    mcr.init(0)
    # 1: the largest query (n = 608, four VaR levels, bf16x3 LDS MFMA kernel)
    mu, sigma = factor_like(608, 20, 6080, T=4000)
    w = np.random.default_rng(1).dirichlet(np.ones(608))
    r1 = both(mcr, mu, sigma, 0.02, 60, 400_000, 11, weights=w, precision=mcr.PREC_BF16X3,
              alphas=(0.95, 0.99, 0.995, 0.999), want_hist=True)
    assert r1.kernel == "mfma_lds" and r1.var_passes == 1
    # 2: a smaller query right after, reading scratch the first one filled:
    # fp32 register kernel with per-path outputs over three 2^18-path chunks,
    # near counts, S histogram and VaR
    mu2, sigma2 = factor_like(40, 4, 400)
    r2 = both(mcr, mu2, sigma2, 0.02, 4, 600_001, 12, alphas=(0.99,), near=True, want_S=True,
              want_loss=True, want_hist=True, path_offset=2**40 + 3)
    assert r2.kernel == "valu_reg" and r2.S.shape == (600_001,)
    assert int(r2.s_hist[4:].sum()) == r2.count
    # 3: smaller again, the bf16x3 register-operand MFMA kernel
    r3 = both(mcr, mu2[:20], sigma2[:20, :20], 0.015, 3, 70_000, 13, precision=mcr.PREC_BF16X3,
              alphas=(0.99,))
    assert r3.kernel == "mfma"


def test_noalloc_refinement_passes_and_tile_kernel(mcr):
    mcr.init(0)
    mu, sigma = factor_like(300, 10, 3000, T=3000)
    # zero-width windows force the histogram + bracket refinement passes,
    # which use the static host histogram and gather buffers
    r = both(mcr, mu, sigma, 0.02, 30, 50_000, 21, precision=mcr.PREC_FP32, alphas=(0.99, 0.999),
             extra_flags=mcr.F_TEST_NARROW)
    assert r.kernel == "valu_tile" and r.var_passes > 1
    r = both(mcr, mu[:7], sigma[:7, :7], 0.02, 2, 1000, 22, precision=mcr.PREC_FP32,
             alphas=(0.5,), want_S=True)
    assert r.kernel == "valu_reg"


def test_noalloc_errors_match_run(mcr):
    mcr.init(0)
    mu, sigma = factor_like(10, 2, 100)
    for kw, code in (({"paths": 0}, mcr.MCR_E_ARG),
                     ({"near": True, "precision": mcr.PREC_BF16X3}, mcr.MCR_E_UNSUPPORTED)):
        args = dict(paths=1000, precision=mcr.PREC_FP32, near=False)
        args.update(kw)
        for fn in (mcr.run, mcr.run_noalloc):
            with pytest.raises(mcr.McrError) as e:
                fn(mu, sigma, 0.02, 2, args["paths"], 1, precision=args["precision"],
                   near=args["near"])
            assert e.value.code == code


def test_noalloc_refuses_more_than_608_assets(mcr):
    """Past 608 assets a query runs through mcr_run only (its buffers are
    sized past the static array's design); the allocation-free entry point
    refuses it before touching the device."""
    mcr.init(0)
    mu, sigma = factor_like(640, 8, 101, T=1500)
    with pytest.raises(mcr.McrError) as e:
        mcr.run_noalloc(mu, sigma, 0.02, 64, 1000, 1)
    assert e.value.code == mcr.MCR_E_UNSUPPORTED
    assert mcr.run(mu, sigma, 0.02, 64, 1000, 1).kernel == "valu_chunk"
