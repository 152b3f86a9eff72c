"""The benchmark workloads at their full sizes (workloads.py builds the same
queries bench.py measures), each anchored to the C oracle.

Bars:
  * the plain n = 20 workload (1M paths, k = 5) and the fused n = 50 one (1e8 paths,
    k = 10): the fp32 kernel's exceedance and near-threshold counts equal the
    oracle's (serial, and OpenMP over all 1e8 paths) on the identical stream;
  * the 256 uneven portfolios: every batched result (count, VaR, tail
    sum, ES) equals the query run alone; the three smallest-n and three
    largest-n queries are within the bf16x3 tolerances of the FP64 oracle on
    the identical stream (whole query where the oracle finishes in about a
    second, else the first paths of the same query run alone);
  * the query stream (1000 queries, Poisson arrivals at 200 QPS, 30 % covariance
    reuse): every served result equals the query run alone, the same anchors
    against FP64, and the latency p99 within the stated bound;
  * the fused n = 50 workload at full size with bf16x3: probability and
    VaR99 / VaR99.9 within 1e-4 relative of the FP64 oracle (all 1e8 paths);
  * at full size, the count and S histogram of one query equal the sums over
    the same paths run as four ranges (fp32 and bf16x3 at n = 50, bf16x3 at
    n = 600, fp32 on the column-chunk kernel at n = 1024);
  * the n = 767 workload (LDS MFMA kernel, 24 row tiles) and the n = 1024
    one (column-chunk kernel, chosen for every precision past 767)
    within the bf16x3 tolerances of the FP64 oracle on the identical stream.
bf16x3 tolerances (as in test_gpu_mfma.py): |count - count_fp64| <=
max(3, 1e-4 count_fp64); VaR99 within 1e-4 relative of the FP64 order
statistic."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import rank_of

pytestmark = pytest.mark.gpu

ORACLE_FLOP_BUDGET = 2e10  # per anchored query: about a second of the oracle on 16 threads


@pytest.fixture(scope="module")
def wl(mcr):
    return mcr.workloads()


def fp32_model(mcr, orc, mu, sigma, x):
    m = mcr.prepare(mu, sigma, x)
    return orc.F32Model(m["L32"], m["mu32"], m["v32"], m["mp32"], m["thr32"])


def alone(mcr, p):
    return mcr.run(p["mu"], p["sigma"], p["x"], p["k"], p["paths"], p["seed"],
                   weights=p["weights"], precision=mcr.PREC_BF16X3, alphas=p["alphas"])


def same(a, b):
    assert a.count == b.count
    assert a.var == b.var
    assert a.tail_sum_fixed == b.tail_sum_fixed
    assert a.es == b.es


def anchor_fp64(mcr, orc, p, result):
    """result (the query's full-size bf16x3 result) against the FP64 oracle on
    the identical stream; over a prefix of the paths when the whole query
    would take the oracle much longer than a second."""
    n = len(p["mu"])
    M = p["paths"]
    P = int(min(M, ORACLE_FLOP_BUDGET // (n * (n + 1))))
    if P < M:
        result = mcr.run(p["mu"], p["sigma"], p["x"], p["k"], P, p["seed"], weights=p["weights"],
                         precision=mcr.PREC_BF16X3, alphas=p["alphas"])
    L, *_ = mcr.factor(p["sigma"])
    ref = orc.simulate_f64(L, p["mu"], p["weights"], p["x"], p["seed"], 0, P, p["k"],
                           want_loss=True, nthreads=16)
    assert abs(result.count - ref["count"]) <= max(3, 1e-4 * ref["count"]), (n, P, result.count,
                                                                              ref["count"])
    v64 = float(np.sort(ref["loss"])[rank_of(p["alphas"][0], P) - 1])
    assert abs(result.var[0] - v64) <= 1e-4 * abs(v64), (n, P, result.var[0], v64)
    return P


def extremes(ps, count=3):
    order = sorted(range(len(ps)), key=lambda i: (len(ps[i]["mu"]), i))
    return order[:count] + order[-count:]


def test_plain_n20_fp32_equals_serial_oracle(mcr, orc, wl):
    C = wl.PLAIN_N20
    mu, sigma = wl.model_of(mcr, C)
    r = mcr.run(mu, sigma, C["x"], C["k"], C["paths"], C["path_seed"], precision=mcr.PREC_FP32,
                near=True)
    assert r.kernel == "valu_reg"
    ref = orc.simulate_f32(fp32_model(mcr, orc, mu, sigma, C["x"]), C["path_seed"], 0,
                           C["paths"], C["k"], nthreads=1)
    print(f"plain n=20: GPU {r.count} (near {r.near_threshold}) vs serial oracle {ref['count']} "
          f"(near {ref['near']})")
    assert (r.count, r.near_threshold) == (ref["count"], ref["near"])
    assert r.p_hat == ref["count"] / C["paths"]


def test_fused_n50_full_size_fp32_equals_openmp_oracle(mcr, orc, wl):
    """All 1e8 paths of the headline stream: the fp32 kernel's counts equal the
    oracle's OpenMP restatement of the serial model (exact integers)."""
    C = wl.FUSED_N50
    mu, sigma = wl.model_of(mcr, C)
    r = mcr.run(mu, sigma, C["x"], C["k"], C["paths"], C["path_seed"], precision=mcr.PREC_FP32,
                near=True)
    ref = orc.simulate_f32(fp32_model(mcr, orc, mu, sigma, C["x"]), C["path_seed"], 0,
                           C["paths"], C["k"], nthreads=16)
    print(f"fused n=50: GPU {r.count} (near {r.near_threshold}) vs oracle {ref['count']} "
          f"(near {ref['near']}) over {C['paths']} paths")
    assert (r.count, r.near_threshold) == (ref["count"], ref["near"])
    b = mcr.run(mu, sigma, C["x"], C["k"], C["paths"], C["path_seed"], precision=mcr.PREC_BF16X3)
    # bf16x3 on the same 1e8 paths: the relative bar on the exceedance probability
    assert abs(b.count - ref["count"]) <= 1e-4 * ref["count"], (b.count, ref["count"])


def test_fused_n50_full_size_bf16x3_within_bar_of_fp64(mcr, orc, wl):
    """All 1e8 paths of the headline stream, bf16x3 MFMA against the FP64
    oracle on the identical stream: the exceedance probability and VaR99 /
    VaR99.9 each within 1e-4 relative (the precision bar), the VaR levels
    against the FP64 losses' exact order statistics."""
    C = wl.FUSED_N50
    n, M = C["n"], C["paths"]
    mu, sigma = wl.model_of(mcr, C)
    r = mcr.run(mu, sigma, C["x"], C["k"], M, C["path_seed"], precision=mcr.PREC_BF16X3,
                alphas=(0.99, 0.999))
    L, *_ = mcr.factor(sigma)
    ref = orc.simulate_f64(L, mu, np.full(n, 1.0 / n), C["x"], C["path_seed"], 0, M, C["k"],
                           want_loss=True, nthreads=16)
    idx = [rank_of(a, M) - 1 for a in (0.99, 0.999)]
    part = np.partition(ref["loss"], idx)
    v64 = [float(part[i]) for i in idx]
    del part
    print(f"fused n=50 bf16x3 vs FP64 over {M} paths: count {r.count} vs {ref['count']}, "
          f"VaR99 {r.var[0]:.9g} vs {v64[0]:.9g}, VaR99.9 {r.var[1]:.9g} vs {v64[1]:.9g}")
    assert r.kernel == "mfma"
    assert abs(r.count - ref["count"]) <= 1e-4 * ref["count"]
    for got, want in zip(r.var, v64):
        assert abs(got - want) <= 1e-4 * abs(want)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_full_size_path_ranges_add_up(mcr, wl, precision):
    """All 1e8 paths of the headline stream in one query equal the sum of the
    same paths run as four uneven ranges (path_offset): the exceedance count
    and the S histogram bin by bin -- the per-path results do not depend on
    how the paths are grouped, which a run split over GPUs or resumed from a
    saved tally relies on.  The large n = 600 workload (1e7 paths) is checked
    the same way with bf16x3, the n = 1024 one (1e6 paths, column-chunk
    kernel) with fp32."""
    prec = mcr.PREC_FP32 if precision == "fp32" else mcr.PREC_BF16X3
    extra = (wl.LARGE_N600,) if precision == "bf16x3" else (wl.LARGEST_N1024,)
    for C in (wl.FUSED_N50,) + extra:
        mu, sigma = wl.model_of(mcr, C)
        M = C["paths"]
        whole = mcr.run(mu, sigma, C["x"], C["k"], M, C["path_seed"], precision=prec,
                        want_hist=True)
        cuts = [0, M // 7, M // 2, M - 12345, M]
        parts = [mcr.run(mu, sigma, C["x"], C["k"], b - a, C["path_seed"], precision=prec,
                         want_hist=True, path_offset=a) for a, b in zip(cuts, cuts[1:])]
        assert all(r.kernel == whole.kernel for r in parts)
        assert whole.count == sum(r.count for r in parts)
        assert np.array_equal(whole.s_hist, np.sum([r.s_hist for r in parts], axis=0))
        assert int(whole.s_hist.sum()) == M


def test_widest_mfma_n767_within_bar_of_fp64(mcr, orc, wl):
    """The n = 767 workload (24 row tiles, the LDS MFMA kernel's largest
    single query) at bf16x3: count and VaR99 within the bf16x3 tolerances of
    the FP64 oracle on the identical stream."""
    C = wl.WIDEST_MFMA_N767
    mu, sigma = wl.model_of(mcr, C)
    p = {"mu": mu, "sigma": sigma, "x": C["x"], "k": C["k"], "paths": C["paths"],
         "seed": C["path_seed"], "weights": np.full(C["n"], 1.0 / C["n"]), "alphas": C["alphas"]}
    r = alone(mcr, p)
    assert r.kernel == "mfma_lds"
    P = anchor_fp64(mcr, orc, p, r)
    assert P >= 10_000


def test_largest_n1024_within_bar_of_fp64(mcr, orc, wl):
    """The n = 1024 workload asked for at bf16x3 runs the column-chunk fp32
    kernel; its count and VaR99 are within the bf16x3 tolerances of the FP64
    oracle on the identical stream (over the paths the oracle finishes in
    about a second)."""
    C = wl.LARGEST_N1024
    mu, sigma = wl.model_of(mcr, C)
    p = {"mu": mu, "sigma": sigma, "x": C["x"], "k": C["k"], "paths": C["paths"],
         "seed": C["path_seed"], "weights": np.full(C["n"], 1.0 / C["n"]), "alphas": C["alphas"]}
    r = alone(mcr, p)
    assert r.kernel == "valu_chunk"
    P = anchor_fp64(mcr, orc, p, r)
    assert P >= 10_000


def test_near_flag_on_bf16x3_is_refused_and_auto_uses_fp32(mcr, wl):
    C = wl.FUSED_N50
    mu, sigma = wl.model_of(mcr, C)
    for kw in ({"precision": mcr.PREC_BF16X3}, {"kernel": mcr.KERNEL_MFMA},
               {"kernel": mcr.KERNEL_MFMA_LDS}):
        with pytest.raises(mcr.McrError) as e:
            mcr.run(mu, sigma, C["x"], C["k"], 1000, 1, near=True, **kw)
        assert e.value.code == mcr.MCR_E_UNSUPPORTED
    a = mcr.run(mu, sigma, C["x"], C["k"], 200_000, 5, precision=mcr.PREC_AUTO, near=True)
    f = mcr.run(mu, sigma, C["x"], C["k"], 200_000, 5, precision=mcr.PREC_FP32, near=True)
    assert a.kernel == "valu_reg"
    assert (a.count, a.near_threshold) == (f.count, f.near_threshold)
    assert mcr.run(mu, sigma, C["x"], C["k"], 1000, 1, precision=mcr.PREC_AUTO).kernel == "mfma"


def test_batch_of_256_uneven_equals_alone_and_fp64(mcr, orc, wl):
    ps = wl.batch_portfolios(mcr)
    assert len(ps) == 256
    res, o = mcr.run_batch(ps, use_graph=True)
    again, o2 = mcr.run_batch(ps, use_graph=True)
    assert o2.graph_replayed == 1
    for i, p in enumerate(ps):
        r = alone(mcr, p)
        same(res[i], r)
        same(again[i], r)
    for i in extremes(ps):
        P = anchor_fp64(mcr, orc, ps[i], res[i])
        print(f"batch query {i}: n={len(ps[i]['mu'])} paths={ps[i]['paths']} anchored on {P}")


def test_stream_of_1000_queries_equals_alone_and_fp64(mcr, orc, wl):
    count = 1000
    qs, gaps, _ = wl.serve_queries(mcr, count)
    with mcr.Server() as srv:
        out, infos, span, wall = wl.replay_stream(srv, qs, gaps)
        st = srv.stats()
    assert st.completed >= count and st.failed == 0
    assert st.cache_hits > 0
    lat = np.array([x.latency_ms for x in infos])
    p99 = float(np.percentile(lat, 99))
    print(f"query stream: {count} queries in {wall:.2f} s (offered 200 QPS, submitted at "
          f"{count / span:.0f}/s); p50 {np.percentile(lat, 50):.2f} ms, p99 {p99:.1f} ms, "
          f"cache hits {st.cache_hits}")
    for q, r in zip(qs, out):
        same(r, alone(mcr, q))
    for i in extremes(qs):
        anchor_fp64(mcr, orc, qs[i], out[i])
    assert p99 <= wl.P99_BOUND_MS, p99
