"""Batched uneven portfolios: batched factor/prepare + one persistent path
launch per kernel class (optionally a replayed hipGraph) against the same
queries run alone.

Bar: every per-query result (count, VaR, tail sum, ES) is identical to the
query run alone with the same kernel (mcr_run, bf16x3, automatic kernel: the
register-operand MFMA kernel for n <= 127, the LDS MFMA kernel above): the
paths are keyed by (seed, path id) and all reductions are integer or exact
order statistics, so batching cannot change them."""
from __future__ import annotations

import os

import numpy as np
import pytest

from helpers import factor_like

pytestmark = pytest.mark.gpu


def portfolios(seed=0, sizes=(1, 5, 17, 50, 64, 100, 128, 200, 333, 450, 600, 608), paths=None):
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(sizes):
        mu, sigma = factor_like(n, max(1, min(20, n // 10)), 500 + 7 * i + seed,
                                T=max(1500, 2 * n))
        m = int(paths[i]) if paths is not None else int(rng.integers(1000, 60000))
        out.append({"mu": mu, "sigma": sigma, "x": float(rng.uniform(0.01, 0.04)),
                    "k": max(1, n // 10), "paths": m, "seed": int(rng.integers(0, 2**63)),
                    "weights": rng.dirichlet(np.ones(n)), "alphas": (0.99,)})
    return out


def alone(mcr, p, extra_flags=0):
    return mcr.run(p["mu"], p["sigma"], p["x"], p["k"], p["paths"], p["seed"],
                   weights=p["weights"], precision=mcr.PREC_BF16X3, alphas=p["alphas"], extra_flags=extra_flags)


def same(a, b):
    assert a.count == b.count
    assert a.var == b.var
    assert a.tail_sum_fixed == b.tail_sum_fixed
    assert a.es == b.es


def test_batch_equals_alone(mcr):
    ps = portfolios(1, paths=[1, 63, 64, 65, 20000, 12345, 30000, 5000, 4000, 7000, 3000, 2500])
    res, o = mcr.run_batch(ps, use_graph=False, chunk_paths=1024)
    assert o.work_items >= len(ps)
    for p, r in zip(ps, res):
        same(r, alone(mcr, p))
        assert r.kernel == ("mfma" if len(p["mu"]) <= 127 else "mfma_lds")


def test_batch_equals_alone_at_class_boundaries(mcr):
    """Sizes on both sides of every kernel-class boundary (the register-operand
    MFMA kernel up to n = 127, the LDS row-tile ranges 5-8, 9-12, 13-16 and
    17-24 tiles, the factor kernel's larger panel from 609, the 767 limit)
    batch to the results of the queries alone."""
    sizes = (127, 128, 255, 256, 383, 384, 511, 512, 607, 608, 609, 700, 767)
    ps = portfolios(11, sizes=sizes, paths=[3000 + 17 * i for i in range(len(sizes))])
    for use_graph in (False, True):
        res, o = mcr.run_batch(ps, use_graph=use_graph, chunk_paths=512)
        for p, r in zip(ps, res):
            same(r, alone(mcr, p))


def test_graph_replay_uses_new_inputs(mcr):
    a = portfolios(2)
    b = portfolios(3, paths=[p["paths"] for p in a])  # same shapes, new data and seeds
    r1, o1 = mcr.run_batch(a, use_graph=True, chunk_paths=2048)
    assert o1.graph_replayed == 0
    # eager work between replays
    eager = alone(mcr, a[2])
    same(eager, r1[2])
    r2, o2 = mcr.run_batch(b, use_graph=True, chunk_paths=2048)
    assert o2.graph_replayed == 1
    r3, o3 = mcr.run_batch(a, use_graph=True, chunk_paths=2048)
    assert o3.graph_replayed == 1
    ref_b, _ = mcr.run_batch(b, use_graph=False, chunk_paths=2048)
    for x, y in zip(r2, ref_b):
        same(x, y)
    for x, y in zip(r3, r1):
        same(x, y)


@pytest.mark.parametrize("chunk,order", [(64, 0), (512, 1), (65536, 0)])
def test_chunking_and_order_do_not_change_results(mcr, chunk, order):
    ps = portfolios(4)
    ref, _ = mcr.run_batch(ps, use_graph=False, chunk_paths=4096, order=0)
    res, _ = mcr.run_batch(ps, use_graph=False, chunk_paths=chunk, order=order)
    for x, y in zip(res, ref):
        same(x, y)


def test_batch_refinement_fallback(mcr):
    ps = portfolios(5, sizes=(20, 40))
    ps[1]["alphas"] = (0.95, 0.99)
    ps[1]["extra_flags"] = mcr.F_TEST_NARROW  # zero-width windows: the bracket must miss
    res, _ = mcr.run_batch(ps, use_graph=False, chunk_paths=1024)
    assert res[1].var_passes > 1 and res[0].var_passes == 1
    same(res[1], alone(mcr, ps[1]))
    same(res[0], alone(mcr, ps[0]))


def test_batch_with_windows_holding_every_loss(mcr):
    """Windows wide enough to hold every loss (every path appends to the
    gather buffer), four VaR levels: each batched result equals the query run
    alone, in one pass."""
    ps = portfolios(12, sizes=(7, 60, 140, 300, 520), paths=[20000, 30000, 15000, 9000, 6000])
    for p in ps:
        p["alphas"] = (0.5, 0.9, 0.99, 0.999)
        p["extra_flags"] = mcr.F_TEST_WIDE
    res, _ = mcr.run_batch(ps, use_graph=False, chunk_paths=1024)
    for p, r in zip(ps, res):
        assert r.var_passes == 1
        same(r, alone(mcr, p, extra_flags=mcr.F_TEST_WIDE))


def test_batch_window_overflow_falls_back_to_refinement(mcr):
    """Every loss of 100k paths in the window overflows a batched query's
    2^16-loss buffer: the query is recomputed alone with the refinement
    passes, its results equal the query alone, and the other queries of the
    batch resolve in one pass."""
    ps = portfolios(4, sizes=(30, 200), paths=[100000, 20000])
    for p in ps:
        p["alphas"] = (0.99,)
        p["extra_flags"] = mcr.F_TEST_WIDE
    res, _ = mcr.run_batch(ps, use_graph=False, chunk_paths=1024)
    for p, r in zip(ps, res):
        assert r.var_passes == (1 if p["paths"] <= 65536 else 2)
        same(r, alone(mcr, p, extra_flags=mcr.F_TEST_WIDE))


def test_per_query_mode_agrees_statistically(mcr):
    ps = portfolios(6, sizes=(30, 90, 250), paths=[200000, 200000, 100000])
    per, o = mcr.run_batch(ps, mode=mcr.BATCH_PER_QUERY)
    bat, _ = mcr.run_batch(ps, use_graph=False)
    for a, b in zip(per, bat):
        assert abs(a.p_hat - b.p_hat) <= 1e-3 + 4 * a.std_error


def test_autotuner_picks_a_setting(mcr):
    ps = portfolios(7, sizes=(16, 48, 150), paths=[30000, 30000, 20000])
    best, tried = mcr.autotune_batch(ps)
    # 4 item sizes x 2 orders on one stream, then 2, 3 and 4 launch streams,
    # then the other workgroup shape of the one row-tile range used (n = 150)
    assert tried == 12 and best.chunk_used in (512, 2048, 8192, 32768)
    assert 1 <= best.streams_used <= 4 and best.big_shapes[0] in (0, 1)
    _, o = mcr.run_batch(ps, use_graph=False)  # chunk_paths = 0, streams = 0: the tuned default
    assert o.chunk_used == best.chunk_used and o.order_used == best.order_used
    assert o.streams_used == best.streams_used
    assert list(o.big_shapes) == list(best.big_shapes)


def test_every_workgroup_shape_gives_the_same_results(mcr):
    """The LDS MFMA batch kernel's alternative workgroup shapes (the
    autotuner's choices) give bit-identical results: each accumulator receives
    the same MFMA sequence whichever wave holds it."""
    ps = portfolios(10, sizes=(130, 250, 260, 380, 390, 500, 520, 608),
                    paths=[9000, 7000, 6000, 5000, 4000, 3000, 3000, 2000])
    ref, o = mcr.run_batch(ps, use_graph=False, chunk_paths=512, shapes=(1, 1, 1, 1))
    assert list(o.big_shapes) == [0, 0, 0, 0]
    alt, o2 = mcr.run_batch(ps, use_graph=True, chunk_paths=512, shapes=(2, 2, 2, 2))
    assert list(o2.big_shapes) == [1, 1, 1, 1]
    for p, x, y in zip(ps, alt, ref):
        same(x, y)
        same(x, alone(mcr, p))


@pytest.mark.parametrize("streams", [2, 4])
def test_forked_class_launches_equal_one_stream(mcr, streams):
    """Class kernels forked over concurrent streams (eager and in a replayed
    graph) give the results of one stream: items are keyed by (query, group)
    and every reduction is an integer."""
    ps = portfolios(9)
    ref, _ = mcr.run_batch(ps, use_graph=False, chunk_paths=1024, streams=1)
    for use_graph in (False, True, True):
        res, o = mcr.run_batch(ps, use_graph=use_graph, chunk_paths=1024, streams=streams)
        assert o.streams_used == streams
        for x, y in zip(res, ref):
            same(x, y)


# MCR_BATCH_RANDOM_CASES raises the number of randomized cases for a longer run
@pytest.mark.parametrize("seed", range(31, 31 + int(os.environ.get("MCR_BATCH_RANDOM_CASES", "3"))))
def test_randomized_batches_equal_alone(mcr, seed):
    """Random batches (2-12 portfolios of 1-767 assets, 1-100k paths, 0-4 VaR
    levels, narrow or wide windows) under random item sizes, orders, launch
    streams, workgroup shapes and graph replay: every result equals the
    query alone."""
    rng = np.random.default_rng(seed)
    count = int(rng.integers(2, 13))
    sizes = tuple(int(v) for v in rng.choice([1, 9, 64, 127, 128, 256, 383, 450, 512, 608,
                                               640, 767], count))
    paths = [int(np.exp(rng.uniform(0.0, np.log(100_000)))) for _ in range(count)]
    ps = portfolios(seed, sizes=sizes, paths=paths)
    for p in ps:
        na = int(rng.integers(0, 5))
        p["alphas"] = tuple(sorted(rng.choice([0.5, 0.9, 0.95, 0.99, 0.999], na, replace=False)))
        p["extra_flags"] = int(rng.choice([0, 0, mcr.F_TEST_NARROW, mcr.F_TEST_WIDE]))
    kw = {"chunk_paths": int(rng.choice([64, 512, 2048, 65536])), "order": int(rng.integers(0, 2)),
          "streams": int(rng.integers(1, 5)),
          "shapes": tuple(int(v) for v in rng.integers(1, 3, 4))}
    for use_graph in (False, True, True):
        res, _ = mcr.run_batch(ps, use_graph=use_graph, **kw)
        for p, r in zip(ps, res):
            same(r, alone(mcr, p, extra_flags=p["extra_flags"]))


def test_batch_rejects_fp32_queries(mcr):
    ps = portfolios(8, sizes=(10,))
    ps[0]["precision"] = mcr.PREC_FP32
    with pytest.raises(mcr.McrError) as e:
        mcr.run_batch(ps)
    assert e.value.code == mcr.MCR_E_UNSUPPORTED


def test_batch_rejects_more_than_767_assets(mcr):
    ps = portfolios(8, sizes=(10, 768), paths=[100, 100])
    with pytest.raises(mcr.McrError) as e:
        mcr.run_batch(ps)
    assert e.value.code == mcr.MCR_E_UNSUPPORTED
