"""Serving loop (mcr_server_*): queries arrive over time, are admitted by
device-memory footprint, share rounds of the persistent kernel with whatever
else is running, and reuse cached Cholesky factors.

Bar: every served result (count, VaR, tail sums, ES) is identical to the same
query run alone with the same kernel (mcr_run, bf16x3, automatic kernel):
paths are keyed by (seed, path id) and reduced into integer / fixed-point
tallies, so the mix, the round boundaries and the cache cannot change them."""
from __future__ import annotations

import os
import threading
import time

import numpy as np
import pytest

from helpers import factor_like

pytestmark = pytest.mark.gpu


def stream(seed, count, sizes=(1, 5, 12, 33, 50, 64, 100, 200, 333, 450, 600, 608), reuse=0.3,
           paths=(2_000, 80_000)):
    """Queries of mixed sizes; a fraction reuses an earlier covariance."""
    rng = np.random.default_rng(seed)
    seen = []
    out = []
    for i in range(count):
        if seen and rng.uniform() < reuse:
            mu, sigma = seen[int(rng.integers(0, len(seen)))]
        else:
            n = int(sizes[int(rng.integers(0, len(sizes)))])
            mu, sigma = factor_like(n, max(1, min(20, n // 10)), 900 + 13 * i + seed,
                                    T=max(1500, 2 * n))
            seen.append((mu, sigma))
        n = mu.shape[0]
        out.append({"mu": mu, "sigma": sigma, "x": float(rng.uniform(0.01, 0.04)),
                    "k": max(1, n // 10), "seed": int(rng.integers(0, 2**63)),
                    "paths": int(np.exp(rng.uniform(np.log(paths[0]), np.log(paths[1])))),
                    "weights": rng.dirichlet(np.ones(n)), "alphas": (0.99,)})
    return out


def alone(mcr, p):
    return mcr.run(p["mu"], p["sigma"], p["x"], p["k"], p["paths"], p["seed"],
                   weights=p["weights"], precision=mcr.PREC_BF16X3, alphas=p.get("alphas", ()),
                   extra_flags=p.get("extra_flags", 0))


def same(a, b):
    assert a.count == b.count
    assert a.var == b.var
    assert a.tail_sum_fixed == b.tail_sum_fixed
    assert a.es == b.es
    assert a.factor_attempts == b.factor_attempts and a.jitter == b.jitter


def serve(mcr, ps, gaps_ms=2.0, seed=0, **opts):
    """Submits ps with exponential gaps (mean gaps_ms), returns results in order."""
    rng = np.random.default_rng(seed)
    with mcr.Server(**opts) as srv:
        tickets = []
        for p in ps:
            tickets.append(srv.submit(p))
            time.sleep(float(rng.exponential(gaps_ms)) / 1e3)
        out = [srv.wait(t) for t in tickets]
        st = srv.stats()
    return out, st


def test_served_results_equal_alone(mcr):
    ps = stream(1, 30)
    t0 = time.perf_counter()
    out, st = serve(mcr, ps, round_ms=0.1)
    elapsed_ms = (time.perf_counter() - t0) * 1e3
    assert st.completed == len(ps) and st.failed == 0
    assert st.cache_hits + st.cache_misses == len(ps)
    # the worker's time per phase: each non-negative, together within the run
    phases = [st.host_admit_ms, st.host_plan_ms, st.host_enqueue_ms, st.host_wait_ms,
              st.host_finalize_ms]
    assert all(v >= 0.0 for v in phases) and 0.0 < sum(phases) <= elapsed_ms
    assert st.cache_hits >= 1
    assert max(info.rounds for _, info in out) > 1  # some query spanned several rounds
    for p, (r, info) in zip(ps, out):
        assert info.status == 0 and info.latency_ms >= info.queue_ms >= 0.0
        assert r.kernel == ("mfma" if len(p["mu"]) <= 127 else "mfma_lds")
        same(r, alone(mcr, p))


def test_cache_does_not_change_results(mcr):
    ps = stream(2, 16, reuse=0.6)
    cached, st1 = serve(mcr, ps, gaps_ms=0.5)
    uncached, st2 = serve(mcr, ps, gaps_ms=0.5, cache_entries=-1)
    assert st1.cache_hits > 0
    assert st2.cache_hits <= st1.cache_hits
    for (a, _), (b, _) in zip(cached, uncached):
        same(a, b)


def test_admission_budget_holds(mcr):
    ps = stream(3, 20, sizes=(100, 333, 600), reuse=0.0, paths=(5_000, 40_000))
    need = max(mcr.server_footprint(p) for p in ps)
    budget = int(need * 2.5)
    out, st = serve(mcr, ps, gaps_ms=0.2, hbm_budget=budget)
    assert st.completed == len(ps)
    assert 0 < st.peak_bytes <= budget
    assert max(info.queue_ms for _, info in out) > 0.0  # someone waited for memory
    for p, (r, _) in zip(ps, out):
        same(r, alone(mcr, p))


def test_query_larger_than_budget_is_refused(mcr):
    p = stream(4, 1, sizes=(600,))[0]
    need = mcr.server_footprint(p)
    with mcr.Server(hbm_budget=need // 2) as srv:
        with pytest.raises(mcr.McrError) as e:
            srv.submit(p)
        assert e.value.code == mcr.MCR_E_NOMEM
        small = stream(5, 1, sizes=(5,))[0]
        r, _ = srv.wait(srv.submit(small))
        same(r, alone(mcr, small))


def test_factor_failure_is_reported_and_not_cached(mcr):
    ps = stream(6, 3, sizes=(20,))
    bad = dict(ps[1])
    bad["sigma"] = bad["sigma"].copy()
    bad["sigma"][3, 4] = np.nan
    with mcr.Server() as srv:
        t = [srv.submit(ps[0]), srv.submit(bad), srv.submit(ps[2])]
        time.sleep(0.05)
        t.append(srv.submit(bad))  # same bad covariance again
        same(srv.wait(t[0])[0], alone(mcr, ps[0]))
        for tb in (t[1], t[3]):
            with pytest.raises(mcr.McrError) as e:
                srv.wait(tb)
            assert e.value.code == mcr.MCR_E_FACTOR
        same(srv.wait(t[2])[0], alone(mcr, ps[2]))
        assert srv.stats().failed == 2


def test_quantile_refinement_fallback(mcr):
    ps = stream(7, 2, sizes=(40,))
    ps[1]["alphas"] = (0.95, 0.99)
    ps[1]["extra_flags"] = mcr.F_TEST_NARROW  # zero-width windows force the fallback
    out, _ = serve(mcr, ps)
    assert out[1][0].var_passes > 1 and out[0][0].var_passes == 1
    for p, (r, _) in zip(ps, out):
        same(r, alone(mcr, p))


def test_windows_holding_every_loss(mcr):
    """Every path's loss appended (windows that hold every loss), three VaR
    levels: each served result equals the query alone, in one pass."""
    ps = stream(9, 6, sizes=(20, 130, 450), paths=(5_000, 40_000))
    for p in ps:
        p["alphas"] = (0.5, 0.99, 0.999)
        p["extra_flags"] = mcr.F_TEST_WIDE
    out, _ = serve(mcr, ps)
    for p, (r, _) in zip(ps, out):
        assert r.var_passes == 1
        same(r, alone(mcr, p))


def test_four_var_levels_fit_their_window_buffers(mcr):
    """Four VaR levels from the median to 99.9 % on up to 1e6 paths: the window
    buffer sized for the widest expected window (the median's) holds every
    level's losses, so each query resolves in one pass, equal to alone."""
    ps = stream(11, 5, sizes=(12, 90, 300), paths=(200_000, 1_000_000))
    for p in ps:
        p["alphas"] = (0.5, 0.9, 0.99, 0.999)
    out, _ = serve(mcr, ps)
    for p, (r, _) in zip(ps, out):
        assert r.var_passes == 1
        same(r, alone(mcr, p))


def test_window_overflow_falls_back_to_refinement(mcr):
    """Windows that hold every loss of 100k-150k paths overflow the served
    query's 2^16-loss window buffer (a single query keeps 2^18): the server
    sends those queries through the refinement passes, and every result still
    equals the query alone (which resolves in one pass)."""
    ps = stream(10, 4, sizes=(20, 140), paths=(100_000, 150_000))
    for p in ps:
        p["alphas"] = (0.5, 0.99)
        p["extra_flags"] = mcr.F_TEST_WIDE
    out, _ = serve(mcr, ps)
    for p, (r, _) in zip(ps, out):
        a = alone(mcr, p)
        assert a.var_passes == 1 and r.var_passes > 1
        same(r, a)


def test_concurrent_submitters_and_poll(mcr):
    ps = stream(8, 24, paths=(1_000, 30_000))
    results = [None] * len(ps)
    with mcr.Server(round_ms=1.0) as srv:
        def client(idx):
            for i in idx:
                t = srv.submit(ps[i])
                while True:
                    got = srv.poll(t)
                    if got is not None:
                        results[i] = got[0]
                        break
                    time.sleep(0.0005)
        threads = [threading.Thread(target=client, args=(range(c, len(ps), 4),)) for c in range(4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=120)
        assert srv.stats().completed == len(ps)
    for p, r in zip(ps, results):
        same(r, alone(mcr, p))


def test_rejects_unservable_queries(mcr):
    p = stream(9, 1, sizes=(10,))[0]
    with mcr.Server() as srv:
        big = {"mu": np.zeros(768), "sigma": np.eye(768) * 1e-4, "weights": None}
        for bad in ({"precision": mcr.PREC_FP32}, {"extra_flags": mcr.F_NEAR}, big):
            q = dict(p, **bad)
            with pytest.raises(mcr.McrError) as e:
                srv.submit(q)
            assert e.value.code == mcr.MCR_E_UNSUPPORTED


def test_age_limit_changes_order_not_results(mcr):
    """With an age limit every query passes at once (1 us), rounds plan the
    queries oldest first instead of least work left first: each result still
    equals the query alone.  A negative or NaN limit is refused."""
    ps = stream(5, 20, paths=(2_000, 40_000))
    out, st = serve(mcr, ps, gaps_ms=0.2, age_limit_ms=0.001)
    assert st.completed == len(ps) and st.failed == 0
    for p, (r, info) in zip(ps, out):
        assert info.status == 0
        same(r, alone(mcr, p))
    for bad in (-1.0, float("nan")):
        with pytest.raises(mcr.McrError) as e:
            mcr.Server(age_limit_ms=bad)
        assert e.value.code == mcr.MCR_E_ARG


def test_queries_past_608_assets_equal_alone(mcr):
    """Queries of 609-767 assets (22 and 24 row tiles; the factor kernel's
    larger panel) served beside small ones equal the queries alone."""
    ps = stream(12, 8, sizes=(20, 609, 700, 767), reuse=0.2, paths=(2_000, 30_000))
    out, st = serve(mcr, ps, gaps_ms=0.5)
    assert st.completed == len(ps) and st.failed == 0
    assert max(len(p["mu"]) for p in ps) > 608
    for p, (r, _) in zip(ps, out):
        same(r, alone(mcr, p))


# MCR_SERVER_RANDOM_CASES raises the number of randomized cases for a longer run
@pytest.mark.parametrize("seed", range(21, 21 + int(os.environ.get("MCR_SERVER_RANDOM_CASES", "3"))))
def test_randomized_queries_and_options_equal_alone(mcr, seed):
    """Random server options (round length, forks, age limit, cache on or
    off) and random queries (1-767 assets, 1-200k paths, 0-4 VaR levels,
    narrow or wide windows, repeated covariances): every served result equals
    the same query alone."""
    rng = np.random.default_rng(seed)
    opts = {"round_ms": float(rng.choice([0.05, 0.5, 4.0])), "streams": int(rng.integers(1, 3)),
            "age_limit_ms": float(rng.choice([0.0, 0.01, 5.0])),
            "cache_entries": int(rng.choice([0, -1, 2]))}
    sizes = (1, 3, 17, 64, 127, 128, 300, 511, 608, 609, 700, 767)
    ps = stream(100 + seed, 14, sizes=sizes, reuse=0.3, paths=(1, 200_000))
    for p in ps:
        na = int(rng.integers(0, 5))
        p["alphas"] = tuple(sorted(rng.choice([0.5, 0.9, 0.95, 0.99, 0.999], na, replace=False)))
        p["extra_flags"] = int(rng.choice([0, 0, mcr.F_TEST_NARROW, mcr.F_TEST_WIDE]))
    out, st = serve(mcr, ps, gaps_ms=float(rng.choice([0.0, 0.3, 2.0])), seed=seed, **opts)
    assert st.completed == len(ps) and st.failed == 0, opts
    for p, (r, _) in zip(ps, out):
        same(r, alone(mcr, p))


def test_forked_round_launches_equal_one_stream(mcr):
    """A round's class kernels forked over two streams give the results of
    one stream (same items, integer reductions)."""
    ps = stream(4, 14, paths=(2_000, 40_000))
    one, _ = serve(mcr, ps, gaps_ms=0.3, streams=1)
    two, _ = serve(mcr, ps, gaps_ms=0.3, streams=2)
    for (a, _), (b, _) in zip(one, two):
        same(a, b)
