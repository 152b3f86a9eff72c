"""Benchmark of the Monte Carlo crash-risk path on one MI355X.

Workload (one step = one full risk query, factor included): n = 50 assets from
a seeded 5-factor model (T = 2000 days), 100,000,000 Philox paths, crash
threshold -2%, k = 10, 99% VaR of the equal-weight portfolio.  Prints one JSON
line: whole-job paths/s, the path kernel's roofline figures (HIP-event timing
of the kernel on the library's stream) and a CPU baseline timed on a bounded
sample of the same workload (the C oracle, serial and OpenMP, host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3    # MI355X published fp32
PEAK_BF16_TFLOPS = 2500.0   # MI355X published dense bf16 MFMA
# HBM bytes per launch of the workload's path kernel, from two rocprofv3 --pmc
# runs of this bench (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 KiB, the factors
# read off a 256 MiB torch.neg calibration kernel in the same runs).
TRAFFIC_BYTES = {"mfma": (178.50 * 2.000 + 1410.66) * 1024}

WORKLOAD = {"name": "fused_n50_100M", "n": 50, "factors": 5, "days": 2000, "paths": 100_000_000,
            "k": 10, "x": 0.02, "alpha": 0.99, "data_seed": 1002, "path_seed": 20251202}
BATCH = {"name": "batch_256_uneven", "count": 256, "n_range": (5, 600), "paths_range": (1e5, 1e7),
         "x_range": (0.01, 0.04), "alpha": 0.99, "seed": 1004}
SERVE = {"name": "serve_poisson_200qps", "qps": 200.0, "seconds": 5.0, "reuse": 0.3,
         "n_range": (5, 600), "paths_range": (1e5, 1e7), "x_range": (0.01, 0.04), "alpha": 0.99,
         "seed": 1005}
LARGE = {"name": "large_n600_10M", "n": 600, "factors": 20, "days": 2500, "paths": 10_000_000,
         "k": 10, "x": 0.02, "alphas": (0.99, 0.999), "data_seed": 1003, "path_seed": 20251203}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=8)
    p.add_argument("--precision", default="auto", choices=["auto", "fp32", "bf16x3"])
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-large", action="store_true", help="skip the n=600 leg")
    p.add_argument("--no-batch", action="store_true", help="skip the 256-portfolio leg")
    p.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    p.add_argument("--serve-seconds", type=float, default=SERVE["seconds"])
    p.add_argument("--copy-mb", type=int, default=0,
                   help="first run a plain device copy of this many MiB (counter calibration)")
    return p.parse_args()


def cpu_baseline(L, mu, x, k, seed, n, budget_s):
    """C oracle (port of the serial/OpenMP models on the same stream) timed on
    a bounded sample; returns the OpenMP figure plus the serial one."""
    import __graft_entry__ as ge
    import numpy as np

    orc = ge.load_oracle()
    w = np.full(n, 1.0 / n)
    model = orc.F32Model.from_f64(L, mu, w, x)
    threads = min(os.cpu_count() or 1, 16)
    # serial rate on a short sample, then size the OpenMP sample to the budget
    t0 = time.perf_counter()
    ns = 20000
    orc.simulate_f32(model, seed, 0, ns, k, nthreads=1)
    serial_rate = ns / (time.perf_counter() - t0)
    npar = int(min(max(serial_rate * threads * budget_s * 0.7, 100000), 50_000_000))
    t0 = time.perf_counter()
    orc.simulate_f32(model, seed, 0, npar, k, nthreads=threads)
    par_rate = npar / (time.perf_counter() - t0)
    return {"value": par_rate, "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": f"{npar} paths of the n=50 workload (OpenMP, {threads} threads); "
                      f"serial 1-core rate {serial_rate:.4g} paths/s on {ns} paths",
            "serial_value": serial_rate}


def batch_portfolios(mcr, np, count=None):
    """Config 4: n log-uniform in [5, 600], paths log-uniform in [1e5, 1e7],
    Dirichlet(1) weights, thresholds U(1%, 4%); Sigma from the seeded factor
    model with K = clamp(n/30, 1, 20) factors over T = 2n + 200 days."""
    B = BATCH
    rng = np.random.default_rng(B["seed"])
    out = []
    for i in range(count or B["count"]):
        n = int(round(np.exp(rng.uniform(np.log(B["n_range"][0]), np.log(B["n_range"][1])))))
        m = int(round(np.exp(rng.uniform(np.log(B["paths_range"][0]), np.log(B["paths_range"][1])))))
        K = int(min(20, max(1, n // 30)))
        mu, sigma = mcr.synth_factor_model(n, K, 2 * n + 200, B["seed"] * 1000 + i)
        out.append({"mu": mu, "sigma": sigma, "x": float(rng.uniform(*B["x_range"])),
                    "k": max(1, (n + 9) // 10), "paths": m, "seed": int(rng.integers(0, 2**62)),
                    "weights": rng.dirichlet(np.ones(n)), "alphas": (B["alpha"],),
                    "precision": mcr.PREC_BF16X3})
    return out


def batch_leg(mcr, np):
    """Config 4: one persistent launch (hipGraph) vs one launch per portfolio."""
    ps = batch_portfolios(mcr, np)
    total = sum(p["paths"] for p in ps)
    best, tried = mcr.autotune_batch(ps)
    r0, o0 = mcr.run_batch(ps, use_graph=True)  # captures
    times = []
    for _ in range(3):
        r1, o1 = mcr.run_batch(ps, use_graph=True)
        times.append(o1.ms_device)
    per, op = mcr.run_batch(ps, mode=mcr.BATCH_PER_QUERY)
    ms_b = float(np.median(times))
    res = {"workload": BATCH["name"], "portfolios": len(ps), "total_paths": total,
           "n_min": min(p["n"] if "n" in p else len(p["mu"]) for p in ps),
           "n_max": max(len(p["mu"]) for p in ps),
           "autotune": {"candidates": tried, "chunk_paths": best.chunk_used,
                        "order": "lpt" if best.order_used == 0 else "natural"},
           "persistent_graph_ms": ms_b, "persistent_paths_per_s": total / (ms_b * 1e-3),
           "graph_replayed": bool(o1.graph_replayed), "work_items": o1.work_items,
           "per_portfolio_ms": op.ms_total, "per_portfolio_device_ms": op.ms_device,
           "per_portfolio_paths_per_s": total / (op.ms_total * 1e-3),
           "speedup_vs_per_portfolio": op.ms_total / ms_b}
    print(f"# batch: persistent+graph {ms_b:.1f} ms, per-portfolio {op.ms_total:.1f} ms "
          f"({total:.3g} paths)", file=sys.stderr, flush=True)
    return res


def accuracy_check(mcr, np):
    """bf16x3 MFMA vs FP64 on the identical stream (the oracle as checker),
    on path samples of the n = 50 and n = 600 workloads: relative error of the
    exceedance probability and of VaR99."""
    import __graft_entry__ as ge

    orc = ge.load_oracle()
    threads = min(os.cpu_count() or 1, 16)
    out = {}
    for name, G, M in (("n50", dict(n=WORKLOAD["n"], K=WORKLOAD["factors"], T=WORKLOAD["days"],
                                    ds=WORKLOAD["data_seed"], seed=WORKLOAD["path_seed"],
                                    k=WORKLOAD["k"], x=WORKLOAD["x"]), 2_000_000),
                       ("n600", dict(n=LARGE["n"], K=LARGE["factors"], T=LARGE["days"],
                                     ds=LARGE["data_seed"], seed=LARGE["path_seed"], k=LARGE["k"],
                                     x=LARGE["x"]), 100_000)):
        n = G["n"]
        mu, sigma = mcr.synth_factor_model(n, G["K"], G["T"], G["ds"])
        r = mcr.run(mu, sigma, G["x"], G["k"], M, G["seed"], precision=mcr.PREC_BF16X3,
                    alphas=(0.99,))
        L, *_ = mcr.factor(sigma)
        ref = orc.simulate_f64(L, mu, np.full(n, 1.0 / n), G["x"], G["seed"], 0, M, G["k"],
                               want_loss=True, nthreads=threads)
        rank = min(max(int(np.ceil(0.99 * M)), 1), M)
        v64 = float(np.sort(ref["loss"])[rank - 1])
        out[name] = {"paths": M, "kernel": r.kernel, "count_bf16x3": r.count,
                     "count_fp64": ref["count"],
                     "rel_err_p_hat": abs(r.count - ref["count"]) / max(ref["count"], 1),
                     "rel_err_var99": abs(r.var[0] - v64) / abs(v64)}
    return out


def serve_queries(mcr, np, count, seed):
    """Config 5 stream: arrival gaps ~ Exp(1/qps); sizes as config 4; a query
    reuses an earlier query's covariance (and mean) with probability 0.3."""
    from concurrent.futures import ThreadPoolExecutor

    S = SERVE
    rng = np.random.default_rng(seed)
    shape, src = [], []
    for i in range(count):
        if i and rng.uniform() < S["reuse"]:
            src.append(src[int(rng.integers(0, i))])
        else:
            src.append(i)
        n = int(round(np.exp(rng.uniform(np.log(S["n_range"][0]), np.log(S["n_range"][1])))))
        shape.append(n)
    distinct = sorted(set(src))

    def make(i):
        n = shape[i]
        return mcr.synth_factor_model(n, int(min(20, max(1, n // 30))), 2 * n + 200,
                                      seed * 1000 + i)

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        data = dict(zip(distinct, ex.map(make, distinct)))
    gaps = rng.exponential(1.0 / S["qps"], count)
    out = []
    for i in range(count):
        mu, sigma = data[src[i]]
        n = mu.shape[0]
        m = int(round(np.exp(rng.uniform(np.log(S["paths_range"][0]), np.log(S["paths_range"][1])))))
        out.append({"mu": mu, "sigma": sigma, "x": float(rng.uniform(*S["x_range"])),
                    "k": max(1, (n + 9) // 10), "paths": m, "seed": int(rng.integers(0, 2**62)),
                    "weights": rng.dirichlet(np.ones(n)), "alphas": (S["alpha"],),
                    "precision": mcr.PREC_BF16X3})
    return out, gaps, len(distinct)


def serve_leg(mcr, np, seconds):
    """Config 5: open-loop Poisson arrivals into the serving loop; latency is
    submit -> result ready as the server records it."""
    import threading

    S = SERVE
    count = max(1, int(S["qps"] * seconds))
    qs, gaps, n_distinct = serve_queries(mcr, np, count, S["seed"])
    warm, _, _ = serve_queries(mcr, np, 8, S["seed"] + 1)
    infos = [None] * count
    with mcr.Server() as srv:
        # warm-up on other queries (kernel load, cost-model calibration)
        for t in [srv.submit(q) for q in warm]:
            srv.wait(t)
        tickets = [None] * count
        ready = threading.Semaphore(0)

        def collect():
            for i in range(count):
                ready.acquire()
                infos[i] = srv.wait(tickets[i])[1]

        th = threading.Thread(target=collect)
        th.start()
        st0 = srv.stats()
        t0 = time.perf_counter()
        due = t0
        for i in range(count):
            due += gaps[i]
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(qs[i])
            ready.release()
        submit_span = time.perf_counter() - t0
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([x.latency_ms for x in infos])
    que = np.array([x.queue_ms for x in infos])
    paths = sum(q["paths"] for q in qs)
    hits = st.cache_hits - st0.cache_hits
    res = {"workload": S["name"], "queries": count, "offered_qps": S["qps"],
           "achieved_qps": count / submit_span, "distinct_covariances": n_distinct,
           "total_paths": paths, "sustained_paths_per_s": paths / wall,
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p90": float(np.percentile(lat, 90)),
                          "p99": float(np.percentile(lat, 99)), "max": float(lat.max()),
                          "mean": float(lat.mean())},
           "queue_ms": {"p50": float(np.percentile(que, 50)), "p99": float(np.percentile(que, 99))},
           "rounds": st.rounds - st0.rounds,
           "device_busy_frac": (st.device_ms - st0.device_ms) / (wall * 1e3),
           "cache_hit_rate": hits / count, "peak_device_bytes": st.peak_bytes,
           "budget_bytes": st.budget_bytes, "failed": st.failed}
    print(f"# serve: {count} queries at {res['achieved_qps']:.0f} QPS, p50 {res['latency_ms']['p50']:.1f} ms,"
          f" p99 {res['latency_ms']['p99']:.1f} ms, {res['sustained_paths_per_s']:.3g} paths/s, "
          f"busy {res['device_busy_frac']:.2f}, cache hits {hits}", file=sys.stderr, flush=True)
    return res


def large_leg(mcr, np):
    """n = 600 (config 3): plain fp32 tile kernel vs bf16x3 MFMA, same stream."""
    G = LARGE
    mu, sigma = mcr.synth_factor_model(G["n"], G["factors"], G["days"], G["data_seed"])
    res = {"workload": G["name"], "n_assets": G["n"], "paths": G["paths"],
           "cond_sigma": float(np.linalg.cond(sigma))}
    for name, prec in (("fp32", mcr.PREC_FP32), ("bf16x3", mcr.PREC_BF16X3)):
        runs = [mcr.run(mu, sigma, G["x"], G["k"], G["paths"], G["path_seed"], precision=prec,
                        alphas=G["alphas"]) for _ in range(3)]
        ms = float(np.mean([r.ms_kernel for r in runs[1:]]))
        r = runs[-1]
        res[name] = {"kernel": r.kernel, "kernel_ms": ms, "paths_per_s": G["paths"] / (ms * 1e-3),
                     "tflops": G["n"] * (G["n"] + 1) * G["paths"] / (ms * 1e-3) / 1e12,
                     "count": r.count, "p_hat": r.p_hat, "var99": r.var[0], "var999": r.var[1]}
    # k = 10 is not a tail event at n = 600 (P ~ 0.6): also the smallest k with
    # P(S >= k) <= 1e-2, read off the in-kernel S histogram of the same stream
    h = mcr.run(mu, sigma, G["x"], G["k"], G["paths"], G["path_seed"], precision=mcr.PREC_BF16X3,
                want_hist=True)
    tail = np.cumsum(h.s_hist[::-1])[::-1] / G["paths"]
    kt = int(np.argmax(tail <= 1e-2))
    t = mcr.run(mu, sigma, G["x"], kt, G["paths"], G["path_seed"], precision=mcr.PREC_BF16X3)
    res["tail_k"] = {"k": kt, "count": t.count, "p_hat": t.p_hat, "std_error": t.std_error,
                     "ci": [t.ci_lower, t.ci_upper], "kernel_ms": t.ms_kernel,
                     "count_equals_histogram_tail": t.count == int(h.s_hist[kt:].sum())}
    f, b = res["fp32"], res["bf16x3"]
    res["speedup_bf16x3_vs_fp32"] = f["kernel_ms"] / b["kernel_ms"]
    res["rel_diff_p_hat"] = abs(b["p_hat"] - f["p_hat"]) / f["p_hat"]
    res["rel_diff_var99"] = abs(b["var99"] - f["var99"]) / f["var99"]
    print(f"# n=600: fp32 {f['kernel_ms']:.2f} ms, bf16x3 {b['kernel_ms']:.2f} ms; tail k={kt}: "
          f"P={t.p_hat:.5f} +- {t.std_error:.5f}", file=sys.stderr, flush=True)
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); across GPUs the paths "
              "would shard by path-id range (mcr_query.path_offset) with one all-reduce of "
              "the integer tallies", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import __graft_entry__ as ge

    mcr = ge.load_package()
    if args.copy_mb:
        import torch

        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)  # one elementwise kernel: copy_mb MiB read, copy_mb MiB written
        torch.cuda.synchronize()
        print(f"# calibration kernel (neg): {args.copy_mb} MiB read + {args.copy_mb} MiB written",
              file=sys.stderr, flush=True)
    W = WORKLOAD
    n = W["n"]
    mu, sigma = mcr.synth_factor_model(n, W["factors"], W["days"], W["data_seed"])
    prec = {"auto": mcr.PREC_AUTO, "fp32": mcr.PREC_FP32, "bf16x3": mcr.PREC_BF16X3}[args.precision]

    def step(precision):
        return mcr.run(mu, sigma, W["x"], W["k"], W["paths"], W["path_seed"], precision=precision,
                       alphas=(W["alpha"],))

    # the first warm-up query pays the library's lazy set-up (HIP context,
    # code objects, device arena, pinned staging): the cold figure a driver
    # that calls run_model once per process sees
    cold_ms = None
    for i in range(args.warmup):
        t_w = time.perf_counter()
        step(prec)
        if i == 0:
            cold_ms = (time.perf_counter() - t_w) * 1e3
    kms, dms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step(prec)
        kms.append(r.ms_kernel)
        dms.append(r.ms_device)
    wall = time.perf_counter() - t0
    # One step = one full query: factor, layouts, all paths with count and VaR
    # reductions.  value: device time of that work with the inputs already in
    # HBM (HIP events from the factor launch to the end of the path kernel);
    # the wall-clock rate, which adds the 20 KB upload, the result copies and
    # the host's order-statistic selection, is reported beside it.
    ms_step = float(np.mean(dms))
    value = W["paths"] / (ms_step * 1e-3)
    wall_ms = 1e3 * wall / args.steps
    k_ms = float(np.mean(kms))
    flop = n * (n + 1) * W["paths"]
    achieved = flop / (k_ms * 1e-3) / 1e12
    is_mfma = r.kernel == "mfma"
    peak = PEAK_BF16_TFLOPS if is_mfma else PEAK_FP32_TFLOPS
    print(f"# step {ms_step:.3f} ms device, {wall_ms:.3f} ms wall, kernel {k_ms:.3f} ms ({r.kernel}), count={r.count} "
          f"p_hat={r.p_hat:.6f} VaR99={r.var[0]:.6f}", file=sys.stderr, flush=True)

    # the plain fp32 fused kernel beside it, same workload
    extra = {}
    if r.kernel != "valu_reg":
        f = [step(mcr.PREC_FP32) for _ in range(3)]
        extra["fp32_kernel_ms"] = float(np.mean([x.ms_kernel for x in f]))
        extra["fp32_paths_per_s"] = W["paths"] / (extra["fp32_kernel_ms"] * 1e-3)
        extra["fp32_count"] = f[-1].count
    out = {
        "metric": "paths/sec at n=50 & n=600 vs plain FP32 kernel; p50/p99 latency at fixed QPS",
        "value": value, "unit": "paths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16" if is_mfma else "f32",
        "dtype_detail": ("bf16x3 split products (hi*hi + hi*lo + lo*hi), fp32 accumulation; "
                         "fp32 normals; FP64 factor") if is_mfma else "fp32; FP64 factor",
        "data": "synthetic (seeded 5-factor model, Philox paths)",
        "config": {"workload": W["name"], "n_assets": n, "paths": W["paths"], "k": W["k"],
                   "x": W["x"], "var_alpha": W["alpha"], "kernel": r.kernel,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma" if is_mfma else "valu", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": TRAFFIC_BYTES.get(r.kernel), "kernel": r.kernel, "kernel_ms": k_ms,
                     "flop_per_path": n * (n + 1),
                     "note": ("achieved counts the mat-vec FLOPs only; the kernel is bound by the "
                              "VALU issue of normal generation (Philox4x32-10, Box-Muller: "
                              "~83 % of its instructions), see profiles/k_mfma_n50_100M.md")},
        "result": {"count": r.count, "p_hat": r.p_hat, "var99": r.var[0], "es99": r.es[0]},
        "wall_ms_per_step": wall_ms,
        "cold_first_query_ms": cold_ms,
        "wall_paths_per_s": W["paths"] / (wall_ms * 1e-3),
        "kernel_paths_per_s": W["paths"] / (k_ms * 1e-3),
        "normals_per_s": W["paths"] * n / (k_ms * 1e-3),
    }
    out.update(extra)
    if not args.no_large:
        out["large"] = large_leg(mcr, np)
    if not args.no_batch:
        out["batch"] = batch_leg(mcr, np)
    if not args.no_serve:
        out["serving"] = serve_leg(mcr, np, args.serve_seconds)
    if not args.no_cpu:
        L, *_ = mcr.factor(sigma)
        out["cpu_baseline"] = cpu_baseline(L, mu, W["x"], W["k"], W["path_seed"], n,
                                           args.cpu_seconds)
        out["accuracy_vs_fp64"] = accuracy_check(mcr, np)
        print(f"# accuracy vs FP64: {out['accuracy_vs_fp64']}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
