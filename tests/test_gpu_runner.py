"""The reference-shaped runner (mcr_runner) end to end on the GPU: the real
15-index model (mu, Sigma, actual_freq of tests/golden/realdata_stats.json,
computed from the reference's dataset) written as data/*.bin, a config.yaml
listing models, the hip model run through ModelFunctions, one CSV row.

Bar: the row's count is exactly the count of the same query through mcr_run
(fp32, the hip model's default precision, same Philox stream), every other
field is the configuration's, and models absent from this build are reported
and skipped as the reference's driver does (main_runner.c:206-240)."""
from __future__ import annotations

import csv
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def setup_run(mcr, tmp_path, M, models="[serial, hip, nonsense]"):
    g = json.load(open(os.path.join(GOLDEN, "realdata_stats.json")))
    mu, S = np.array(g["mu"]), np.array(g["Sigma"])
    mcr.write_bins(str(tmp_path / "data"), mu, S, g["actual_freq"])
    c = g["config"]
    idx = "\n".join(f'  {k}: "{v}"' for k, v in c["indices"].items())
    (tmp_path / "config.yaml").write_text(
        f"models: {models}\nstart: \"{c['start']}\"\nend: \"{c['end']}\"\nindices:\n{idx}\n"
        f"train_ratio: {c['train_ratio']}\nM: {M}\nrandom_seed: {c['random_seed']}\n"
        f"x: {c['x']}\nk: {c['k']}\ncomment: \"from config\"\n")
    return g, mu, S


def read_rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def test_runner_real_data_row_matches_oracle(mcr, orc, tmp_path):
    """The CSV row's count against the C oracle (the serial model restated) on
    the same real-data query and Philox stream, with the device's fp32 model:
    identical integers, and the oracle's own FP64 statistics."""
    M = 2_000_000
    g, mu, S = setup_run(mcr, tmp_path, M, models="[hip]")
    r = subprocess.run([mcr.RUNNER_PATH], cwd=tmp_path, input="oracle check\n",
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    row = read_rows(tmp_path / "results" / "simulation_results.csv")[0]
    c = g["config"]
    m = mcr.prepare(mu, S, c["x"])
    model = orc.F32Model(m["L32"], m["mu32"], m["v32"], m["mp32"], m["thr32"])
    ref = orc.simulate_f32(model, c["random_seed"], 0, M, c["k"], nthreads=16)
    assert int(row["count"]) == ref["count"], (row["count"], ref["count"])
    p, se, lo, hi = orc.stats(ref["count"], M)
    assert row["P_hat"] == f"{p:.6f}" and row["std_error"] == f"{se:.6f}"
    assert row["ci_lower"] == f"{lo:.6f}" and row["ci_upper"] == f"{hi:.6f}"


def test_runner_real_data_row_matches_mcr_run(mcr, tmp_path):
    M = 4_000_000
    g, mu, S = setup_run(mcr, tmp_path, M)
    r = subprocess.run([mcr.RUNNER_PATH], cwd=tmp_path, input="real data, gpu\n",
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Error: serial model not available in this build" in r.stderr
    assert "Error: Unknown model type 'nonsense'" in r.stderr
    rows = read_rows(tmp_path / "results" / "simulation_results.csv")
    assert len(rows) == 1 and len(rows[0]) == 22
    row = rows[0]
    c = g["config"]
    ref = mcr.run(mu, S, c["x"], c["k"], M, c["random_seed"])
    assert row["model_name"] == "hip"
    assert int(row["count"]) == ref.count
    assert row["P_hat"] == f"{ref.p_hat:.6f}"
    assert row["std_error"] == f"{ref.std_error:.6f}"
    assert row["ci_lower"] == f"{ref.ci_lower:.6f}" and row["ci_upper"] == f"{ref.ci_upper:.6f}"
    assert (int(row["M"]), int(row["k"]), row["x"], int(row["seed"])) == (
        M, c["k"], f"{c['x']:.4f}", c["random_seed"])
    assert row["actual_freq"] == f"{g['actual_freq']:.6f}"
    assert row["comment"] == "real data, gpu"  # the typed comment wins over the config's
    assert row["indices"] == ";".join(f"{k}:{v}" for k, v in c["indices"].items())
    assert (row["start_date"], row["end_date"]) == (c["start"], c["end"])
    assert int(row["execution_time_ms"]) >= 0
    # the model's probability and the held-out frequency are the same order
    assert 0.2 * g["actual_freq"] < ref.p_hat < 5 * g["actual_freq"]


def test_runner_appends_and_uses_config_comment(mcr, tmp_path):
    setup_run(mcr, tmp_path, 200_000, models="[hip, hip]")
    env = dict(os.environ, MCR_HIP_PRECISION="bf16x3")
    r = subprocess.run([mcr.RUNNER_PATH], cwd=tmp_path, input="", capture_output=True,
                       text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    rows = read_rows(tmp_path / "results" / "simulation_results.csv")
    assert len(rows) == 2
    assert rows[0]["count"] == rows[1]["count"]  # same query, same stream
    assert all(x["comment"] == "from config" for x in rows)
