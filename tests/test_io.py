"""Data formats around the risk path (CPU): mu/sigma/params binaries, the
YAML configuration, the results CSV row, the real-data preprocessing and the
reference-shaped runner's failure paths.

What pins them:
  * binaries: byte layout written independently with numpy exactly as the
    reference's writer describes it (int32 header + float64 data), and the
    reader's validation cases (load_binary.c:20-24, :65-76, :113-118);
  * config: the reference's own config.yaml (when the reference tree is
    present), plus YAML texts exercising comments, quoting, flow and block
    collections, defaults and the atoi/atof conversions (load_config.c:11-25,
    :243-268);
  * CSV: the 22-column header and row formats (csv_writer.c:40-45, :282-315)
    restated with Python's % formatting, quoting cases, header written once;
  * preprocessing: a pandas computation through the reference's library calls
    on synthetic index files with reversed order, bad dates, gaps and missing
    files, and on the reference's dataset (golden realdata_stats.json made by
    tests/golden/make_realdata_golden.py; the dataset itself is read only
    where the reference tree is present)."""
from __future__ import annotations

import json
import os
import struct
import subprocess
import time

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = "/root/reference"


# ---------------------------------------------------------------- binaries ---
def test_bins_layout_and_round_trip(mcr, tmp_path):
    rng = np.random.default_rng(3)
    n = 7
    mu = rng.normal(size=n)
    a = rng.normal(size=(n, n))
    sigma = a @ a.T
    mcr.write_bins(str(tmp_path / "d"), mu, sigma, 0.125)
    raw_mu = (tmp_path / "d" / "mu.bin").read_bytes()
    assert raw_mu == np.array([n], np.int32).tobytes() + mu.astype(np.float64).tobytes()
    raw_s = (tmp_path / "d" / "sigma.bin").read_bytes()
    assert raw_s == np.array([n, n], np.int32).tobytes() + sigma.astype(np.float64).tobytes()
    assert (tmp_path / "d" / "params.bin").read_bytes() == struct.pack("<d", 0.125)
    assert np.array_equal(mcr.read_mu_bin(tmp_path / "d" / "mu.bin"), mu)
    assert np.array_equal(mcr.read_sigma_bin(tmp_path / "d" / "sigma.bin"), sigma)
    assert mcr.read_params_bin(tmp_path / "d" / "params.bin") == 0.125


def test_bin_reader_rejects_what_the_reference_rejects(mcr, tmp_path):
    def put(name, data):
        p = tmp_path / name
        p.write_bytes(data)
        return p

    bad = [
        (mcr.read_mu_bin, put("n0.bin", np.array([0], np.int32).tobytes())),
        (mcr.read_mu_bin, put("nbig.bin", np.array([1000001], np.int32).tobytes())),
        (mcr.read_mu_bin, put("short.bin", np.array([4], np.int32).tobytes() + b"\0" * 24)),
        (mcr.read_mu_bin, put("empty.bin", b"")),
        (mcr.read_mu_bin, tmp_path / "missing.bin"),
        (mcr.read_sigma_bin, put("neg.bin", np.array([-1, 3], np.int32).tobytes())),
        (mcr.read_sigma_bin, put("huge.bin", np.array([50000, 50000], np.int32).tobytes())),
        (mcr.read_sigma_bin, put("sshort.bin", np.array([2, 2], np.int32).tobytes() + b"\0" * 24)),
        (mcr.read_params_bin, put("p_neg.bin", struct.pack("<d", -0.1))),
        (mcr.read_params_bin, put("p_big.bin", struct.pack("<d", 1.5))),
        (mcr.read_params_bin, put("p_short.bin", b"\0\0\0")),
    ]
    for fn, path in bad:
        with pytest.raises(mcr.McrError) as e:
            fn(path)
        assert e.value.code == mcr.MCR_E_ARG, path
    # rectangular matrices and the edges of the valid ranges load
    rect = put("rect.bin", np.array([2, 3], np.int32).tobytes() + np.arange(6.0).tobytes())
    assert mcr.read_sigma_bin(rect).shape == (2, 3)
    assert mcr.read_params_bin(put("p0.bin", struct.pack("<d", 0.0))) == 0.0
    assert mcr.read_params_bin(put("p1.bin", struct.pack("<d", 1.0))) == 1.0
    # the reference's range test lets a NaN through
    assert np.isnan(mcr.read_params_bin(put("pnan.bin", struct.pack("<d", float("nan")))))


# ------------------------------------------------------------------ config ---
def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_config_defaults(mcr, tmp_path):
    c = mcr.load_config(write(tmp_path, "c.yaml", "# nothing set\n"))
    assert (c["x"], c["N"], c["M"], c["k"], c["random_seed"], c["train_ratio"]) == (
        0.02, 0, 100000, 5, 42, 0.8)
    assert c["start"] is None and c["end"] is None and c["comment"] is None
    assert c["models"] == [] and c["indices"] == {}


def test_config_block_style_like_the_reference(mcr, tmp_path):
    text = """# header comment
models:
  #- openmp
  - hip   # trailing comment
  - "serial"
start: "2015-01-01"  # quoted date
end: '2025-12-31'
indices:
  AEX: "AEX.csv"
  Bel20: Bel20.csv
train_ratio: 0.75
M: 2000000
random_seed: 67
x: 0.03
k: 4
comment: "Run # with a hash, and a comma"
"""
    c = mcr.load_config(write(tmp_path, "c.yaml", text))
    assert c["models"] == ["hip", "serial"]
    assert c["start"] == "2015-01-01" and c["end"] == "2025-12-31"
    assert c["indices"] == {"AEX": "AEX.csv", "Bel20": "Bel20.csv"}
    assert list(c["indices"]) == ["AEX", "Bel20"]
    assert (c["train_ratio"], c["M"], c["random_seed"], c["x"], c["k"]) == (0.75, 2000000, 67,
                                                                          0.03, 4)
    assert c["comment"] == "Run # with a hash, and a comma"


def test_config_flow_style_and_conversions(mcr, tmp_path):
    text = ("models: [hip, 'openmp']\nindices: {A: a.csv, \"B\": 'b.csv'}\n"
            "M: 1e8\nk: 7.9\nrandom_seed: 12abc\nx: .5e-1\nN: 15\nunknown_key: 3\n")
    c = mcr.load_config(write(tmp_path, "c.yaml", text))
    assert c["models"] == ["hip", "openmp"]
    assert c["indices"] == {"A": "a.csv", "B": "b.csv"}
    # atoi / atol / atof as the reference converts scalars
    assert (c["M"], c["k"], c["random_seed"], c["x"], c["N"]) == (1, 7, 12, 0.05, 15)


def test_config_errors(mcr, tmp_path):
    for text in ("models:\n\t- hip\n", "just a line\n", "  - orphan\n", "start: \"open\n",
                 "models: [a, b\n"):
        with pytest.raises(mcr.McrError) as e:
            mcr.load_config(write(tmp_path, "bad.yaml", text))
        assert e.value.code == mcr.MCR_E_ARG, text
    with pytest.raises(mcr.McrError):
        mcr.load_config(tmp_path / "missing.yaml")


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "config.yaml")),
                    reason="reference tree not present")
def test_config_reference_file(mcr):
    c = mcr.load_config(os.path.join(REF, "config.yaml"))
    g = json.load(open(os.path.join(GOLDEN, "realdata_stats.json")))["config"]
    assert c["models"] == ["cuda"]
    assert c["indices"] == g["indices"] and list(c["indices"]) == list(g["indices"])
    assert (c["start"], c["end"], c["x"], c["k"], c["train_ratio"], c["M"], c["random_seed"]) == (
        g["start"], g["end"], g["x"], g["k"], g["train_ratio"], g["M"], g["random_seed"])


# --------------------------------------------------------------------- CSV ---
HEADER = ("iterationnumber,timestamp,execution_time_ms,comment,start_date,end_date,train_ratio,"
          "M,k,x,model_name,seed,nodes,threads,processes,indices,actual_freq,P_hat,count,"
          "std_error,ci_lower,ci_upper")


def row_fields(**over):
    f = dict(iteration_id=3, timestamp=1700000000, execution_time_ms=1234, comment="plain",
             start_date="2015-01-01", end_date="2025-12-31", train_ratio=0.8, M=100000000, k=5,
             x=0.02, model_name="hip", seed=67, nodes=1, threads=16, processes=1,
             indices={"AEX": "AEX.csv", "Bel20": "Bel20.csv"}, actual_freq=0.014,
             P_hat=0.0123456789, count=1234568, std_error=1.1e-5, ci_lower=0.0123,
             ci_upper=0.0124)
    f.update(over)
    return f


def expected_line(f):
    ts = time.strftime("%Y-%m-%d %H:%M:%S", time.localtime(f["timestamp"]))
    idx = ";".join(f"{k}:{v}" for k, v in f["indices"].items())

    def esc(s):
        if not s:
            return ""
        if any(c in s for c in ',"\n\r'):
            return '"' + s.replace('"', '""') + '"'
        return s

    return ("%d,%s,%d,%s,%s,%s,%.4f,%d,%d,%.4f,%s,%d,%d,%d,%d,%s,%.6f,%.6f,%d,%.6f,%.6f,%.6f" % (
        f["iteration_id"], ts, f["execution_time_ms"], esc(f["comment"]), f["start_date"] or "",
        f["end_date"] or "", f["train_ratio"], f["M"], f["k"], f["x"], f["model_name"],
        f["seed"], max(f["nodes"], 0), max(f["threads"], 0), max(f["processes"], 0), esc(idx),
        f["actual_freq"], f["P_hat"], f["count"], f["std_error"], f["ci_lower"], f["ci_upper"]))


def test_csv_header_has_22_columns(mcr):
    assert mcr.csv_header() == HEADER
    assert len(HEADER.split(",")) == 22


@pytest.mark.parametrize("over", [
    {},
    {"comment": 'said "hi", then left'},
    {"comment": "line1\nline2"},
    {"comment": ""},
    {"start_date": None, "end_date": None, "nodes": 0, "threads": -4, "processes": 0},
    {"indices": {"A,B": "a.csv"}},
    {"indices": {}},
    {"actual_freq": float("nan"), "P_hat": 1.0, "count": 0},
])
def test_csv_row_format(mcr, over):
    f = row_fields(**over)
    line = mcr.csv_format_row(**f)
    assert line == expected_line(f)


def test_csv_append_writes_header_once(mcr, tmp_path):
    path = tmp_path / "results" / "sub" / "simulation_results.csv"
    mcr.append_csv(path, **row_fields())
    mcr.append_csv(path, **row_fields(model_name="hip2", comment="x,y"))
    lines = path.read_text().splitlines()
    assert lines[0] == HEADER
    assert lines[1] == expected_line(row_fields())
    assert lines[2] == expected_line(row_fields(model_name="hip2", comment="x,y"))
    assert len(lines) == 3


# ----------------------------------------------------------- preprocessing ---
def make_index_csv(path, rows):
    with open(path, "w") as f:
        f.write("Date,Return\n")
        for d, r in rows:
            f.write(f"{d},{r}\n")


def pandas_pipeline(indices, ddir, start, end, train_ratio, x, k):
    """The reference's library calls (fetch_split.py, preprocessing_stats.py,
    evaluation.py) in pandas, as an independent check."""
    pd = pytest.importorskip("pandas")
    series = {}
    for name, fname in indices.items():
        p = os.path.join(ddir, fname)
        if not os.path.exists(p):
            continue
        df = pd.read_csv(p)
        if "Date" not in df.columns or "Return" not in df.columns:
            continue
        df["Date"] = pd.to_datetime(df["Date"], format="%m/%d/%Y", errors="coerce")
        series[name] = df.dropna(subset=["Date"]).set_index("Date")["Return"]
    r = pd.DataFrame(series).loc[pd.to_datetime(start):pd.to_datetime(end)].dropna()
    s = int(len(r) * train_ratio)
    tr, te = r.iloc[:s], r.iloc[s:]
    return (list(r.columns), tr.mean().values, tr.cov().values,
            float(np.mean((te < -x).sum(axis=1) >= k)), len(tr), len(te))


def test_preprocessing_matches_pandas_on_ragged_files(mcr, tmp_path):
    pp = mcr.preprocessing()
    rng = np.random.default_rng(11)
    days = [np.datetime64("2019-01-01") + np.timedelta64(i, "D") for i in range(400)]
    fmt = lambda d: d.astype(object).strftime("%m/%d/%Y")  # noqa: E731
    ddir = tmp_path / "dataset"
    ddir.mkdir()
    indices = {}
    for j in range(5):
        rows = []
        for i, d in enumerate(days):
            if (i + 3 * j) % 17 == 0:
                continue  # a gap in this index
            rows.append((fmt(d), f"{rng.normal(0, 0.015):.4f}"))
        rows.reverse()  # newest first, as the reference's files
        if j == 2:
            rows[20] = (rows[20][0], "")               # missing return: the day is dropped
            rows.insert(5, ("2019-13-45", "0.01"))     # unparseable date: row dropped
            rows.insert(9, ("02/30/2019", "0.02"))     # impossible date: row dropped
        make_index_csv(ddir / f"I{j}.csv", rows)
        indices[f"I{j}"] = f"I{j}.csv"
    (ddir / "nocols.csv").write_text("Day,Value\n01/02/2019,0.1\n")
    indices["NoCols"] = "nocols.csv"
    indices["Missing"] = "missing.csv"
    cfg = {"indices": indices, "start": "2019-02-01", "end": "2019-12-31", "train_ratio": 0.7,
           "x": 0.02, "k": 2}
    got = pp.run_pipeline(cfg, str(ddir), str(tmp_path / "data"), log=lambda *a: None)
    names, mu, sigma, freq, ntr, nte = pandas_pipeline(indices, str(ddir), cfg["start"],
                                                       cfg["end"], cfg["train_ratio"], cfg["x"],
                                                       cfg["k"])
    assert got["names"] == names == ["I0", "I1", "I2", "I3", "I4"]
    assert (got["train_days"], got["test_days"]) == (ntr, nte)
    assert np.allclose(got["mu"], mu, rtol=1e-12, atol=1e-18)
    assert np.allclose(got["Sigma"], sigma, rtol=1e-10, atol=1e-18)
    assert got["actual_freq"] == freq
    assert np.array_equal(mcr.read_mu_bin(tmp_path / "data" / "mu.bin"), got["mu"])
    assert np.array_equal(mcr.read_sigma_bin(tmp_path / "data" / "sigma.bin"), got["Sigma"])
    assert mcr.read_params_bin(tmp_path / "data" / "params.bin") == got["actual_freq"]


def test_preprocessing_errors(mcr, tmp_path):
    pp = mcr.preprocessing()
    ddir = tmp_path / "ds"
    ddir.mkdir()
    make_index_csv(ddir / "A.csv", [("01/%02d/2020" % d, "0.01") for d in range(1, 29)])
    base = {"indices": {"A": "A.csv"}, "start": "2020-01-01", "end": "2020-12-31",
            "train_ratio": 0.8, "x": 0.02, "k": 1}
    with pytest.raises(pp.PreprocessError):  # fewer than 100 days
        pp.run_pipeline(base, str(ddir), log=lambda *a: None)
    with pytest.raises(pp.PreprocessError):
        pp.run_pipeline(dict(base, indices={"Z": "none.csv"}), str(ddir), log=lambda *a: None)
    with pytest.raises(pp.PreprocessError):
        pp.run_pipeline(dict(base, start="2021-01-01", end="2021-02-01"), str(ddir),
                        log=lambda *a: None)
    with pytest.raises(TypeError):
        pp.run_pipeline(dict(base, indices=["A.csv"]), str(ddir), log=lambda *a: None)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "dataset")),
                    reason="reference dataset not present")
def test_preprocessing_reference_dataset(mcr, tmp_path):
    pp = mcr.preprocessing()
    g = json.load(open(os.path.join(GOLDEN, "realdata_stats.json")))
    cfg = mcr.load_config(os.path.join(REF, "config.yaml"))
    got = pp.run_pipeline(cfg, os.path.join(REF, "dataset"), str(tmp_path), log=lambda *a: None)
    assert got["names"] == g["names"] and len(g["names"]) == 15
    assert (got["train_days"], got["test_days"]) == (g["train_days"], g["test_days"]) == (2000, 500)
    assert str(got["first_date"]) == g["first_date"] and str(got["last_date"]) == g["last_date"]
    assert got["actual_freq"] == g["actual_freq"] == 0.014
    assert np.allclose(got["mu"], g["mu"], rtol=1e-12, atol=1e-18)
    assert np.allclose(got["Sigma"], g["Sigma"], rtol=1e-10, atol=1e-18)
    assert abs(np.linalg.cond(got["Sigma"]) - 1.66e3) < 0.01e3


def test_golden_real_data_is_a_valid_model(mcr):
    g = json.load(open(os.path.join(GOLDEN, "realdata_stats.json")))
    S = np.array(g["Sigma"])
    assert S.shape == (15, 15) and np.allclose(S, S.T)
    assert np.all(np.linalg.eigvalsh(S) > 0)
    assert g["days"] == g["train_days"] + g["test_days"]


# ------------------------------------------------------------------ runner ---
def run_runner(mcr, cwd, *args, stdin=""):
    return subprocess.run([mcr.RUNNER_PATH, *args], cwd=cwd, input=stdin, capture_output=True,
                          text=True, timeout=120)


def test_runner_missing_inputs_exit_1(mcr, tmp_path):
    r = run_runner(mcr, tmp_path)
    assert r.returncode == 1
    assert "Failed to load binary files" in r.stderr


def test_runner_bad_config_exit_1(mcr, tmp_path):
    mcr.write_bins(str(tmp_path / "data"), np.zeros(3), np.eye(3), 0.5)
    (tmp_path / "config.yaml").write_text("models:\n\t- hip\n")
    r = run_runner(mcr, tmp_path)
    assert r.returncode == 1 and "config.yaml" in r.stderr


def test_runner_reports_models_missing_from_this_build(mcr, tmp_path):
    mcr.write_bins(str(tmp_path / "data"), np.zeros(3), np.eye(3), 0.5)
    (tmp_path / "config.yaml").write_text("models: [serial, openmp, cuda, gpu, mpi, nonsense]\n")
    r = run_runner(mcr, tmp_path)
    assert r.returncode == 0
    for m in ("serial", "openmp", "cuda", "gpu", "mpi"):
        assert f"Error: {m} model not available in this build" in r.stderr
    assert "Error: Unknown model type 'nonsense'" in r.stderr
    assert not (tmp_path / "results").exists()
